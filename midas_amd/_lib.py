"""ctypes binding of libmidas_caking.so (the C ABI in include/midas_caking.h).

There is no CPU fallback: when the library is missing or no GPU is present,
every compute entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

# MIDAS_AMD_LIB selects another build of the library (same ABI).
LIB_PATH = Path(os.environ.get("MIDAS_AMD_LIB") or
                Path(__file__).resolve().with_name("libmidas_caking.so"))

MC_DTYPE_U8, MC_DTYPE_U16, MC_DTYPE_U32, MC_DTYPE_I64 = 0, 1, 2, 3
MC_DTYPE_F32, MC_DTYPE_F64, MC_DTYPE_HYBRID, MC_DTYPE_I32 = 4, 5, 6, 7
MC_MODE_FLOOR, MC_MODE_BILINEAR, MC_MODE_GRADIENT = 0, 1, 2
MC_F_NORMALIZE, MC_F_MASK_NAN = 1, 2
MC_FIT_PV, MC_FIT_TCH, MC_FIT_PV2, MC_FIT_OUT = 0, 1, 2, 8
MC_FIT_MAX_PEAKS = 8


def MC_FIT_TCH_PEAKS(n: int) -> int:
    """Model code of the TCH fit with n peaks on one window."""
    return MC_FIT_TCH if n == 1 else n + 1

MODES = {"floor": MC_MODE_FLOOR, "bilinear": MC_MODE_BILINEAR, "gradient": MC_MODE_GRADIENT}


class MidasError(RuntimeError):
    """A status code other than MC_OK returned by the library."""

    def __init__(self, rc: int, msg: str):
        super().__init__(f"[{rc}] {msg}")
        self.rc = rc


class Geometry(C.Structure):
    """mc_geometry: the DetectorMapper parameter set."""
    _fields_ = [
        ("NrPixelsY", C.c_int32), ("NrPixelsZ", C.c_int32),
        ("pxY", C.c_double), ("pxZ", C.c_double), ("yCen", C.c_double),
        ("zCen", C.c_double), ("Lsd", C.c_double), ("RhoD", C.c_double),
        ("tx", C.c_double), ("ty", C.c_double), ("tz", C.c_double),
        ("p", C.c_double * 15),
        ("RMin", C.c_double), ("RMax", C.c_double), ("RBinSize", C.c_double),
        ("EtaMin", C.c_double), ("EtaMax", C.c_double), ("EtaBinSize", C.c_double),
        ("QMin", C.c_double), ("QMax", C.c_double), ("QBinSize", C.c_double),
        ("Wavelength", C.c_double),
        ("NrTransOpt", C.c_int32), ("TransOpt", C.c_int32 * 10),
        ("SubPixelLevel", C.c_int32), ("SubPixelCardinalWidth", C.c_double),
        ("Parallax", C.c_double),
        ("SolidAngleCorrection", C.c_int32), ("PolarizationCorrection", C.c_int32),
        ("PolarizationFraction", C.c_double),
    ]


class PanelC(C.Structure):
    _fields_ = [("yMin", C.c_int32), ("yMax", C.c_int32), ("zMin", C.c_int32),
                ("zMax", C.c_int32), ("dY", C.c_double), ("dZ", C.c_double),
                ("dTheta", C.c_double), ("dLsd", C.c_double), ("dP2", C.c_double),
                ("centerY", C.c_double), ("centerZ", C.c_double)]


class MapInfo(C.Structure):
    _fields_ = [("nnz", C.c_int64), ("n_r", C.c_int32), ("n_eta", C.c_int32),
                ("ny", C.c_int32), ("nz", C.c_int32), ("n_flagged", C.c_int32),
                ("q_mode", C.c_int32), ("hbm_bytes", C.c_int64),
                ("skipped", C.c_int64 * 3)]


class ServerStats(C.Structure):
    _fields_ = [("frames_done", C.c_int64), ("batches", C.c_int64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("max_ms", C.c_double),
                ("frames_per_s", C.c_double), ("admitted_bytes", C.c_int64),
                ("peak_bytes", C.c_int64), ("rejected", C.c_int64)]


class OutConfig(C.Structure):
    """mc_out_config: an output set's shape, files and axes."""
    _fields_ = [("n_r", C.c_int32), ("n_eta", C.c_int32), ("files", C.c_uint32),
                ("omega_sum_frames", C.c_int32), ("px", C.c_double), ("lsd", C.c_double),
                ("wavelength", C.c_double), ("q_mode", C.c_int32), ("q_min", C.c_double),
                ("q_bin_size", C.c_double), ("reta", C.POINTER(C.c_double)),
                ("r_centers", C.POINTER(C.c_double)), ("eta_lo", C.POINTER(C.c_double)),
                ("eta_hi", C.POINTER(C.c_double)), ("instrument", C.POINTER(C.c_double)),
                ("source", C.c_char_p)]


class PeakRow(C.Structure):
    """mc_peak_row: one per-eta fit (7 PeakFit parameters)."""
    _fields_ = [("frame", C.c_int32), ("eta_bin", C.c_int32), ("peak", C.c_int32),
                ("reserved", C.c_int32), ("p", C.c_double * 7)]


MC_OUT_CAKED_HDF, MC_OUT_PER_FRAME, MC_OUT_SUM_FRAMES, MC_OUT_OMEGAS = 0x01, 0x02, 0x04, 0x08
MC_OUT_LINEOUT, MC_OUT_STREAM, MC_OUT_INT2D, MC_OUT_PEAKS = 0x10, 0x20, 0x40, 0x80
MC_OUT_CAKED_ZARR = 0x100


HOST_FN = C.CFUNCTYPE(None, C.c_void_p)
RESULT_CB = C.CFUNCTYPE(None, C.c_void_p, C.c_int64, C.c_int32,
                        C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int32)

# Every symbol the header declares, with (restype, argtypes).
_P, _V, _I, _I32, _I64, _U32, _D = (C.POINTER, C.c_void_p, C.c_int, C.c_int32,
                                    C.c_int64, C.c_uint32, C.c_double)
_PF = C.POINTER(C.c_float)
_PD = C.POINTER(C.c_double)
SIGNATURES = {
    "mc_version": (_I, [_P(_I), _P(_I), _P(_I)]),
    "mc_last_error": (C.c_char_p, []),
    "mc_device_count": (_I, [_P(_I)]),
    "mc_ctx_create": (_I, [_I, _P(_V)]),
    "mc_ctx_destroy": (_I, [_V]),
    "mc_ctx_sync": (_I, [_V]),
    "mc_ctx_stream": (_V, [_V]),
    "mc_dev_alloc": (_I, [_V, C.c_uint64, _P(_V)]),
    "mc_dev_free": (_I, [_V, _V]),
    "mc_memcpy_h2d": (_I, [_V, _V, _V, C.c_uint64]),
    "mc_memcpy_d2h": (_I, [_V, _V, _V, C.c_uint64]),
    "mc_map_build": (_I, [_V, _P(Geometry), _PD, _PD, _PD, _PD, _P(PanelC), _I32, _P(_V)]),
    "mc_map_write": (_I, [_V, C.c_char_p, C.c_char_p]),
    "mc_map_load": (_I, [_V, C.c_char_p, C.c_char_p, C.c_char_p, _I32, _I32, _I32, _I32,
                         _P(_V)]),
    "mc_map_from_host": (_I, [_V, _V, _I64, _V, _V, _I32, _I32, _I32, _I32, _P(_V)]),
    "mc_map_info_get": (_I, [_V, _P(MapInfo)]),
    "mc_map_download": (_I, [_V, _V, _V, _V]),
    "mc_map_reta": (_I, [_V, _PD, _PD, _PD, _PD, _D, _D, _D, _PD]),
    "mc_map_reta_stream": (_I, [_V, _PD, _PD, _PD, _PD, _D, _D, _D, _PD]),
    "mc_map_destroy": (_I, [_V]),
    "mc_map_set_dark": (_I, [_V, _V, _I32, _I32, _D, _I32]),
    "mc_map_set_beam_center": (_I, [_V, _D, _D]),
    "mc_map_set_tuning": (_I, [_V, _I32, _I32, _I32]),
    "mc_map_get_tuning": (_I, [_V, _P(_I32), _P(_I32), _P(_I32)]),
    "mc_map_autotune": (_I, [_V, _I32, _PD]),
    "mc_ctx_set_map_verify": (_I, [_V, _I64]),
    "mc_map_verify_info": (_I, [_V, _P(_I64)]),
    "mc_map_set_plain_lanes": (_I, [_V, _I32]),
    "mc_map_get_plain_lanes": (_I, [_V, C.POINTER(C.c_int32)]),
    "mc_map_plan_key": (_I, [_V, _I32, C.c_char_p, C.c_char_p, _I32]),
    "mc_map_autotune_plan": (_I, [_V, _I32, C.c_char_p, C.c_char_p, _PD,
                                  C.POINTER(C.c_int32)]),
    "mc_map_layout_info": (_I, [_V, _I32, _I32, _P(_I64)]),
    "mc_integrate": (_I, [_V, _V, _I32, _I32, _I32, _I32, _U32, _V, _I32, _V, _I32]),
    "mc_profile": (_I, [_V, _V, _I32, _I32, _V, _I32]),
    "mc_profile_variance": (_I, [_V, _V, _V, _I32, _I32, _I32, _V, _V, _I32]),
    "mc_integrate_variance": (_I, [_V, _V, _I32, _I32, _I32, _V, _I32, _I32, _U32, _V, _V,
                                   _I32]),
    "mc_time_integrate": (_I, [_V, _V, _I32, _I32, _I32, _U32, _I32, _I32, _I32, _V, _V,
                               _PD, _PD]),
    "mc_graph_create": (_I, [_V, _V, _I32, _I32, _I32, _U32, _V, _V, _P(_V)]),
    "mc_graph_launch": (_I, [_V]),
    "mc_graph_destroy": (_I, [_V]),
    "mc_fit_windows": (_I, [_V, _I32, _I32, _I32, _V, _V, _V, _D, _I32, _V, _I32]),
    "mc_fit_windows_noalloc": (_I, [_V, _I32, _I32, _I32, _V, _V, _V, _D, _I32, _V]),
    "mc_map_fuse": (_I, [_V, _I32, _V, _I32, _P(_V)]),
    "mc_blosc_decompress": (_I, [_V, _I64, _V, _I64, _P(_I64)]),
    "mc_blosc_decompress_many": (_I, [_V, _V, _V, _I64, _I32, _I32, _V]),
    "mc_smooth_eta": (_I, [_V, _V, _I32, _I32, _I32, _D, _V, _I32, _I32, _V]),
    "mc_median_eta": (_I, [_V, _V, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _V]),
    "mc_fit_cake_rings": (_I, [_V, _V, _I32, _PD, _PD, _I32, _I32, _I32, _I32, _V, _I32]),
    "mc_server_create": (_I, [_V, _I64, _I32, _I32, RESULT_CB, _V, _P(_V)]),
    "mc_server_add_geometry": (_I, [_V, _V, _I32, _I32, _U32, _P(_I32)]),
    "mc_server_submit": (_I, [_V, _I32, _V, _I64, _I32]),
    "mc_server_acquire_buffer": (_I, [_V, _I32, _P(_V)]),
    "mc_server_submit_buffer": (_I, [_V, _I32, _V, _I64, _I32]),
    "mc_server_release_buffer": (_I, [_V, _I32, _V]),
    "mc_map_set_tile_shape": (_I, [_V, _I32, _I32]),
    "mc_map_get_tile_shape": (_I, [_V, _P(_I32), _P(_I32)]),
    "mc_integrate_async": (_I, [_V, _V, _I32, _I32, _I32, _U32, _V, _V]),
    "mc_event_record": (_I, [_V, _P(_V)]),
    "mc_event_query": (_I, [_V, _P(_I32)]),
    "mc_event_wait": (_I, [_V]),
    "mc_event_destroy": (_I, [_V]),
    "mc_ctx_callback": (_I, [_V, HOST_FN, _V]),
    "mc_server_drain": (_I, [_V, _I32]),
    "mc_server_set_batch_window": (_I, [_V, _I64]),
    "mc_server_stats_get": (_I, [_V, _P(ServerStats)]),
    "mc_server_destroy": (_I, [_V]),
    "mc_out_open": (_I, [C.c_char_p, C.c_char_p, _P(OutConfig), _P(_V)]),
    "mc_out_frame": (_I, [_V, _I32, _D, _PF, _PF, _PD]),
    "mc_out_lineout_fits": (_I, [_V, _I32, _I32, _PD]),
    "mc_out_peaks": (_I, [_V, _I32, _P(PeakRow)]),
    "mc_out_close": (_I, [_V]),
}

_lib = None
_lock = threading.Lock()


def lib() -> C.CDLL:
    """The loaded library.  Raises when it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with `python -m midas_amd.build` "
                    "(there is no CPU fallback)")
            # torch bundles its own HIP runtime, loaded by the name
            # libamdhip64.so; its soname is libamdhip64.so.7, which is what
            # this library needs.  Loading torch first makes both share that
            # one runtime, so device pointers pass freely between them; two
            # runtimes in one process do not see each other's devices.
            if os.environ.get("MIDAS_AMD_NO_TORCH", "0") != "1":
                try:
                    import torch  # noqa: F401
                except ImportError:
                    pass
            L = C.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name, None)
                if fn is None:
                    continue
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def last_error() -> str:
    msg = lib().mc_last_error()
    return msg.decode(errors="replace") if msg else ""


def check(rc: int) -> None:
    if rc != 0:
        raise MidasError(rc, last_error())


def device_count() -> int:
    n = C.c_int()
    check(lib().mc_device_count(C.byref(n)))
    return n.value
