"""GPU caking API: the reference ``midas_integrate`` calls (load_map,
build_csr, integrate, profile_1d) backed by libmidas_caking.so.

Arrays may be numpy (host) or torch CUDA tensors (device); device inputs
stay on the device and device outputs are returned for device inputs.
"""
from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import check, lib
from .bin_io import PXLIST_DTYPE, PixelMap, header_for, write_map
from .params import IntegrationParams, bin_edges

AREA_THRESHOLD = 1e-9

NP_TO_CODE = {
    np.dtype(np.uint8): _lib.MC_DTYPE_U8, np.dtype(np.uint16): _lib.MC_DTYPE_U16,
    np.dtype(np.uint32): _lib.MC_DTYPE_U32, np.dtype(np.int64): _lib.MC_DTYPE_I64,
    np.dtype(np.float32): _lib.MC_DTYPE_F32, np.dtype(np.float64): _lib.MC_DTYPE_F64,
    np.dtype(np.int32): _lib.MC_DTYPE_I32,
}


def _torch():
    try:
        import torch
        return torch
    except ImportError:  # pragma: no cover
        return None


def _is_cuda(x) -> bool:
    t = _torch()
    return t is not None and isinstance(x, t.Tensor) and x.is_cuda


def _dtype_code(x) -> int:
    if _is_cuda(x):
        t = _torch()
        m = {t.uint8: 0, t.uint16: 1, t.uint32: 2, t.int64: 3, t.float32: 4,
             t.float64: 5, t.int32: 7}
        if x.dtype not in m:
            raise TypeError(f"unsupported frame dtype {x.dtype}")
        return m[x.dtype]
    dt = np.asarray(x).dtype
    if dt not in NP_TO_CODE:
        raise TypeError(f"unsupported frame dtype {dt}")
    return NP_TO_CODE[dt]


class Context:
    """One GPU context (a HIP stream on one device)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib().mc_ctx_create(device, C.byref(h)))
        self.handle = h
        self.device = device

    def sync(self):
        check(lib().mc_ctx_sync(self.handle))

    def set_map_verify(self, n_pixels: int):
        """Pixels of each later GPU map build recomputed on the host and
        compared (mc_ctx_set_map_verify; default 4096, 0 = off)."""
        check(lib().mc_ctx_set_map_verify(self.handle, int(n_pixels)))

    def close(self):
        if self.handle:
            lib().mc_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[Context] = None
_ctx_lock = threading.Lock()


def default_context() -> Context:
    global _default_ctx
    with _ctx_lock:
        if _default_ctx is None:
            _default_ctx = Context(0)
    return _default_ctx


def _geometry(P: IntegrationParams) -> _lib.Geometry:
    g = _lib.Geometry()
    g.NrPixelsY, g.NrPixelsZ = P.NrPixelsY, P.NrPixelsZ
    g.pxY, g.pxZ, g.yCen, g.zCen = P.pxY, P.pxZ, P.BC_y, P.BC_z
    g.Lsd, g.RhoD, g.tx, g.ty, g.tz = P.Lsd, P.RhoD, P.tx, P.ty, P.tz
    for i in range(15):
        g.p[i] = P.p[i]
    g.RMin, g.RMax, g.RBinSize = P.RMin, P.RMax, P.RBinSize
    g.EtaMin, g.EtaMax, g.EtaBinSize = P.EtaMin, P.EtaMax, P.EtaBinSize
    g.QMin, g.QMax, g.QBinSize, g.Wavelength = P.QMin, P.QMax, P.QBinSize, P.Wavelength
    g.NrTransOpt = len(P.TransOpt)
    for i, t in enumerate(P.TransOpt[:10]):
        g.TransOpt[i] = t
    g.SubPixelLevel = P.SubPixelLevel
    g.SubPixelCardinalWidth = P.SubPixelCardinalWidth
    g.Parallax = P.Parallax
    g.SolidAngleCorrection = P.SolidAngleCorrection
    g.PolarizationCorrection = P.PolarizationCorrection
    g.PolarizationFraction = P.PolarizationFraction
    return g


def generate_panels(P: IntegrationParams, shifts=None):
    """Panel grid of Panel.c:3-57 (+ optional rows id,dY,dZ,dTheta,dLsd,dP2
    as LoadPanelShifts reads them, Panel.c:59-91)."""
    if P.NPanelsY <= 0 or P.NPanelsZ <= 0:
        return None
    gy = list(P.PanelGapsY) + [0] * 64
    gz = list(P.PanelGapsZ) + [0] * 64
    out = []
    y0 = 0
    for a in range(P.NPanelsY):
        y1 = y0 + P.PanelSizeY - 1
        z0 = 0
        for b in range(P.NPanelsZ):
            z1 = z0 + P.PanelSizeZ - 1
            out.append([y0, y1, z0, z1, 0.0, 0.0, 0.0, 0.0, 0.0, (y0 + y1) / 2.0, (z0 + z1) / 2.0])
            if b < P.NPanelsZ - 1:
                z0 = z1 + 1 + gz[b]
        if a < P.NPanelsY - 1:
            y0 = y1 + 1 + gy[a]
    for row in (shifts or []):
        i = int(row[0])
        if 0 <= i < len(out):
            for k, v in enumerate(row[1:6]):
                out[i][4 + k] = float(v)
    arr = (_lib.PanelC * len(out))()
    for i, r in enumerate(out):
        arr[i] = _lib.PanelC(*r)
    return arr


def _dp(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(
        C.POINTER(C.c_double))


class CSRGeometry:
    """A map resident on the GPU (the reference's CSRGeometry,
    kernels.py:47-116): the canonical Map.bin records, the per-mode CSR of
    the one-frame kernel, the tile layout of the batched kernel, the dark
    term and the per-bin normalisation."""

    def __init__(self, handle, ctx: Context, n_r, n_eta, ny, nz, params=None):
        self.handle = handle
        self.ctx = ctx
        self.n_r, self.n_eta = n_r, n_eta
        self.n_pixels_y, self.n_pixels_z = ny, nz
        self.params = params
        self.fused = False  # set by build_fused_geometry
        self._lock = threading.Lock()

    # -- info -----------------------------------------------------------------
    @property
    def n_bins(self) -> int:
        return self.n_r * self.n_eta

    def info(self) -> _lib.MapInfo:
        i = _lib.MapInfo()
        check(lib().mc_map_info_get(self.handle, C.byref(i)))
        return i

    def download(self) -> PixelMap:
        i = self.info()
        rec = np.zeros(i.nnz, dtype=PXLIST_DTYPE)
        pairs = np.zeros((self.n_bins, 2), dtype=np.int32)
        flags = np.zeros(self.n_bins, dtype=np.int32)
        check(lib().mc_map_download(self.handle, rec.ctypes.data, pairs.ctypes.data,
                                    flags.ctypes.data))
        return PixelMap(pxList=rec, counts=pairs[:, 0].copy(), offsets=pairs[:, 1].copy(),
                        mask_flags=flags)

    def write(self, directory, params: Optional[IntegrationParams] = None):
        P = params or self.params
        hdr = header_for(P).to_bytes() if P is not None else None
        from pathlib import Path
        Path(directory).mkdir(parents=True, exist_ok=True)
        check(lib().mc_map_write(self.handle, str(directory).encode(), hdr))

    def reta_map(self, params: Optional[IntegrationParams] = None, lam: float | None = None,
                 in_bounds: bool = False):
        """REtaMap [5, n_r, n_eta]: R, 2theta, eta, sum areaWeight, Q.
        in_bounds=True: the GPU stream's RTthEtaAreaMap rows (the area sums
        only entries on the detector)."""
        P = params or self.params
        RLo, RHi, ELo, EHi = bin_edges(P)
        out = np.zeros((5, self.n_r, self.n_eta))
        fn = lib().mc_map_reta_stream if in_bounds else lib().mc_map_reta
        check(fn(self.handle, _dp(RLo), _dp(RHi), _dp(ELo), _dp(EHi), P.pxY,
                                P.Lsd, P.Lam if lam is None else lam,
                                out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    # -- configuration --------------------------------------------------------
    def set_dark(self, dark, n_frames: int | None = None, scale: float = 1.0):
        if dark is None:
            check(lib().mc_map_set_dark(self.handle, None, 0, 0, 1.0, 0))
            return
        on_dev = _is_cuda(dark)
        code = _dtype_code(dark)
        npx = self.n_pixels_y * self.n_pixels_z
        if on_dev:
            _torch().cuda.current_stream().synchronize()
            ptr = dark.data_ptr()
            n = dark.numel() // npx
        else:
            dark = np.ascontiguousarray(dark)
            ptr = dark.ctypes.data
            n = dark.size // npx
        if n_frames is None:
            n_frames = n
        check(lib().mc_map_set_dark(self.handle, C.c_void_p(ptr), code, n_frames, scale,
                                    int(on_dev)))

    def set_beam_center(self, bc_y: float, bc_z: float):
        check(lib().mc_map_set_beam_center(self.handle, bc_y, bc_z))

    def set_tuning(self, batch=0, rows_per_group=0, lds_bytes=0):
        check(lib().mc_map_set_tuning(self.handle, batch, rows_per_group, lds_bytes))

    def set_tile_shape(self, bins=0, entries=0):
        """Bin records and staged entries a batched-kernel tile holds in LDS
        (0 keeps a value)."""
        check(lib().mc_map_set_tile_shape(self.handle, bins, entries))

    def tile_shape(self):
        b, e = C.c_int32(), C.c_int32()
        check(lib().mc_map_get_tile_shape(self.handle, C.byref(b), C.byref(e)))
        return b.value, e.value

    def tuning(self):
        b, r, l = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().mc_map_get_tuning(self.handle, C.byref(b), C.byref(r), C.byref(l)))
        return b.value, r.value, l.value

    def verify_info(self) -> dict:
        """Host verification of the GPU build (mc_map_verify_info)."""
        out = (C.c_int64 * 4)()
        check(lib().mc_map_verify_info(self.handle, out))
        return {"pixels": out[0], "records": out[1], "membership_differs": out[2],
                "values_differ": out[3]}

    def autotune(self, dtype=np.uint16) -> float:
        """Times the candidate plans (batched kernel: batch, rows per band,
        LDS budget, tile shape; one-frame kernel: lanes per bin) on frames
        of this geometry, keeps the fastest and returns its frames/s."""
        fps = C.c_double()
        check(lib().mc_map_autotune(self.handle, NP_TO_CODE[np.dtype(dtype)], C.byref(fps)))
        return fps.value

    def plan_key(self, dtype=np.uint16) -> str:
        """Key of this geometry's tuned plan (the map's content, the pixel
        type, the device and the map header's parameter hash)."""
        out = C.create_string_buffer(32)
        check(lib().mc_map_plan_key(self.handle, NP_TO_CODE[np.dtype(dtype)],
                                    self._header_hash(), out, 32))
        return out.value.decode()

    def autotune_plan(self, plan_dir, dtype=np.uint16) -> dict:
        """mc_map_autotune_plan: applies plan_dir/<key>.plan when it exists,
        else tunes and writes it.  Returns {key, loaded, frames_per_s,
        batch, rows_per_group, lds_bytes, tile_bins, tile_entries,
        plain_lanes}."""
        fps, loaded = C.c_double(), C.c_int32()
        check(lib().mc_map_autotune_plan(self.handle, NP_TO_CODE[np.dtype(dtype)],
                                         str(plan_dir).encode(), self._header_hash(),
                                         C.byref(fps), C.byref(loaded)))
        b, r, l = self.tuning()
        tb, te = self.tile_shape()
        return {"key": self.plan_key(dtype), "loaded": bool(loaded.value),
                "frames_per_s": fps.value, "batch": b, "rows_per_group": r, "lds_bytes": l,
                "tile_bins": tb, "tile_entries": te, "plain_lanes": self.plain_lanes()}

    def _header_hash(self):
        if self.params is None:
            return None
        from .bin_io import header_for
        return header_for(self.params).param_hash.hex().encode()

    def set_plain_lanes(self, lanes: int):
        """Lanes per bin of the one-frame kernels: 4, 8 or 16."""
        check(lib().mc_map_set_plain_lanes(self.handle, lanes))

    def plain_lanes(self) -> int:
        v = C.c_int32()
        check(lib().mc_map_get_plain_lanes(self.handle, C.byref(v)))
        return v.value

    def layout_info(self, mode: str = "bilinear", dtype=np.uint16) -> dict:
        """Tile layout of the batched kernel under the current tuning."""
        out = (C.c_int64 * 9)()
        check(lib().mc_map_layout_info(self.handle, _lib.MODES[mode],
                                       NP_TO_CODE[np.dtype(dtype)], out))
        keys = ("tiles", "chunks", "direct_bins", "staged_px", "entries", "lds_bytes", "bands",
                "vec", "px_blocks")
        return dict(zip(keys, (int(v) for v in out)))

    def close(self):
        if self.handle:
            lib().mc_map_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def build_map(params: IntegrationParams, *, mask=None, dist_y=None, dist_z=None,
              residual=None, panel_shifts=None, ctx: Context | None = None,
              verify_pixels: int | None = None) -> CSRGeometry:
    """GPU FP64 DetectorMapper (MapperCore.c:45-424) for ``params``.
    ``mask``: 1.0 = masked pixel, raw detector orientation (NZ x NY).
    ``verify_pixels``: pixels recomputed on the host after the build (the
    context's setting when None, else this sets it; see
    ``CSRGeometry.verify_info``)."""
    ctx = ctx or default_context()
    if verify_pixels is not None:
        ctx.set_map_verify(verify_pixels)
    g = _geometry(params)
    panels = generate_panels(params, panel_shifts)
    h = C.c_void_p()
    npx = params.NrPixelsY * params.NrPixelsZ

    def chk(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != npx:
            raise ValueError(f"image of {a.size} pixels, detector has {npx}")
        return a
    mask, dist_y, dist_z, residual = chk(mask), chk(dist_y), chk(dist_z), chk(residual)
    check(lib().mc_map_build(ctx.handle, C.byref(g), _dp(mask), _dp(dist_y), _dp(dist_z),
                             _dp(residual), panels, len(panels) if panels is not None else 0,
                             C.byref(h)))
    m = CSRGeometry(h, ctx, params.n_r_bins, params.n_eta_bins, params.NrPixelsY,
                    params.NrPixelsZ, params)
    m.set_beam_center(params.BC_y, params.BC_z)
    return m


def build_csr(pixmap: PixelMap, *, n_r: int, n_eta: int, n_pixels_y: int, n_pixels_z: int,
              device="cuda", dtype=None, bc_y: float = 0.0, bc_z: float = 0.0,
              build_modes: Sequence[str] = ("floor", "bilinear", "gradient"),
              compute_variance: bool = False, mask_flags=None,
              ctx: Context | None = None) -> CSRGeometry:
    """Uploads a loaded map (kernels.py:242-375 of the reference package).
    ``device`` must be a GPU: there is no CPU path."""
    if str(device).startswith("cpu"):
        raise ValueError("midas_amd integrates on the GPU only (device='cuda')")
    # compute_variance: the reference precomputes squared weights for its
    # variance SpMV; the variance kernel here squares each weight as it
    # reads it, so every geometry supports integrate_with_variance and the
    # flag needs no extra buffer.
    if pixmap.n_bins != n_r * n_eta:
        raise ValueError(f"PixelMap has {pixmap.n_bins} bins; n_r * n_eta = {n_r * n_eta}")
    ctx = ctx or default_context()
    rec = np.ascontiguousarray(pixmap.pxList, dtype=PXLIST_DTYPE)
    pairs = pixmap.pairs()
    flags = mask_flags if mask_flags is not None else pixmap.mask_flags
    fl = np.ascontiguousarray(flags, dtype=np.int32) if flags is not None else None
    h = C.c_void_p()
    check(lib().mc_map_from_host(ctx.handle, rec.ctypes.data, rec.size, pairs.ctypes.data,
                                 fl.ctypes.data if fl is not None else None, n_pixels_y,
                                 n_pixels_z, n_r, n_eta, C.byref(h)))
    m = CSRGeometry(h, ctx, n_r, n_eta, n_pixels_y, n_pixels_z)
    m.set_beam_center(bc_y, bc_z)
    return m


def load_map_gpu(map_path, nmap_path, mask_path, *, n_r, n_eta, ny, nz,
                 ctx: Context | None = None) -> CSRGeometry:
    ctx = ctx or default_context()
    h = C.c_void_p()
    check(lib().mc_map_load(ctx.handle, str(map_path).encode(), str(nmap_path).encode(),
                            str(mask_path).encode() if mask_path else None, ny, nz, n_r,
                            n_eta, C.byref(h)))
    return CSRGeometry(h, ctx, n_r, n_eta, ny, nz)


def _flags(normalize: bool, mask_nan: bool) -> int:
    return (_lib.MC_F_NORMALIZE if normalize else 0) | (_lib.MC_F_MASK_NAN if mask_nan else 0)


def integrate_batch(frames, geom: CSRGeometry, *, mode: str = "bilinear",
                    normalize: bool = True, mask_nan: bool = True, lineouts: bool = True):
    """Cakes [F, n_r, n_eta] (+ lineouts [F, 3, n_r]) of F frames.  One frame
    runs the one-frame kernel, more run the batched kernel."""
    if mode not in _lib.MODES:
        raise ValueError(f"unknown integration mode {mode!r}; expected floor|bilinear|gradient")
    npx = geom.n_pixels_y * geom.n_pixels_z
    code = _dtype_code(frames)
    fl = _flags(normalize, mask_nan)
    if _is_cuda(frames):
        t = _torch()
        frames = frames.contiguous()
        if frames.numel() % npx:
            raise ValueError(f"frames numel {frames.numel()} is not a multiple of {npx}")
        F = frames.numel() // npx
        cakes = t.empty((F, geom.n_r, geom.n_eta), dtype=t.float32, device=frames.device)
        lo = t.empty((F, 3, geom.n_r), dtype=t.float32, device=frames.device) if lineouts else None
        t.cuda.current_stream(frames.device).synchronize()
        check(lib().mc_integrate(geom.handle, C.c_void_p(frames.data_ptr()), code, F, 1,
                                 _lib.MODES[mode], fl, C.c_void_p(cakes.data_ptr()), 1,
                                 C.c_void_p(lo.data_ptr()) if lo is not None else None, 1))
        return cakes, lo
    a = np.ascontiguousarray(frames)
    if a.size % npx:
        raise ValueError(f"frames size {a.size} is not a multiple of {npx}")
    F = a.size // npx
    cakes = np.empty((F, geom.n_r, geom.n_eta), dtype=np.float32)
    lo = np.empty((F, 3, geom.n_r), dtype=np.float32) if lineouts else None
    check(lib().mc_integrate(geom.handle, a.ctypes.data, code, F, 0, _lib.MODES[mode], fl,
                             cakes.ctypes.data, 0, lo.ctypes.data if lo is not None else None, 0))
    return cakes, lo


class Event:
    """Completion of the work queued on a context's stream before it
    (mc_event_*): ``done`` polls, ``wait`` blocks."""

    def __init__(self, ctx: Context):
        h = C.c_void_p()
        check(lib().mc_event_record(ctx.handle, C.byref(h)))
        self.handle = h

    @property
    def done(self) -> bool:
        d = C.c_int32()
        check(lib().mc_event_query(self.handle, C.byref(d)))
        return bool(d.value)

    def wait(self):
        check(lib().mc_event_wait(self.handle))

    def close(self):
        if getattr(self, "handle", None):
            lib().mc_event_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def integrate_async(frames, geom: CSRGeometry, cakes, lineouts=None, *, mode: str = "bilinear",
                    normalize: bool = True, mask_nan: bool = True) -> Event:
    """Queues the integration of device frames (a torch CUDA tensor
    [F, NZ, NY]) into device cakes [F, n_r, n_eta] (and lineouts
    [F, 3, n_r]) on the map's context stream and returns at once; the
    returned Event completes with it (mc_integrate_async).  The frames must
    be ready: the caller's torch stream is synchronised first."""
    if mode not in _lib.MODES:
        raise ValueError(f"unknown integration mode {mode!r}; expected floor|bilinear|gradient")
    if not _is_cuda(frames) or not _is_cuda(cakes):
        raise ValueError("integrate_async takes device tensors (frames, cakes)")
    t = _torch()
    npx = geom.n_pixels_y * geom.n_pixels_z
    frames = frames.contiguous()
    F = frames.numel() // npx
    if frames.numel() != F * npx or cakes.numel() != F * geom.n_bins:
        raise ValueError("frames / cakes sizes do not match the geometry")
    if lineouts is not None and lineouts.numel() != F * 3 * geom.n_r:
        raise ValueError("lineouts size does not match the geometry")
    t.cuda.current_stream(frames.device).synchronize()
    check(lib().mc_integrate_async(geom.handle, C.c_void_p(frames.data_ptr()),
                                   _dtype_code(frames), F, _lib.MODES[mode],
                                   _flags(normalize, mask_nan), C.c_void_p(cakes.data_ptr()),
                                   C.c_void_p(lineouts.data_ptr()) if lineouts is not None
                                   else None))
    return Event(geom.ctx)


def hybrid_message(frame: np.ndarray) -> bytes:
    """One frame in the hybrid wire layout (dtype 6): u32 count, the frame's
    low 16 bits as uint16, then (i32 pixel, i64 value) for every pixel
    outside 0..65535."""
    f = np.ascontiguousarray(frame).reshape(-1).astype(np.int64)
    over = np.flatnonzero((f < 0) | (f > 65535)).astype(np.int32)
    base = (f & 0xFFFF).astype(np.uint16)
    return (np.uint32(over.size).tobytes() + base.tobytes() + over.tobytes()
            + f[over].astype(np.int64).tobytes())


def integrate_hybrid(messages, n_frames: int, geom: CSRGeometry, *, mode: str = "bilinear",
                     normalize: bool = True, mask_nan: bool = True, lineouts: bool = True):
    """Cakes (and lineouts) of n_frames hybrid messages back to back
    (bytes, or a torch CUDA uint8 tensor), decoded on the GPU
    (mc_integrate with MC_DTYPE_HYBRID)."""
    if mode not in _lib.MODES:
        raise ValueError(f"unknown integration mode {mode!r}; expected floor|bilinear|gradient")
    cakes = np.empty((n_frames, geom.n_r, geom.n_eta), dtype=np.float32)
    lo = np.empty((n_frames, 3, geom.n_r), dtype=np.float32) if lineouts else None
    fl = _flags(normalize, mask_nan)
    if _is_cuda(messages):
        _torch().cuda.current_stream(messages.device).synchronize()
        ptr, on_dev = C.c_void_p(messages.data_ptr()), 1
    else:
        buf = np.frombuffer(messages, dtype=np.uint8)
        ptr, on_dev = buf.ctypes.data, 0
    check(lib().mc_integrate(geom.handle, ptr, _lib.MC_DTYPE_HYBRID, n_frames, on_dev,
                             _lib.MODES[mode], fl, cakes.ctypes.data, 0,
                             lo.ctypes.data if lo is not None else None, 0))
    return cakes, lo


def integrate(image, geom: CSRGeometry, *, mode: str = "bilinear", normalize: bool = True,
              mask_nan: bool = False):
    """One frame -> (n_r, n_eta) cake (kernels.py:416-473).  The dark set on
    ``geom`` is subtracted."""
    cakes, _ = integrate_batch(image, geom, mode=mode, normalize=normalize,
                               mask_nan=mask_nan, lineouts=False)
    return cakes[0]


def integrate_with_variance(image, geom: CSRGeometry, *, mode: str = "bilinear",
                            normalize: bool = True, variance_image=None,
                            mask_nan: bool = False):
    """Cake and propagated variance of one or more frames
    (integrate_with_variance, kernels.py:476-561).  ``variance_image``
    (same shape as ``image``) gives the per-pixel variance; None assumes
    Poisson statistics (the image minus the dark is its own variance).
    Per bin the variance is sum(w^2 * var) over the merged (bin, pixel)
    weights, divided by (sum A)^2 when normalising.  Returns
    (intensity, variance), each (n_r, n_eta) for one frame or
    (F, n_r, n_eta)."""
    if mode not in _lib.MODES:
        raise ValueError(f"unknown integration mode {mode!r}; expected floor|bilinear|gradient")
    npx = geom.n_pixels_y * geom.n_pixels_z
    code = _dtype_code(image)
    fl = _flags(normalize, mask_nan)
    single = (image.ndim == 2)
    if _is_cuda(image):
        t = _torch()
        fr = image.contiguous()
        F = fr.numel() // npx
        if fr.numel() != F * npx:
            raise ValueError(f"image numel {fr.numel()} is not a multiple of {npx}")
        v = None
        if variance_image is not None:
            v = variance_image.to(device=fr.device, dtype=t.float32).contiguous()
            if v.numel() != fr.numel():
                raise ValueError("variance_image must match image")
        cakes = t.empty((F, geom.n_r, geom.n_eta), dtype=t.float32, device=fr.device)
        var = t.empty_like(cakes)
        t.cuda.current_stream(fr.device).synchronize()
        check(lib().mc_integrate_variance(
            geom.handle, C.c_void_p(fr.data_ptr()), code, F, 1,
            C.c_void_p(v.data_ptr()) if v is not None else None, 1, _lib.MODES[mode], fl,
            C.c_void_p(cakes.data_ptr()), C.c_void_p(var.data_ptr()), 1))
    else:
        a = np.ascontiguousarray(image)
        F = a.size // npx
        if a.size != F * npx:
            raise ValueError(f"image size {a.size} is not a multiple of {npx}")
        v = None
        if variance_image is not None:
            v = np.ascontiguousarray(variance_image, dtype=np.float32)
            if v.size != a.size:
                raise ValueError("variance_image must match image")
        cakes = np.empty((F, geom.n_r, geom.n_eta), dtype=np.float32)
        var = np.empty_like(cakes)
        check(lib().mc_integrate_variance(geom.handle, a.ctypes.data, code, F, 0,
                                          v.ctypes.data if v is not None else None, 0,
                                          _lib.MODES[mode], fl, cakes.ctypes.data,
                                          var.ctypes.data, 0))
    return (cakes[0], var[0]) if single else (cakes, var)


_PROFILE_ROWS = {"cpu_mean": 0, "area_weighted": 1, "simple_mean": 2}


def profile_1d(int2d, geom: CSRGeometry, *, mode: str = "area_weighted"):
    """(n_r, n_eta) or (F, n_r, n_eta) -> (n_r,) / (F, n_r) lineout
    (kernels.py:567-601): 'area_weighted', 'simple_mean' or 'cpu_mean'."""
    if mode not in _PROFILE_ROWS:
        raise ValueError(f"unknown profile_1d mode {mode!r}")
    row = _PROFILE_ROWS[mode]
    squeeze = (int2d.ndim == 2)
    if _is_cuda(int2d):
        t = _torch()
        c = int2d.reshape(-1, geom.n_r, geom.n_eta).to(t.float32).contiguous()
        F = c.shape[0]
        lo = t.empty((F, 3, geom.n_r), dtype=t.float32, device=c.device)
        t.cuda.current_stream(c.device).synchronize()
        check(lib().mc_profile(geom.handle, C.c_void_p(c.data_ptr()), F, 1,
                               C.c_void_p(lo.data_ptr()), 1))
    else:
        c = np.ascontiguousarray(int2d, dtype=np.float32).reshape(-1, geom.n_r, geom.n_eta)
        F = c.shape[0]
        lo = np.empty((F, 3, geom.n_r), dtype=np.float32)
        check(lib().mc_profile(geom.handle, c.ctypes.data, F, 0, lo.ctypes.data, 0))
    out = lo[:, row, :]
    return out[0] if squeeze else out


def profile_1d_with_variance(int2d, var2d, geom: CSRGeometry, *, mode: str = "area_weighted"):
    """(I_1d, Var_1d) of a cake and its propagated variance
    (profile_1d_with_variance, kernels.py:604-644): 'area_weighted' gives
    sum(A I)/sum(A) and sum(A^2 Var)/sum(A)^2, 'simple_mean' the mean of I
    and sum(Var)/N^2, over the bins with A > 1e-9 (NaN bins missing).
    (n_r, n_eta) -> (n_r,) each, (F, n_r, n_eta) -> (F, n_r)."""
    if mode not in ("area_weighted", "simple_mean"):
        raise ValueError(f"unknown profile_1d_with_variance mode {mode!r}")
    code = _PROFILE_ROWS[mode]
    squeeze = (int2d.ndim == 2)
    if _is_cuda(int2d):
        t = _torch()
        c = int2d.reshape(-1, geom.n_r, geom.n_eta).to(t.float32).contiguous()
        v = var2d.reshape(-1, geom.n_r, geom.n_eta).to(device=c.device, dtype=t.float32)
        v = v.contiguous()
        if v.shape != c.shape:
            raise ValueError("var2d must match int2d")
        F = c.shape[0]
        prof = t.empty((F, geom.n_r), dtype=t.float32, device=c.device)
        pvar = t.empty_like(prof)
        t.cuda.current_stream(c.device).synchronize()
        check(lib().mc_profile_variance(geom.handle, C.c_void_p(c.data_ptr()),
                                        C.c_void_p(v.data_ptr()), F, 1, code,
                                        C.c_void_p(prof.data_ptr()),
                                        C.c_void_p(pvar.data_ptr()), 1))
    else:
        c = np.ascontiguousarray(int2d, dtype=np.float32).reshape(-1, geom.n_r, geom.n_eta)
        v = np.ascontiguousarray(var2d, dtype=np.float32).reshape(-1, geom.n_r, geom.n_eta)
        if v.shape != c.shape:
            raise ValueError("var2d must match int2d")
        F = c.shape[0]
        prof = np.empty((F, geom.n_r), dtype=np.float32)
        pvar = np.empty_like(prof)
        check(lib().mc_profile_variance(geom.handle, c.ctypes.data, v.ctypes.data, F, 0, code,
                                        prof.ctypes.data, pvar.ctypes.data, 0))
    return (prof[0], pvar[0]) if squeeze else (prof, pvar)


class IntegrationGraph:
    """The batched passes over one device frame buffer, captured once in a
    HIP graph and replayed by :meth:`launch`.  The caller refills
    ``frames`` (a CUDA tensor [F, NZ, NY]) between launches; each launch
    writes ``cakes`` [F, n_r, n_eta] and, with ``lineouts``, ``lineouts``
    [F, 3, n_r] in place.  Results equal :func:`integrate_batch` on the same
    frames (the same kernels run).  A new dark is picked up by the next
    launch; a new tuning leaves the graph on its captured layout; a new beam
    centre in gradient mode invalidates it: :meth:`launch` then raises."""

    def __init__(self, geom: CSRGeometry, frames, *, mode: str = "bilinear",
                 normalize: bool = True, mask_nan: bool = True, lineouts: bool = True):
        if mode not in _lib.MODES:
            raise ValueError(f"unknown integration mode {mode!r}")
        if not _is_cuda(frames):
            raise ValueError("IntegrationGraph reads device frames (a CUDA tensor)")
        t = _torch()
        npx = geom.n_pixels_y * geom.n_pixels_z
        if not frames.is_contiguous() or frames.numel() % npx or frames.numel() == 0:
            raise ValueError(f"frames must be a contiguous [F, {geom.n_pixels_z}, "
                             f"{geom.n_pixels_y}] tensor")
        F = frames.numel() // npx
        self.geom, self.frames = geom, frames
        self.cakes = t.empty((F, geom.n_r, geom.n_eta), dtype=t.float32, device=frames.device)
        self.lineouts = (t.empty((F, 3, geom.n_r), dtype=t.float32, device=frames.device)
                         if lineouts else None)
        self.handle = C.c_void_p()
        t.cuda.current_stream(frames.device).synchronize()
        check(lib().mc_graph_create(
            geom.handle, C.c_void_p(frames.data_ptr()), _dtype_code(frames), F,
            _lib.MODES[mode], _flags(normalize, mask_nan), C.c_void_p(self.cakes.data_ptr()),
            C.c_void_p(self.lineouts.data_ptr()) if lineouts else None,
            C.byref(self.handle)))
        geom.ctx.sync()

    def launch(self, sync: bool = True):
        """Replays the graph on the map's stream after the work already
        queued on torch's current stream (the writes into ``frames``)."""
        _torch().cuda.current_stream(self.frames.device).synchronize()
        check(lib().mc_graph_launch(self.handle))
        if sync:
            self.geom.ctx.sync()
        return self.cakes, self.lineouts

    def close(self):
        if self.handle:
            lib().mc_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def time_integrate(geom: CSRGeometry, frames_dev, *, kind: str, mode="bilinear",
                   normalize=True, mask_nan=True, iters=10, warmup=2, cakes_dev=None,
                   lineouts_dev=None):
    """HIP-event timing of the kernels on device-resident frames.
    Returns (ms per launch, ms per pass over all frames)."""
    t = _torch()
    npx = geom.n_pixels_y * geom.n_pixels_z
    F = frames_dev.numel() // npx
    if cakes_dev is None:
        cakes_dev = t.empty((F, geom.n_r, geom.n_eta), dtype=t.float32, device=frames_dev.device)
    a, b = C.c_double(), C.c_double()
    t.cuda.current_stream(frames_dev.device).synchronize()
    check(lib().mc_time_integrate(
        geom.handle, C.c_void_p(frames_dev.data_ptr()), _dtype_code(frames_dev), F,
        _lib.MODES[mode], _flags(normalize, mask_nan),
        {"plain": 0, "batched": 1, "graph": 2}[kind], iters,
        warmup, C.c_void_p(cakes_dev.data_ptr()),
        C.c_void_p(lineouts_dev.data_ptr()) if lineouts_dev is not None else None,
        C.byref(a), C.byref(b)))
    return a.value, b.value
