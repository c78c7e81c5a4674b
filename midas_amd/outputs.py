"""Result files of an integration run (the mc_out_* C ABI).

``ResultFiles`` writes, from the writer thread of the native library, the
files the reference integrators write:

* ``<name>.caked.hdf`` (IntegratorZarrOMP.c:1671-1687, 2084-2168):
  ``/IntegrationResult/FrameNr_<i>`` with attributes ``omega``, ``Header``,
  ``Units``; ``/OmegaSumFrame/LastFrameNumber_<i>`` sums of
  ``OmegaSumFrames`` frames; ``/REtaMap`` [5, n_r, n_eta]; ``/Omegas``;
  ``/SumFrames``; ``/InstrumentParameters/{Polariz, Lam, SH_L, U, V, W, X,
  Y, Z, Distance}``,
* ``<prefix>_lineout.bin`` / ``<prefix>_lineout.xy``
  (IntegratorZarrOMP.c:1829-1863),
* ``lineout.bin``, ``lineout_simple_mean.bin``, ``Int2D.bin``
  (IntegratorFitPeaksGPUStream.cu:506-565),
* ``<prefix>_fit.bin``, ``<prefix>_fit_per_eta.csv`` and
  ``<prefix>_caked_peaks.h5`` (IntegratorZarrOMP.c:1530-1549, 2038-2058,
  PeakFitIO.c:52-267),
* with ``"caked_zarr"``: ``<name>.caked.hdf.zarr.zip``, the caked file's
  groups, datasets and attributes as a Zarr zip store (the conversion of
  FF_HEDM/workflows/integrator.py:379-426); :func:`enrich_zarr` then adds
  the input zip's metadata.

``<prefix>`` is the data file name without its extension.  The HDF5 files
are written by the library's own HDF5 writer (no libhdf5 dependency).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import Optional

import numpy as np

from . import _lib
from ._lib import OutConfig, PeakRow, check, lib
from .params import IntegrationParams, bin_edges

FILES = {
    "caked_hdf": _lib.MC_OUT_CAKED_HDF, "per_frame": _lib.MC_OUT_PER_FRAME,
    "sum_frames": _lib.MC_OUT_SUM_FRAMES, "omegas": _lib.MC_OUT_OMEGAS,
    "lineout": _lib.MC_OUT_LINEOUT, "stream": _lib.MC_OUT_STREAM,
    "int2d": _lib.MC_OUT_INT2D, "peaks": _lib.MC_OUT_PEAKS,
    "caked_zarr": _lib.MC_OUT_CAKED_ZARR,
}


def default_files(P: IntegrationParams) -> set:
    """The files IntegratorZarrOMP writes for a parameter set: the caked
    file (per-frame results when SaveIndividualFrames, /SumFrames when
    SumImages, /Omegas when an omega step is set), the lineouts, and the
    peak files when DoPeakFit."""
    f = {"caked_hdf", "lineout"}
    if P.SaveIndividualFrames:
        f.add("per_frame")
    if P.SumImages:
        f.add("sum_frames")
    if P.OmegaStep != 0.0:
        f.add("omegas")
    if P.DoPeakFit:
        f.add("peaks")
    return f


class ResultFiles:
    """One output set.  ``frame()`` queues a frame's cake (and lineouts) to
    the writer thread; ``close()`` (or leaving a ``with`` block) writes the
    closing datasets and returns once every file is complete."""

    def __init__(self, directory, data_name: str, params: IntegrationParams, *,
                 files=None, omega_sum_frames: Optional[int] = None, reta=None,
                 source: Optional[str] = None, reta_stream=None):
        P = params
        self.params = P
        RLo, RHi, ELo, EHi = bin_edges(P)
        self.n_r, self.n_eta = RLo.size, ELo.size
        self.r_centers = (RLo + RHi) / 2.0
        self._eta_lo = np.ascontiguousarray(ELo)
        self._eta_hi = np.ascontiguousarray(EHi)
        names = default_files(P) if files is None else set(files)
        unknown = names - set(FILES)
        if unknown:
            raise ValueError(f"unknown output files {sorted(unknown)}")
        mask = 0
        for n in names:
            mask |= FILES[n]
        self._inst = np.array([P.Polariz, P.Lam, P.SHpL, P.U, P.V, P.W, P.X, P.Y, P.Z, P.Lsd],
                              dtype=np.float64)
        self._reta = None
        if reta is not None:
            self._reta = np.ascontiguousarray(reta, dtype=np.float64)
            if self._reta.shape != (5, self.n_r, self.n_eta):
                raise ValueError(f"reta must be [5, {self.n_r}, {self.n_eta}]")
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        self._src = (source if source is not None else str(data_name)).encode()
        cfg = OutConfig(
            n_r=self.n_r, n_eta=self.n_eta, files=mask,
            omega_sum_frames=P.OmegaSumFrames if omega_sum_frames is None else omega_sum_frames,
            px=P.pxY, lsd=P.Lsd, wavelength=P.Lam, q_mode=int(P.q_mode), q_min=P.QMin,
            q_bin_size=P.QBinSize, reta=dp(self._reta) if self._reta is not None else None,
            r_centers=dp(self.r_centers), eta_lo=dp(self._eta_lo), eta_hi=dp(self._eta_hi),
            instrument=dp(self._inst), source=self._src)
        Path(directory).mkdir(parents=True, exist_ok=True)
        if reta_stream is not None and "stream" in names:
            # RTthEtaAreaMap.bin of the GPU stream (.cu:2443-2462): the five
            # [n_r * n_eta] float64 rows of reta_map(in_bounds=True)
            rs = np.ascontiguousarray(reta_stream, dtype="<f8")
            if rs.shape != (5, self.n_r, self.n_eta):
                raise ValueError(f"reta_stream must be [5, {self.n_r}, {self.n_eta}]")
            rs.tofile(Path(directory) / "RTthEtaAreaMap.bin")
        self.handle = C.c_void_p()
        check(lib().mc_out_open(str(directory).encode(), str(data_name).encode(), C.byref(cfg),
                                C.byref(self.handle)))
        self.files = names

    def frame(self, frame_nr: int, cake, lineouts=None, omega: Optional[float] = None,
              meta=None) -> None:
        """Queues one frame: cake [n_r, n_eta], lineouts [3, n_r] (rows
        MC_LINEOUT_*), omega in degrees (default OmegaStart + i * OmegaStep),
        meta = (Temperature, Pressure, I, I0)."""
        c = np.ascontiguousarray(_host(cake), dtype=np.float32)
        if c.shape != (self.n_r, self.n_eta):
            raise ValueError(f"cake must be [{self.n_r}, {self.n_eta}]")
        lo = None
        if lineouts is not None:
            lo = np.ascontiguousarray(_host(lineouts), dtype=np.float32)
            if lo.shape != (3, self.n_r):
                raise ValueError(f"lineouts must be [3, {self.n_r}]")
        if omega is None:
            omega = self.params.OmegaStart + frame_nr * self.params.OmegaStep
        m = None if meta is None else np.ascontiguousarray(meta, dtype=np.float64)
        if m is not None and m.size != 4:
            raise ValueError("meta is (Temperature, Pressure, I, I0)")
        fp = C.POINTER(C.c_float)
        check(lib().mc_out_frame(self.handle, int(frame_nr), float(omega), c.ctypes.data_as(fp),
                                 lo.ctypes.data_as(fp) if lo is not None else None,
                                 m.ctypes.data_as(C.POINTER(C.c_double)) if m is not None
                                 else None))

    def lineout_fits(self, frame_nr: int, params) -> None:
        """fitPeaks results of one frame's lineout, [n_peaks, 7]."""
        p = np.ascontiguousarray(params, dtype=np.float64).reshape(-1, 7)
        check(lib().mc_out_lineout_fits(self.handle, int(frame_nr), p.shape[0],
                                        p.ctypes.data_as(C.POINTER(C.c_double))))

    def peaks(self, frames, eta_bins, peak_nrs, params) -> None:
        """Per-eta fit rows: params [n, 7] (area, center, sig, gam, FWHM,
        eta, chi2 or SNR); rows with center 0 are skipped."""
        p = np.asarray(params, dtype=np.float64).reshape(-1, 7)
        n = p.shape[0]
        rows = (PeakRow * max(n, 1))()
        f, e, k = (np.broadcast_to(np.asarray(a, dtype=np.int32), (n,)) for a in
                   (frames, eta_bins, peak_nrs))
        for i in range(n):
            rows[i].frame, rows[i].eta_bin, rows[i].peak = int(f[i]), int(e[i]), int(k[i])
            rows[i].p[:] = p[i].tolist()
        check(lib().mc_out_peaks(self.handle, n, rows))

    def peaks_from_ring_fits(self, fits, frame0: int = 0, model: str = "pv") -> None:
        """Rows from fit_cake_rings output [F, n_eta, n_rings, 8].  As in
        IntegratorZarrOMP.c:2013-2026, a pV singlet keeps only its centre
        and SNR, and only when the fit succeeded with SNR > 1; TCH fits keep
        all seven parameters."""
        a = np.asarray(fits, dtype=np.float64)
        F, E, K, _ = a.shape
        ok = a[..., 7] >= 0
        if model == "pv":
            ok &= a[..., 6] > 1.0
        fi, ei, ki = np.nonzero(ok)
        if fi.size == 0:
            return
        p = a[fi, ei, ki, :7].copy()
        if model == "pv":
            keep = p[:, [1, 6]].copy()
            p[:] = 0.0
            p[:, 1], p[:, 6] = keep[:, 0], keep[:, 1]
        self.peaks(fi + frame0, ei, ki, p)

    def close(self) -> None:
        if self.handle:
            h, self.handle = self.handle, None
            check(lib().mc_out_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def _host(a):
    if hasattr(a, "is_cuda") and a.is_cuda:
        return a.detach().cpu().numpy()
    return np.asarray(a)
