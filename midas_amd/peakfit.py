"""Batched peak fits on the GPU (mc_fit_windows / mc_fit_cake_rings).

Names follow the reference package (packages/midas_integrate/midas_integrate/
peakfit.py) and the C module PeakFit.h: PF_PARAMS_PER_PEAK = 7 values per
peak [area, center, sig, gam, FWHM, eta, chi2].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib
from .caking import CSRGeometry, Context, _is_cuda, _torch, default_context

PF_PARAMS_PER_PEAK = 7
MODELS = {"pv": _lib.MC_FIT_PV, "tch": _lib.MC_FIT_TCH}


def fit_windows(x, y, peak_guess, rstep: float, *, model: str = "pv", max_iter: int = 100,
                ctx: Context | None = None) -> np.ndarray:
    """Fits W windows of N samples each ([W, N] arrays) in one launch.
    Returns [W, 8]: area, center, sig, gam, FWHM, eta, chi2 (SNR for pV),
    status (iterations; negative when the window could not be fitted)."""
    ctx = ctx or default_context()
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    g = np.ascontiguousarray(peak_guess, dtype=np.float32).reshape(-1)
    if x.shape != y.shape or x.ndim != 2 or g.size != x.shape[0]:
        raise ValueError("x, y must be [W, N] and peak_guess [W]")
    out = np.zeros((x.shape[0], _lib.MC_FIT_OUT), dtype=np.float64)
    check(lib().mc_fit_windows(ctx.handle, MODELS[model], x.shape[0], x.shape[1],
                               x.ctypes.data, y.ctypes.data, g.ctypes.data, rstep, max_iter,
                               out.ctypes.data, 0))
    return out


def fit_cake_rings(cakes, geom: CSRGeometry, r_centers, rings, *, pad: int = 20,
                   model: str = "pv", max_iter: int = 100):
    """Per-eta fits of every ring in every cake, in one launch
    (IntegratorZarrOMP.c:1903-2070).  ``cakes`` is a CUDA tensor
    [F, n_r, n_eta]; returns [F, n_eta, n_rings, 8] float64 (host)."""
    if not _is_cuda(cakes):
        raise TypeError("cakes must be a device tensor (the output of integrate_batch)")
    t = _torch()
    c = cakes.contiguous()
    F = c.shape[0]
    rc = np.ascontiguousarray(r_centers, dtype=np.float64)
    rg = np.ascontiguousarray(rings, dtype=np.float64)
    out = np.zeros((F, geom.n_eta, rg.size, _lib.MC_FIT_OUT), dtype=np.float64)
    t.cuda.current_stream(c.device).synchronize()
    check(lib().mc_fit_cake_rings(geom.handle, C.c_void_p(c.data_ptr()), F,
                                  rc.ctypes.data_as(C.POINTER(C.c_double)),
                                  rg.ctypes.data_as(C.POINTER(C.c_double)), rg.size, pad,
                                  MODELS[model], max_iter, out.ctypes.data, 0))
    return out


def fit_peaks(x, intensity, peak_locations, x_bin_size: float, fit_roi_padding: int = 20,
              *, max_iter: int = 200, ctx: Context | None = None) -> np.ndarray:
    """fitPeaks (PeakFit.c:275-544) on a 1-D profile: the TCH model on the
    ROI of +-fit_roi_padding bins around the bin nearest each location
    (locations farther than 2 bins from every sample are skipped, as in
    the reference).  Returns [n_peaks, 7]; rows of skipped peaks are 0.
    Each peak is one window (the reference merges overlapping ROIs into a
    multi-peak job; here each is fitted on its own)."""
    x = np.asarray(x, dtype=np.float64)
    yv = np.asarray(intensity, dtype=np.float64)
    locs = np.atleast_1d(np.asarray(peak_locations, dtype=np.float64))
    out = np.zeros((locs.size, PF_PARAMS_PER_PEAK))
    wins, keep = [], []
    for i, p in enumerate(locs):
        k = int(np.argmin(np.abs(x - p)))
        if abs(x[k] - p) >= 2.0 * x_bin_size:
            continue
        lo, hi = max(0, k - fit_roi_padding), min(x.size - 1, k + fit_roi_padding)
        wins.append((lo, hi, x[k]))
        keep.append(i)
    if not wins:
        return out
    # One launch per distinct window length (ROIs clipped at the profile
    # ends are shorter).
    by_len: dict[int, list] = {}
    for (lo, hi, c), i in zip(wins, keep):
        by_len.setdefault(hi - lo + 1, []).append((lo, hi, c, i))
    for L, items in by_len.items():
        X = np.stack([x[lo:hi + 1] for lo, hi, _, _ in items])
        Y = np.stack([yv[lo:hi + 1] for lo, hi, _, _ in items])
        G = np.array([c for _, _, c, _ in items])
        r = fit_windows(X, Y, G, x_bin_size, model="tch", max_iter=max_iter, ctx=ctx)
        for row, (_, _, _, i) in zip(r, items):
            if row[7] > 0:
                out[i] = row[:PF_PARAMS_PER_PEAK]
    return out


def pseudo_voigt(x, center, mu, fwhm, imax, bg):
    """Height-normalised pV of pf_pv_singlet_obj (PeakFit.c:882-920)."""
    x = np.asarray(x, dtype=np.float64)
    d2 = (x - center) ** 2
    L = 1.0 / (1.0 + 4.0 * d2 / fwhm ** 2)
    G = np.exp(-4.0 * np.log(2.0) * d2 / fwhm ** 2)
    return bg + imax * (mu * L + (1.0 - mu) * G)
