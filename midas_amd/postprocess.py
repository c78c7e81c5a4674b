"""Post-process filters on cakes along eta, on the GPU (mc_smooth_eta /
mc_median_eta).

Mirrors the reference's packages/midas_integrate/midas_integrate/
postprocess.py:36-142 (gauss_smooth_eta, median_filter_eta): same
arguments, same results (FP64 arithmetic in the reference's order); a cake
is ``(n_r, n_eta)`` or a batch ``(F, n_r, n_eta)``, numpy or a torch CUDA
tensor (the result has the input's kind and dtype float32).
"""
from __future__ import annotations

import numpy as np

from ._lib import check, lib
from .caking import Context, _is_cuda, _torch, default_context


def _run(fn, cake, area, ctx, args):
    ctx = ctx or default_context()
    if _is_cuda(cake):
        t = _torch()
        c = cake.contiguous().to(t.float32)
        shape = c.shape
        if c.dim() not in (2, 3):
            raise ValueError(f"cake must be (n_r, n_eta) or (F, n_r, n_eta), got {tuple(shape)}")
        n_c = 1 if c.dim() == 2 else shape[0]
        n_r, n_eta = shape[-2], shape[-1]
        a = None
        if area is not None:
            a = (area if _is_cuda(area) else t.as_tensor(np.asarray(area), device=c.device))
            a = a.reshape(n_r, n_eta).contiguous().to(t.float32)
        out = t.empty_like(c)
        t.cuda.current_stream().synchronize()
        check(fn(ctx.handle, c.data_ptr(), n_c, n_r, n_eta, *args(a.data_ptr() if a is not None
                                                                  else None), 1, out.data_ptr()))
        return out
    c = np.ascontiguousarray(cake, dtype=np.float32)
    if c.ndim not in (2, 3):
        raise ValueError(f"cake must be (n_r, n_eta) or (F, n_r, n_eta), got {c.shape}")
    n_c = 1 if c.ndim == 2 else c.shape[0]
    n_r, n_eta = c.shape[-2], c.shape[-1]
    a = None
    if area is not None:
        a = np.asarray(area.detach().cpu().numpy() if _is_cuda(area) else area)
        a = np.ascontiguousarray(a.reshape(n_r, n_eta), dtype=np.float32)
    out = np.empty_like(c)
    check(fn(ctx.handle, c.ctypes.data, n_c, n_r, n_eta,
             *args(a.ctypes.data if a is not None else None), 0, out.ctypes.data))
    return out


def gauss_smooth_eta(cake, sigma_bins: float, *, area_per_bin=None, periodic: bool = True,
                     ctx: Context | None = None):
    """1-D Gaussian smoothing along eta (postprocess.py:36-103).

    sigma_bins: width in eta bins (taps: max(1, round(4 sigma)) per side).
    area_per_bin: optional ``(n_r * n_eta,)`` or ``(n_r, n_eta)`` areas; the
    smoothing is then area-weighted and keeps each ring's sum of I*A.
    periodic: wrap the eta axis; otherwise reflect at its ends."""
    return _run(lib().mc_smooth_eta, cake, area_per_bin, ctx,
                lambda a: (float(sigma_bins), a, int(bool(periodic))))


def median_filter_eta(cake, window: int = 3, *, periodic: bool = True,
                      preserve_ring_flux: bool = True, ctx: Context | None = None):
    """1-D median filter along eta (postprocess.py:106-142); with
    preserve_ring_flux each row is rescaled to its input sum."""
    if window < 1 or window % 2 == 0:
        raise ValueError(f"window must be a positive odd integer, got {window}")
    return _run(lib().mc_median_eta, cake, None, ctx,
                lambda a: (int(window), int(bool(periodic)), int(bool(preserve_ring_flux))))
