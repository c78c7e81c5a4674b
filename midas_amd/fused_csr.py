"""Multi-map averaging fused into one map (mc_map_fuse).

Mirrors build_fused_geometry of the reference package
(packages/midas_integrate/midas_integrate/fused_csr.py:126-260): N maps
with R edges shifted by k * RBinSize / N are built (on the GPU), and
M = (1/N) sum_k P_k D_k M_k is folded into one matrix, so one sparse pass
per frame gives the averaged cake on map 0's R axis.  The fused geometry
has unit bin areas: normalising divides by 1.
"""
from __future__ import annotations

import ctypes as C
from copy import deepcopy

import numpy as np

from . import _lib
from ._lib import check, lib
from .caking import CSRGeometry, Context, build_map, default_context
from .params import IntegrationParams


def r_axis(*, n_r: int, RMin: float, RBinSize: float) -> np.ndarray:
    """R at bin centres (kernels.py r_axis)."""
    return RMin + RBinSize * (np.arange(n_r, dtype=np.float64) + 0.5)


def build_fused_geometry(params: IntegrationParams, *, n_shifts: int = 4,
                         mode: str = "gradient", device="cuda", dtype=None,
                         verbose: bool = False, ctx: Context | None = None) -> CSRGeometry:
    """Fused multi-map geometry (fused_csr.py:126-260).  ``mode`` selects the
    per-shift matrices ('bilinear' or 'gradient'); integrate the result in
    either mode, with or without normalisation (the areas are 1)."""
    if n_shifts < 1:
        raise ValueError(f"n_shifts must be >= 1; got {n_shifts}")
    if mode not in ("bilinear", "gradient"):
        raise ValueError(f"mode must be 'bilinear' or 'gradient'; got {mode!r}")
    if str(device).startswith("cpu"):
        raise ValueError("midas_amd integrates on the GPU only (device='cuda')")
    rbin = float(params.RBinSize)
    if rbin <= 0:
        raise ValueError(f"params.RBinSize must be > 0; got {rbin}")
    ctx = ctx or default_context()
    geoms, axes = [], []
    for k in range(n_shifts):
        p = deepcopy(params)
        p.RMin = params.RMin + rbin * k / n_shifts
        g = build_map(p, ctx=ctx)
        geoms.append(g)
        axes.append(r_axis(n_r=p.n_r_bins, RMin=p.RMin, RBinSize=p.RBinSize))
        if verbose:
            print(f"[fused] shift {k}/{n_shifts} RMin+={rbin * k / n_shifts:.4f} "
                  f"({p.n_r_bins} R bins)")
    r_all = np.ascontiguousarray(np.concatenate(axes), dtype=np.float64)
    handles = (C.c_void_p * n_shifts)(*[g.handle for g in geoms])
    h = C.c_void_p()
    check(lib().mc_map_fuse(handles, n_shifts, r_all.ctypes.data, _lib.MODES[mode], C.byref(h)))
    out = CSRGeometry(h, ctx, geoms[0].n_r, geoms[0].n_eta, params.NrPixelsY,
                      params.NrPixelsZ, params)
    out.fused = True
    return out
