"""Restatement of the reference's fused multi-map averaging (numpy, FP64).

TEST INFRASTRUCTURE ONLY.  build_fused_geometry
(packages/midas_integrate/midas_integrate/fused_csr.py:126-260) folds
M = (1/N) sum_k P_k D_k M_k into one matrix; applied to an image this is
(1/N) sum_k P_k (D_k (M_k img)), which is what these functions compute:
per map the raw bin sums M_k img (the oracle's integrate without
normalisation), divided by the in-bounds area where it exceeds 1e-9
(D_k, kernels.py:311-318), interpolated linearly along R onto map 0's axis
(P_k, fused_csr.py:64-99), and averaged.
"""
from __future__ import annotations

import numpy as np

AREA_THRESHOLD = 1e-9


def r_interp(src, tgt):
    """(j, w_lo, w_hi) per target R (fused_csr.py:79-93): clamped to the
    first pair below src[0] and to the last pair above src[-1]."""
    src = np.asarray(src, dtype=np.float64)
    n = len(src)
    j = np.empty(len(tgt), dtype=np.int64)
    wl = np.empty(len(tgt))
    wh = np.empty(len(tgt))
    for i, r in enumerate(tgt):
        if r <= src[0]:
            j[i], wl[i], wh[i] = 0, 1.0, 0.0
        elif r >= src[-1]:
            j[i], wl[i], wh[i] = n - 2, 0.0, 1.0
        else:
            k = int(np.searchsorted(src, r, side="right") - 1)
            k = max(0, min(k, n - 2))
            d = src[k + 1] - src[k]
            t = (r - src[k]) / d if d > 0 else 0.0
            j[i], wl[i], wh[i] = k, 1.0 - t, t
    return j, wl, wh


def fused_cake(raw_cakes, areas, r_axes):
    """raw_cakes[k], areas[k]: [n_r_k, n_eta] un-normalised sums and in-bounds
    areas of map k; r_axes[k]: its R bin centres.  Returns the fused cake
    on r_axes[0]."""
    out = None
    for raw, A, R in zip(raw_cakes, areas, r_axes):
        X = np.where(A > AREA_THRESHOLD, raw / np.maximum(A, AREA_THRESHOLD), 0.0)
        j, wl, wh = r_interp(R, r_axes[0])
        Y = wl[:, None] * X[j] + wh[:, None] * X[j + 1]
        out = Y if out is None else out + Y
    return out / len(raw_cakes)
