"""Restatement of the reference's 1-D lineout with variance (numpy, FP64).

TEST INFRASTRUCTURE ONLY: the GPU kernel behind mc_profile_variance is
checked against this function, which restates profile_1d_with_variance,
packages/midas_integrate/midas_integrate/kernels.py:604-644, bin by bin:
the bins of a row that count are those whose in-bounds area exceeds the
threshold (AREA_THRESHOLD, 1e-9) and whose value is not NaN (the cakes of
this build mark masked bins with NaN; the reference's own cakes hold none).
"""
from __future__ import annotations

import numpy as np

AREA_THRESHOLD = 1e-9


def profile_with_variance(cake, var, area, mode="area_weighted"):
    """cake, var: [n_r, n_eta]; area: [n_r, n_eta] in-bounds areas.
    area_weighted (kernels.py:625-634): I = sum A I / sum A,
    V = sum A^2 Var / (sum A)^2, 0 where sum A <= 1e-9;
    simple_mean (kernels.py:635-644): I = sum I / N, V = sum Var / N^2,
    0 where N = 0.  Returns (I, V), each [n_r], FP64."""
    cake = np.asarray(cake, np.float64)
    var = np.asarray(var, np.float64)
    area = np.asarray(area, np.float64)
    n_r, n_eta = cake.shape
    prof = np.zeros(n_r)
    pvar = np.zeros(n_r)
    for r in range(n_r):
        s1 = s2 = sw = 0.0
        for e in range(n_eta):
            if not (area[r, e] > AREA_THRESHOLD) or np.isnan(cake[r, e]):
                continue
            w = area[r, e] if mode == "area_weighted" else 1.0
            s1 += w * cake[r, e]
            s2 += w * w * var[r, e]
            sw += w
        ok = sw > AREA_THRESHOLD if mode == "area_weighted" else sw > 0
        if ok:
            prof[r] = s1 / sw
            pvar[r] = s2 / (sw * sw)
    return prof, pvar
