"""Restatement of the reference's eta filters on cakes (numpy, FP64).

TEST INFRASTRUCTURE ONLY: the GPU filters (mc_smooth_eta, mc_median_eta)
are checked against these functions, which restate
packages/midas_integrate/midas_integrate/postprocess.py:36-142 with an
explicit index map for the padded eta axis instead of array padding.
"""
from __future__ import annotations

import numpy as np


def eta_index(i, n, periodic):
    """Bin read at padded position i of an eta axis of n bins: wrap when
    periodic (postprocess.py:84-86), else numpy's "reflect" (mirror without
    the edge, postprocess.py:87-89)."""
    i = np.asarray(i)
    if periodic:
        return np.mod(i, n)
    if n == 1:
        return np.zeros_like(i)
    p = 2 * (n - 1)
    m = np.mod(i, p)
    return np.where(m < n, m, p - m)


def gauss_weights(sigma_bins):
    """Taps k = -h..h and normalised weights (postprocess.py:71-74); h uses
    Python's round (half to even)."""
    half = max(1, int(round(4 * sigma_bins)))
    k = np.arange(-half, half + 1, dtype=np.float64)
    w = np.exp(-(k ** 2) / (2.0 * sigma_bins ** 2))
    return half, k, w / w.sum()


def gauss_smooth_eta(cake, sigma_bins, area_per_bin=None, periodic=True):
    """postprocess.py:36-103: out = sum_k w_k (I A)[j+k] / sum_k w_k A[j+k],
    the input where the weighted area is 0; FP64, cast back to the input
    dtype."""
    c = np.asarray(cake)
    n_r, n_eta = c.shape
    half, k, w = gauss_weights(sigma_bins)
    A = (np.ones((n_r, n_eta)) if area_per_bin is None
         else np.asarray(area_per_bin).reshape(n_r, n_eta).astype(np.float64))
    IA = c.astype(np.float64) * A if area_per_bin is not None else c.astype(np.float64)
    j = np.arange(n_eta)
    sia = np.zeros((n_r, n_eta))
    sa = np.zeros((n_r, n_eta))
    for kk, wk in zip(k.astype(int), w):
        q = eta_index(j + kk, n_eta, periodic)
        sia = sia + wk * IA[:, q]
        sa = sa + wk * A[:, q]
    with np.errstate(divide="ignore", invalid="ignore"):
        out = np.where(sa > 0, sia / sa, c.astype(np.float64))
    return out.astype(c.dtype)


def median_filter_eta(cake, window=3, periodic=True, preserve_ring_flux=True):
    """postprocess.py:106-142: the median of each window (NaN if it holds a
    NaN), then rows rescaled by sum(in)/sum(out) (1 where sum(out) == 0)."""
    c = np.asarray(cake)
    if window < 1 or window % 2 == 0:
        raise ValueError("window must be a positive odd integer")
    n_r, n_eta = c.shape
    half = window // 2
    j = np.arange(n_eta)
    cols = np.stack([c[:, eta_index(j + kk, n_eta, periodic)] for kk in range(-half, half + 1)],
                    axis=2)
    # a C-ordered result like the reference's np.empty_like: numpy's float32
    # row sums below depend on the memory layout (pairwise over contiguous
    # rows)
    out = np.empty_like(c, order="C")
    out[...] = np.median(cols, axis=2)
    if preserve_ring_flux:
        s_in = c.sum(axis=1, keepdims=True)
        s_out = out.sum(axis=1, keepdims=True)
        with np.errstate(divide="ignore", invalid="ignore"):
            scale = np.where(s_out != 0, s_in / s_out, 1.0)
        out = out * scale
    return out
