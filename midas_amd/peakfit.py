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
MODELS = {"pv": _lib.MC_FIT_PV, "tch": _lib.MC_FIT_TCH, "pv2": _lib.MC_FIT_PV2}
MODELS.update({f"tch{n}": _lib.MC_FIT_TCH_PEAKS(n) for n in range(2, _lib.MC_FIT_MAX_PEAKS + 1)})


def _guesses(model: str) -> int:
    """Guesses per window (= result rows for the TCH models)."""
    if model in ("pv", "tch"):
        return 1
    if model == "pv2":
        return 2
    return int(model[3:])


def fit_windows(x, y, peak_guess, rstep: float, *, model: str = "pv", max_iter: int = 100,
                ctx: Context | None = None) -> np.ndarray:
    """Fits W windows of N samples each ([W, N] arrays) in one launch.
    Returns [W, 8]: area, center, sig, gam, FWHM, eta, chi2 (SNR for pV),
    status (iterations; negative when the window could not be fitted).
    model "pv2" (pV doublet, PeakFit.c:984-1101) takes peak_guess [W, 2]
    (the two centres) and returns R1, R2, SNR1, SNR2, Gamma1, Gamma2, mu,
    status.  model "tchK" (K = 2..8 TCH peaks on each window, a merged
    fitPeaks job, PeakFit.c:321-351) takes peak_guess [W, K] and returns
    [W, K, 8], one row per peak."""
    if model not in MODELS:
        raise ValueError(f"unknown model {model!r}; expected one of {sorted(MODELS)}")
    ctx = ctx or default_context()
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    g = np.ascontiguousarray(peak_guess, dtype=np.float32).reshape(-1)
    per = _guesses(model)
    if x.shape != y.shape or x.ndim != 2 or g.size != per * x.shape[0]:
        raise ValueError(f"x, y must be [W, N] and peak_guess [W{', %d' % per if per > 1 else ''}]")
    rows = per if model.startswith("tch") else 1
    out = np.zeros((x.shape[0], rows, _lib.MC_FIT_OUT), dtype=np.float64)
    check(lib().mc_fit_windows(ctx.handle, MODELS[model], x.shape[0], x.shape[1],
                               x.ctypes.data, y.ctypes.data, g.ctypes.data, rstep, max_iter,
                               out.ctypes.data, 0))
    return out if rows > 1 else out[:, 0]


def fit_cake_rings(cakes, geom: CSRGeometry, r_centers, rings, *, pad: int = 20,
                   model: str = "pv", max_iter: int = 100, doublet_separation: float = 0.0):
    """Per-eta fits of every ring in every cake, in one launch
    (IntegratorZarrOMP.c:1903-2070).  ``cakes`` is a CUDA tensor
    [F, n_r, n_eta]; returns [F, n_eta, n_rings, 8] float64 (host).
    With doublet_separation > 0, adjacent rings closer than it are fitted
    jointly as doublets on one ROI spanning both (:1912-1923, 1960-2016):
    pV doublets (pf_fit_doublet_pv; the reference's per-eta pass always
    uses this mode, :1997), or for model "tch" TCH doublets, the TCH mode
    of pf_fit_doublet_peak (PeakFit.c:1183-1195) that the reference's
    calibrant integrator selects (pf_fit_doublet_tch, :1138-1170: SNR =
    area / (1.064 FWHM) / sqrt(chi2/dof), falling back to the pV doublet
    where the TCH fit fails);
    their rows then hold the centre [1], FWHM [4], SNR [6] and status [7]
    of the doublet fit, the other fields 0."""
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
    pairs = doublet_pairs(rg, doublet_separation)
    if pairs:
        _fit_ring_doublets(c, rc, rg, pairs, pad, max_iter, out, geom, model)
    return out


def _fit_ring_doublets(c, rc, rg, pairs, pad, max_iter, out, geom, model="pv"):
    """The doublet pass of the per-eta fits (IntegratorZarrOMP.c:1960-2016):
    ROI [min - pad bins, max + pad bins] over both rings, skipped when under
    6 bins or when the eta profile was skipped (status -3)."""
    F, n_eta = out.shape[0], out.shape[1]
    step = float(rc[1] - rc[0]) if rc.size > 1 else 1.0
    host = None
    for i, j in pairs:
        lo_r, hi_r = (i, j) if rg[i] < rg[j] else (j, i)
        plo, phi = rg[lo_r], rg[hi_r]
        roi_lo, roi_hi = plo - pad * step, phi + pad * step
        ge = np.nonzero(rc >= roi_lo)[0]
        le = np.nonzero(rc <= roi_hi)[0]
        out[:, :, i, :7] = 0.0
        out[:, :, j, :7] = 0.0
        out[:, :, i, 7] = out[:, :, j, 7] = -1.0
        if ge.size == 0 or le.size == 0 or le[-1] - ge[0] < 5:
            continue
        lo, hi = int(ge[0]), int(le[-1])
        if host is None:
            host = c.float().cpu().numpy()
        ok = out[:, :, lo_r, 7] != -3  # eta profiles the singlet pass kept
        ok &= out[:, :, hi_r, 7] != -3
        fi, ei = np.nonzero(ok)
        if fi.size == 0:
            continue
        Y = host[fi, lo:hi + 1, ei]                       # [W, roiN]
        X = np.broadcast_to(rc[lo:hi + 1], Y.shape)
        G = np.tile([plo, phi], (fi.size, 1))
        res = fit_windows(X, Y, G, step, model="pv2", max_iter=max_iter, ctx=geom.ctx)
        good = res[:, 7] > 0
        for ring, col_r, col_snr, col_fw in ((lo_r, 0, 2, 4), (hi_r, 1, 3, 5)):
            out[fi, ei, ring, 1] = np.where(good, res[:, col_r], 0.0)
            out[fi, ei, ring, 4] = np.where(good, res[:, col_fw], 0.0)
            out[fi, ei, ring, 6] = np.where(good, res[:, col_snr], 0.0)
            out[fi, ei, ring, 7] = res[:, 7]
        if model == "tch":
            # pf_fit_doublet_tch: both TCH peaks fitted on the window; where
            # both centres come out positive its rows replace the pV rows
            t = fit_windows(X, Y, G, step, model="tch2", max_iter=max_iter, ctx=geom.ctx)
            ok = (t[:, 0, 7] > 0) & (t[:, 1, 7] > 0) & (t[:, 0, 1] > 0) & (t[:, 1, 1] > 0)
            for ring, k in ((lo_r, 0), (hi_r, 1)):
                area, fw, chi = t[:, k, 0], t[:, k, 4], t[:, k, 6]
                h = np.where(fw > 0, area / np.where(fw > 0, fw, 1.0) / 1.064, area)
                snr = np.where(chi > 0, h / np.sqrt(np.where(chi > 0, chi, 1.0)), 1.0)
                out[fi, ei, ring, 1] = np.where(ok, t[:, k, 1], out[fi, ei, ring, 1])
                out[fi, ei, ring, 4] = np.where(ok, fw, out[fi, ei, ring, 4])
                out[fi, ei, ring, 6] = np.where(ok, snr, out[fi, ei, ring, 6])
                out[fi, ei, ring, 7] = np.where(ok, t[:, k, 7], out[fi, ei, ring, 7])


def fit_jobs(n_bins: int, peak_bins, fit_roi_padding: int) -> list:
    """The fit jobs of fitPeaks (PeakFit.c:321-351): peaks sorted by bin,
    ROI [bin - pad, bin + pad] clipped to the profile; a peak whose ROI
    starts at or before the current job's end joins it and moves the end to
    its own ROI end.  Returns [(start, end, [positions in peak_bins])].
    Jobs of more than MC_FIT_MAX_PEAKS peaks are split into consecutive
    groups of at most that many, each on [first - pad, last + pad] (the
    reference fits any number jointly)."""
    order = sorted(range(len(peak_bins)), key=lambda i: peak_bins[i])
    jobs = []
    for i in order:
        k = int(peak_bins[i])
        start = max(0, k - fit_roi_padding)
        end = min(n_bins - 1, k + fit_roi_padding)
        if jobs and start <= jobs[-1][1]:
            jobs[-1][1] = end
            jobs[-1][2].append(i)
        else:
            jobs.append([start, end, [i]])
    out = []
    cap = _lib.MC_FIT_MAX_PEAKS
    for start, end, members in jobs:
        if len(members) <= cap:
            out.append((start, end, members))
            continue
        for c in range(0, len(members), cap):
            part = members[c:c + cap]
            lo = max(0, int(peak_bins[part[0]]) - fit_roi_padding)
            hi = min(n_bins - 1, int(peak_bins[part[-1]]) + fit_roi_padding)
            out.append((lo, hi, part))
    return out


def fit_peaks(x, intensity, peak_locations, x_bin_size: float, fit_roi_padding: int = 20,
              *, max_iter: int = 200, ctx: Context | None = None) -> np.ndarray:
    """fitPeaks (PeakFit.c:275-544) on a 1-D profile: each location snaps to
    the nearest sample (skipped when it is 2 bins or more away); ROIs of
    +-fit_roi_padding bins that overlap are merged into one job fitted as a
    sum of TCH peaks over one Chebyshev background (fit_jobs); jobs of
    equal length and peak count go to the GPU in one launch.  Returns
    [n_peaks, 7] in the order of peak_locations, chi2/dof shared by the
    peaks of a job; rows of skipped or failed peaks are 0.  (The reference
    writes the fitted rows in order of their bins.)"""
    x = np.asarray(x, dtype=np.float64)
    yv = np.asarray(intensity, dtype=np.float64)
    locs = np.atleast_1d(np.asarray(peak_locations, dtype=np.float64))
    out = np.zeros((locs.size, PF_PARAMS_PER_PEAK))
    bins, keep = [], []
    for i, p in enumerate(locs):
        k = int(np.argmin(np.abs(x - p)))
        if abs(x[k] - p) >= 2.0 * x_bin_size:
            continue
        bins.append(k)
        keep.append(i)
    if not bins:
        return out
    groups: dict[tuple, list] = {}
    for start, end, members in fit_jobs(x.size, bins, fit_roi_padding):
        groups.setdefault((end - start + 1, len(members)), []).append((start, end, members))
    for (L, npk), jobs in groups.items():
        X = np.stack([x[s:e + 1] for s, e, _ in jobs])
        Y = np.stack([yv[s:e + 1] for s, e, _ in jobs])
        G = np.array([[x[bins[m]] for m in mem] for _, _, mem in jobs])
        model = "tch" if npk == 1 else f"tch{npk}"
        r = fit_windows(X, Y, G if npk > 1 else G[:, 0], x_bin_size, model=model,
                        max_iter=max_iter, ctx=ctx)
        if npk == 1:
            r = r[:, None, :]
        for rows, (_, _, mem) in zip(r, jobs):
            for row, m in zip(rows, mem):
                if row[7] > 0:
                    out[keep[m]] = row[:PF_PARAMS_PER_PEAK]
    return out


# ---- automatic peak detection (PeakFit.c:148-246, 549-868) ----

_SG = {5: ([-3, 12, 17, 12, -3], 35.0), 7: ([-2, 3, 6, 7, 6, 3, -2], 21.0),
       9: ([-21, 14, 39, 54, 59, 54, 39, 14, -21], 231.0)}


def smooth_data(y, window: int = 7) -> np.ndarray:
    """Savitzky-Golay smoothing of pf_smoothData (PeakFit.c:148-213):
    windows 5, 7 or 9 (others copy); the W//2 samples at each end are
    copied."""
    y = np.asarray(y, dtype=np.float64)
    out = y.copy()
    if window not in _SG or y.size < window:
        return out
    c, norm = _SG[window]
    c = [v / norm for v in c]
    H = window // 2
    n = y.size
    acc = np.zeros(n - 2 * H)
    for j in range(window):  # the reference's summation order
        acc = acc + c[j] * y[j:n - 2 * H + j]
    out[H:n - H] = acc
    return out


def estimate_initial_params(y, peak_idx: int):
    """pf_estimate_initial_params (PeakFit.c:218-246): (FWHM in samples,
    background, amplitude) from the window's ends and the half-maximum
    crossing."""
    y = np.asarray(y, dtype=np.float64)
    n = y.size
    bw = max(int(min(5, n // 4)), 1)
    s = 0.0
    for i in range(bw):  # sum order of the reference
        s += y[i]
        s += y[n - 1 - i]
    bg = s / (2.0 * bw)
    amp = y[peak_idx] - bg
    if amp <= 0:
        amp = y[peak_idx]
    half = bg + amp / 2.0
    lo = peak_idx
    while lo > 0 and y[lo] > half:
        lo -= 1
    hi = peak_idx
    while hi < n - 1 and y[hi] > half:
        hi += 1
    fwhm = float(hi - lo)
    return (fwhm if fwhm > 1.0 else 2.0), bg, amp


def snip_background(y, n_iter: int = 50) -> np.ndarray:
    """SNIP background with the LLS transform (pf_snip_background,
    PeakFit.c:549-592): clipping windows n_iter down to 1."""
    y = np.asarray(y, dtype=np.float64)
    n = y.size
    if n == 0 or n_iter <= 0:
        return np.zeros(n)
    v = np.log(np.log(np.sqrt(np.maximum(y, 0.0) + 1.0) + 1.0) + 1.0)
    for p in range(n_iter, 0, -1):
        if n - 2 * p <= 0:
            continue
        avg = (v[:n - 2 * p] + v[2 * p:]) / 2.0
        v[p:n - p] = np.minimum(v[p:n - p], avg)
    w = np.exp(np.exp(v) - 1.0) - 1.0
    return np.maximum(w * w - 1.0, 0.0)


def detect_peaks(x, corrected, max_peaks: int, min_separation: float) -> np.ndarray:
    """pf_detect_peaks (PeakFit.c:620-712): local maxima of the negative
    second difference of the SG-7 smoothed signal, ranked by prominence,
    picked greedily at least max(3, min_separation / bin) bins apart, in x
    order."""
    x = np.asarray(x, dtype=np.float64)
    c = np.asarray(corrected, dtype=np.float64)
    n = c.size
    if n < 10 or max_peaks <= 0:
        return np.zeros(0)
    sm = smooth_data(c, 7)
    nd2 = np.zeros(n)
    nd2[1:n - 1] = -(sm[2:] - 2.0 * sm[1:n - 1] + sm[:n - 2])
    bw = abs(x[1] - x[0]) if n > 1 else 1.0
    msep = int(max(3.0, min_separation / bw))
    i = np.arange(2, n - 2)
    ok = (nd2[i] > 0) & (nd2[i] >= nd2[i - 1]) & (nd2[i] >= nd2[i + 1]) & (c[i] > 0)
    cands = []
    for k in i[ok][: n // 2]:
        lm = nd2[max(0, k - msep):k + 1].min()
        rm = nd2[k:min(n, k + msep + 1)].min()
        prom = nd2[k] - max(lm, rm)
        cands.append((int(k), prom if prom >= 0 else nd2[k]))
    cands.sort(key=lambda t: -t[1])  # stable, as glibc's merge-sort qsort
    sel: list[int] = []
    for k, _ in cands:
        if len(sel) >= max_peaks:
            break
        if all(abs(k - s) >= msep for s in sel):
            sel.append(k)
    return x[np.array(sorted(sel), dtype=np.int64)] if sel else np.zeros(0)


def select_auto_fits(rows, max_peaks: int, x_bin_size: float) -> list:
    """Quality filter of fitPeaksAutoDetect (PeakFit.c:781-860): indices of
    the fits kept, in detection order."""
    ent = [(i, abs(r[0]), r[4]) for i, r in enumerate(rows) if abs(r[0]) > 0 and r[4] > 0]
    if len(ent) > 2:
        f = sorted(e[2] for e in ent)
        mx = max(5.0 * f[len(ent) // 2], x_bin_size * 2.0)
        ent = [e for e in ent if e[2] <= mx]
    if len(ent) > max_peaks:
        for a in range(len(ent) - 1):  # the reference's exchange sort, by area
            for b in range(a + 1, len(ent)):
                if ent[b][1] > ent[a][1]:
                    ent[a], ent[b] = ent[b], ent[a]
        ent = ent[:max_peaks]
    return sorted(e[0] for e in ent)


def fit_peaks_auto(x, intensity, max_peaks: int, x_bin_size: float,
                   fit_roi_padding: int = 20, snip_iterations: int = 50, *,
                   max_iter: int = 200, ctx: Context | None = None):
    """fitPeaksAutoDetect (PeakFit.c:718-868): SNIP background, detection of
    up to 3 x max_peaks candidates 3 bins apart, TCH fits of the
    background-subtracted profile on the GPU (:func:`fit_peaks`), then the
    quality filter.  Returns ([max_peaks, 7] zero-padded, count)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(intensity, dtype=np.float64)
    out = np.zeros((max(max_peaks, 0), PF_PARAMS_PER_PEAK))
    n = y.size
    if n == 0 or max_peaks <= 0:
        return out, 0
    corrected = np.maximum(y - snip_background(y, snip_iterations), 0.0)
    over = min(max_peaks * 3, n // 2)
    locs = detect_peaks(x, corrected, over, x_bin_size * 3.0)
    if locs.size == 0:
        return out, 0
    rows = fit_peaks(x, corrected, locs, x_bin_size, fit_roi_padding, max_iter=max_iter,
                     ctx=ctx)
    keep = select_auto_fits(rows, max_peaks, x_bin_size)
    for k, i in enumerate(keep):
        out[k] = rows[i]
    return out, len(keep)


def doublet_pairs(peak_locations, separation: float) -> list:
    """Pairs of adjacent peak locations closer than `separation`, each peak
    in at most one pair, scanned left to right (IntegratorZarrOMP.c
    :1912-1923); fit each pair with ``fit_windows(model="pv2")``."""
    locs = [float(v) for v in peak_locations]
    taken = [False] * len(locs)
    pairs = []
    if separation <= 0.0:
        return pairs
    for i in range(len(locs) - 1):
        if taken[i]:
            continue
        if abs(locs[i + 1] - locs[i]) < separation:
            taken[i] = taken[i + 1] = True
            pairs.append((i, i + 1))
    return pairs


def pseudo_voigt(x, center, mu, fwhm, imax, bg):
    """Height-normalised pV of pf_pv_singlet_obj (PeakFit.c:882-920)."""
    x = np.asarray(x, dtype=np.float64)
    d2 = (x - center) ** 2
    L = 1.0 / (1.0 + 4.0 * d2 / fwhm ** 2)
    G = np.exp(-4.0 * np.log(2.0) * d2 / fwhm ** 2)
    return bg + imax * (mu * L + (1.0 - mu) * G)
