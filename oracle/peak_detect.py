"""Oracle: automatic peak detection of the reference's peak fitter.

TEST INFRASTRUCTURE ONLY.  Plain-loop restatements of FF_HEDM/src/PeakFit.c,
used by tests/ to check midas_amd.peakfit; nothing in the product imports
this module.  PeakFit.c itself is not buildable here (it includes nlopt.h;
NLopt is absent), so these functions are pinned by the properties tests/
check: Savitzky-Golay filters reproduce cubics, SNIP returns a constant
background unchanged and clips narrow peaks to it, detected peaks land on
the generating bins.
"""
from __future__ import annotations

import math

SG = {5: ([-3, 12, 17, 12, -3], 35.0), 7: ([-2, 3, 6, 7, 6, 3, -2], 21.0),
      9: ([-21, 14, 39, 54, 59, 54, 39, 14, -21], 231.0)}


def smooth_data(inp, W):
    """pf_smoothData, PeakFit.c:148-213: Savitzky-Golay of window 5, 7 or 9;
    the first and last W//2 samples are copied."""
    n = len(inp)
    if W < 3 or W % 2 == 0 or W not in SG:
        return list(inp)
    c, norm = SG[W]
    c = [v / norm for v in c]
    H = W // 2
    out = [0.0] * n
    for i in range(n):
        if i < H or i >= n - H:
            out[i] = inp[i]
        else:
            val = 0.0
            for j in range(W):
                val += c[j] * inp[i - H + j]
            out[i] = val
    return out


def estimate_initial_params(data, peak_idx):
    """pf_estimate_initial_params, PeakFit.c:218-246: (fwhm in samples, bg,
    amp)."""
    n = len(data)
    bw = int(min(5, n // 4))
    bw = max(bw, 1)
    s = 0.0
    for i in range(bw):
        s += data[i]
        s += data[n - 1 - i]
    bg = s / (2.0 * bw)
    amp = data[peak_idx] - bg
    if amp <= 0:
        amp = data[peak_idx]
    half = bg + amp / 2.0
    lo = peak_idx
    while lo > 0 and data[lo] > half:
        lo -= 1
    hi = peak_idx
    while hi < n - 1 and data[hi] > half:
        hi += 1
    fwhm = float(hi - lo)
    return (fwhm if fwhm > 1.0 else 2.0), bg, amp


def snip_background(y, n_iter):
    """pf_snip_background, PeakFit.c:549-592 (LLS transform, windows n_iter
    down to 1)."""
    n = len(y)
    if n <= 0 or n_iter <= 0:
        return [0.0] * n
    v = [math.log(math.log(math.sqrt(max(t, 0.0) + 1.0) + 1.0) + 1.0) for t in y]
    for p in range(n_iter, 0, -1):
        tmp = list(v)
        for i in range(p, n - p):
            avg = (tmp[i - p] + tmp[i + p]) / 2.0
            if v[i] > avg:
                v[i] = avg
    out = []
    for t in v:
        w = math.exp(math.exp(t) - 1.0) - 1.0
        out.append(max(w * w - 1.0, 0.0))
    return out


def detect_peaks(x, corrected, max_peaks, min_separation):
    """pf_detect_peaks, PeakFit.c:620-712: peaks of the negative second
    difference of the SG-7 smoothed signal, greedy by prominence with a
    minimum separation, returned in x order."""
    n = len(corrected)
    if n < 10 or max_peaks <= 0:
        return []
    sm = smooth_data(corrected, 7)
    nd2 = [0.0] * n
    for i in range(1, n - 1):
        nd2[i] = -(sm[i + 1] - 2.0 * sm[i] + sm[i - 1])
    bw = abs(x[1] - x[0]) if n > 1 else 1.0
    msep = int(max(3, min_separation / bw))
    max_c = n // 2
    cands = []
    for i in range(2, n - 2):
        if nd2[i] > 0 and nd2[i] >= nd2[i - 1] and nd2[i] >= nd2[i + 1] and corrected[i] > 0:
            lm = rm = nd2[i]
            j = i - 1
            while j >= 0 and j >= i - msep:
                lm = min(lm, nd2[j])
                j -= 1
            j = i + 1
            while j < n and j <= i + msep:
                rm = min(rm, nd2[j])
                j += 1
            prom = nd2[i] - max(lm, rm)
            if prom < 0:
                prom = nd2[i]
            if len(cands) < max_c:
                cands.append((i, prom))
    # stable order for equal prominences (qsort leaves it unspecified)
    cands.sort(key=lambda c: -c[1])
    sel = []
    for i, _ in cands:
        if len(sel) >= max_peaks:
            break
        if all(abs(i - s) >= msep for s in sel):
            sel.append(i)
    return [x[i] for i in sorted(sel)]


def select_fits(rows, max_peaks, x_bin_size):
    """Steps 4-5 of fitPeaksAutoDetect, PeakFit.c:781-860: drop fits with
    no area or FWHM, then FWHMs above max(5 x median, 2 bins), keep the
    max_peaks largest areas, in detection order.  rows: [n][7]."""
    ent = [(i, abs(r[0]), r[4]) for i, r in enumerate(rows) if abs(r[0]) > 0 and r[4] > 0]
    if len(ent) > 2:
        f = sorted(e[2] for e in ent)
        mx = max(5.0 * f[len(ent) // 2], x_bin_size * 2.0)
        ent = [e for e in ent if e[2] <= mx]
    if len(ent) > max_peaks:
        for i in range(len(ent) - 1):
            for j in range(i + 1, len(ent)):
                if ent[j][1] > ent[i][1]:
                    ent[i], ent[j] = ent[j], ent[i]
        ent = ent[:max_peaks]
    ent.sort(key=lambda e: e[0])
    return [e[0] for e in ent]
