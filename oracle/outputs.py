"""Restatement of the reference's result-file contents (pure Python, FP64).

TEST INFRASTRUCTURE ONLY: the tests compare the files written by the
native output set (mc_out_*) with what these functions say the reference
writes for the same frames.  Each function cites the reference lines it
follows.  The line formats reuse the reference's printf formats (Python's
%-formatting rounds like C's printf).
"""
from __future__ import annotations

import math

import numpy as np


def lineout_cpu_mean(cake):
    """Mean over eta of the values that are nonzero and not NaN, 0 when
    there is none (IntegratorZarrOMP.c:1829-1843)."""
    c = np.asarray(cake, dtype=np.float64)
    ok = (c != 0) & ~np.isnan(c)
    n = ok.sum(axis=1)
    s = np.where(ok, c, 0.0).sum(axis=1)
    return np.where(n > 0, s / np.maximum(n, 1), 0.0)


def lineout_xy(l1, r_centers, px, lsd, q_mode=False, q_min=0.0, q_bin_size=0.0):
    """Text of one frame in <prefix>_lineout.xy (IntegratorZarrOMP.c:1849-1862)."""
    out = []
    for j, v in enumerate(l1):
        if math.isnan(v):
            continue
        if q_mode:
            x = q_min + q_bin_size * (j + 0.5)
        else:
            x = math.atan(r_centers[j] * px / lsd) * 180.0 / math.pi
        out.append("%.6f  %.6f\n" % (x, v))
    return "".join(out)


def omega_sum_chunks(frame_nrs, omegas, cakes, n_sum, metas=None):
    """/OmegaSumFrame datasets (IntegratorZarrOMP.c:1633-1642, 2093-2132):
    a chunk starts at every n_sum-th frame (in processing order), sums the
    cakes and is written at its last frame or at the last frame of the run.
    The meta values (Temperature, Pressure, I, I0) are chunk means.  The
    number summed is the count of frames in the chunk (the reference derives
    it from the omega step, which gives the same count for evenly stepped
    omegas)."""
    out = []
    nf = len(frame_nrs)
    for i in range(nf):
        if i % n_sum == 0:
            acc = np.zeros_like(np.asarray(cakes[i], dtype=np.float64))
            first, cnt = omegas[i], 0
            m = np.zeros(4)
        acc = acc + np.asarray(cakes[i], dtype=np.float64)
        if metas is not None:
            m = m + np.asarray(metas[i], dtype=np.float64)
        cnt += 1
        if (i + 1) % n_sum == 0 or i == nf - 1:
            out.append({"name": f"LastFrameNumber_{frame_nrs[i]}", "data": acc,
                        "LastFrameNumber": frame_nrs[i], "Number Of Frames Summed": cnt,
                        "FirstOme": first, "LastOme": omegas[i],
                        "Temperature": m[0] / cnt, "Pressure": m[1] / cnt,
                        "I": m[2] / cnt, "I0": m[3] / cnt})
    return out


def fit_per_eta_line(frame, eta_bin, eta_cen, peak, p, px, lsd):
    """One row of <prefix>_fit_per_eta.csv, or None when the peak was not
    fitted (IntegratorZarrOMP.c:2036-2053); p = the 7 PeakFit parameters."""
    center_px = p[1]
    if center_px == 0:
        return None
    r_um = center_px * px
    tth = math.atan(r_um / lsd) * 180.0 / math.pi
    return ("%d,%d,%.4f,%d,%.6f,%.4f,%.6f,%.4f,%.6f,%.6f,%.4f,%.4f,%.4f,%.4f\n"
            % (frame, eta_bin, eta_cen, peak, center_px, r_um, tth, p[0], p[2], p[3], p[4],
               p[5], p[6], p[0]))


FIT_PER_ETA_HEADER = ("Frame,EtaBin,EtaCen,PeakIdx,R_px,R_um,TwoTheta_deg,"
                      "Area,Sig_centideg2,Gam_centideg,FWHM_deg,Eta,ChiSq,Area2\n")


def pfio_row(frame, eta_bin, peak, eta_deg, p, px, lsd, lam):
    """PeakH5Row of pfio_make_row (PeakFitIO.c:52-86)."""
    tth = math.atan(p[1] * px / lsd) * 180.0 / math.pi
    theta = tth * math.pi / 360.0
    d = lam / (2.0 * math.sin(theta)) if theta > 1e-10 else 0.0
    return {"frame_idx": frame, "eta_idx": eta_bin, "peak_nr": peak, "eta_deg": eta_deg,
            "center_2theta": tth, "area": p[0], "sig": p[2], "gam": p[3], "FWHM_deg": p[4],
            "eta_mix": p[5], "d_spacing_A": d, "chi_sq": p[6]}


def stream_pairs(r_centers, profile):
    """One frame of lineout.bin / lineout_simple_mean.bin: (R, I) double
    pairs (IntegratorFitPeaksGPUStream.cu:534-551)."""
    out = np.empty((len(r_centers), 2))
    out[:, 0] = r_centers
    out[:, 1] = profile
    return out
