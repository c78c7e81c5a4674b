"""Pins the oracle's REtaMap rows (orc_reta_map, IntegratorZarrOMP.c:1754-1763)
by physics the rows must satisfy, computed here by other routes:

* Bragg's law: a bin centred on the ring of a cubic lattice plane (hkl),
  d = a / sqrt(h^2 + k^2 + l^2), lies at R = Lsd * tan(2 theta) / px with
  sin(theta) = lambda / (2 d); its 2theta row must be that 2 theta (degrees)
  and its Q row 2 pi / d.  The test derives 2 theta with asin, the row with
  atan: a dropped factor 2 or a sin(2 theta) for sin(theta) fails.
* The Q-mode bin edges of the mapper, R(Q) = (Lsd/px) tan(2 asin(Q lambda /
  4 pi)) (DetectorMapper.c:714-746): a bin with both edges at R(Q) has Q row
  Q.  The oracle's bin_edges compute those edges, so the Q row inverts them.
* R and eta rows are the bin centres, the area row is the area passed in,
  and lambda <= 0 gives Q = 0.
"""
import math

import numpy as np

from oracle import oracle as orc

CEO2_A = 5.411102  # lattice constant of CeO2 (Angstrom), the calibrant of FF_HEDM/Example


def _reta(r_lo, r_hi, e_lo, e_hi, area, px, lsd, lam):
    r_lo, r_hi = np.asarray(r_lo, float), np.asarray(r_hi, float)
    e_lo, e_hi = np.asarray(e_lo, float), np.asarray(e_hi, float)
    nR, nE = len(r_lo), len(e_lo)
    out = np.zeros((5, nR, nE))
    a = np.ascontiguousarray(area, dtype=np.float64)
    orc.lib().orc_reta_map(orc._dp(r_lo), orc._dp(r_hi), orc._dp(e_lo), orc._dp(e_hi), nR, nE,
                           orc._dp(a), px, lsd, lam, orc._dp(out))
    return out


def test_bragg_rings_of_ceo2():
    px, lsd = 200.0, 1.0e6  # micrometres
    for lam in (0.172979, 0.22291, 0.1):
        hkls = [(1, 1, 1), (2, 0, 0), (2, 2, 0), (3, 1, 1), (4, 2, 0), (5, 3, 1)]
        d = np.array([CEO2_A / math.sqrt(h * h + k * k + l * l) for h, k, l in hkls])
        tth = 2.0 * np.arcsin(lam / (2.0 * d))        # radians
        R = lsd * np.tan(tth) / px                   # ring radius in pixels
        out = _reta(R - 0.125, R + 0.125, [-2.5, 10.0], [2.5, 20.0], np.ones((len(R), 2)),
                    px, lsd, lam)
        np.testing.assert_allclose(out[0, :, 0], R, rtol=1e-15)
        np.testing.assert_allclose(out[1, :, 1], np.degrees(tth), rtol=1e-12)
        np.testing.assert_allclose(out[4, :, 0], 2.0 * math.pi / d, rtol=1e-12)
        np.testing.assert_array_equal(out[2, 0], [0.0, 15.0])


def test_q_row_inverts_qmode_edges():
    P = orc.make_params(NrPixelsY=96, NrPixelsZ=96, pxY=172.0, pxZ=172.0, Lsd=4.5e5,
                        yCen=48.0, zCen=48.0, QMin=0.6, QMax=9.0, QBinSize=0.3,
                        Wavelength=0.172979, EtaMin=-180.0, EtaMax=180.0, EtaBinSize=30.0)
    RLo, RHi, _, _, _ = orc.bin_edges(P)
    q_lo = 0.6 + 0.3 * np.arange(len(RLo))
    for edges, q in ((RLo, q_lo), (RHi, q_lo + 0.3)):
        out = _reta(edges, edges, [0.0], [1.0], np.zeros((len(edges), 1)), P.pxY, P.Lsd,
                    P.Wavelength)
        np.testing.assert_allclose(out[4, :, 0], q, rtol=1e-12)


def test_area_row_and_no_wavelength():
    rng = np.random.default_rng(4)
    area = rng.random((3, 4))
    out = _reta([1.0, 2.0, 3.0], [1.5, 2.5, 3.5], [-180, -90, 0, 90], [-90, 0, 90, 180], area,
                100.0, 5e5, 0.0)
    np.testing.assert_array_equal(out[3], area)
    np.testing.assert_array_equal(out[4], 0.0)
    np.testing.assert_array_equal(out[0][:, 0], [1.25, 2.25, 3.25])
    np.testing.assert_array_equal(out[2][0], [-135.0, -45.0, 45.0, 135.0])
    assert np.all(np.diff(out[1][:, 0]) > 0)
