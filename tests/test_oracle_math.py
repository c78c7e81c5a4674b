"""What the mathematics fixes about the oracle: closed forms, invariants,
limit cases.  Each check is chosen so a plausible slip (a sign, a dropped
term, a swapped axis) fails it."""
import math

import numpy as np
import pytest

from midas_amd.params import IntegrationParams
from helpers import small_cases, to_oracle


def test_flat_geometry_closed_form(oracle_lib):
    P = to_oracle(small_cases()["flat"])
    rng = np.random.default_rng(0)
    Y, Z = rng.uniform(0, 96, 500), rng.uniform(0, 80, 500)
    R, E = oracle_lib.pixel_to_reta(P, Y, Z)
    np.testing.assert_allclose(R, np.hypot(Y - P.yCen, Z - P.zCen), rtol=1e-13)
    np.testing.assert_allclose(E, np.degrees(np.arctan2(Y - P.yCen, Z - P.zCen)), atol=1e-11)


def _rot(tx, ty, tz):
    a, b, c = np.radians([tx, ty, tz])
    Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    Rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return Rx @ Ry @ Rz


def test_tilt_distortion_parallax_closed_form(oracle_lib):
    Q = small_cases()["tilt"]
    P = to_oracle(Q)
    rng = np.random.default_rng(1)
    Y, Z = rng.uniform(0, 80, 300), rng.uniform(0, 64, 300)
    R, E = oracle_lib.pixel_to_reta(P, Y, Z)
    T = _rot(Q.tx, Q.ty, Q.tz)
    v = T @ np.vstack([np.zeros_like(Y), (Q.BC_y - Y) * Q.pxY, (Z - Q.BC_z) * Q.pxY])
    rad = Q.Lsd / (Q.Lsd + v[0]) * np.hypot(v[1], v[2])
    eta = np.degrees(np.arctan2(-v[1], v[2]))
    rn, et = rad / Q.RhoD, np.radians(90 - eta)
    p = Q.p
    D = (1 + p[0] * rn**2 * np.cos(2 * et + np.radians(p[6]))
         + p[1] * rn**4 * np.cos(4 * et + np.radians(p[3])) + p[2] * rn**2
         + p[4] * rn**6 + p[5] * rn**4 + p[7] * rn**4 * np.cos(et + np.radians(p[8]))
         + p[9] * rn**3 * np.cos(3 * et + np.radians(p[10]))
         + p[11] * rn**5 * np.cos(5 * et + np.radians(p[12]))
         + p[13] * rn**6 * np.cos(6 * et + np.radians(p[14])))
    Rt = rad * D / Q.pxY + Q.Parallax * np.sin(np.arctan(rad / Q.Lsd)) / Q.pxY
    np.testing.assert_allclose(R, Rt, rtol=1e-11)
    np.testing.assert_allclose(E, eta, atol=1e-10)


def test_bin_edges_closed_form(oracle_lib):
    Q = small_cases()["qmode"]
    RLo, RHi, ELo, EHi, _ = oracle_lib.bin_edges(to_oracle(Q))
    n = math.ceil((Q.QMax - Q.QMin) / Q.QBinSize)
    q = Q.QMin + Q.QBinSize * np.arange(n + 1)
    Rq = Q.Lsd / Q.pxY * np.tan(2 * np.arcsin(q * Q.Wavelength / (4 * np.pi)))
    np.testing.assert_allclose(RLo, Rq[:-1], rtol=1e-13)
    np.testing.assert_allclose(RHi, Rq[1:], rtol=1e-13)
    assert len(ELo) == math.ceil((Q.EtaMax - Q.EtaMin) / Q.EtaBinSize)
    np.testing.assert_allclose(EHi - ELo, Q.EtaBinSize)
    from midas_amd.params import bin_edges
    for a, b in zip(bin_edges(Q), (RLo, RHi, ELo, EHi)):
        assert a.tobytes() == b.tobytes()


def _plain_map(oracle_lib, **kw):
    base = dict(NrPixelsY=64, NrPixelsZ=60, pxY=200.0, pxZ=200.0, Lsd=1e6, BC_y=31.7,
                BC_z=29.2, RMin=1.0, RMax=80.0, RBinSize=0.25, EtaBinSize=5.0)
    base.update(kw)
    Q = IntegrationParams(**base)
    return Q, oracle_lib.build_map(to_oracle(Q))


def test_area_conservation_per_pixel(oracle_lib):
    """Untilted pixels tile their own area: the overlaps of one pixel with all
    bins sum to 1, less only the dropped slivers (< 1e-5 each)."""
    Q, m = _plain_map(oracle_lib)
    e = m.entries
    key = e["z"].astype(np.int64) * Q.NrPixelsY + e["y"].astype(np.int64)
    tot = np.bincount(key, weights=e["areaWeight"].astype(np.float64),
                      minlength=Q.NrPixelsY * Q.NrPixelsZ).reshape(Q.NrPixelsZ, Q.NrPixelsY)
    z, y = np.mgrid[0:Q.NrPixelsZ, 0:Q.NrPixelsY]
    r = np.hypot(y - Q.BC_y, z - Q.BC_z)
    # Pixels that straddle the eta = +-180 deg seam (the column through the
    # beam centre, on the -Z side) lose part of their area in the reference
    # mapper (its eta wrap test, MapperCore.c:186-209 / 255-262); the oracle
    # reproduces that, so they are checked separately below.
    seam = (np.abs(y - Q.BC_y) < 0.5 + 1e-9) & (z < Q.BC_z)
    inner = (r > 3) & (r < 28)
    assert (inner & ~seam).sum() > 1000
    assert np.all(tot[inner] <= 1 + 1e-6)
    assert np.all(tot[inner & ~seam] >= 1 - 2e-4)
    lost = 1 - tot[inner & seam]
    assert lost.max() > 1e-3 and lost.max() < 0.1


def test_constant_image_gives_constant_cake(oracle_lib):
    Q, m = _plain_map(oracle_lib)
    c, _ = oracle_lib.integrate(m, np.full((Q.NrPixelsZ, Q.NrPixelsY), 7.25), Q.NrPixelsY,
                                Q.NrPixelsZ)
    nz = m.counts.reshape(c.shape) > 0
    np.testing.assert_allclose(c[nz], 7.25, rtol=1e-6)  # areaWeight is float32
    assert np.all(c[~nz] == 0)


def test_solid_angle_and_polarisation_factors(oracle_lib):
    kw = dict(SolidAngleCorrection=1, PolarizationCorrection=1, PolarizationFraction=0.9,
              Lsd=2e4)
    Q, on = _plain_map(oracle_lib, **kw)
    _, off = _plain_map(oracle_lib, Lsd=2e4)
    np.testing.assert_array_equal(on.pairs, off.pairs)
    e = on.entries
    R, Eta = oracle_lib.pixel_to_reta(to_oracle(Q), e["y"].astype(np.float64),
                                      e["z"].astype(np.float64))
    tth = np.arctan(R * Q.pxY / Q.Lsd)
    factor = np.cos(tth) ** 3 * (1 - 0.9 * np.sin(tth) ** 2 * np.cos(np.radians(Eta)) ** 2)
    np.testing.assert_allclose(off.entries["frac"] / on.entries["frac"], factor, rtol=1e-10)


def test_masked_pixel_flags_its_bins_and_emits_nothing(oracle_lib):
    Q = IntegrationParams(NrPixelsY=32, NrPixelsZ=32, pxY=200.0, pxZ=200.0, Lsd=1e6,
                          BC_y=15.5, BC_z=15.5, RMin=1.0, RMax=22.0, RBinSize=0.5,
                          EtaBinSize=10.0)
    mask = np.zeros((32, 32))
    mask[20, 25] = 1.0
    m0 = oracle_lib.build_map(to_oracle(Q))
    m1 = oracle_lib.build_map(to_oracle(Q), mask=mask)
    e0 = m0.entries
    hit = (e0["y"] == 25) & (e0["z"] == 20)
    assert m1.entries.size == m0.entries.size - hit.sum()
    bins0 = np.repeat(np.arange(m0.counts.size), m0.counts)[hit]
    assert set(np.nonzero(m1.flags)[0]) >= set(bins0)


def test_integration_brute_force(oracle_lib):
    """A pure-Python loop over the records (bilinear read, IntegratorZarrOMP
    semantics) on a tiny map."""
    Q = small_cases()["subpix"]
    m = oracle_lib.build_map(to_oracle(Q))
    rng = np.random.default_rng(3)
    img = rng.normal(50, 10, (Q.NrPixelsZ, Q.NrPixelsY))
    c, _ = oracle_lib.integrate(m, img, Q.NrPixelsY, Q.NrPixelsZ)
    NY, NZ = Q.NrPixelsY, Q.NrPixelsZ
    for b in rng.choice(np.nonzero(m.counts)[0], 60, replace=False):
        s = a = 0.0
        for r in m.entries[m.offsets[b]:m.offsets[b] + m.counts[b]]:
            y, z = float(r["y"]), float(r["z"])
            iy, iz = math.floor(y), math.floor(z)
            fy, fz = y - iy, z - iz
            if iy < 0: iy, fy = 0, 0.0
            if iy >= NY - 1: iy, fy = NY - 2, 1.0
            if iz < 0: iz, fz = 0, 0.0
            if iz >= NZ - 1: iz, fz = NZ - 2, 1.0
            v = ((1 - fy) * (1 - fz) * img[iz, iy] + fy * (1 - fz) * img[iz, iy + 1]
                 + (1 - fy) * fz * img[iz + 1, iy] + fy * fz * img[iz + 1, iy + 1])
            s += v * r["frac"]
            a += float(r["areaWeight"])
        assert c.reshape(-1)[b] == pytest.approx(s / a, rel=1e-12)


def test_lineout_definitions(oracle_lib):
    cake = np.array([[0.0, 2.0, np.nan, 4.0], [1.0, 1.0, 1.0, 1.0]])
    area = np.array([[1.0, 1.0, 1.0, 3.0], [0.0, 2.0, 2.0, 0.0]])
    lc, la, lm = oracle_lib.lineouts(cake, area)
    assert lc.tolist() == [3.0, 1.0]                     # nonzero, non-NaN mean
    assert la[0] == pytest.approx((0 * 1 + 2 * 1 + 4 * 3) / 5)  # NaN bin is missing
    assert lm[0] == pytest.approx((0 + 2 + 4) / 3)
    assert la[1] == 1.0 and lm[1] == 1.0


# ------------------------------------------------------------------ peaks
def test_tch_limits(oracle_lib):
    f, e = oracle_lib.tch_eta_fwhm(400.0, 1e-9)  # no Lorentzian: eta -> 0
    assert e == pytest.approx(0.0, abs=1e-5)
    assert f == pytest.approx(math.sqrt(8 * math.log(2) * 400.0) / 100, rel=1e-6)
    f, e = oracle_lib.tch_eta_fwhm(1e-14, 30.0)  # no Gaussian: eta -> 1
    assert e == pytest.approx(1.0, abs=1e-3)
    assert f == pytest.approx(0.30, rel=1e-6)


@pytest.mark.parametrize("sig,gam", [(300.0, 20.0), (50.0, 80.0), (900.0, 1.0)])
def test_tch_is_area_normalised(oracle_lib, sig, gam):
    x = np.linspace(-200, 200, 400001)
    y = oracle_lib.tch_model([3.5, 0.0, sig, gam, 0.0, 0.0], x)
    area = np.trapezoid(y, x)
    fwhm, _ = oracle_lib.tch_eta_fwhm(sig, gam)
    # Lorentzian tails beyond +-200: (2/pi) * hwhm / 200 of the Lorentz part
    assert area == pytest.approx(3.5, rel=2e-3)


def test_pv_model_values(oracle_lib):
    p = [10.0, 0.3, 2.0, 100.0, 5.0]  # centre, mu, FWHM, Imax, BG
    x = np.array([10.0, 11.0, 9.0, 30.0])
    y = oracle_lib.pv_model(p, x)
    assert y[0] == pytest.approx(105.0)
    assert y[1] == pytest.approx(y[2])  # symmetric
    assert y[1] == pytest.approx(5 + 100 * (0.3 * 0.5 + 0.7 * 0.5))  # half max at FWHM/2
    assert y[3] == pytest.approx(5.0, abs=0.1)


@pytest.mark.parametrize("model", [0, 1])
def test_lm_recovers_noise_free_peaks(oracle_lib, model):
    x = np.arange(40) * 0.25 + 100.0
    if model == 0:
        truth = np.array([104.83, 0.37, 1.3, 800.0, 40.0])
        y = oracle_lib.pv_model(truth, x)
    else:
        truth = np.array([1500.0, 104.83, 250.0, 40.0, 30.0, 4.0])
        y = oracle_lib.tch_model(truth, x)
    p0, lo, hi = oracle_lib.window_init(model, x, y, 0.25, 104.9)
    p, chi, it = oracle_lib.lm_fit(model, x, y, p0, lo, hi, max_iter=300)
    assert chi < 1e-12 * np.sum(y ** 2)
    assert abs(p[0 if model == 0 else 1] - truth[0 if model == 0 else 1]) < 1e-6


def _tch_code(n):
    return 1 if n == 1 else n + 1  # model codes of the N-peak TCH fits


def test_multi_peak_tch_is_a_sum_of_single_peaks(oracle_lib):
    """The merged-job model (PeakFit.c:160-215) is the single-peak model
    summed over peaks with one Chebyshev background."""
    x = np.arange(120) * 0.25 + 50.0
    a = [900.0, 60.0, 120.0, 30.0]
    b = [1400.0, 68.5, 200.0, 10.0]
    bg = [25.0, -3.0]
    both = oracle_lib.tch_model(a + b + bg, x)
    one = oracle_lib.tch_model(a + bg, x) + oracle_lib.tch_model(b + [0.0, 0.0], x)
    np.testing.assert_allclose(both, one, rtol=1e-13, atol=1e-11)


@pytest.mark.parametrize("n_peaks", [2, 3, 5, 8])
def test_lm_recovers_merged_multi_peak_jobs(oracle_lib, n_peaks):
    """fitPeaks' merged job: N TCH peaks 3-5 px apart plus a sloped
    background, noise-free, started from the fitPeaks init at the snapped
    bins; every centre comes back within 1e-6."""
    rng = np.random.default_rng(n_peaks)
    step = 0.25
    centres = 60.0 + np.cumsum(rng.uniform(3.0, 5.0, n_peaks))
    x = np.arange(int((centres[-1] - centres[0] + 10) / step) + 1) * step + centres[0] - 5.0
    truth = []
    for c in centres:
        truth += [rng.uniform(800, 2000), c, rng.uniform(50, 300), rng.uniform(5, 60)]
    truth += [30.0, 4.0]
    y = oracle_lib.tch_model(truth, x)
    locs = x[np.argmin(np.abs(x[None, :] - centres[:, None]), axis=1)]
    p0, lo, hi = oracle_lib.tchn_init(x, y, step, locs)
    assert np.all(lo <= p0) and np.all(p0 <= hi)
    p, chi, it = oracle_lib.lm_fit(_tch_code(n_peaks), x, y, p0, lo, hi, max_iter=500)
    assert it > 0
    np.testing.assert_allclose(p[1:4 * n_peaks:4], centres, atol=1e-6)
    np.testing.assert_allclose(p[0:4 * n_peaks:4], truth[0:4 * n_peaks:4], rtol=1e-5)
