"""The eta-filter restatement (oracle/postprocess.py) pinned by closed
forms and invariants: a constant cake is a fixed point; a unit impulse
returns the normalised Gaussian taps (and h = round-half-even(4 sigma));
periodic smoothing keeps the row sum, and with areas keeps
sum_j out_j * (w * A)_j = sum_j I_j A_j; the reflect padding agrees with
numpy.pad(mode="reflect") used as the textbook routine; the median removes
an isolated spike, is the identity for window 1, propagates NaN, and the
flux-preserving rescale restores each row's sum."""
import numpy as np
import pytest

from oracle import postprocess as pp


def test_constant_cake_is_fixed_point():
    c = np.full((3, 36), 7.25, np.float32)
    for periodic in (True, False):
        np.testing.assert_array_equal(pp.gauss_smooth_eta(c, 1.3, periodic=periodic), c)
        np.testing.assert_array_equal(pp.median_filter_eta(c, 5, periodic=periodic), c)


@pytest.mark.parametrize("sigma,half", [(0.625, 2), (0.875, 4), (1.0, 4), (0.1, 1), (2.2, 9)])
def test_impulse_response_is_the_gaussian(sigma, half):
    n = 40
    c = np.zeros((1, n))
    c[0, 20] = 1.0
    out = pp.gauss_smooth_eta(c, sigma)
    k = np.arange(-half, half + 1)
    w = np.exp(-k.astype(float) ** 2 / (2 * sigma * sigma))
    w /= w.sum()
    np.testing.assert_allclose(out[0, 20 + k], w, rtol=1e-15)
    assert np.count_nonzero(out) == 2 * half + 1


def test_periodic_smoothing_keeps_flux():
    rng = np.random.default_rng(1)
    c = rng.random((4, 72))
    out = pp.gauss_smooth_eta(c, 1.7)
    np.testing.assert_allclose(out.sum(axis=1), c.sum(axis=1), rtol=1e-13)
    A = rng.random((4, 72)) + 0.2
    outa = pp.gauss_smooth_eta(c, 1.7, area_per_bin=A)
    half, k, w = pp.gauss_weights(1.7)
    Abar = sum(wk * np.roll(A, -kk, axis=1) for kk, wk in zip(k.astype(int), w))
    np.testing.assert_allclose((outa * Abar).sum(axis=1), (c * A).sum(axis=1), rtol=1e-12)


def test_reflect_padding_matches_numpy_pad():
    rng = np.random.default_rng(2)
    for n, sigma in ((9, 0.9), (5, 1.6), (3, 1.2)):  # pads shorter and longer than the axis
        c = rng.random((2, n))
        half, k, w = pp.gauss_weights(sigma)
        padded = np.pad(c, ((0, 0), (half, half)), mode="reflect")
        want = sum(wk * padded[:, half + kk: half + kk + n] for kk, wk in zip(k.astype(int), w))
        np.testing.assert_allclose(pp.gauss_smooth_eta(c, sigma, periodic=False), want,
                                   rtol=1e-14)


def test_median_properties():
    c = np.ones((2, 20), np.float32)
    c[0, 7] = 100.0
    out = pp.median_filter_eta(c, 3, preserve_ring_flux=False)
    np.testing.assert_array_equal(out, np.ones_like(c))
    rng = np.random.default_rng(3)
    r = rng.random((3, 30)).astype(np.float32)
    np.testing.assert_array_equal(pp.median_filter_eta(r, 1, preserve_ring_flux=False), r)
    kept = pp.median_filter_eta(r, 5)
    np.testing.assert_allclose(kept.sum(axis=1), r.sum(axis=1), rtol=1e-5)
    # window 1 with the rescale is the identity (equal sums, factor 1)
    np.testing.assert_array_equal(pp.median_filter_eta(r, 1), r)
    ramp = np.arange(10, dtype=np.float32)[None, :]
    m = pp.median_filter_eta(ramp, 3, periodic=False, preserve_ring_flux=False)
    # reflect: [1,0,1] -> 1 at the left end, [8,9,8] -> 8 at the right
    np.testing.assert_array_equal(m[0], [1, 1, 2, 3, 4, 5, 6, 7, 8, 8])
    r[1, 4] = np.nan
    nan_out = pp.median_filter_eta(r, 3, preserve_ring_flux=False)
    assert np.isnan(nan_out[1, 3:6]).all() and not np.isnan(nan_out[1, 7:]).any()
    assert np.isnan(pp.median_filter_eta(r, 3)[1]).all()  # the row's sum is NaN
