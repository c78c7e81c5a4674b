"""Pins the oracle's lineout with variance (oracle/profile.py, a
restatement of kernels.py:604-644) by properties it must have:

* a constant cake I = c with constant variance v: the lineout is c in both
  modes; the variance is v / N when every counted bin has the same area
  (the area-weighted mean is then the plain mean, and the variance of a
  mean of N independent bins is v / N: a denominator not squared gives v,
  a weight not squared gives v / N only by accident of equal areas, so
  unequal areas are checked against sum A^2 / (sum A)^2 computed from the
  areas' moments);
* bins at or below the area threshold and NaN bins do not count, and a row
  without counted bins gives 0;
* the lineout is linear in the cake and the variance linear in the
  variance cake."""
import numpy as np

from oracle.profile import profile_with_variance


def test_constant_cake_equal_areas():
    n_r, n_eta = 3, 12
    cake = np.full((n_r, n_eta), 7.5)
    var = np.full((n_r, n_eta), 2.0)
    area = np.full((n_r, n_eta), 0.8)
    for mode in ("area_weighted", "simple_mean"):
        p, v = profile_with_variance(cake, var, area, mode)
        np.testing.assert_allclose(p, 7.5, rtol=1e-14)
        np.testing.assert_allclose(v, 2.0 / n_eta, rtol=1e-14)


def test_unequal_areas_variance_of_weighted_mean():
    rng = np.random.default_rng(0)
    area = rng.uniform(0.2, 1.5, (2, 30))
    var = np.full_like(area, 3.0)
    cake = np.full_like(area, -4.0)
    p, v = profile_with_variance(cake, var, area, "area_weighted")
    np.testing.assert_allclose(p, -4.0, rtol=1e-14)
    # Var of sum a_i X_i / sum a_i for iid X_i of variance 3
    m1 = area.mean(axis=1)
    m2 = (area ** 2).mean(axis=1)
    np.testing.assert_allclose(v, 3.0 * m2 / (m1 * m1) / area.shape[1], rtol=1e-12)
    # simple_mean ignores the areas
    _, vs = profile_with_variance(cake, var, area, "simple_mean")
    np.testing.assert_allclose(vs, 3.0 / area.shape[1], rtol=1e-14)


def test_threshold_nan_and_empty_rows():
    cake = np.array([[1.0, 100.0, np.nan, 3.0], [5.0, 6.0, 7.0, 8.0]])
    var = np.array([[1.0, 50.0, 9.0, 1.0], [1.0, 1.0, 1.0, 1.0]])
    area = np.array([[1.0, 1e-9, 1.0, 1.0], [0.0, 1e-10, -1.0, 0.0]])
    for mode in ("area_weighted", "simple_mean"):
        p, v = profile_with_variance(cake, var, area, mode)
        np.testing.assert_allclose(p[0], 2.0, rtol=1e-14)    # bins 0 and 3
        np.testing.assert_allclose(v[0], 0.5, rtol=1e-14)    # (1 + 1) / 2^2
        assert p[1] == 0.0 and v[1] == 0.0


def test_linearity():
    rng = np.random.default_rng(1)
    area = rng.uniform(0.0, 1.0, (4, 20))
    c1, c2 = rng.normal(size=(2, 4, 20))
    v1, v2 = rng.uniform(0.1, 2.0, (2, 4, 20))
    for mode in ("area_weighted", "simple_mean"):
        p1, w1 = profile_with_variance(c1, v1, area, mode)
        p2, w2 = profile_with_variance(c2, v2, area, mode)
        p, w = profile_with_variance(2.0 * c1 - 3.0 * c2, 0.5 * v1 + v2, area, mode)
        np.testing.assert_allclose(p, 2.0 * p1 - 3.0 * p2, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(w, 0.5 * w1 + w2, rtol=1e-12)
