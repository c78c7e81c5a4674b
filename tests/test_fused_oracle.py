"""The fused multi-map restatement (oracle/fused.py) pinned by closed
forms: shifted axes give the weights (k/N, 1 - k/N) on rows (i-1, i); the
interpolation reproduces any function linear in R inside the axis and
clamps at both ends; one map is the area-normalised cake itself; bins with
area <= 1e-9 contribute zero."""
import numpy as np

from oracle import fused as fu


def test_shift_weights_closed_form():
    dR, N = 0.25, 4
    R0 = 10 + dR * (np.arange(20) + 0.5)
    for k in range(N):
        Rk = R0 + dR * k / N
        j, wl, wh = fu.r_interp(Rk, R0)
        if k == 0:
            np.testing.assert_array_equal(wl[:-1], 1.0)
            assert j[-1] == 18 and wh[-1] == 1.0
            continue
        assert j[0] == 0 and wl[0] == 1.0 and wh[0] == 0.0  # below the shifted axis
        np.testing.assert_array_equal(j[1:], np.arange(19))
        np.testing.assert_allclose(wl[1:], k / N, rtol=1e-12)
        np.testing.assert_allclose(wh[1:], 1 - k / N, rtol=1e-12)


def test_linear_in_r_is_reproduced():
    dR, N, n_eta = 0.5, 3, 4
    rng = np.random.default_rng(0)
    a, b = rng.random(n_eta), rng.random(n_eta)
    R0 = 5 + dR * (np.arange(12) + 0.5)
    axes = [R0 + dR * k / N for k in range(N)]
    raws = [a[None, :] + b[None, :] * R[:, None] for R in axes]
    areas = [np.ones((12, n_eta))] * N
    f = fu.fused_cake(raws, areas, axes)
    np.testing.assert_allclose(f[1:], a[None, :] + b[None, :] * R0[1:, None], rtol=1e-12)


def test_single_map_and_area_threshold():
    rng = np.random.default_rng(1)
    raw = rng.random((6, 5))
    A = rng.random((6, 5)) + 0.5
    A[2, 3] = 1e-10
    R = np.arange(6) + 0.5
    f = fu.fused_cake([raw], [A], [R])
    want = raw / A
    want[2, 3] = 0.0
    np.testing.assert_allclose(f, want, rtol=1e-15)
