"""GPU eta filters (mc_smooth_eta / mc_median_eta) vs the FP64 restatement
of postprocess.py:36-142.

The smoothing runs the reference's FP64 operations in its order; the only
difference is the weights' exp() (C library vs numpy), so values agree to
1e-6 relative (float32 output; most bins are bit-identical).  The median is
a selection and the flux rescale uses numpy's float32 pairwise sums: bit
identical."""
import numpy as np
import pytest

import midas_amd as ma
from oracle import postprocess as pp

pytestmark = pytest.mark.gpu


def _cakes(seed, f=3, n_r=17, n_eta=72, nan=True):
    rng = np.random.default_rng(seed)
    c = (rng.random((f, n_r, n_eta)) * 1000).astype(np.float32)
    if nan:
        c[rng.random(c.shape) < 0.01] = np.nan
    return c


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("sigma", [0.625, 1.5, 4.0])
def test_smooth_matches_oracle(periodic, sigma):
    c = _cakes(1)
    rng = np.random.default_rng(2)
    A = (rng.random((17, 72)) + 0.1).astype(np.float32)
    print(f"[post] smooth sigma={sigma} periodic={periodic}", flush=True)
    for area in (None, A):
        got = ma.gauss_smooth_eta(c, sigma, area_per_bin=area, periodic=periodic)
        for i in range(len(c)):
            want = pp.gauss_smooth_eta(c[i], sigma, area_per_bin=area, periodic=periodic)
            np.testing.assert_allclose(got[i], want, rtol=1e-6, atol=0, equal_nan=True)
            assert np.array_equal(np.isnan(got[i]), np.isnan(want))
        # 2-D input gives the same as the batch's first cake
        np.testing.assert_array_equal(
            ma.gauss_smooth_eta(c[0], sigma, area_per_bin=area, periodic=periodic), got[0])


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("window", [1, 3, 7])
def test_median_matches_oracle(periodic, window):
    c = _cakes(3)
    c[1, 5, :] = 0.0  # a zero row: factor 1
    print(f"[post] median window={window} periodic={periodic}", flush=True)
    for keep in (True, False):
        got = ma.median_filter_eta(c, window, periodic=periodic, preserve_ring_flux=keep)
        for i in range(len(c)):
            want = pp.median_filter_eta(c[i], window, periodic=periodic, preserve_ring_flux=keep)
            np.testing.assert_array_equal(got[i], want)


def test_torch_tensors_and_errors():
    torch = pytest.importorskip("torch")
    c = _cakes(4, nan=False)
    print("[post] torch tensors", flush=True)
    d = ma.gauss_smooth_eta(torch.from_numpy(c).cuda(), 2.0)
    np.testing.assert_array_equal(d.cpu().numpy(), ma.gauss_smooth_eta(c, 2.0))
    m = ma.median_filter_eta(torch.from_numpy(c).cuda(), 5)
    np.testing.assert_array_equal(m.cpu().numpy(), ma.median_filter_eta(c, 5))
    with pytest.raises(ValueError):
        ma.median_filter_eta(c, 4)
    with pytest.raises(RuntimeError):
        ma.gauss_smooth_eta(c, -1.0)
    with pytest.raises(RuntimeError):
        ma.gauss_smooth_eta(c[:, :, :5], 3.0)  # 12 taps per side over 5 periodic bins
