"""Variance propagation (mc_integrate_variance) vs the FP64 restatement of
integrate_with_variance (kernels.py:476-561).

For maps whose entries sit on whole pixels (SubPixelLevel 1) every map
entry is one (bin, pixel) weight, so the oracle is direct:
Var_b = s_b^2 * sum_k frac_k^2 * V[pixel_k], with s_b the value's scale
(1 / sum areaWeight when normalising).  With merged weights (bilinear
corners) the result must be linear in the variance image.  Tolerance
1e-5 relative (positive terms, FP32 Kahan sums vs FP64)."""
import numpy as np
import pytest

import midas_amd as ma
from helpers import small_cases, to_oracle

pytestmark = pytest.mark.gpu


def _oracle_var(om, V, NY, normalize=True):
    e = om.entries
    cnt, off = om.counts, om.offsets
    nb = cnt.size
    out = np.zeros(nb)
    for b in range(nb):
        s = slice(off[b], off[b] + cnt[b])
        y = e["y"][s].astype(np.int64)
        z = e["z"][s].astype(np.int64)
        fr = e["frac"][s]
        raw = np.sum(fr * fr * V[z, y])
        if normalize:
            a = np.sum(e["areaWeight"][s].astype(np.float64))
            raw = raw / (a * a) if a > 0 else 0.0
        out[b] = raw
    return out.reshape(om.n_r, om.n_eta)


def test_poisson_and_given_variance_match_oracle(oracle_lib):
    P = small_cases()["flat"]
    assert P.SubPixelLevel == 1
    rng = np.random.default_rng(3)
    frames = rng.poisson(200.0, (2, P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    dark = rng.poisson(20.0, (2, P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    print("[variance] flat, floor, Poisson with dark", flush=True)
    g = ma.build_map(P)
    g.set_dark(dark)
    om = oracle_lib.build_map(to_oracle(P))
    dmean = dark.astype(np.float64).mean(axis=0)
    cakes, var = ma.integrate_with_variance(frames, g, mode="floor")
    for i in range(len(frames)):
        # the one-frame kernel's values, bit for bit (same summation order)
        one, _ = ma.integrate_batch(frames[i], g, mode="floor", mask_nan=False, lineouts=False)
        np.testing.assert_array_equal(cakes[i], one[0])
    for i in range(len(frames)):
        want = _oracle_var(om, frames[i].astype(np.float64) - dmean, P.NrPixelsY)
        np.testing.assert_allclose(var[i], want, rtol=1e-5, atol=1e-6)
    # an explicit variance image, unnormalised
    V = rng.random((2, P.NrPixelsZ, P.NrPixelsY)).astype(np.float32) * 50
    _, var2 = ma.integrate_with_variance(frames, g, mode="floor", variance_image=V,
                                         normalize=False)
    for i in range(2):
        want = _oracle_var(om, V[i].astype(np.float64), P.NrPixelsY, normalize=False)
        np.testing.assert_allclose(var2[i], want, rtol=1e-5, atol=1e-6)


def test_linear_in_variance_image_with_merged_weights():
    """Bilinear corners on a sub-pixel map: the weights are merged per
    (bin, pixel); the propagated variance is linear in the variance image,
    non-negative for a non-negative image, and NaN exactly where the value
    is NaN."""
    P = small_cases()["subpix"]
    rng = np.random.default_rng(4)
    mask = (rng.random((P.NrPixelsZ, P.NrPixelsY)) < 0.01).astype(np.float64)
    frames = rng.poisson(100.0, (1, P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    print("[variance] subpix, bilinear, linearity", flush=True)
    g = ma.build_map(P, mask=mask)
    V1 = rng.random(frames.shape).astype(np.float32) * 10
    V2 = rng.random(frames.shape).astype(np.float32) * 10
    c, a = ma.integrate_with_variance(frames, g, variance_image=V1, mask_nan=True)
    _, b = ma.integrate_with_variance(frames, g, variance_image=V2, mask_nan=True)
    _, ab = ma.integrate_with_variance(frames, g, variance_image=V1 + 2 * V2, mask_nan=True)
    assert np.array_equal(np.isnan(a), np.isnan(c)) and np.isnan(c).any()
    ok = ~np.isnan(a)
    np.testing.assert_allclose(ab[ok], (a + 2 * b)[ok], rtol=2e-5, atol=1e-7)
    assert (a[ok] >= 0).all()
    # Poisson without a dark: the variance image is the frame itself
    _, p = ma.integrate_with_variance(frames, g, mask_nan=True)
    _, q = ma.integrate_with_variance(frames, g, variance_image=frames.astype(np.float32),
                                      mask_nan=True)
    np.testing.assert_allclose(p[ok], q[ok], rtol=1e-6)


def test_profile_with_variance_matches_oracle(oracle_lib):
    """profile_1d_with_variance (mc_profile_variance) on cakes with
    propagated variance, NaN (masked) bins included, against the FP64
    restatement of kernels.py:604-644 with the oracle map's in-bounds
    areas: 1e-5 of the same reduction of |cake| (intensity) and 1e-5
    relative (variance, positive terms); host arrays and torch tensors."""
    from oracle.profile import profile_with_variance
    P = small_cases()["tilt"]
    rng = np.random.default_rng(8)
    mask = (rng.random((P.NrPixelsZ, P.NrPixelsY)) < 0.01).astype(np.float64)
    frames = rng.poisson(300.0, (3, P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    dark = rng.poisson(30.0, (P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    print("[variance] tilt, profile with variance", flush=True)
    g = ma.build_map(P, mask=mask)
    g.set_dark(dark)
    om = oracle_lib.build_map(to_oracle(P), mask=mask)
    area = oracle_lib.area_inbounds(om, P.NrPixelsY, P.NrPixelsZ).reshape(P.n_r_bins,
                                                                         P.n_eta_bins)
    cakes, var = ma.integrate_with_variance(frames, g, mode="bilinear", mask_nan=True)
    assert np.isnan(cakes).any()
    worst = [0.0, 0.0]
    for mode in ("area_weighted", "simple_mean"):
        prof, pvar = ma.profile_1d_with_variance(cakes, var, g, mode=mode)
        for i in range(len(frames)):
            want_p, want_v = profile_with_variance(cakes[i], var[i], area, mode)
            scale, _ = profile_with_variance(np.abs(cakes[i]), var[i], area, mode)
            ep = np.abs(prof[i] - want_p) / np.maximum(scale, 1e-30)
            ev = np.abs(pvar[i] - want_v) / np.maximum(np.abs(want_v), 1e-30)
            worst = [max(worst[0], ep.max()), max(worst[1], ev.max())]
    print(f"[variance] profile worst {worst[0]:.2e}, variance worst {worst[1]:.2e}", flush=True)
    assert worst[0] <= 1e-5 and worst[1] <= 1e-5
    torch = pytest.importorskip("torch")
    tp, tv = ma.profile_1d_with_variance(torch.as_tensor(cakes[0], device="cuda"),
                                         torch.as_tensor(var[0], device="cuda"), g)
    hp, hv = ma.profile_1d_with_variance(cakes[0], var[0], g)
    np.testing.assert_array_equal(tp.cpu().numpy(), hp)
    np.testing.assert_array_equal(tv.cpu().numpy(), hv)


@pytest.mark.parametrize("dtype", [np.uint16, np.float32])
def test_batched_poisson_variance(oracle_lib, dtype):
    """Three frames or more with Poisson statistics take the batched kernel
    (32-frame passes summing w x and w^2 x from the same staged pixel
    rows): values and variances against the FP64 oracle on a whole-pixel
    map, and against the one-frame kernel frame by frame on a tilted
    bilinear map with masked bins, over two passes (40 frames); 1e-5
    relative (of the |frame - dark| cake for the values)."""
    P = small_cases()["flat"]
    rng = np.random.default_rng(4)
    frames = rng.poisson(200.0, (5, P.NrPixelsZ, P.NrPixelsY)).astype(dtype)
    dark = rng.poisson(20.0, (P.NrPixelsZ, P.NrPixelsY)).astype(dtype)
    print(f"[variance] batched Poisson, flat/floor, {np.dtype(dtype).name}", flush=True)
    g = ma.build_map(P)
    g.set_dark(dark)
    om = oracle_lib.build_map(to_oracle(P))
    cakes, var = ma.integrate_with_variance(frames, g, mode="floor")
    for i in range(len(frames)):
        img = frames[i].astype(np.float64) - dark
        want = _oracle_var(om, img, P.NrPixelsY)
        np.testing.assert_allclose(var[i], want, rtol=1e-5, atol=1e-6)
        # values: the one-frame kernel (checked against the oracle elsewhere)
        # (frames - dark > 0 everywhere here, so |cake| is the cake of |frame - dark|)
        one, _ = ma.integrate_batch(frames[i], g, mode="floor", mask_nan=False, lineouts=False)
        assert (img > 0).all()
        assert (np.abs(cakes[i] - one[0]) / np.maximum(np.abs(one[0]), 1e-30)).max() <= 1e-5
    P = small_cases()["tilt"]
    mask = (rng.random((P.NrPixelsZ, P.NrPixelsY)) < 0.01).astype(np.float64)
    frames = rng.poisson(300.0, (40, P.NrPixelsZ, P.NrPixelsY)).astype(dtype)
    dark = rng.poisson(30.0, (P.NrPixelsZ, P.NrPixelsY)).astype(dtype)
    print(f"[variance] batched Poisson, tilt/bilinear, masked, 40 frames", flush=True)
    g = ma.build_map(P, mask=mask)
    g.set_dark(dark)
    cakes, var = ma.integrate_with_variance(frames, g, mode="bilinear", mask_nan=True)
    worst = [0.0, 0.0]
    for i in range(0, 40, 3):
        c1, v1 = ma.integrate_with_variance(frames[i], g, mode="bilinear", mask_nan=True)
        assert np.array_equal(np.isnan(cakes[i]), np.isnan(c1))
        assert np.array_equal(np.isnan(var[i]), np.isnan(v1))
        assert (frames[i].astype(np.float64) - dark > 0).all()
        ok = ~np.isnan(c1)
        ec = np.abs(cakes[i][ok] - c1[ok]) / np.maximum(np.abs(c1[ok]), 1e-30)
        ev = np.abs(var[i][ok] - v1[ok]) / np.maximum(np.abs(v1[ok]), 1e-30)
        worst = [max(worst[0], ec.max()), max(worst[1], ev.max())]
    print(f"[variance] batched vs one-frame: values {worst[0]:.2e}, variances {worst[1]:.2e}",
          flush=True)
    assert worst[0] <= 1e-5 and worst[1] <= 1e-5
