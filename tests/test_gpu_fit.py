"""Batched LM fits on the GPU vs the FP64 CPU LM of the oracle, and exact
recovery of noise-free peaks.  Centres must agree within 1e-4 px."""
import numpy as np
import pytest

import midas_amd as ma
from midas_amd import peakfit, synth

pytestmark = pytest.mark.gpu
CENTER_TOL = 1e-4


def _pv_windows(n_win, n_pts, seed, noise):
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(50, 950, n_win)
    x = x0[:, None] + 0.25 * np.arange(n_pts)[None, :]
    c = x0 + 0.25 * (n_pts // 2) + rng.uniform(-0.6, 0.6, n_win)
    mu = rng.uniform(0, 1, n_win)
    fw = rng.uniform(1.0, 3.0, n_win)  # FWHM 1-3 px
    imax = rng.uniform(200, 2000, n_win)
    bg = rng.uniform(10, 100, n_win)
    y = np.stack([peakfit.pseudo_voigt(x[i], c[i], mu[i], fw[i], imax[i], bg[i])
                  for i in range(n_win)])
    if noise:
        y = y + rng.normal(0, 1, y.shape) * imax[:, None] / rng.uniform(20, 200, n_win)[:, None]
    return x, y, c


def test_pv_noise_free_recovery():
    x, y, c = _pv_windows(2000, 41, 0, noise=False)
    out = peakfit.fit_windows(x, y, c + 0.1, 0.25, model="pv", max_iter=200)
    assert (out[:, 7] > 0).all()
    err = np.abs(out[:, 1] - c)
    assert np.quantile(err, 0.99) < CENTER_TOL, np.sort(err)[-5:]


@pytest.mark.parametrize("model", ["pv", "tch"])
def test_gpu_lm_matches_fp64_oracle(model, oracle_lib):
    m = 0 if model == "pv" else 1
    rng = np.random.default_rng(4)
    n_win, n_pts = 300, 41
    x = 100.0 + 0.25 * np.arange(n_pts)[None, :] + rng.uniform(0, 500, (n_win, 1))
    xs, ys, gs = [], [], []
    for i in range(n_win):
        ctr = x[i, 20] + rng.uniform(-0.5, 0.5)
        if model == "pv":
            p = [ctr, rng.uniform(0, 1), rng.uniform(0.5, 1.5), rng.uniform(500, 3000), 50.0]
            yy = oracle_lib.pv_model(p, x[i])
            amp = p[3]
        else:
            p = [rng.uniform(500, 3000), ctr, rng.uniform(50, 400), rng.uniform(5, 60), 40.0, 3.0]
            yy = oracle_lib.tch_model(p, x[i])
            amp = yy.max() - 40
        yy = yy + rng.normal(0, amp / rng.uniform(20, 200), n_pts)
        xs.append(x[i]); ys.append(yy); gs.append(x[i, 20])
    X, Y, G = np.array(xs), np.array(ys), np.array(gs)
    out = peakfit.fit_windows(X, Y, G, 0.25, model=model, max_iter=300)
    cpu = []
    for i in range(n_win):
        p0, lo, hi = oracle_lib.window_init(m, X[i], Y[i].astype(np.float32).astype(np.float64),
                                            0.25, G[i])
        p, chi, it = oracle_lib.lm_fit(m, X[i], Y[i].astype(np.float32).astype(np.float64),
                                       p0, lo, hi, max_iter=300)
        cpu.append(p[0] if model == "pv" else p[1])
    d = np.abs(out[:, 1] - np.array(cpu))
    ok = out[:, 7] > 0
    assert ok.mean() > 0.99
    assert np.quantile(d[ok], 0.98) < CENTER_TOL, np.sort(d[ok])[-8:]


def test_fit_cake_rings_finds_the_rings():
    torch = pytest.importorskip("torch")
    P = synth.plain_1024()
    frames, dark, mask = synth.frames(P, 2, seed=3)
    g = ma.build_map(P, mask=mask)
    g.set_dark(dark)
    cakes, _ = ma.integrate_batch(torch.from_numpy(frames).cuda(), g, mode="bilinear",
                                  mask_nan=False)
    RLo, RHi, _, _ = ma.bin_edges(P)
    rc = (RLo + RHi) / 2
    rings = synth.ring_radii(P, 8)
    out = peakfit.fit_cake_rings(cakes, g, rc, rings, pad=20, model="pv")
    assert out.shape == (2, P.n_eta_bins, 8, 8)
    ok = out[..., 7] > 0
    assert ok.mean() > 0.95
    # noisy rings of sigma 1.5 px: centres within a small fraction of a bin
    assert np.median(np.abs(out[..., 1][ok] - np.broadcast_to(rings, ok.shape)[ok])) < 0.05


def test_fit_peaks_on_a_lineout():
    x = 10.0 + 0.25 * np.arange(400)
    locs = [30.0, 55.5, 80.25]
    y = np.full_like(x, 20.0)
    for c in locs:
        y += peakfit.pseudo_voigt(x, c, 0.4, 1.2, 500.0, 0.0)
    out = peakfit.fit_peaks(x, y, locs, 0.25, 20)
    assert out.shape == (3, 7)
    np.testing.assert_allclose(out[:, 1], locs, atol=0.02)


def _lineout(x, locs, slope=0.05, seed=None):
    y = 20.0 + slope * (x - x[0])
    for c in locs:
        y = y + peakfit.pseudo_voigt(x, c, 0.4, 1.2, 500.0, 0.0)
    if seed is not None:
        y = y + np.random.default_rng(seed).normal(0, 0.3, x.size)
    return y


def test_fit_peaks_auto_finds_the_peaks():
    """fitPeaksAutoDetect: SNIP removes the sloped background, detection
    finds the three peaks, the fits put the centres on them."""
    print("[fit] auto-detect", flush=True)
    x = 10.0 + 0.25 * np.arange(400)
    locs = [30.0, 55.5, 80.25]
    y = _lineout(x, locs, seed=4)
    out, n = peakfit.fit_peaks_auto(x, y, 3, 0.25, 20, 50)
    assert out.shape == (3, 7) and n == 3  # the largest areas: the three peaks
    np.testing.assert_allclose(out[:, 1], locs, atol=0.05)
    # over-selection also fits noise bumps; the peaks are among the five kept
    out5, n5 = peakfit.fit_peaks_auto(x, y, 5, 0.25, 20, 50)
    assert 3 <= n5 <= 5 and np.all(out5[n5:] == 0)
    assert np.all(np.diff(out5[:n5, 1]) > 0)
    for c in locs:
        assert np.min(np.abs(out5[:n5, 1] - c)) < 0.05
    out1, n1 = peakfit.fit_peaks_auto(x, _lineout(x, locs, seed=4), 2, 0.25, 20, 50)
    assert n1 == 2  # the two largest areas (all equal here: any two), in x order
    assert out1[0, 1] < out1[1, 1]


def test_stream_fitter_batches_frames(tmp_path):
    """Specified peaks over many frames fitted in shared launches give the
    per-frame fit_peaks rows; fit.bin holds them in frame order; a live
    update switches the peak list."""
    from midas_amd.stream_fits import StreamPeakFitter
    print("[fit] stream fitter", flush=True)
    x = 10.0 + 0.25 * np.arange(400)
    locs = [30.0, 55.5, 80.25]
    frames = [_lineout(x, [c + 0.01 * k for c in locs], seed=k) for k in range(12)]
    with StreamPeakFitter(x, 0.25, locs, out_dir=tmp_path, max_batch=8) as f:
        for k, y in enumerate(frames):
            f.submit(k, y)
    for k, y in enumerate(frames):
        want = peakfit.fit_peaks(x, y, [x[np.argmin(np.abs(x - c))] for c in locs], 0.25, 20)
        np.testing.assert_allclose(f.fitted[k], want, rtol=1e-6, atol=1e-9)
    raw = np.fromfile(tmp_path / "fit.bin", "<f8").reshape(12, 3, 7)
    np.testing.assert_array_equal(raw, np.stack([f.fitted[k] for k in range(12)]))
    (tmp_path / "peak_update.txt").write_text("replace\n55.5\n")
    with StreamPeakFitter(x, 0.25, locs, out_dir=tmp_path, fit_roi_auto=True) as g:
        g.submit(0, frames[0])
    assert g.locations == [55.5] and g.fitted[0].shape == (1, 7)
    assert abs(g.fitted[0][0, 1] - 55.5) < 0.05


def test_pv_doublet_recovers_both_centres():
    """pf_fit_doublet_pv (PeakFit.c:984-1101): two pV peaks 1.2-2 px apart
    sharing mu and background, noise-free, recovered from starting centres
    0.3 px off; SNR and widths come back with them."""
    print("[fit] pV doublet", flush=True)
    rng = np.random.default_rng(9)
    W, N, step = 64, 41, 0.25
    x = 100.0 + step * np.arange(N)[None, :] + rng.uniform(0, 50, (W, 1)).round(2)
    c1 = x[:, 15] + rng.uniform(-0.1, 0.1, W)
    c2 = c1 + rng.uniform(1.2, 2.0, W)
    mu, g1, g2 = 0.4, 0.9, 1.1
    i1 = rng.uniform(200, 400, W)
    i2 = rng.uniform(200, 400, W)
    y = (20.0 + peakfit.pseudo_voigt(x, c1[:, None], mu, g1, i1[:, None], 0.0)
         + peakfit.pseudo_voigt(x, c2[:, None], mu, g2, i2[:, None], 0.0))
    guess = np.stack([c1 - 0.3, c2 + 0.3], axis=1)
    out = peakfit.fit_windows(x, y, guess, step, model="pv2", max_iter=200)
    assert np.all(out[:, 7] > 0)
    np.testing.assert_allclose(out[:, 0], c1, atol=2e-3)
    np.testing.assert_allclose(out[:, 1], c2, atol=2e-3)
    np.testing.assert_allclose(out[:, 4], g1, rtol=2e-2)
    np.testing.assert_allclose(out[:, 5], g2, rtol=2e-2)
    np.testing.assert_allclose(out[:, 6], mu, atol=2e-2)
    assert np.all(out[:, 2] > 100) and np.all(out[:, 3] > 100)  # noise-free: large SNR
    with pytest.raises(ValueError):
        peakfit.fit_windows(x, y, c1, step, model="pv2")


def test_fit_cake_rings_with_a_doublet():
    """Per-eta doublet pass (IntegratorZarrOMP.c:1912-2016): two rings
    1.5 px apart in synthetic cakes, fitted jointly, come back on their own
    centres; a ring far from both stays a singlet fit."""
    torch = pytest.importorskip("torch")
    print("[fit] cake rings with a doublet", flush=True)
    P = synth.plain_1024()
    g = ma.build_map(P)
    RLo, RHi, _, _ = ma.bin_edges(P)
    rc = (RLo + RHi) / 2
    nr, ne = P.n_r_bins, P.n_eta_bins
    step = float(rc[1] - rc[0])
    r1, r2, r3 = rc[nr // 3] + 0.1, rc[nr // 3] + 0.1 + 6 * step, rc[2 * nr // 3]
    prof = 30.0 + peakfit.pseudo_voigt(rc, r1, 0.3, 3 * step, 400.0, 0.0) \
        + peakfit.pseudo_voigt(rc, r2, 0.3, 3 * step, 300.0, 0.0) \
        + peakfit.pseudo_voigt(rc, r3, 0.3, 3 * step, 500.0, 0.0)
    cakes = np.repeat(prof[None, :, None], ne, axis=2).repeat(2, axis=0).astype(np.float32)
    dev = torch.from_numpy(cakes).cuda()
    rings = [r1, r2, r3]
    single = peakfit.fit_cake_rings(dev, g, rc, rings, pad=20, model="pv")
    both = peakfit.fit_cake_rings(dev, g, rc, rings, pad=20, model="pv",
                                  doublet_separation=8 * step)
    np.testing.assert_array_equal(both[..., 2, :], single[..., 2, :])
    assert np.all(both[..., :2, 7] > 0)
    np.testing.assert_allclose(both[..., 0, 1], r1, atol=2e-3)
    np.testing.assert_allclose(both[..., 1, 1], r2, atol=2e-3)
    assert np.all(both[..., :2, 6] > 1.0)
    # the singlet fits of the close pair are pulled by the neighbour
    assert np.abs(single[..., 0, 1] - r1).max() > np.abs(both[..., 0, 1] - r1).max()
