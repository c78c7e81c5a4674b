"""A high dark pedestal: 1024x1024 frames with a dark of 3000 + Poisson(5)
counts (the mean of three dark frames, so the dark has fractional parts)
under a signal of 1-5 counts a pixel plus weak rings, through the one-frame
kernel, the batched kernel (uint16 and FP32 pixel-major batches) and the
variance kernel, each bin within 1e-5 of the FP64 oracle cake of
|frame - dark| (IntegratorZarrOMP.c:1646-1669: Image = raw - dark, in FP64).

Summing raw counts in FP32 and subtracting the dark per bin afterwards loses
about log2(dark / signal) bits; this case fails that way by an order of
magnitude, which is why the kernels subtract the dark's integer part from
each pixel before the product."""
import numpy as np
import pytest

import midas_amd as ma
from midas_amd import synth
from helpers import rel_err, to_oracle

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def pedestal(oracle_lib):
    P = synth.plain_1024()
    rng = np.random.default_rng(21)
    shape = (P.NrPixelsZ, P.NrPixelsY)
    darks = (3000 + rng.poisson(5.0, (3,) + shape)).astype(np.uint16)
    signal = rng.uniform(1.0, 5.0, shape) + 0.02 * synth.ring_image(P, synth.ring_radii(P, 8))
    frames = np.stack([darks[0] + rng.poisson(signal) for _ in range(6)]).astype(np.uint16)
    mask = (rng.random(shape) < 0.001).astype(np.float64)
    print("[dark] map of the 1024^2 workload", flush=True)
    g = ma.build_map(P, mask=mask)
    g.set_dark(darks)
    pm = g.download()
    pairs = np.stack([pm.counts, pm.offsets], axis=1).astype(np.int32)
    om = oracle_lib.MapArrays(np.ascontiguousarray(pm.pxList).view(oracle_lib.ENTRY_DTYPE),
                              pairs, pm.mask_flags.astype(np.int32), g.n_r, g.n_eta)
    dmean = darks.astype(np.float64).mean(axis=0)
    assert np.any(dmean != np.rint(dmean))
    yield P, frames, dmean, g, om
    g.close()


def _ref(orc, om, P, frame, dmean):
    img = frame.astype(np.float64) - dmean
    ref, _ = orc.integrate(om, img, P.NrPixelsY, P.NrPixelsZ)
    sc, _ = orc.integrate(om, np.abs(img), P.NrPixelsY, P.NrPixelsZ)
    return ref, sc


@pytest.mark.parametrize("dtype", [np.uint16, np.float32, np.int32])
def test_dark_pedestal_batched_and_one_frame(pedestal, oracle_lib, dtype):
    P, frames, dmean, g, om = pedestal
    f = frames.astype(dtype)
    print(f"[dark] {np.dtype(dtype).name}: batched pass of 6, one-frame of frame 2", flush=True)
    cb, _ = ma.integrate_batch(f, g, mode="bilinear", mask_nan=True)
    c1, _ = ma.integrate_batch(f[2], g, mode="bilinear", mask_nan=True)
    worst = 0.0
    for i in range(len(f)):
        ref, sc = _ref(oracle_lib, om, P, frames[i], dmean)
        e = rel_err(cb[i], ref, sc)
        worst = max(worst, float(e.max()))
        assert e.max() <= TOL, f"batched frame {i}: {e.max():.3g}"
        if i == 2:
            e1 = rel_err(c1[0], ref, sc)
            assert e1.max() <= TOL, f"one-frame: {e1.max():.3g}"
    print(f"[dark] worst {worst:.3g}", flush=True)


def _oracle_var(om, V):
    """sum_k frac_k^2 V[pixel_k] / (sum areaWeight)^2 per bin (whole-pixel
    entries, floor mode), FP64."""
    e = om.entries
    bins = np.repeat(np.arange(om.n_r * om.n_eta), om.counts)
    y, z = e["y"].astype(np.int64), e["z"].astype(np.int64)
    fr = e["frac"].astype(np.float64)
    raw = np.zeros(om.n_r * om.n_eta)
    np.add.at(raw, bins, fr * fr * V[z, y])
    a = np.zeros(om.n_r * om.n_eta)
    np.add.at(a, bins, e["areaWeight"].astype(np.float64))
    out = np.where(a > 0, raw / np.maximum(a, 1e-300) ** 2, 0.0)
    return out.reshape(om.n_r, om.n_eta)


def test_dark_pedestal_variance(pedestal, oracle_lib):
    """Poisson variance of the dark-subtracted frame, sum w^2 (frame - dark)
    over whole-pixel entries (floor mode), and the value beside it, each
    within 1e-5 of FP64 (scale: the same sums of absolute terms)."""
    P, frames, dmean, g, om = pedestal
    print("[dark] variance kernel, floor mode", flush=True)
    cake, var = ma.integrate_with_variance(frames[1], g, mode="floor", mask_nan=True)
    img = frames[1].astype(np.float64) - dmean
    ref, _ = oracle_lib.integrate(om, img, P.NrPixelsY, P.NrPixelsZ, semantics=1)
    sc, _ = oracle_lib.integrate(om, np.abs(img), P.NrPixelsY, P.NrPixelsZ, semantics=1)
    flags = om.flags.reshape(ref.shape).astype(bool)
    assert np.array_equal(np.isnan(cake), flags)
    ok = ~flags
    e = np.abs(cake[ok] - ref[ok]) / np.maximum(np.abs(sc[ok]), 1e-30)
    assert e.max() <= TOL, e.max()
    vref = _oracle_var(om, img)
    vsc = _oracle_var(om, np.abs(img))
    ev = np.abs(var[ok] - vref[ok]) / np.maximum(vsc[ok], 1e-30)
    assert ev.max() <= TOL, ev.max()
