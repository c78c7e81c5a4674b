"""The Q-mode workload (Q 0.5-12 A^-1 by 0.001, eta by 0.5 deg: 11500 x 720
bins; SubPixelLevel 2, tilts ty 0.2 / tz -0.15 deg, parallax 200 um, float32
frames, gradient mode) on the GPU:

* on a 320x288 crop of its detector: bin membership of the GPU mapper
  against the compiled reference mapper's record (tests/golden/
  qmode_crop.json), the oracle's gradient-mode cake of the seeded frame
  against the reference cake summary, and the batched and one-frame kernels
  against the FP64 oracle within 1e-5 of the |frame - dark| cake;
* on the full 2880x2880 detector: a batched pass of float32 frames in
  gradient mode against the FP64 oracle integrated on the GPU-built map."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import midas_amd as ma
from midas_amd import synth
from helpers import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5
REC = json.loads((Path(__file__).parent / "golden" / "qmode_crop.json").read_text())


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _oracle_map(orc, g):
    pm = g.download()
    pairs = np.stack([pm.counts, pm.offsets], axis=1).astype(np.int32)
    om = orc.MapArrays(np.ascontiguousarray(pm.pxList).view(orc.ENTRY_DTYPE), pairs,
                       pm.mask_flags.astype(np.int32), g.n_r, g.n_eta)
    return pm, om


def _oracle(orc, om, P, frame, dark):
    img = frame.astype(np.float64) - dark.astype(np.float64)
    ref, _ = orc.integrate(om, img, P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z),
                           gradient=True)
    sc, _ = orc.integrate(om, np.abs(img), P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z),
                          gradient=True)
    return ref, sc


def test_qmode_crop_membership_and_cakes(oracle_lib):
    P = synth.qmode_crop()
    frames, dark = synth.qmode_frames(P, 5, seed=0)
    assert _digest(frames[0]) == REC["sha256_frame"] and _digest(dark) == REC["sha256_dark"]
    print("[qmode] crop map", flush=True)
    g = ma.build_map(P)
    g.set_dark(dark)
    pm, om = _oracle_map(oracle_lib, g)
    assert pm.n_entries == REC["nnz"]
    assert _digest(om.pairs) == REC["sha256_pairs"]
    assert _digest(om.flags) == REC["sha256_flags"]
    assert _digest(pm.pxList["y"]) == REC["sha256_y"]
    assert _digest(pm.pxList["z"]) == REC["sha256_z"]
    np.testing.assert_allclose(np.sum(pm.pxList["frac"], dtype=np.float64), REC["sum_frac"],
                               rtol=1e-10)
    np.testing.assert_allclose(np.sum(pm.pxList["deltaR"], dtype=np.float64),
                               REC["sum_deltaR"], rtol=1e-6, atol=1e-3)
    ref0, _ = _oracle(oracle_lib, om, P, frames[0], dark)
    # the oracle on the GPU-built map against the reference cake: the map's
    # FP64 areas and float32 deltaR may differ from glibc's in the last bit,
    # which moves a gradient-mode sample by ~1e-8
    np.testing.assert_allclose(np.nansum(ref0), REC["cake_nansum"], rtol=1e-9)
    flat = ref0.reshape(-1)
    for k, v in REC["cake_sample"].items():
        np.testing.assert_allclose(flat[int(k)], v, rtol=1e-6)
    print("[qmode] crop: batched pass of 5, one-frame of frame 3", flush=True)
    cb, _ = ma.integrate_batch(frames, g, mode="gradient", mask_nan=True)
    c1, _ = ma.integrate_batch(frames[3], g, mode="gradient", mask_nan=True)
    for i in range(len(frames)):
        ref, sc = _oracle(oracle_lib, om, P, frames[i], dark)
        e = rel_err(cb[i], ref, sc)
        assert e.max() <= TOL, f"batched frame {i}: {e.max():.3g}"
        if i == 3:
            e1 = rel_err(c1[0], ref, sc)
            assert e1.max() <= TOL, f"one-frame: {e1.max():.3g}"
    g.close()


def test_qmode_2880_batched_pass(oracle_lib):
    """The full 2880x2880 detector (about 92 M map records): a 64-frame
    batched pass of float32 frames in gradient mode; three lanes and the
    one-frame kernel against the FP64 oracle on the GPU-built map."""
    P = synth.qmode_2880()
    print("[qmode] 2880^2 map on the GPU", flush=True)
    g = ma.build_map(P)
    frames, dark = synth.qmode_frames(P, 4, seed=1)
    g.set_dark(dark)
    batch = np.concatenate([frames + np.float32(k) for k in range(16)])
    print("[qmode] 2880^2 batched pass of 64 float32 frames (gradient)", flush=True)
    cb, _ = ma.integrate_batch(batch, g, mode="gradient", mask_nan=True)
    c1, _ = ma.integrate_batch(batch[9], g, mode="gradient", mask_nan=True)
    print("[qmode] oracle on the downloaded map", flush=True)
    pm, om = _oracle_map(oracle_lib, g)
    assert pm.n_entries > 80_000_000
    for lane in (0, 9, 63):
        ref, sc = _oracle(oracle_lib, om, P, batch[lane], dark)
        e = rel_err(cb[lane], ref, sc)
        print(f"[qmode] lane {lane}: max rel err {e.max():.3g}", flush=True)
        assert e.max() <= TOL, f"lane {lane}: {e.max():.3g}"
        if lane == 9:
            e1 = rel_err(c1[0], ref, sc)
            assert e1.max() <= TOL, f"one-frame: {e1.max():.3g}"
    g.close()
