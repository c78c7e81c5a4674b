"""The 2048x2048 batching workload (2048^2 uint16 frames, 12 rings, R 10-1000
by 0.25 px, eta by 1 deg: 3960 x 360 bins, solid angle + polarisation 0.99,
0.1 % hot pixels masked) on the GPU:

* bin membership of the GPU mapper against the compiled reference mapper's
  record (tests/golden/batch_2048.json: SHA-256 of the nMap pairs, the mask
  flags and the y/z of every record in order; FP64 sums of frac and
  areaWeight);
* a full 64-frame batched pass and the one-frame kernel against the FP64
  oracle integrated on the same map (oracle cake of frame 0 checked against
  the reference cake summary of the record), per bin within 1e-5 of the
  oracle cake of |frame - dark|, NaN bins at the same places, lineouts
  within 1e-5 of the lineouts of the |frame - dark| cake.
"""
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
REC = json.loads((Path(__file__).parent / "golden" / "batch_2048.json").read_text())


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def c2(oracle_lib):
    P = synth.batch_2048()
    print("[c2] generating 8 seeded 2048^2 frames", flush=True)
    frames, dark, mask = synth.frames(P, 8, seed=0, n_rings=12)
    assert _digest(mask) == REC["sha256_mask"] and _digest(dark) == REC["sha256_dark"]
    print("[c2] GPU map", flush=True)
    g = ma.build_map(P, mask=mask)
    g.set_dark(dark)
    pm = g.download()
    pairs = np.stack([pm.counts, pm.offsets], axis=1).astype(np.int32)
    om = oracle_lib.MapArrays(np.ascontiguousarray(pm.pxList).view(oracle_lib.ENTRY_DTYPE),
                              pairs, pm.mask_flags.astype(np.int32), g.n_r, g.n_eta)
    yield P, frames, dark, g, pm, om
    g.close()


def test_c2_map_membership(c2):
    P, frames, dark, g, pm, om = c2
    # the build's host verification pass (4096 pixels spread over the detector)
    v = g.verify_info()
    print(f"[c2] host verification of the GPU build: {v}", flush=True)
    assert v["pixels"] >= 4096 and v["records"] > 0
    assert v["membership_differs"] == 0 and v["values_differ"] == 0, v
    assert pm.n_entries == REC["nnz"]
    assert _digest(om.pairs) == REC["sha256_pairs"]
    assert _digest(om.flags) == REC["sha256_flags"]
    assert _digest(pm.pxList["y"]) == REC["sha256_y"]
    assert _digest(pm.pxList["z"]) == REC["sha256_z"]
    assert int(om.flags.sum()) == REC["n_flagged"]
    np.testing.assert_allclose(np.sum(pm.pxList["frac"], dtype=np.float64), REC["sum_frac"],
                               rtol=1e-10)
    np.testing.assert_allclose(np.sum(pm.pxList["areaWeight"], dtype=np.float64),
                               REC["sum_area"], rtol=1e-7)


def _oracle(orc, om, P, frame, dark):
    img = frame.astype(np.float64) - dark
    cake, _ = orc.integrate(om, img, P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z))
    scale, _ = orc.integrate(om, np.abs(img), P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z))
    return cake, scale


def test_c2_batched_and_one_frame_match_oracle(c2, oracle_lib):
    P, frames, dark, g, pm, om = c2
    assert _digest(frames[0]) == REC["sha256_frame"]
    # 64 lanes: the 8 seeded frames, each repeated with a per-copy offset of
    # 0..7 counts (distinct values in every lane of the pass)
    batch = np.concatenate([frames + np.uint16(k) for k in range(8)])
    print("[c2] one 64-frame batched pass", flush=True)
    cakes, lo = ma.integrate_batch(batch, g, mode="bilinear", mask_nan=True)
    print("[c2] one-frame kernel", flush=True)
    c1, l1 = ma.integrate_batch(batch[5], g, mode="bilinear", mask_nan=True)
    ainb = oracle_lib.area_inbounds(om, P.NrPixelsY, P.NrPixelsZ)
    ref0, _ = _oracle(oracle_lib, om, P, frames[0], dark)
    # the oracle on the GPU-built map reproduces the reference cake of frame
    # 0 (the map's FP64 areas may differ from glibc's in the last bits)
    np.testing.assert_allclose(np.nansum(ref0), REC["cake_nansum"], rtol=1e-9)
    assert int(np.isnan(ref0).sum()) == REC["cake_n_nan"]
    flat = ref0.reshape(-1)
    for k, v in REC["cake_sample"].items():
        np.testing.assert_allclose(flat[int(k)], v, rtol=1e-8)
    worst = 0.0
    for lane in (0, 5, 13, 22, 31, 40, 47, 54, 63):
        ref, sc = _oracle(oracle_lib, om, P, batch[lane], dark)
        e = rel_err(cakes[lane], ref, sc)
        worst = max(worst, float(e.max()))
        assert e.max() <= TOL, f"lane {lane}: max rel err {e.max():.3g}"
        got = [lo[lane, r] for r in range(3)]
        if lane == 5:
            e1 = rel_err(c1[0], ref, sc)
            assert e1.max() <= TOL, f"one-frame kernel: max rel err {e1.max():.3g}"
            got1 = [l1[0, r] for r in range(3)]
        want = oracle_lib.lineouts(ref, ainb)
        scale = oracle_lib.lineouts(np.where(np.isnan(ref), np.nan, sc), ainb)
        for r in range(3):
            el = np.abs(got[r] - want[r]) / np.maximum(np.abs(scale[r]), 1e-30)
            assert el.max() <= TOL, f"lane {lane} lineout {r}: {el.max():.3g}"
            if lane == 5:
                el1 = np.abs(got1[r] - want[r]) / np.maximum(np.abs(scale[r]), 1e-30)
                assert el1.max() <= TOL, f"one-frame lineout {r}: {el1.max():.3g}"
    print(f"[c2] worst cake rel err over the checked lanes {worst:.3g}", flush=True)


@pytest.mark.parametrize("batch", [16, 32])
def test_c2_narrow_passes_match_oracle(c2, oracle_lib, batch):
    """The 16- and 32-frame pass kernels (one and two frames per lane) on
    the full 2048^2 layout: 32 frames in passes of `batch`, lanes checked
    against the FP64 oracle at 1e-5 (cakes and lineouts)."""
    P, frames, dark, g, pm, om = c2
    saved = g.tuning()
    g.set_tuning(batch=batch, rows_per_group=16, lds_bytes=saved[2])
    try:
        batch32 = np.concatenate([frames + np.uint16(k) for k in range(4)])
        print(f"[c2] 32 frames in passes of {batch}", flush=True)
        cakes, lo = ma.integrate_batch(batch32, g, mode="bilinear", mask_nan=True)
    finally:
        g.set_tuning(*saved)
    ainb = oracle_lib.area_inbounds(om, P.NrPixelsY, P.NrPixelsZ)
    for lane in (0, 9, 17, 31):
        ref, sc = _oracle(oracle_lib, om, P, batch32[lane], dark)
        e = rel_err(cakes[lane], ref, sc)
        assert e.max() <= TOL, f"width {batch} lane {lane}: max rel err {e.max():.3g}"
        want = oracle_lib.lineouts(ref, ainb)
        scale = oracle_lib.lineouts(np.where(np.isnan(ref), np.nan, sc), ainb)
        for r in range(3):
            el = np.abs(lo[lane, r] - want[r]) / np.maximum(np.abs(scale[r]), 1e-30)
            assert el.max() <= TOL, f"width {batch} lane {lane} lineout {r}: {el.max():.3g}"
