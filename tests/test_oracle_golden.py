"""The oracle against the fixtures written by the reference's own code
(tests/golden/make_golden.py).  Runs anywhere: the fixtures are data."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from midas_amd.params import IntegrationParams
from helpers import to_oracle

GOLD = Path(__file__).resolve().parent / "golden"
TINY = sorted(p.stem for p in GOLD.glob("tiny_*.npz"))


def _load(name):
    z = np.load(GOLD / f"{name}.npz", allow_pickle=False)
    P = IntegrationParams(**json.loads(str(z["params"])))
    sh = json.loads(str(z["panel_shifts"]))
    return z, P, sh


@pytest.mark.parametrize("name", TINY)
def test_oracle_map_is_byte_identical_to_reference(name, oracle_lib):
    z, P, sh = _load(name)
    om = oracle_lib.build_map(to_oracle(P), mask=z["mask"], panel_shifts=sh)
    np.testing.assert_array_equal(om.pairs, z["pairs"])
    np.testing.assert_array_equal(om.flags, z["flags"])
    assert om.entries.view(np.uint8).tobytes() == z["entries"].tobytes()


@pytest.mark.parametrize("name", TINY)
def test_oracle_cake_matches_reference(name, oracle_lib):
    z, P, sh = _load(name)
    om = oracle_lib.build_map(to_oracle(P), mask=z["mask"], panel_shifts=sh)
    img = z["frame"] - z["dark"]
    cake, area = oracle_lib.integrate(om, img, P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z))
    np.testing.assert_array_equal(area, z["area"])
    np.testing.assert_allclose(cake, z["cake"], rtol=1e-12, atol=1e-9, equal_nan=True)
    cg, _ = oracle_lib.integrate(om, img, P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z),
                                 gradient=True)
    np.testing.assert_allclose(cg, z["cake_gradient"], rtol=1e-12, atol=1e-9, equal_nan=True)


def test_oracle_plain_1024_digests(oracle_lib):
    from midas_amd import synth
    rec = json.loads((GOLD / "plain_1024.json").read_text())
    P = synth.plain_1024()
    frames, dark, mask = synth.frames(P, 1, seed=0)
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    assert sha(frames[0]) == rec["sha256_frame"] and sha(mask) == rec["sha256_mask"]
    om = oracle_lib.build_map(to_oracle(P), mask=mask)
    assert om.entries.size == rec["nnz"]
    assert sha(om.pairs) == rec["sha256_pairs"]
    assert sha(om.flags) == rec["sha256_flags"]
    assert sha(om.entries) == rec["sha256_entries"]
    cake, _ = oracle_lib.integrate(om, frames[0].astype(np.float64) - dark, P.NrPixelsY,
                                   P.NrPixelsZ)
    assert int(np.isnan(cake).sum()) == rec["cake_n_nan"]
    flat = cake.reshape(-1)
    for k, v in rec["cake_sample"].items():
        assert flat[int(k)] == pytest.approx(v, rel=1e-12, nan_ok=True)


@pytest.mark.slow
def test_oracle_batch_2048_digests(oracle_lib):
    """The 2048x2048 workload: the oracle's map has the reference's bin
    membership (pairs, flags, y/z of every record) and the same FP64 area
    sums; its cake of frame 0 matches the reference cake summary."""
    from midas_amd import synth
    rec = json.loads((GOLD / "batch_2048.json").read_text())
    P = synth.batch_2048()
    frames, dark, mask = synth.frames(P, 1, seed=0, n_rings=12)
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
    assert sha(frames[0]) == rec["sha256_frame"] and sha(mask) == rec["sha256_mask"]
    om = oracle_lib.build_map(to_oracle(P), mask=mask)
    assert om.entries.size == rec["nnz"]
    assert sha(om.pairs) == rec["sha256_pairs"]
    assert sha(om.flags) == rec["sha256_flags"]
    assert sha(om.entries["y"]) == rec["sha256_y"] and sha(om.entries["z"]) == rec["sha256_z"]
    assert float(np.sum(om.entries["frac"], dtype=np.float64)) == pytest.approx(
        rec["sum_frac"], rel=1e-13)
    cake, _ = oracle_lib.integrate(om, frames[0].astype(np.float64) - dark, P.NrPixelsY,
                                   P.NrPixelsZ, bc=(P.BC_y, P.BC_z))
    assert int(np.isnan(cake).sum()) == rec["cake_n_nan"]
    assert float(np.nansum(cake)) == pytest.approx(rec["cake_nansum"], rel=1e-12)
    flat = cake.reshape(-1)
    for k, v in rec["cake_sample"].items():
        assert flat[int(k)] == pytest.approx(v, rel=1e-12, nan_ok=True)
