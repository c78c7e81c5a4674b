"""The autotuner: a tuned plan is written to the plan directory keyed by
the map's content, pixel type, device and map header hash, and applied
from there without timing; every candidate the tuner may pick (one-frame
lanes per bin, batched tile shapes) integrates within 1e-5 of the FP64
oracle."""
import numpy as np
import pytest

import midas_amd as ma
from helpers import rel_err, small_cases, to_oracle

pytestmark = pytest.mark.gpu
TOL = 1e-5


def test_plan_is_tuned_once_then_loaded(tmp_path):
    P = small_cases()["tilt"]
    print("[tune] tuning a small map", flush=True)
    g = ma.build_map(P)
    plan = g.autotune_plan(tmp_path)
    assert not plan["loaded"] and plan["frames_per_s"] > 0
    files = list(tmp_path.glob("*.plan"))
    assert [f.name for f in files] == [plan["key"] + ".plan"]
    text = files[0].read_text()
    assert "# batched B" in text and "# one-frame lanes" in text
    assert plan["batch"] in (32, 64) and plan["rows_per_group"] in (16, 32)
    assert (plan["tile_bins"], plan["tile_entries"]) in ((256, 2048), (128, 1024), (512, 4096))
    assert plan["plain_lanes"] in (4, 8, 16)
    # a second map of the same geometry: same key, plan read back
    g2 = ma.build_map(P)
    assert g2.plan_key() == plan["key"]
    again = g2.autotune_plan(tmp_path)
    assert again["loaded"]
    for k in ("batch", "rows_per_group", "lds_bytes", "tile_bins", "tile_entries",
              "plain_lanes"):
        assert again[k] == plan[k], k
    # another pixel type or another geometry: another key
    assert g2.plan_key(np.int32) != plan["key"]
    Q = small_cases()["flat"]
    assert ma.build_map(Q).plan_key() != plan["key"]


@pytest.mark.parametrize("lanes", [4, 8, 16])
def test_one_frame_lanes_match_oracle(lanes, oracle_lib):
    P = small_cases()["tilt"]
    rng = np.random.default_rng(3)
    frame = rng.poisson(200.0, (P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    dark = rng.poisson(20.0, (P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    g = ma.build_map(P)
    g.set_dark(dark)
    g.set_plain_lanes(lanes)
    assert g.plain_lanes() == lanes
    cake, _ = ma.integrate_batch(frame, g, mode="bilinear", mask_nan=True)
    om = oracle_lib.build_map(to_oracle(P))
    img = frame.astype(np.float64) - dark
    ref, _ = oracle_lib.integrate(om, img, P.NrPixelsY, P.NrPixelsZ, semantics=0)
    sc, _ = oracle_lib.integrate(om, np.abs(img), P.NrPixelsY, P.NrPixelsZ, semantics=0)
    assert rel_err(cake[0], ref, sc).max() <= TOL
    with pytest.raises(ma._lib.MidasError):
        g.set_plain_lanes(32)
