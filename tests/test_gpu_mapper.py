"""GPU FP64 mapper vs the oracle mapper (which is byte-identical to the
reference mapper_build_map, see test_oracle_vs_reference.py)."""
import numpy as np
import pytest

import midas_amd as ma
from helpers import assert_maps_equal, panel_shifts_for, small_cases, to_oracle

pytestmark = pytest.mark.gpu


def _mask(P, seed, frac=0.01):
    rng = np.random.default_rng(seed)
    return (rng.random((P.NrPixelsZ, P.NrPixelsY)) < frac).astype(np.float64)


@pytest.mark.parametrize("name", list(small_cases()))
def test_mapper_matches_oracle(name, oracle_lib):
    P = small_cases()[name]
    print(f"[mapper] case={name} {P.NrPixelsY}x{P.NrPixelsZ}", flush=True)
    mask = _mask(P, 7)
    sh = panel_shifts_for(name)
    npx = P.NrPixelsY * P.NrPixelsZ
    g = ma.build_map(P, mask=mask, panel_shifts=sh, verify_pixels=npx)
    # the build's own host pass recomputed every pixel and agreed
    v = g.verify_info()
    assert v["pixels"] == npx and v["records"] > 0, v
    assert v["membership_differs"] == 0 and v["values_differ"] == 0, v
    g.ctx.set_map_verify(4096)  # back to the default for later tests
    pm = g.download()
    om = oracle_lib.build_map(to_oracle(P), mask=mask, panel_shifts=sh)
    assert pm.n_entries == om.entries.size > 0
    assert_maps_equal(pm, om)
    # Rejected candidates in total (the reason can differ at rounding ties:
    # a degenerate polygon fails the vertex count on one side, the unique
    # vertex count on the other).
    info = g.info()
    assert sum(info.skipped) == sum(om.skipped)
    g.close()


def test_mapper_with_distortion_maps(oracle_lib):
    P = small_cases()["tilt"]
    rng = np.random.default_rng(3)
    dy = rng.normal(0, 0.05, (P.NrPixelsZ, P.NrPixelsY))
    dz = rng.normal(0, 0.05, (P.NrPixelsZ, P.NrPixelsY))
    res = rng.normal(0, 0.02, (P.NrPixelsZ, P.NrPixelsY))
    g = ma.build_map(P, dist_y=dy, dist_z=dz, residual=res)
    om = oracle_lib.build_map(to_oracle(P), dist_y=dy, dist_z=dz, residual=res)
    assert_maps_equal(g.download(), om)


def test_mapper_plain_1024_membership(oracle_lib):
    """The 1024x1024 workload: every bin's count and pixel list."""
    from midas_amd import synth
    P = synth.plain_1024()
    _, _, mask = synth.frames(P, 1, seed=0)
    print("[mapper] plain_1024", flush=True)
    g = ma.build_map(P, mask=mask, verify_pixels=P.NrPixelsY * P.NrPixelsZ)
    v = g.verify_info()
    print(f"[mapper] host verification: {v}", flush=True)
    assert v["pixels"] == P.NrPixelsY * P.NrPixelsZ
    assert v["membership_differs"] == 0 and v["values_differ"] == 0, v
    g.ctx.set_map_verify(4096)  # back to the default for later tests
    pm = g.download()
    om = oracle_lib.build_map(to_oracle(P), mask=mask)
    assert_maps_equal(pm, om)


def test_map_files_round_trip(tmp_path):
    P = small_cases()["flat"]
    g = ma.build_map(P)
    g.write(tmp_path, P)
    pm = ma.load_map(tmp_path / "Map.bin", tmp_path / "nMap.bin", tmp_path / "maskMap.bin")
    assert pm.map_header is not None and pm.map_header.param_hash == ma.header_for(P).param_hash
    ref = g.download()
    assert pm.pxList.tobytes() == ref.pxList.tobytes()
    g2 = ma.load_map_gpu(tmp_path / "Map.bin", tmp_path / "nMap.bin", tmp_path / "maskMap.bin",
                         n_r=P.n_r_bins, n_eta=P.n_eta_bins, ny=P.NrPixelsY, nz=P.NrPixelsZ)
    assert g2.download().pxList.tobytes() == ref.pxList.tobytes()
