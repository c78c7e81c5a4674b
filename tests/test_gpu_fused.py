"""Fused multi-map averaging (mc_map_fuse) vs the FP64 restatement of
build_fused_geometry (oracle/fused.py).

The oracle applies the factors to the image (per map: raw sums from the
oracle's mapper and integrator, divided by the in-bounds areas,
interpolated along R, averaged); the product folds them into one matrix
first, in FP64, stored as FP32 weights.  Tolerance 1e-5 relative (the
frames are positive, so the cake is its own scale)."""
from copy import deepcopy

import numpy as np
import pytest

import midas_amd as ma
from midas_amd.fused_csr import r_axis
from helpers import small_cases, to_oracle
from oracle import fused as fu

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _case(name):
    P = small_cases()[name if name != "flat_uneven" else "flat"]
    if name == "flat_uneven":
        P.RMax = 60.2  # shifted maps get one R bin fewer
    return P


@pytest.mark.parametrize("name,mode,n", [("flat", "bilinear", 4), ("tilt", "gradient", 3),
                                         ("flat_uneven", "bilinear", 3), ("tilt", "bilinear", 1)])
def test_fused_matches_oracle(oracle_lib, name, mode, n):
    P = _case(name)
    rng = np.random.default_rng(7)
    frames = rng.poisson(300.0, (3, P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    print(f"[fused] {name} {mode} n_shifts={n}", flush=True)
    fused = ma.build_fused_geometry(P, n_shifts=n, mode=mode)
    assert fused.n_r == P.n_r_bins
    cakes, _ = ma.integrate_batch(frames, fused, mode=mode, normalize=False, mask_nan=False)
    normed, _ = ma.integrate_batch(frames, fused, mode=mode, normalize=True, mask_nan=False)
    np.testing.assert_array_equal(cakes, normed)  # unit areas
    one, _ = ma.integrate_batch(frames[1], fused, mode=mode, normalize=False, mask_nan=False)
    np.testing.assert_allclose(one[0], cakes[1], rtol=3e-6, atol=1e-4)
    oms, axes = [], []
    for k in range(n):
        Pk = deepcopy(P)
        Pk.RMin = P.RMin + P.RBinSize * k / n
        oms.append(oracle_lib.build_map(to_oracle(Pk)))
        axes.append(r_axis(n_r=Pk.n_r_bins, RMin=Pk.RMin, RBinSize=Pk.RBinSize))
    if name == "flat_uneven":
        assert min(len(a) for a in axes) == len(axes[0]) - 1
    for i in range(len(frames)):
        img = frames[i].astype(np.float64)
        raws, areas = [], []
        for om in oms:
            raw, _ = oracle_lib.integrate(om, img, P.NrPixelsY, P.NrPixelsZ,
                                          bc=(P.BC_y, P.BC_z), gradient=mode == "gradient",
                                          normalize=False, semantics=1)
            raws.append(raw)
            areas.append(oracle_lib.area_inbounds(om, P.NrPixelsY, P.NrPixelsZ))
        want = fu.fused_cake(raws, areas, axes)
        err = np.abs(cakes[i] - want) / np.maximum(np.abs(want), 1e-30)
        err[(want == 0) & (cakes[i] == 0)] = 0
        assert err.max() <= TOL, (name, i, err.max())


def test_fused_map_refusals(tmp_path):
    P = small_cases()["flat"]
    fused = ma.build_fused_geometry(P, n_shifts=2, mode="bilinear")
    frame = np.ones((P.NrPixelsZ, P.NrPixelsY), np.uint16)
    with pytest.raises(RuntimeError):
        ma.integrate_batch(frame, fused, mode="floor")
    with pytest.raises(RuntimeError):
        fused.write(tmp_path)
    with pytest.raises(ValueError):
        ma.build_fused_geometry(P, n_shifts=0)
