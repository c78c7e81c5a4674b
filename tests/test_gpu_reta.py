"""REtaMap (IntegratorZarrOMP.c:1754-1763, 2073-2083) and the GPU stream's
RTthEtaAreaMap (IntegratorFitPeaksGPUStream.cu:817-899, 2443-2462): all five
rows of mc_map_reta and mc_map_reta_stream against the oracle's REtaMap
(pinned by Bragg's law and the Q-mode edges in test_oracle_reta.py), with
the oracle's per-bin areas: the sum of areaWeight over all of a bin's
entries, and over its on-detector entries only."""
import numpy as np
import pytest

import midas_amd as ma
from helpers import panel_shifts_for, small_cases, to_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["tilt", "qmode", "subpix", "panels"])
def test_reta_rows_match_oracle(name, oracle_lib):
    P = small_cases()[name]
    sh = panel_shifts_for(name)
    print(f"[reta] case={name}", flush=True)
    g = ma.build_map(P, panel_shifts=sh)
    O = to_oracle(P)
    om = oracle_lib.build_map(O, panel_shifts=sh)
    lam = 0.172979
    area_all = np.zeros(om.n_r * om.n_eta)
    np.add.at(area_all, np.repeat(np.arange(om.n_r * om.n_eta), om.counts),
              om.entries["areaWeight"].astype(np.float64))
    area_inb = oracle_lib.area_inbounds(om, P.NrPixelsY, P.NrPixelsZ)
    for in_bounds, area in ((False, area_all.reshape(om.n_r, om.n_eta)), (True, area_inb)):
        got = g.reta_map(P, lam=lam, in_bounds=in_bounds)
        want = oracle_lib.reta_map(O, area, lam)
        assert got.shape == want.shape == (5, om.n_r, om.n_eta)
        # R, 2theta, eta and Q: the same FP64 closed forms on the host
        np.testing.assert_allclose(got[[0, 1, 2, 4]], want[[0, 1, 2, 4]], rtol=1e-14, atol=0)
        # areas of the GPU-built map: FP64 overlaps within 1e-6 of a pixel
        np.testing.assert_allclose(got[3], want[3], rtol=1e-6, atol=1e-9)
    g.close()
