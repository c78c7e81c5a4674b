"""Result files fed by the GPU path: cakes and lineouts from the batched
kernel go through an output set and are read back."""
import numpy as np
import pytest

import midas_amd as ma
from helpers import rel_err, small_cases
from oracle import outputs as ref_out
from tests import h5lite

pytestmark = pytest.mark.gpu


def test_integrate_then_write(tmp_path):
    P = small_cases()["tilt"]
    P.OmegaStart, P.OmegaStep = 5.0, 0.5
    rng = np.random.default_rng(9)
    frames = rng.poisson(150.0, (5, P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    print("[outputs] tilt, 5 frames", flush=True)
    g = ma.build_map(P)
    cakes, lo = ma.integrate_batch(frames, g, mode="bilinear", mask_nan=True)
    reta = g.reta_map()
    files = {"caked_hdf", "per_frame", "omegas", "lineout", "stream", "int2d"}
    with ma.ResultFiles(tmp_path, "tilt_000001.zip", P, files=files, omega_sum_frames=2,
                        reta=reta) as out:
        for i in range(len(frames)):
            out.frame(i, cakes[i], lo[i])
    f = h5lite.File(tmp_path / "tilt_000001.zip.caked.hdf")
    for i in range(len(frames)):
        np.testing.assert_array_equal(f.dataset(f"/IntegrationResult/FrameNr_{i}").read(),
                                      cakes[i].astype(np.float64))
    np.testing.assert_array_equal(f.dataset("/REtaMap").read(), reta)
    assert sorted(f.keys("/OmegaSumFrame")) == ["LastFrameNumber_1", "LastFrameNumber_3",
                                                "LastFrameNumber_4"]
    nr = g.n_r
    st = np.fromfile(tmp_path / "lineout.bin").reshape(len(frames), nr, 2)
    np.testing.assert_array_equal(st[:, :, 1], lo[:, 1].astype(np.float64))
    lb = np.fromfile(tmp_path / "tilt_000001_lineout.bin").reshape(len(frames), nr)
    for i in range(len(frames)):
        np.testing.assert_allclose(lb[i], ref_out.lineout_cpu_mean(cakes[i]), rtol=1e-14)
        # the GPU's FP32 CPU-mean row agrees with the FP64 one written
        assert rel_err(lo[i, 0], lb[i], lb[i]).max() <= 1e-5


def test_rtth_eta_area_map_of_the_stream(tmp_path, oracle_lib):
    """RTthEtaAreaMap.bin (IntegratorFitPeaksGPUStream.cu:817-899, 2443-2462):
    the REtaMap rows with the area summed over on-detector entries only
    (the oracle's in-bounds area), five float64 rows of n_r * n_eta."""
    from helpers import panel_shifts_for, to_oracle
    from midas_amd.outputs import ResultFiles
    P = small_cases()["panels"]
    print("[outputs] RTthEtaAreaMap", flush=True)
    g = ma.build_map(P, panel_shifts=panel_shifts_for("panels"))
    om = oracle_lib.build_map(to_oracle(P), panel_shifts=panel_shifts_for("panels"))
    full = g.reta_map(P)
    inb = g.reta_map(P, in_bounds=True)
    np.testing.assert_array_equal(inb[[0, 1, 2, 4]], full[[0, 1, 2, 4]])
    want = oracle_lib.area_inbounds(om, P.NrPixelsY, P.NrPixelsZ)
    np.testing.assert_allclose(inb[3], want, rtol=1e-6, atol=1e-9)
    assert np.all(inb[3] <= full[3] + 1e-9)
    with ResultFiles(tmp_path, "s.zip", P, files={"stream"}, reta_stream=inb):
        pass
    raw = np.fromfile(tmp_path / "RTthEtaAreaMap.bin", "<f8")
    np.testing.assert_array_equal(raw, inb.reshape(-1))
