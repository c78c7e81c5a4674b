"""Frames and dark read from a Zarr zip (midas_amd.zarr_zip) integrate on
the GPU to exactly what the same arrays give when passed directly."""
import numpy as np
import pytest

import midas_amd as ma
from midas_amd.zarr_zip import ZarrZip
from helpers import small_cases
from test_zarr_zip import write_zip

pytestmark = pytest.mark.gpu


def test_zip_frames_integrate_like_direct(tmp_path):
    P = small_cases()["flat"]
    rng = np.random.default_rng(8)
    frames = rng.poisson(200.0, (9, P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    darks = rng.poisson(20.0, (3, P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    p = tmp_path / "flat.MIDAS.zip"
    write_zip(p, frames, darks, {"SkipFrame": np.int32(1), "RMin": P.RMin, "RMax": P.RMax,
                                 "RBinSize": P.RBinSize, "EtaBinSize": P.EtaBinSize,
                                 "Lsd": P.Lsd, "PixelSize": P.pxY},
              {"start": np.array([0.0]), "step": np.array([0.5])})
    print("[zarr] 9 frames through the zip", flush=True)
    with ZarrZip(p) as z:
        Pz = z.integration_params(small_cases()["flat"])
        g = ma.build_map(Pz)
        g.set_dark(z.dark_mean()[None])
        cz, lz = ma.integrate_batch(z.frames(), g, mode="bilinear", mask_nan=True)
    g2 = ma.build_map(P)
    g2.set_dark(((darks[1].astype(np.float64) / 2) + darks[2].astype(np.float64) / 2)[None])
    cd, ld = ma.integrate_batch(frames[1:], g2, mode="bilinear", mask_nan=True)
    np.testing.assert_array_equal(cz, cd)
    np.testing.assert_array_equal(lz, ld)
