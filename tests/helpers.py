"""Shared cases and comparisons for the parity tests."""
from __future__ import annotations

import numpy as np

from midas_amd.params import IntegrationParams


def to_oracle(P: IntegrationParams):
    """The oracle's parameter struct for product parameters."""
    from oracle import oracle as orc
    kw = dict(NrPixelsY=P.NrPixelsY, NrPixelsZ=P.NrPixelsZ, pxY=P.pxY, pxZ=P.pxZ,
              yCen=P.BC_y, zCen=P.BC_z, Lsd=P.Lsd, RhoD=P.RhoD, tx=P.tx, ty=P.ty, tz=P.tz,
              p=list(P.p), RMin=P.RMin, RMax=P.RMax, RBinSize=P.RBinSize, EtaMin=P.EtaMin,
              EtaMax=P.EtaMax, EtaBinSize=P.EtaBinSize, QMin=P.QMin, QMax=P.QMax,
              QBinSize=P.QBinSize, Wavelength=P.Wavelength, TransOpt=list(P.TransOpt),
              SubPixelLevel=P.SubPixelLevel, SubPixelCardinalWidth=P.SubPixelCardinalWidth,
              Parallax=P.Parallax, SolidAngleCorrection=P.SolidAngleCorrection,
              PolarizationCorrection=P.PolarizationCorrection,
              PolarizationFraction=P.PolarizationFraction, NPanelsY=P.NPanelsY,
              NPanelsZ=P.NPanelsZ, PanelSizeY=P.PanelSizeY, PanelSizeZ=P.PanelSizeZ,
              PanelGapsY=list(P.PanelGapsY), PanelGapsZ=list(P.PanelGapsZ))
    return orc.make_params(**kw)


def small_cases() -> dict[str, IntegrationParams]:
    """Small geometries exercising every branch of the mapper."""
    c = {}
    c["flat"] = IntegrationParams(NrPixelsY=96, NrPixelsZ=80, pxY=200.0, pxZ=200.0, Lsd=1e6,
                                  BC_y=47.3, BC_z=39.6, RMin=2.0, RMax=60.0, RBinSize=0.5,
                                  EtaBinSize=10.0)
    c["centred"] = IntegrationParams(NrPixelsY=64, NrPixelsZ=64, pxY=200.0, pxZ=200.0, Lsd=1e6,
                                     BC_y=32.0, BC_z=32.0, RMin=1.0, RMax=40.0,
                                     RBinSize=0.25, EtaBinSize=5.0)
    c["tilt"] = IntegrationParams(NrPixelsY=80, NrPixelsZ=64, pxY=150.0, pxZ=150.0, Lsd=5e5,
                                  BC_y=38.2, BC_z=30.9, tx=0.2, ty=0.4, tz=-0.3,
                                  p=[1e-4, 2e-4, -1e-4, 10.0, 3e-4, 2e-5, 5.0, 1e-4, 30.0,
                                     -2e-4, 10.0, 1e-4, 20.0, -1e-4, 45.0],
                                  RhoD=5e4, RMin=3.0, RMax=45.0, RBinSize=0.25, EtaBinSize=5.0,
                                  Parallax=100.0, SolidAngleCorrection=1,
                                  PolarizationCorrection=1)
    c["subpix"] = IntegrationParams(NrPixelsY=48, NrPixelsZ=48, pxY=200.0, pxZ=200.0, Lsd=1e6,
                                    BC_y=24.0, BC_z=24.0, RMin=2.0, RMax=30.0, RBinSize=0.25,
                                    EtaBinSize=5.0, SubPixelLevel=2, TransOpt=[1, 2])
    c["qmode"] = IntegrationParams(NrPixelsY=96, NrPixelsZ=96, pxY=200.0, pxZ=200.0, Lsd=3e5,
                                   BC_y=48.5, BC_z=47.5, QBinSize=0.05, QMin=0.5, QMax=6.0,
                                   Wavelength=0.172979, EtaBinSize=10.0, EtaMin=-185.0,
                                   EtaMax=175.0)
    c["panels"] = IntegrationParams(NrPixelsY=100, NrPixelsZ=90, pxY=172.0, pxZ=172.0,
                                    Lsd=4e5, BC_y=51.1, BC_z=44.7, RMin=2.0, RMax=70.0,
                                    RBinSize=0.5, EtaBinSize=10.0, NPanelsY=2, NPanelsZ=2,
                                    PanelSizeY=48, PanelSizeZ=42, PanelGapsY=[4],
                                    PanelGapsZ=[6])
    return c


def panel_shifts_for(name: str):
    if name == "panels":
        return [[0, 0.3, -0.2, 0.05, 20.0, 1e-5], [3, -0.25, 0.1, -0.1, -15.0, 0.0]]
    return None


def assert_maps_equal(gpu_pm, orc_map, frac_atol=1e-6):
    """Bin membership bit-exact (per-bin counts, offsets, mask flags, the
    pixel y/z of every record in the same order); the FP64 overlap areas to
    frac_atol in pixel-area units.  The GPU's sin/cos/atan2 may round the
    last bit differently from glibc: the shoelace sum of a clipped polygon
    turns that into ~1e-12 absolute differences of the area, and where a
    vertex sits exactly on a bin's arc (beam centre on a pixel corner) the
    1e-6 relative arc test of the area routine can take the other branch,
    a chord/arc difference below 1e-6 of a pixel area."""
    np.testing.assert_array_equal(gpu_pm.counts, orc_map.counts, "per-bin counts")
    np.testing.assert_array_equal(gpu_pm.offsets, orc_map.offsets, "per-bin offsets")
    np.testing.assert_array_equal(gpu_pm.mask_flags, orc_map.flags, "mask flags")
    g, o = gpu_pm.pxList, orc_map.entries
    np.testing.assert_array_equal(g["y"], o["y"], "pixel y")
    np.testing.assert_array_equal(g["z"], o["z"], "pixel z")
    np.testing.assert_allclose(g["frac"], o["frac"], rtol=1e-9, atol=frac_atol)
    np.testing.assert_allclose(g["areaWeight"], o["areaWeight"], rtol=1e-6, atol=frac_atol)
    np.testing.assert_allclose(g["deltaR"], o["deltaR"], rtol=0, atol=1e-5)
    exact = float(np.mean(g["frac"] == o["frac"]))
    loose = int(np.sum(np.abs(g["frac"] - o["frac"]) > 1e-10))
    print(f"[map] {g.size} records, frac bit-identical for {exact:.4f} of them, "
          f"{loose} beyond 1e-10", flush=True)
    # A systematic error (a dropped term, a wrong corner) moves most records;
    # last-bit transcendental differences move a small fraction of them.
    assert loose <= max(10, g.size // 100)


def rel_err(gpu, ref, scale):
    """|gpu - ref| / scale with NaNs required at the same places."""
    gpu = np.asarray(gpu, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    assert np.array_equal(np.isnan(gpu), np.isnan(ref)), "NaN (masked-bin) positions differ"
    ok = ~np.isnan(ref)
    s = np.maximum(np.abs(np.asarray(scale, dtype=np.float64)), 1e-30)
    e = np.zeros_like(ref)
    e[ok] = np.abs(gpu[ok] - ref[ok]) / s[ok]
    return e
