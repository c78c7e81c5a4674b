"""Parameter-file parsing: the product parser against the oracle's C
restatement of DetectorMapper.c:116-375, and the integrator's exact-key
options (IntegratorZarrOMP.c:391-526)."""
import numpy as np
import pytest

import midas_amd as ma
from midas_amd.params import parse_integrator_params, parse_mapper_params

TRICKY = """\
# comment tx 99
NrPixels 1024
NrPixelsZ 1000
px 172
pxZ 150
BC 512.5 511.25
Lsd 1000000
Lsd abc
tx 0.1
txx 5
ty -0.2
tz 0.3
p0 1e-4
p1 2e-4
p10 5
p14 3
p6 12.5
RhoD 200000
MaxRingRad 180000
RMin 10
RMax 500.5
RBinSize 0.25
EtaMin -185
EtaMax 175
EtaBinSize 5
ImTransOpt 1
ImTransOpt 2
SubPixelLevel 2
SubPixelCardinalWidth 7.5
Wavelength 0.172979
QBinSize 0.001
QMin 0.5
QMax 12
Parallax 100
SolidAngleCorrection 1
PolarizationCorrection 1
PolarizationFraction 0.95
NPanelsY 3
NPanelsZ 2
PanelSizeY 300
PanelSizeZ 400
PanelGapsY 5 7
PanelGapsZ 11
DistortionOrder 6
MaskFile mask.tif
"""


def test_mapper_parser_matches_oracle(tmp_path, oracle_lib):
    f = tmp_path / "params.txt"
    f.write_text(TRICKY)
    P = parse_mapper_params(f)
    O = oracle_lib.parse_params(f)
    assert (P.NrPixelsY, P.NrPixelsZ) == (O.NrPixelsY, O.NrPixelsZ) == (1024, 1000)
    assert (P.pxY, P.pxZ) == (O.pxY, O.pxZ) == (172.0, 150.0)
    assert (P.BC_y, P.BC_z) == (O.yCen, O.zCen)
    for a, b in [("Lsd", "Lsd"), ("RhoD", "RhoD"), ("tx", "tx"), ("ty", "ty"), ("tz", "tz"),
                 ("RMin", "RMin"), ("RMax", "RMax"), ("RBinSize", "RBinSize"),
                 ("EtaMin", "EtaMin"), ("EtaMax", "EtaMax"), ("EtaBinSize", "EtaBinSize"),
                 ("QBinSize", "QBinSize"), ("QMin", "QMin"), ("QMax", "QMax"),
                 ("Wavelength", "Wavelength"), ("SubPixelLevel", "SubPixelLevel"),
                 ("SubPixelCardinalWidth", "SubPixelCardinalWidth"), ("Parallax", "Parallax"),
                 ("SolidAngleCorrection", "SolidAngleCorrection"),
                 ("PolarizationCorrection", "PolarizationCorrection"),
                 ("PolarizationFraction", "PolarizationFraction"), ("NPanelsY", "NPanelsY"),
                 ("NPanelsZ", "NPanelsZ"), ("PanelSizeY", "PanelSizeY"),
                 ("PanelSizeZ", "PanelSizeZ")]:
        assert getattr(P, a) == getattr(O, b), a
    assert list(P.p) == list(O.p)
    assert P.p[1] == 2e-4 and P.p[10] == 5.0  # "p1 " does not catch "p10"
    assert P.RhoD == 180000.0                 # MaxRingRad is an alias of RhoD
    assert P.Lsd == 1e6                       # an unparsable value leaves it
    assert P.tx == 0.1                        # "txx" and the comment are ignored
    assert P.TransOpt == list(O.TransOpt)[:O.NrTransOpt] == [1, 2]
    assert P.PanelGapsY == list(O.PanelGapsY)[:2] == [5, 7]
    assert P.PanelGapsZ == list(O.PanelGapsZ)[:1] == [11]
    assert P.MaskFile == "mask.tif"
    assert ma.canonical_string(P) == oracle_lib.header_string(O)


def test_mapper_defaults_match_oracle(tmp_path, oracle_lib):
    f = tmp_path / "empty.txt"
    f.write_text("\n")
    P = parse_mapper_params(f)
    O = oracle_lib.parse_params(f)
    assert (P.NrPixelsY, P.pxY, P.BC_y, P.Lsd, P.RhoD, P.RMax, P.RMin, P.RBinSize,
            P.EtaBinSize, P.SubPixelCardinalWidth, P.PolarizationFraction) == (
        O.NrPixelsY, O.pxY, O.yCen, O.Lsd, O.RhoD, O.RMax, O.RMin, O.RBinSize,
        O.EtaBinSize, O.SubPixelCardinalWidth, O.PolarizationFraction)


def test_integrator_options(tmp_path):
    f = tmp_path / "p.txt"
    f.write_text("Normalize 0\nGradientCorrection 1\nPeakLocation 120.5\nPeakLocation 200\n"
                 "FitROIPadding 15\nImTransOpt 1 2\nOmegaSumFrames 0\nWavelength 0.2\n"
                 "NormalizeX 1\nMinEta -170\n")
    P = parse_integrator_params(f)
    assert P.Normalize == 0 and P.GradientCorrection == 1
    assert P.PeakLocations == [120.5, 200.0] and P.DoPeakFit == 1
    assert P.FitROIPadding == 15 and P.TransOpt == [1, 2]
    assert P.OmegaSumFrames == 1  # 0 is read as 1 (IntegratorZarrOMP.c:527-528)
    assert P.Lam == 0.2 and P.EtaMin == -170.0
    g = tmp_path / "q.txt"
    g.write_text("AutoDetectPeaks 5\n")
    assert parse_integrator_params(g).DoPeakFit == 1


def test_parse_params_keeps_mapper_geometry(tmp_path):
    f = tmp_path / "p.txt"
    f.write_text("px 172\nImTransOpt 1\nImTransOpt 2\nWavelength 0.17\nQBinSize 0.01\n"
                 "QMin 1\nQMax 5\n")
    P = ma.parse_params(f)
    assert P.TransOpt == [1, 2] and P.pxY == 172.0 and P.Wavelength == 0.17 and P.q_mode


def test_bin_counts(tmp_path):
    f = tmp_path / "p.txt"
    f.write_text("RMin 10\nRMax 500\nRBinSize 0.25\nEtaBinSize 5\n")
    P = parse_mapper_params(f)
    assert P.n_r_bins == 1960 and P.n_eta_bins == 72
