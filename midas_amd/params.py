"""MIDAS parameter files, with the two parsers of the caking path.

``parse_mapper_params`` follows DetectorMapper.c:116-375: every line is
tested against each key as a *prefix* ("tx " ...), values are read like
``sscanf("%s %lf")`` and left unchanged when they do not parse, ``px``
sets pxY and pxZ, ``NrPixels`` sets both dimensions, ``MaxRingRad`` is an
alias of RhoD, ``ImTransOpt`` lines append, ``PanelGapsY/Z`` read
NPanels-1 integers.  Defaults are DetectorMapper.c:84-114.

``parse_integrator_params`` follows the exact-key parser of
IntegratorZarrOMP.c:391-526 (defaults :371-389) for the integration and
fit options: Normalize, GradientCorrection, PeakLocation (enables the fit),
AutoDetectPeaks (implies DoPeakFit), FitROIPadding, ...

``IntegrationParams`` carries the union, named like the reference Python
package (packages/midas_integrate/midas_integrate/params.py:15-200).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from pathlib import Path
from typing import List

import numpy as np


@dataclass
class IntegrationParams:
    NrPixelsY: int = 2048
    NrPixelsZ: int = 2048
    pxY: float = 200.0
    pxZ: float = 200.0
    Lsd: float = 1e6
    BC_y: float = 1024.0
    BC_z: float = 1024.0
    RhoD: float = 2e5
    tx: float = 0.0
    ty: float = 0.0
    tz: float = 0.0
    p: List[float] = field(default_factory=lambda: [0.0] * 15)
    RMin: float = 10.0
    RMax: float = 1524.0
    RBinSize: float = 0.25
    EtaMin: float = -180.0
    EtaMax: float = 180.0
    EtaBinSize: float = 5.0
    QMin: float = 0.0
    QMax: float = 0.0
    QBinSize: float = 0.0
    Wavelength: float = 0.0
    NrTransOpt: int = 0
    TransOpt: List[int] = field(default_factory=list)
    SubPixelLevel: int = 1
    SubPixelCardinalWidth: float = 10.0
    Parallax: float = 0.0
    SolidAngleCorrection: int = 0
    PolarizationCorrection: int = 0
    PolarizationFraction: float = 0.99
    NPanelsY: int = 0
    NPanelsZ: int = 0
    PanelSizeY: int = 0
    PanelSizeZ: int = 0
    PanelGapsY: List[int] = field(default_factory=list)
    PanelGapsZ: List[int] = field(default_factory=list)
    PanelShiftsFile: str = ""
    MaskFile: str = ""
    DistortionFile: str = ""
    ResidualCorrectionMap: str = ""
    # integrator options (IntegratorZarrOMP.c:371-389)
    Normalize: int = 1
    GradientCorrection: int = 0
    Lam: float = 0.172978
    SkipFrame: int = 0
    OmegaStart: float = 0.0
    OmegaStep: float = 0.0
    OmegaSumFrames: int = 1
    SumImages: int = 0
    SaveIndividualFrames: int = 1
    DoPeakFit: int = 0
    FitROIPadding: int = 20
    FitROIAuto: int = 0          # GPU stream: ROI from the widest FWHM estimate
    DoSmoothing: int = 0         # GPU stream flags, read and kept
    MultiplePeaks: int = 0
    AutoDetectPeaks: int = 0
    SNIPIterations: int = 50
    DoubletSeparation: float = 0.0
    PeakLocations: List[float] = field(default_factory=list)
    Polariz: float = 0.99
    SHpL: float = 0.002
    U: float = 1.163
    V: float = -0.126
    W: float = 0.063
    X: float = 0.0
    Y: float = 0.0
    Z: float = 0.0

    @property
    def q_mode(self) -> bool:
        return self.QBinSize > 0 and self.Wavelength > 0

    @property
    def n_eta_bins(self) -> int:
        return int(math.ceil((self.EtaMax - self.EtaMin) / self.EtaBinSize))

    @property
    def n_r_bins(self) -> int:
        if self.q_mode:
            return int(math.ceil((self.QMax - self.QMin) / self.QBinSize))
        return int(math.ceil((self.RMax - self.RMin) / self.RBinSize))


def _sscanf_double(line: str, idx: int = 1):
    toks = line.split()
    if len(toks) <= idx:
        return None
    try:
        return float(_c_number_prefix(toks[idx], float_ok=True))
    except ValueError:
        return None


def _sscanf_int(line: str, idx: int = 1):
    toks = line.split()
    if len(toks) <= idx:
        return None
    try:
        return int(_c_number_prefix(toks[idx], float_ok=False))
    except ValueError:
        return None


def _c_number_prefix(tok: str, float_ok: bool) -> str:
    """The longest prefix sscanf would convert ("%lf" / "%d")."""
    import re
    pat = (r"[+-]?(?:inf(?:inity)?|nan|(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)"
           if float_ok else r"[+-]?\d+")
    m = re.match(pat, tok, re.IGNORECASE)
    if not m:
        raise ValueError(tok)
    return m.group(0)


def _strtol_list(rest: str, n: int) -> List[int]:
    out, toks = [], rest.split()
    for i in range(n):
        try:
            out.append(int(_c_number_prefix(toks[i], False)) if i < len(toks) else 0)
        except ValueError:
            out.append(0)
    return out


_MAPPER_DOUBLES = {
    "tx ": "tx", "ty ": "ty", "tz ": "tz", "pxY ": "pxY", "pxZ ": "pxZ",
    "Lsd ": "Lsd", "RhoD ": "RhoD", "MaxRingRad ": "RhoD",
    "EtaBinSize ": "EtaBinSize", "RBinSize ": "RBinSize", "RMax ": "RMax",
    "RMin ": "RMin", "EtaMax ": "EtaMax", "EtaMin ": "EtaMin",
    "QBinSize ": "QBinSize", "SubPixelCardinalWidth ": "SubPixelCardinalWidth",
    "QMin ": "QMin", "QMax ": "QMax", "Wavelength ": "Wavelength",
    "Parallax ": "Parallax", "PolarizationFraction ": "PolarizationFraction",
}
_MAPPER_INTS = {
    "SubPixelLevel ": "SubPixelLevel", "NrPixelsY ": "NrPixelsY",
    "NrPixelsZ ": "NrPixelsZ", "SolidAngleCorrection ": "SolidAngleCorrection",
    "PolarizationCorrection ": "PolarizationCorrection",
}


def parse_mapper_params(path: str | Path, params: IntegrationParams | None = None
                        ) -> IntegrationParams:
    """DetectorMapper.c:116-375 semantics (prefix keys, sscanf values)."""
    P = params if params is not None else IntegrationParams()
    with open(path, "r", errors="replace") as f:
        for line in f:
            def st(k):
                return line.startswith(k)
            if st("NPanelsY "):
                v = _sscanf_int(line); P.NPanelsY = v if v is not None else P.NPanelsY
                continue
            if st("NPanelsZ "):
                v = _sscanf_int(line); P.NPanelsZ = v if v is not None else P.NPanelsZ
                continue
            if st("PanelSizeY "):
                v = _sscanf_int(line); P.PanelSizeY = v if v is not None else P.PanelSizeY
                continue
            if st("PanelSizeZ "):
                v = _sscanf_int(line); P.PanelSizeZ = v if v is not None else P.PanelSizeZ
                continue
            if st("PanelShiftsFile "):
                t = line.split()
                if len(t) > 1:
                    P.PanelShiftsFile = t[1]
                continue
            if st("PanelGapsY "):
                if P.NPanelsY > 1:
                    P.PanelGapsY = _strtol_list(line[len("PanelGapsY "):], P.NPanelsY - 1)
                continue
            if st("PanelGapsZ "):
                if P.NPanelsZ > 1:
                    P.PanelGapsZ = _strtol_list(line[len("PanelGapsZ "):], P.NPanelsZ - 1)
                continue
            if st("DistortionOrder "):
                continue
            if st("ImTransOpt "):
                v = _sscanf_int(line)
                P.TransOpt.append(v if v is not None else 0)
                P.NrTransOpt = len(P.TransOpt)
                continue
            for k, attr in _MAPPER_DOUBLES.items():
                if st(k):
                    v = _sscanf_double(line)
                    if v is not None:
                        setattr(P, attr, v)
            for k, attr in _MAPPER_INTS.items():
                if st(k):
                    v = _sscanf_int(line)
                    if v is not None:
                        setattr(P, attr, v)
            for i in range(15):
                if st(f"p{i} "):
                    v = _sscanf_double(line)
                    if v is not None:
                        P.p[i] = v
            if st("px "):
                v = _sscanf_double(line)
                if v is not None:
                    P.pxY = v
                P.pxZ = P.pxY
            if st("BC "):
                t = line.split()
                vals = []
                for tok in t[1:3]:
                    try:
                        vals.append(float(_c_number_prefix(tok, True)))
                    except ValueError:
                        break
                if len(vals) > 0:
                    P.BC_y = vals[0]
                if len(vals) > 1:
                    P.BC_z = vals[1]
            if st("NrPixels "):
                v = _sscanf_int(line)
                if v is not None:
                    P.NrPixelsY = P.NrPixelsZ = v
            if st("DistortionFile "):
                t = line.split()
                if len(t) > 1:
                    P.DistortionFile = t[1]
            if st("MaskFile "):
                t = line.split()
                if len(t) > 1:
                    P.MaskFile = t[1]
            if st("ResidualCorrectionMap "):
                t = line.split()
                if len(t) > 1:
                    P.ResidualCorrectionMap = t[1]
    return P


def parse_integrator_params(path: str | Path, params: IntegrationParams | None = None
                            ) -> IntegrationParams:
    """IntegratorZarrOMP.c:391-526 semantics (exact keys) for the options
    the integrator reads; geometry keys it shares are read too."""
    P = params if params is not None else IntegrationParams()
    with open(path, "r", errors="replace") as f:
        for line in f:
            if not line or line[0] in "#\n\r":
                continue
            t = line.split(None, 1)
            if len(t) < 2:
                continue
            key, rest = t[0], t[1].strip()

            def fl():
                try:
                    return float(_c_number_prefix(rest.split()[0], True))
                except (ValueError, IndexError):
                    return 0.0

            def it():
                try:
                    return int(_c_number_prefix(rest.split()[0], False))
                except (ValueError, IndexError):
                    return 0

            if key == "RMin": P.RMin = fl()
            elif key == "RMax": P.RMax = fl()
            elif key == "RBinSize": P.RBinSize = fl()
            elif key == "BC":
                v = rest.split()
                if len(v) >= 2:
                    P.BC_y, P.BC_z = float(v[0]), float(v[1])
            elif key in ("EtaMin", "MinEta"): P.EtaMin = fl()
            elif key == "EtaMax": P.EtaMax = fl()
            elif key == "EtaBinSize": P.EtaBinSize = fl()
            elif key in ("Lsd", "DetDist"): P.Lsd = fl()
            elif key in ("px", "PixelSize", "YPixelSize"): P.pxY = P.pxZ = fl()
            elif key == "Wavelength": P.Lam = fl()
            elif key == "QBinSize": P.QBinSize = fl()
            elif key == "QMin": P.QMin = fl()
            elif key == "QMax": P.QMax = fl()
            elif key == "NrPixels": P.NrPixelsY = P.NrPixelsZ = it()
            elif key == "NrPixelsY": P.NrPixelsY = it()
            elif key == "NrPixelsZ": P.NrPixelsZ = it()
            elif key == "ImTransOpt":
                P.TransOpt = [int(x) for x in rest.split()[:10]]
                P.NrTransOpt = len(P.TransOpt)
            elif key == "Normalize": P.Normalize = it()
            elif key == "GradientCorrection": P.GradientCorrection = it()
            elif key == "SkipFrame": P.SkipFrame = it()
            elif key in ("OmegaStart", "OmegaFirstFile"): P.OmegaStart = fl()
            elif key == "OmegaStep": P.OmegaStep = fl()
            elif key == "OmegaSumFrames": P.OmegaSumFrames = it()
            elif key == "SumImages": P.SumImages = it()
            elif key == "SaveIndividualFrames": P.SaveIndividualFrames = it()
            elif key in ("MaskFile", "MaskFN"): P.MaskFile = rest.split()[0]
            elif key == "Polariz": P.Polariz = fl()
            elif key == "SHpL": P.SHpL = fl()
            elif key in ("U", "V", "W", "X", "Y", "Z"): setattr(P, key, fl())
            elif key == "DoPeakFit": P.DoPeakFit = it()
            elif key == "FitROIPadding": P.FitROIPadding = it()
            elif key == "FitROIAuto": P.FitROIAuto = it()
            elif key == "DoSmoothing": P.DoSmoothing = it()
            elif key == "MultiplePeaks": P.MultiplePeaks = it()
            elif key == "AutoDetectPeaks":
                P.AutoDetectPeaks = it()
                P.DoPeakFit = 1
            elif key == "SNIPIterations": P.SNIPIterations = it()
            elif key == "DoubletSeparation": P.DoubletSeparation = fl()
            elif key == "PeakLocation":
                if len(P.PeakLocations) < 200:
                    P.PeakLocations.append(fl())
                    P.DoPeakFit = 1
    if P.OmegaSumFrames == 0:
        P.OmegaSumFrames = 1
    return P


def parse_params(path: str | Path) -> IntegrationParams:
    """Both parsers over one file: geometry as DetectorMapper reads it, then
    the integrator's options (the two binaries read the same file)."""
    P = parse_mapper_params(path)
    lam = P.Wavelength
    keep = (P.pxY, P.pxZ, P.TransOpt[:], P.NrTransOpt)
    parse_integrator_params(path, P)
    # Geometry stays as the mapper read it (the integrator's px alias and
    # single-line ImTransOpt differ from DetectorMapper's).
    P.pxY, P.pxZ, P.TransOpt, P.NrTransOpt = keep[0], keep[1], keep[2], keep[3]
    P.Wavelength = lam
    return P


def bin_edges(P: IntegrationParams):
    """(RLo, RHi, EtaLo, EtaHi) in FP64 as DetectorMapper.c:712-755 computes
    them (Q-mode: R(Q) = Lsd/px * tan(2 asin(Q lambda / 4 pi)), evaluated
    with the C library's tan/asin through ``math``)."""
    nR, nE = P.n_r_bins, P.n_eta_bins
    e = np.arange(nE, dtype=np.float64)
    ELo = P.EtaBinSize * e + P.EtaMin
    EHi = P.EtaBinSize * (e + 1) + P.EtaMin
    if P.q_mode:
        f = lambda q: (P.Lsd / P.pxY) * math.tan(2.0 * math.asin(q * P.Wavelength / (4.0 * math.pi)))
        RLo = np.array([f(P.QMin + P.QBinSize * r) for r in range(nR)])
        RHi = np.array([f(P.QMin + P.QBinSize * (r + 1)) for r in range(nR)])
    else:
        r = np.arange(nR, dtype=np.float64)
        RLo = P.RBinSize * r + P.RMin
        RHi = P.RBinSize * (r + 1) + P.RMin
    return RLo, RHi, ELo, EHi


def effective_r_limits(P: IntegrationParams):
    """RMin/RMax after the Q-mode rewrite (DetectorMapper.c:716-719)."""
    if not P.q_mode:
        return P.RMin, P.RMax
    f = lambda q: (P.Lsd / P.pxY) * math.tan(2.0 * math.asin(q * P.Wavelength / (4.0 * math.pi)))
    rmin = f(P.QMin) if P.QMin > 0 else P.RMin
    rmax = f(P.QMax) if P.QMax > 0 else P.RMax
    return rmin, rmax
