"""Seeded synthetic detector data for the caking workloads.

Configurations (all distances in micrometres, pixels and degrees):
  plain_1024  : 1024x1024 uint16, px 200, Lsd 1e6, BC (512, 512), no tilt,
                lambda 0.172979; R 10-500 by 0.25, eta -180..180 by 5;
                8 Gaussian rings at R 60..480 (sigma 1.5, amplitude 5000,
                x (1 + 0.3 cos 2 eta)); dark = 100 + Poisson(5);
                frame = dark + Poisson(50 + rings); 0.1 % hot pixels masked.
  batch_2048  : 2048x2048 uint16, same generator with 12 rings up to 1000 px,
                BC (1024, 1024), R 10-1000 by 0.25, eta by 1 deg, solid-angle
                and polarisation (0.99) corrections.
  pilatus     : 1679x1475 int32 (Pilatus 6x8 modules of 243x195, gaps Y
                1 7 1 7 1, Z 17 as FF_HEDM/Example/Calibration/parameters.txt),
                gap pixels masked.
  qmode_2880  : 2880x2880 float32, px 150, Lsd 900 mm, BC (1440.3, 1439.6),
                tilts ty 0.2 deg, tz -0.15 deg, parallax 200 um, lambda
                0.172979; Q 0.5-12 A^-1 by 0.001 (11500 bins), eta by 0.5 deg
                (720 bins), SubPixelLevel 2, GradientCorrection (about 9 map
                entries per pixel); 16 Gaussian-blurred rings at Q 1.0-8.0.
  qmode_crop  : a 320x288 crop of qmode_2880 (pixels y 1900-2219, z
                1300-1587 of the full detector) with the same binning: the
                case the reference mapper maps in seconds.
"""
from __future__ import annotations

import math

import numpy as np

from .params import IntegrationParams


def plain_1024() -> IntegrationParams:
    return IntegrationParams(NrPixelsY=1024, NrPixelsZ=1024, pxY=200.0, pxZ=200.0, Lsd=1e6,
                             BC_y=512.0, BC_z=512.0, RMin=10.0, RMax=500.0, RBinSize=0.25,
                             EtaMin=-180.0, EtaMax=180.0, EtaBinSize=5.0, Lam=0.172979)


def batch_2048() -> IntegrationParams:
    return IntegrationParams(NrPixelsY=2048, NrPixelsZ=2048, pxY=200.0, pxZ=200.0, Lsd=1e6,
                             BC_y=1024.0, BC_z=1024.0, RMin=10.0, RMax=1000.0, RBinSize=0.25,
                             EtaMin=-180.0, EtaMax=180.0, EtaBinSize=1.0, Lam=0.172979,
                             SolidAngleCorrection=1, PolarizationCorrection=1,
                             PolarizationFraction=0.99)


def pilatus() -> IntegrationParams:
    return IntegrationParams(NrPixelsY=1475, NrPixelsZ=1679, pxY=172.0, pxZ=172.0, Lsd=650000.0,
                             BC_y=700.3, BC_z=840.6, RMin=10.0, RMax=800.0, RBinSize=0.25,
                             EtaMin=-180.0, EtaMax=180.0, EtaBinSize=5.0, Lam=0.172979,
                             NPanelsY=6, NPanelsZ=8, PanelSizeY=243, PanelSizeZ=195,
                             PanelGapsY=[1, 7, 1, 7, 1], PanelGapsZ=[17] * 7)


def qmode_2880() -> IntegrationParams:
    return IntegrationParams(NrPixelsY=2880, NrPixelsZ=2880, pxY=150.0, pxZ=150.0, Lsd=900000.0,
                             BC_y=1440.3, BC_z=1439.6, ty=0.2, tz=-0.15, Parallax=200.0,
                             QMin=0.5, QMax=12.0, QBinSize=0.001, Wavelength=0.172979,
                             Lam=0.172979, EtaMin=-180.0, EtaMax=180.0, EtaBinSize=0.5,
                             SubPixelLevel=2, GradientCorrection=1, RMin=1.0, RMax=2100.0)


QMODE_CROP = (1900, 1300, 320, 288)  # y0, z0, width, height in the full detector


def qmode_crop() -> IntegrationParams:
    P = qmode_2880()
    y0, z0, w, h = QMODE_CROP
    P.NrPixelsY, P.NrPixelsZ = w, h
    P.BC_y, P.BC_z = P.BC_y - y0, P.BC_z - z0
    return P


def ring_q() -> np.ndarray:
    """The 16 ring positions of the Q-mode workload (A^-1)."""
    return np.linspace(1.0, 8.0, 16)


def qmode_ring_image(P: IntegrationParams, sigma: float = 2.0, amplitude: float = 3000.0):
    """Noise-free Gaussian-blurred rings at the Q positions of ring_q(),
    placed at R = Lsd tan(2 theta) / px on a flat detector (the tilts move
    them by a fraction of a pixel), x (1 + 0.3 cos 2 eta)."""
    z, y = np.mgrid[0:P.NrPixelsZ, 0:P.NrPixelsY].astype(np.float64)
    dy, dz = y - P.BC_y, z - P.BC_z
    R = np.hypot(dy, dz)
    eta = np.arctan2(dy, dz)
    img = np.zeros_like(R)
    for q in ring_q():
        r0 = P.Lsd * math.tan(2.0 * math.asin(q * P.Wavelength / (4.0 * math.pi))) / P.pxY
        img += amplitude * np.exp(-0.5 * ((R - r0) / sigma) ** 2)
    return img * (1.0 + 0.3 * np.cos(2.0 * eta))


def qmode_frames(P: IntegrationParams, n: int, seed: int = 0):
    """(frames [n, NZ, NY] float32, dark [NZ, NY] float32): dark = 100 +
    N(0, 2); frame = dark + 40 + rings + N(0, sqrt(40 + rings)), seeded per
    frame."""
    rng = np.random.default_rng(seed)
    dark = (100.0 + rng.normal(0.0, 2.0, (P.NrPixelsZ, P.NrPixelsY))).astype(np.float32)
    base = 40.0 + qmode_ring_image(P)
    sd = np.sqrt(base)
    out = np.empty((n, P.NrPixelsZ, P.NrPixelsY), dtype=np.float32)
    for i in range(n):
        fr = np.random.default_rng(seed * 1000003 + i + 1)
        out[i] = dark + base + fr.normal(0.0, 1.0, base.shape) * sd
    return out, dark


def torch_qmode_frames(P: IntegrationParams, n: int, seed: int = 0, device="cuda"):
    """The same generator on the GPU with torch (for large batches)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    base = torch.as_tensor(40.0 + qmode_ring_image(P), device=device, dtype=torch.float32)
    sd = base.sqrt()
    dark = 100.0 + 2.0 * torch.randn(base.shape, generator=g, device=device)
    out = torch.empty((n,) + tuple(base.shape), dtype=torch.float32, device=device)
    for i in range(n):
        out[i] = dark + base + torch.randn(base.shape, generator=g, device=device) * sd
    return out, dark


def pilatus_gap_mask(P: IntegrationParams) -> np.ndarray:
    """1.0 on pixels outside every module (the gaps)."""
    m = np.ones((P.NrPixelsZ, P.NrPixelsY))
    y0 = 0
    for a in range(P.NPanelsY):
        z0 = 0
        for b in range(P.NPanelsZ):
            m[z0:z0 + P.PanelSizeZ, y0:y0 + P.PanelSizeY] = 0.0
            if b < P.NPanelsZ - 1:
                z0 += P.PanelSizeZ + P.PanelGapsZ[b]
        if a < P.NPanelsY - 1:
            y0 += P.PanelSizeY + P.PanelGapsY[a]
    return m


def ring_radii(P: IntegrationParams, n_rings: int, r_max: float | None = None) -> np.ndarray:
    lo = 60.0
    hi = r_max if r_max is not None else (480.0 if P.NrPixelsY <= 1024 else 1000.0)
    if n_rings == 8 and hi == 480.0:
        return np.linspace(60.0, 480.0, 8)
    return np.linspace(lo, hi * 0.96, n_rings)


def ring_image(P: IntegrationParams, radii, sigma=1.5, amplitude=5000.0) -> np.ndarray:
    """Noise-free ring intensity at the pixel centres (flat detector)."""
    z, y = np.mgrid[0:P.NrPixelsZ, 0:P.NrPixelsY].astype(np.float64)
    dy, dz = y - P.BC_y, z - P.BC_z
    R = np.hypot(dy, dz)
    eta = np.arctan2(dy, dz)
    img = np.zeros_like(R)
    for r0 in radii:
        img += amplitude * np.exp(-0.5 * ((R - r0) / sigma) ** 2)
    return img * (1.0 + 0.3 * np.cos(2.0 * eta))


def dark_frame(P: IntegrationParams, rng: np.random.Generator) -> np.ndarray:
    return (100 + rng.poisson(5.0, (P.NrPixelsZ, P.NrPixelsY))).astype(np.uint16)


def frames(P: IntegrationParams, n: int, seed: int = 0, n_rings: int = 8,
           hot_fraction: float = 0.001, dtype=np.uint16):
    """(frames [n, NZ, NY], dark [NZ, NY], mask [NZ, NY] with 1.0 = hot)."""
    rng = np.random.default_rng(seed)
    dark = dark_frame(P, rng)
    base = 50.0 + ring_image(P, ring_radii(P, n_rings))
    npx = P.NrPixelsY * P.NrPixelsZ
    hot = rng.choice(npx, size=int(round(hot_fraction * npx)), replace=False)
    mask = np.zeros(npx)
    mask[hot] = 1.0
    mask = mask.reshape(P.NrPixelsZ, P.NrPixelsY)
    out = np.empty((n, P.NrPixelsZ, P.NrPixelsY), dtype=dtype)
    hi = np.iinfo(dtype).max if np.issubdtype(dtype, np.integer) else 65535
    for i in range(n):
        fr = np.random.default_rng(seed * 1000003 + i + 1)
        f = dark.astype(np.float64) + fr.poisson(base)
        f = np.minimum(f, hi)
        f.reshape(-1)[hot] = hi
        out[i] = f.astype(dtype)
    return out, dark, mask


def torch_frames(P: IntegrationParams, n: int, seed: int = 0, n_rings: int = 12,
                 device="cuda", dtype=None):
    """The same generator on the GPU with torch (for large batches)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    dt = dtype or torch.uint16
    base = torch.as_tensor(50.0 + ring_image(P, ring_radii(P, n_rings)), device=device,
                           dtype=torch.float32)
    dark = 100.0 + torch.poisson(torch.full_like(base, 5.0), generator=g)
    out = torch.empty((n, P.NrPixelsZ, P.NrPixelsY), dtype=dt, device=device)
    for i in range(n):
        f = dark + torch.poisson(base, generator=g)
        out[i] = f.clamp_(max=65535.0).to(torch.int32).to(dt)
    return out, dark.to(torch.int32).to(dt)
