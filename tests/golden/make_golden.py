"""Writes the golden fixtures of tests/golden/ from the reference's own
mapper and integration code (oracle/_ref/libmidasref.so, compiled in place
from /root/reference by oracle/Makefile).  Run in the build container:

    python tests/golden/make_golden.py

Fixtures are data only: for tiny geometries the full Map.bin / nMap.bin /
maskMap.bin content, a seeded frame and the reference cake; for the
1024x1024 and 2048x2048 workloads, SHA-256 digests of the map arrays and
summary values.

    python tests/golden/make_golden.py batch_2048   # only the 2048^2 record
    python tests/golden/make_golden.py qmode_crop   # only the Q-mode crop record
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from oracle import oracle as orc  # noqa: E402
from midas_amd import synth  # noqa: E402
from midas_amd.params import IntegrationParams  # noqa: E402
from helpers import to_oracle  # noqa: E402


def tiny_cases() -> dict[str, IntegrationParams]:
    return {
        "tiny_flat": IntegrationParams(NrPixelsY=40, NrPixelsZ=36, pxY=200.0, pxZ=200.0,
                                       Lsd=1e6, BC_y=19.3, BC_z=17.6, RMin=1.0, RMax=24.0,
                                       RBinSize=0.5, EtaBinSize=10.0),
        "tiny_tilt": IntegrationParams(NrPixelsY=36, NrPixelsZ=40, pxY=150.0, pxZ=150.0,
                                       Lsd=3e5, BC_y=17.2, BC_z=20.9, tx=0.3, ty=0.5, tz=-0.4,
                                       p=[2e-4, 1e-4, -2e-4, 15.0, 1e-4, 3e-5, 7.0] + [0.0] * 8,
                                       RhoD=3e4, RMin=2.0, RMax=22.0, RBinSize=0.25,
                                       EtaBinSize=7.5, Parallax=150.0, SolidAngleCorrection=1,
                                       PolarizationCorrection=1, PolarizationFraction=0.95),
        "tiny_subpix": IntegrationParams(NrPixelsY=32, NrPixelsZ=32, pxY=200.0, pxZ=200.0,
                                         Lsd=1e6, BC_y=16.0, BC_z=16.0, RMin=1.0, RMax=20.0,
                                         RBinSize=0.25, EtaBinSize=5.0, SubPixelLevel=3,
                                         SubPixelCardinalWidth=12.0, TransOpt=[2, 1, 3]),
        "tiny_qmode": IntegrationParams(NrPixelsY=48, NrPixelsZ=40, pxY=200.0, pxZ=200.0,
                                        Lsd=2e5, BC_y=24.5, BC_z=20.5, QBinSize=0.08, QMin=0.6,
                                        QMax=9.0, Wavelength=0.172979, EtaBinSize=15.0,
                                        EtaMin=-190.0, EtaMax=170.0),
        "tiny_panels": IntegrationParams(NrPixelsY=44, NrPixelsZ=40, pxY=172.0, pxZ=172.0,
                                         Lsd=2e5, BC_y=22.4, BC_z=19.1, RMin=1.0, RMax=30.0,
                                         RBinSize=0.5, EtaBinSize=12.0, NPanelsY=2, NPanelsZ=2,
                                         PanelSizeY=20, PanelSizeZ=18, PanelGapsY=[4],
                                         PanelGapsZ=[4]),
    }


PANEL_SHIFTS = {"tiny_panels": [[0, 0.3, -0.2, 0.5, 40.0, 2e-5], [3, -0.2, 0.15, -0.3, -30.0, 0.0]]}


def seeded_inputs(P: IntegrationParams, seed: int):
    rng = np.random.default_rng(seed)
    mask = (rng.random((P.NrPixelsZ, P.NrPixelsY)) < 0.02).astype(np.float64)
    frame = rng.poisson(300.0, (P.NrPixelsZ, P.NrPixelsY)).astype(np.float64)
    dark = rng.poisson(30.0, (P.NrPixelsZ, P.NrPixelsY)).astype(np.float64)
    return mask, frame, dark


def params_dict(P: IntegrationParams) -> dict:
    return {k: v for k, v in P.__dict__.items()}


def make_tiny():
    for name, P in tiny_cases().items():
        O = to_oracle(P)
        mask, frame, dark = seeded_inputs(P, 11)
        sh = PANEL_SHIFTS.get(name)
        m = orc.ref_build_map(O, mask=mask, panel_shifts=sh)
        img = frame - dark
        cake, area = orc.ref_integrate(m, img, P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z))
        cake_g, _ = orc.ref_integrate(m, img, P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z),
                                      gradient=True)
        np.savez_compressed(
            HERE / f"{name}.npz", entries=m.entries.view(np.uint8), pairs=m.pairs,
            flags=m.flags, mask=mask, frame=frame, dark=dark, cake=cake, area=area,
            cake_gradient=cake_g, params=json.dumps(params_dict(P)),
            panel_shifts=json.dumps(sh))
        print(name, m.entries.size, "records")


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def make_plain_1024():
    P = synth.plain_1024()
    frames, dark, mask = synth.frames(P, 1, seed=0)
    m = orc.ref_build_map(to_oracle(P), mask=mask)
    img = frames[0].astype(np.float64) - dark
    cake, area = orc.ref_integrate(m, img, P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z))
    lo = orc.lineouts(cake, orc.area_inbounds(m, P.NrPixelsY, P.NrPixelsZ))[0]
    rec = {
        "workload": "plain_1024 (midas_amd.synth.plain_1024, synth.frames seed 0)",
        "nnz": int(m.entries.size),
        "sha256_entries": digest(m.entries),
        "sha256_pairs": digest(m.pairs),
        "sha256_flags": digest(m.flags),
        "n_flagged": int(m.flags.sum()),
        "sha256_mask": digest(mask),
        "sha256_frame": digest(frames[0]),
        "sha256_dark": digest(dark),
        "cake_nansum": float(np.nansum(cake)),
        "cake_n_nan": int(np.isnan(cake).sum()),
        "cake_sample": {str(i): float(cake.reshape(-1)[i]) for i in range(0, cake.size, 9973)},
        "lineout_cpu_sample": {str(i): float(lo[i]) for i in range(0, lo.size, 97)},
    }
    (HERE / "plain_1024.json").write_text(json.dumps(rec, indent=1))
    print("plain_1024", rec["nnz"], "records")


def make_batch_2048():
    """The 2048x2048 batching workload: bin membership digests (pairs,
    flags, the y/z of every record in order), FP64 sums of the areas, and
    the reference cake and lineout of frame 0 of the seeded generator."""
    P = synth.batch_2048()
    frames, dark, mask = synth.frames(P, 1, seed=0, n_rings=12)
    m = orc.ref_build_map(to_oracle(P), mask=mask)
    img = frames[0].astype(np.float64) - dark
    cake, _ = orc.ref_integrate(m, img, P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z))
    lo = orc.lineouts(cake, orc.area_inbounds(m, P.NrPixelsY, P.NrPixelsZ))[0]
    flat = cake.reshape(-1)
    rec = {
        "workload": "batch_2048 (midas_amd.synth.batch_2048, synth.frames seed 0, 12 rings)",
        "nnz": int(m.entries.size),
        "sha256_pairs": digest(m.pairs),
        "sha256_flags": digest(m.flags),
        "sha256_y": digest(m.entries["y"]),
        "sha256_z": digest(m.entries["z"]),
        "sum_frac": float(np.sum(m.entries["frac"], dtype=np.float64)),
        "sum_area": float(np.sum(m.entries["areaWeight"], dtype=np.float64)),
        "n_flagged": int(m.flags.sum()),
        "sha256_mask": digest(mask),
        "sha256_frame": digest(frames[0]),
        "sha256_dark": digest(dark),
        "cake_nansum": float(np.nansum(cake)),
        "cake_n_nan": int(np.isnan(cake).sum()),
        "cake_sample": {str(i): float(flat[i]) for i in range(0, flat.size, 49999)},
        "lineout_cpu_sample": {str(i): float(lo[i]) for i in range(0, lo.size, 131)},
    }
    (HERE / "batch_2048.json").write_text(json.dumps(rec, indent=1))
    print("batch_2048", rec["nnz"], "records")


def make_qmode_crop():
    """The Q-mode workload's binning (Q 0.5-12 by 0.001, eta by 0.5 deg,
    SubPixelLevel 2, tilts, parallax) on a 320x288 crop of its detector:
    membership digests, area sums, and the reference's gradient-mode cake of
    a seeded float32 frame (IntegratorZarrOMP.c:1724-1732)."""
    P = synth.qmode_crop()
    frames, dark = synth.qmode_frames(P, 1, seed=0)
    m = orc.ref_build_map(to_oracle(P))
    img = frames[0].astype(np.float64) - dark.astype(np.float64)
    cake, _ = orc.ref_integrate(m, img, P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z),
                                gradient=True)
    flat = cake.reshape(-1)
    nz = np.flatnonzero(np.nan_to_num(flat) != 0)
    rec = {
        "workload": "qmode_crop (midas_amd.synth.qmode_crop, synth.qmode_frames seed 0, "
                    "gradient mode)",
        "nnz": int(m.entries.size),
        "sha256_pairs": digest(m.pairs),
        "sha256_flags": digest(m.flags),
        "sha256_y": digest(m.entries["y"]),
        "sha256_z": digest(m.entries["z"]),
        "sum_frac": float(np.sum(m.entries["frac"], dtype=np.float64)),
        "sum_area": float(np.sum(m.entries["areaWeight"], dtype=np.float64)),
        "sum_deltaR": float(np.sum(m.entries["deltaR"], dtype=np.float64)),
        "sha256_frame": digest(frames[0]),
        "sha256_dark": digest(dark),
        "cake_nansum": float(np.nansum(cake)),
        "cake_n_nonzero": int(nz.size),
        "cake_sample": {str(int(i)): float(flat[i]) for i in nz[::max(1, nz.size // 40)]},
    }
    (HERE / "qmode_crop.json").write_text(json.dumps(rec, indent=1))
    print("qmode_crop", rec["nnz"], "records")


if __name__ == "__main__":
    if orc.ref_lib() is None:
        sys.exit("the reference sources are needed to write the fixtures")
    only = {"batch_2048": make_batch_2048, "qmode_crop": make_qmode_crop}
    if len(sys.argv) == 2 and sys.argv[1] in only:
        only[sys.argv[1]]()
        sys.exit(0)
    make_tiny()
    make_plain_1024()
    make_batch_2048()
    make_qmode_crop()
