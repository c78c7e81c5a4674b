"""Builds libmidas_caking.so (HIP, gfx950) in-tree with hipcc.

    python -m midas_amd.build          # compile every source, link the library

Each source is compiled to an object of its own (the mapper with FP
contraction off, so its FP64 geometry rounds like the CPU mapper), then all
objects are linked into ``midas_amd/libmidas_caking.so``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
INCLUDE = PKG.parent / "include"
LIB = PKG / "libmidas_caking.so"
OBJDIR = PKG / "_build"
ARCH = os.environ.get("MIDAS_AMD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions off: some packed forms have produced wrong
    # results on this hardware, and the kernels do not need them
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", "-Wno-unused-command-line-argument",
]
PER_FILE = {
    "mc_mapper.hip": ["-ffp-contract=off"],
}
SOURCES = ["mc_core.cpp", "mc_kernels.hip", "mc_mapper.hip", "mc_fit.hip", "mc_server.cpp",
           "mc_h5.cpp", "mc_out.cpp", "mc_post.hip", "mc_blosc.cpp", "mc_zarr.cpp"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build midas_amd)")


def _compile(src: Path) -> Path:
    obj = OBJDIR / (src.name + ".o")
    lang = ["-x", "hip"] if src.suffix == ".hip" else []
    cmd = [hipcc(), *COMMON, *PER_FILE.get(src.name, []), *lang, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False) -> Path:
    OBJDIR.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES if (CSRC / s).exists()]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
           *map(str, objs), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
