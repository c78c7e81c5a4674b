"""Map.bin / nMap.bin / maskMap.bin on disk (the DetectorMapper layout).

    Map.bin     64-byte header + n x 24-byte records {f32 y, f32 z, f64 frac,
                f32 deltaR, f32 areaWeight}          (MapperCore.h:19-25)
    nMap.bin    64-byte header + n_bins x {i32 count, i32 offset}
    maskMap.bin 64-byte header + n_bins x i32 flag   (DetectorMapper.c:862-879)

The header (MapHeader.h:44-55) is magic "MAP0", version 3, the SHA-256 of
the canonical parameter string of map_header_compute (MapHeader.h:216-234:
only p0-p4 and p6 of the distortion terms, Wavelength only in Q-mode), the
q_mode byte and the wavelength.  Readers detect the header by its magic and
also accept legacy files without one.
"""
from __future__ import annotations

import hashlib
import struct
from dataclasses import dataclass
from pathlib import Path
from typing import Optional

import numpy as np

from .params import IntegrationParams, effective_r_limits

MAP_HEADER_MAGIC = 0x3050414D
MAP_HEADER_VERSION = 3
MAP_HEADER_SIZE = 64

PXLIST_DTYPE = np.dtype([("y", "<f4"), ("z", "<f4"), ("frac", "<f8"),
                         ("deltaR", "<f4"), ("areaWeight", "<f4")])
assert PXLIST_DTYPE.itemsize == 24
NMAP_PAIR_DTYPE = np.dtype([("count", "<i4"), ("offset", "<i4")])


@dataclass
class MapHeader:
    magic: int = MAP_HEADER_MAGIC
    version: int = MAP_HEADER_VERSION
    param_hash: bytes = b"\x00" * 32
    q_mode: int = 0
    gradient_mode: int = 0
    wavelength: float = 0.0

    def to_bytes(self) -> bytes:
        out = bytearray(MAP_HEADER_SIZE)
        struct.pack_into("<II", out, 0, self.magic, self.version)
        out[8:40] = self.param_hash
        out[40] = self.q_mode & 0xFF
        out[41] = self.gradient_mode & 0xFF
        struct.pack_into("<d", out, 48, self.wavelength)
        return bytes(out)

    @classmethod
    def from_bytes(cls, raw: bytes) -> Optional["MapHeader"]:
        if len(raw) < MAP_HEADER_SIZE or struct.unpack_from("<I", raw, 0)[0] != MAP_HEADER_MAGIC:
            return None
        return cls(magic=MAP_HEADER_MAGIC, version=struct.unpack_from("<I", raw, 4)[0],
                   param_hash=bytes(raw[8:40]), q_mode=raw[40], gradient_mode=raw[41],
                   wavelength=struct.unpack_from("<d", raw, 48)[0])


def canonical_string(P: IntegrationParams) -> str:
    """The string map_header_compute hashes (C printf %.6f / %.8f)."""
    rmin, rmax = effective_r_limits(P)
    s = ("BC=%.6f,%.6f|EtaBinSize=%.6f|EtaMax=%.6f|EtaMin=%.6f|Lsd=%.6f|"
         "NrPixelsY=%d|NrPixelsZ=%d|RBinSize=%.6f|RMax=%.6f|RMin=%.6f|RhoD=%.6f|"
         "TransOpt=%d" % (P.BC_y, P.BC_z, P.EtaBinSize, P.EtaMax, P.EtaMin, P.Lsd,
                          P.NrPixelsY, P.NrPixelsZ, P.RBinSize, rmax, rmin, P.RhoD,
                          len(P.TransOpt)))
    for t in P.TransOpt[:10]:
        s += ",%d" % t
    s += ("|p0=%.6f|p1=%.6f|p2=%.6f|p3=%.6f|p4=%.6f|p6=%.6f|pxY=%.6f|pxZ=%.6f|"
          "tx=%.6f|ty=%.6f|tz=%.6f" % (P.p[0], P.p[1], P.p[2], P.p[3], P.p[4], P.p[6],
                                       P.pxY, P.pxZ, P.tx, P.ty, P.tz))
    if P.q_mode:
        s += "|qMode=1|Wavelength=%.8f" % P.Wavelength
    return s


def header_for(P: IntegrationParams) -> MapHeader:
    return MapHeader(param_hash=hashlib.sha256(canonical_string(P).encode("ascii")).digest(),
                     q_mode=1 if P.q_mode else 0,
                     wavelength=P.Wavelength if P.q_mode else 0.0)


@dataclass
class PixelMap:
    """Host view of a Map.bin / nMap.bin pair (and maskMap.bin flags)."""
    pxList: np.ndarray
    counts: np.ndarray
    offsets: np.ndarray
    map_header: Optional[MapHeader] = None
    nmap_header: Optional[MapHeader] = None
    mask_flags: Optional[np.ndarray] = None

    @property
    def n_bins(self) -> int:
        return int(self.counts.shape[0])

    @property
    def n_entries(self) -> int:
        return int(self.pxList.shape[0])

    def pairs(self) -> np.ndarray:
        p = np.empty((self.n_bins, 2), dtype=np.int32)
        p[:, 0] = self.counts
        p[:, 1] = self.offsets
        return p


def _split(path) -> tuple[Optional[MapHeader], np.ndarray]:
    raw = np.fromfile(path, dtype=np.uint8)
    hdr = MapHeader.from_bytes(raw[:MAP_HEADER_SIZE].tobytes())
    return hdr, raw[MAP_HEADER_SIZE if hdr is not None else 0:]


def load_map(map_path, nmap_path, mask_path=None) -> PixelMap:
    """Reads Map.bin / nMap.bin (header auto-detected), like ReadBins,
    IntegratorZarrOMP.c:97-191, and bin_io.py:206 of the reference package."""
    mh, mb = _split(map_path)
    if mb.nbytes % PXLIST_DTYPE.itemsize:
        raise ValueError(f"{map_path}: {mb.nbytes} data bytes is not a multiple of 24 "
                         "(legacy 16-byte map? regenerate it)")
    nh, nbuf = _split(nmap_path)
    if nbuf.nbytes % 8:
        raise ValueError(f"{nmap_path}: {nbuf.nbytes} bytes is not a multiple of 8")
    pairs = nbuf.view(NMAP_PAIR_DTYPE)
    flags = None
    if mask_path is not None and Path(mask_path).exists():
        _, fb = _split(mask_path)
        flags = fb.view("<i4").copy()
    return PixelMap(pxList=mb.view(PXLIST_DTYPE), counts=pairs["count"].copy(),
                    offsets=pairs["offset"].copy(), map_header=mh, nmap_header=nh,
                    mask_flags=flags)


def write_map(directory, *, pxList, counts, offsets, mask_flags=None,
              header: Optional[MapHeader] = None) -> None:
    """Writes Map.bin, nMap.bin and (when a bin is flagged) maskMap.bin with
    the same header on each, as DetectorMapper.c:836-881."""
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    hb = header.to_bytes() if header is not None else b""
    pairs = np.empty(len(counts), dtype=NMAP_PAIR_DTYPE)
    pairs["count"] = counts
    pairs["offset"] = offsets
    with open(d / "Map.bin", "wb") as f:
        f.write(hb)
        f.write(np.ascontiguousarray(pxList, dtype=PXLIST_DTYPE).tobytes())
    with open(d / "nMap.bin", "wb") as f:
        f.write(hb)
        f.write(pairs.tobytes())
    if mask_flags is not None and np.any(mask_flags):
        with open(d / "maskMap.bin", "wb") as f:
            f.write(hb)
            f.write(np.ascontiguousarray(mask_flags, dtype="<i4").tobytes())
