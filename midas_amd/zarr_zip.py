"""Zarr-zip input of the far-field integrator (IntegratorZarrOMP.c:933-1454).

A MIDAS far-field zip is a Zarr v2 store in a zip archive:
``exchange/data`` [frames, NZ, NY] (one frame per chunk), optional
``exchange/dark`` and ``exchange/flood``, scalar analysis parameters under
``analysis/process/analysis_parameters/<Key>`` and scan parameters under
``measurement/process/scan_parameters/<Key>``.  Chunks are Blosc1
compressed (the reference's zipper writes zstd + bitshuffle,
midas_zipper/ff_zip.py:44); they are decoded natively
(mc_blosc_decompress, many frames on host threads at once).

The reader follows the reference's semantics:
* ``SkipFrame`` drops the first frames and darks (:1329-1331); omegas
  start at ``start + SkipFrame * step`` (:1402);
* the dark is the mean of darks SkipFrame..nDarks-1, each divided by
  (nDarks - SkipFrame) and summed in order (:1416-1454);
* the pixel type is the scan parameter ``datatype`` when present, else the
  data array's own dtype (:1106-1134, 1239-1263).
"""
from __future__ import annotations

import ctypes as C
import json
import zipfile
import zlib
from pathlib import Path

import numpy as np

from ._lib import check, lib
from .params import IntegrationParams

ANALYSIS = "analysis/process/analysis_parameters/"
SCAN = "measurement/process/scan_parameters/"
_DATATYPES = {"uint16": np.uint16, "uint32": np.uint32, "int32": np.int32,
              "float32": np.float32, "float64": np.float64}


class ZarrZip:
    """Reader of one far-field Zarr zip.  ``frames(a, b)`` returns frames
    a..b-1 after SkipFrame as an array [b - a, NZ, NY] of the file's pixel
    type, ready for :func:`midas_amd.integrate_batch`."""

    def __init__(self, path, *, threads: int = 0):
        self.path = Path(path)
        self.threads = threads
        self._z = zipfile.ZipFile(self.path)
        self._names = set(self._z.namelist())
        if "exchange/data/.zarray" not in self._names:
            raise ValueError(f"{path}: no exchange/data array")
        self.data_meta = self._meta("exchange/data")
        shape = self.data_meta["shape"]
        if len(shape) != 3:
            raise ValueError(f"{path}: exchange/data has shape {shape}, expected 3-D")
        self.n_frames_total, self.nz, self.ny = (int(x) for x in shape)
        self.analysis = self._group(ANALYSIS)
        self.scan = self._group(SCAN)
        dt = self.scan.get("datatype")
        if isinstance(dt, str) and dt:
            self.dtype = np.dtype(_DATATYPES.get(dt.lower(), np.uint16))
        else:
            self.dtype = np.dtype(self.data_meta["dtype"]).newbyteorder("=")
        self.skip_frame = int(self.analysis.get("SkipFrame", 0) or 0)
        self.n_frames = max(0, self.n_frames_total - self.skip_frame)

    # ---- zarr plumbing ----
    def _meta(self, array: str) -> dict:
        return json.loads(self._z.read(f"{array}/.zarray"))

    def _chunk_bytes(self, array: str, meta: dict, key: str) -> bytes | None:
        name = f"{array}/{key}"
        if name not in self._names:
            return None
        return self._z.read(name)

    def _decode(self, raw: bytes, meta: dict, nbytes: int) -> bytes:
        comp = meta.get("compressor")
        if meta.get("filters"):
            raise ValueError(f"zarr filters {meta['filters']} are not supported")
        if comp is None:
            return raw
        cid = comp.get("id")
        if cid == "blosc":
            out = np.empty(nbytes, np.uint8)
            n = C.c_int64()
            check(lib().mc_blosc_decompress(raw, len(raw), out.ctypes.data, nbytes, C.byref(n)))
            return out[:n.value].tobytes()
        if cid in ("zlib", "gzip"):
            return zlib.decompress(raw, 47 if cid == "gzip" else 15)
        raise ValueError(f"zarr compressor {cid!r} is not supported")

    def _read_array(self, array: str):
        meta = self._meta(array)
        dt = np.dtype(meta["dtype"])
        shape = tuple(meta["shape"]) or (1,)
        chunks = tuple(meta["chunks"])
        if tuple(chunks) != tuple(shape):
            # parameter arrays are single-chunk; read chunk by chunk otherwise
            parts = []
            for c0 in range(0, shape[0], chunks[0]):
                key = ".".join([str(c0 // chunks[0])] + ["0"] * (len(shape) - 1))
                raw = self._chunk_bytes(array, meta, key)
                n = min(chunks[0], shape[0] - c0)
                nb = int(np.prod(chunks)) * dt.itemsize
                a = (np.frombuffer(self._decode(raw, meta, nb), dt).reshape(chunks)[:n]
                     if raw is not None else np.full((n,) + chunks[1:], meta.get("fill_value") or 0,
                                                     dt))
                parts.append(a)
            return np.concatenate(parts)
        key = ".".join(["0"] * len(shape))
        raw = self._chunk_bytes(array, meta, key)
        nb = int(np.prod(shape)) * dt.itemsize
        if raw is None:
            return np.full(shape, meta.get("fill_value") or 0, dt)
        return np.frombuffer(self._decode(raw, meta, nb), dt).reshape(shape)

    def _group(self, prefix: str) -> dict:
        out = {}
        keys = sorted({n[len(prefix):].split("/")[0] for n in self._names
                       if n.startswith(prefix) and n.endswith("/.zarray")})
        for k in keys:
            a = self._read_array(prefix + k)
            if a.dtype.kind == "S":
                vals = [v.split(b"\0")[0].decode(errors="replace") for v in a.reshape(-1)]
                out[k] = vals[0] if len(vals) == 1 else vals
            elif a.dtype.kind == "U":
                vals = [str(v) for v in a.reshape(-1)]
                out[k] = vals[0] if len(vals) == 1 else vals
            else:
                out[k] = a.reshape(-1)[0].item() if a.size == 1 else a.reshape(-1).copy()
        return out

    # ---- frames ----
    def _frame_chunks(self, array: str, meta: dict, idx):
        chunks = meta["chunks"]
        if list(chunks[1:]) != [self.nz, self.ny]:
            raise ValueError(f"{array}: chunks {chunks} do not hold whole frames")
        per = int(chunks[0])
        dt = np.dtype(meta["dtype"])
        nb = per * self.nz * self.ny * dt.itemsize
        ids = sorted({i // per for i in idx})
        raws = {c: self._chunk_bytes(array, meta, f"{c}.0.0") for c in ids}
        comp = meta.get("compressor")
        dec = {}
        blosc = [c for c in ids if raws[c] is not None and comp is not None
                 and comp.get("id") == "blosc"]
        if blosc:
            keep = [np.frombuffer(raws[c], np.uint8) for c in blosc]
            outs = [np.empty(nb, np.uint8) for _ in blosc]
            srcs = (C.c_void_p * len(blosc))(*[k.ctypes.data for k in keep])
            sizes = (C.c_int64 * len(blosc))(*[k.size for k in keep])
            dsts = (C.c_void_p * len(blosc))(*[o.ctypes.data for o in outs])
            got = (C.c_int64 * len(blosc))()
            check(lib().mc_blosc_decompress_many(srcs, sizes, dsts, nb, len(blosc),
                                                 self.threads, got))
            for c, o, n in zip(blosc, outs, got):
                dec[c] = o[:n]
        for c in ids:
            if c in dec:
                continue
            if raws[c] is None:
                dec[c] = np.full(nb // dt.itemsize, meta.get("fill_value") or 0, dt).view(np.uint8)
            else:
                dec[c] = np.frombuffer(self._decode(raws[c], meta, nb), np.uint8)
        out = np.empty((len(idx), self.nz, self.ny), dt)
        for j, i in enumerate(idx):
            a = dec[i // per].view(dt).reshape(per, self.nz, self.ny)
            out[j] = a[i % per]
        return out

    def frames(self, start: int = 0, stop: int | None = None) -> np.ndarray:
        """Frames start..stop-1 (after SkipFrame), native byte order, the
        pixel type of the file."""
        stop = self.n_frames if stop is None else stop
        if not (0 <= start <= stop <= self.n_frames):
            raise IndexError(f"frames [{start}, {stop}) of {self.n_frames}")
        idx = list(range(self.skip_frame + start, self.skip_frame + stop))
        a = self._frame_chunks("exchange/data", self.data_meta, idx)
        return a.astype(self.dtype.newbyteorder("="), copy=False)

    def frame(self, i: int) -> np.ndarray:
        return self.frames(i, i + 1)[0]

    def dark_mean(self) -> np.ndarray:
        """[NZ, NY] FP64 dark: sum over darks SkipFrame..nDarks-1 of
        dark / (nDarks - SkipFrame), in order (IntegratorZarrOMP.c:1416-1454);
        zeros without darks."""
        out = np.zeros((self.nz, self.ny))
        if "exchange/dark/.zarray" not in self._names:
            return out
        meta = self._meta("exchange/dark")
        n_darks = int(meta["shape"][0])
        n = n_darks - self.skip_frame
        if n <= 0:
            return out
        d = self._frame_chunks("exchange/dark", meta, list(range(self.skip_frame, n_darks)))
        for k in range(len(d)):
            out += d[k].astype(np.float64) / n
        return out

    def omegas(self) -> np.ndarray:
        """Omega of each frame after SkipFrame: start + (SkipFrame + i) * step."""
        start = float(self.scan.get("start", 0.0) or 0.0)
        step = float(self.scan.get("step", 0.0) or 0.0)
        start += self.skip_frame * step
        return start + step * np.arange(self.n_frames)

    def integration_params(self, base: IntegrationParams | None = None) -> IntegrationParams:
        """Binning and detector values of the zip over ``base`` (the keys
        the integrator reads: RMin/RMax/RBinSize, EtaMin/EtaMax/EtaBinSize,
        QMin/QMax/QBinSize, Lsd, PixelSize[Y|Z], Wavelength, Normalize,
        OmegaSumFrames, SkipFrame) and the detector size of the data."""
        P = base if base is not None else IntegrationParams()
        a = self.analysis
        for key in ("RMin", "RMax", "RBinSize", "EtaMin", "EtaMax", "EtaBinSize", "QMin",
                    "QMax", "QBinSize", "Lsd"):
            if key in a:
                setattr(P, key, float(a[key]))
        for key in ("PixelSize", "PixelSizeY", "PixelSizeZ"):
            if key in a:
                P.pxY = P.pxZ = float(a[key])
        if "Wavelength" in a:  # the integrator's wavelength (REtaMap Q)
            P.Lam = float(a["Wavelength"])
        P.NrPixelsY, P.NrPixelsZ = self.ny, self.nz
        return P

    def close(self):
        self._z.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
