"""Metadata of the caked Zarr zip (FF_HEDM/workflows/integrator.py:428-781).

After the caked results are written as ``<name>.caked.hdf.zarr.zip``
(``ResultFiles(..., files={..., "caked_zarr"})``), the reference copies the
input zip's metadata into it (``_enrich_zarr_with_metadata``, :682-781):

* ``measurement/process/scan_parameters`` except ``datatype``, ``start`` and
  ``step`` (:718-729),
* every other top-level group or array of the input, except the groups the
  integrator owns (``exchange``, ``analysis``, ``processed``,
  ``measurement``, ``InstrumentParameters``, ``IntegrationResult``,
  ``REtaMap``, ``Omegas``, ``SumFrames``; :741-773),
* arrays coerced by ``_coerce_metadata_array`` (:428-488): numeric arrays
  whose first dimension is the input's frame count are averaged over
  OmegaSumFrames-sized chunks; string arrays of that length are tiled to
  the output count when all values agree; everything else is copied as is;
  arrays over 256 MiB are skipped (:490-511),
* version attributes on the root group (``stamp_zarr``: ``midas_version``,
  ``midas_git_commit``, ``midas_git_date``, ``creation_timestamp``).

Arrays are written as one chunk each, with the compressor entry zarr gives
new arrays by default (Blosc lz4, level 5, byte shuffle) and the chunk
stored in Blosc's uncompressed form.  The bright-field lineout of :778-781
needs a bright file and the Python integrator, and is not written here.
"""
from __future__ import annotations

import datetime
import json
import math
import os
import shutil
import struct
import tempfile
import zipfile
from pathlib import Path

import numpy as np

from .zarr_zip import ZarrZip

INTEGRATOR_OWNED = {"exchange", "analysis", "processed", "measurement", "InstrumentParameters",
                    "IntegrationResult", "REtaMap", "Omegas", "SumFrames"}
SCAN = "measurement/process/scan_parameters"
MAX_METADATA_BYTES = 256 * 1024 * 1024


def coerce_metadata_array(a, total_frames: int, omega_sum_frames: int, n_output: int):
    """_coerce_metadata_array (integrator.py:428-488)."""
    if not isinstance(a, np.ndarray) or a.size <= 1:
        return a
    aligned = total_frames > 0 and a.shape[0] == total_frames
    if np.issubdtype(a.dtype, np.number):
        if not aligned or omega_sum_frames <= 1:
            return a
        n = math.ceil(total_frames / omega_sum_frames)
        out = np.zeros((n,) + a.shape[1:], dtype=a.dtype)
        for i in range(n):
            s = i * omega_sum_frames
            out[i] = np.mean(a[s:min(s + omega_sum_frames, total_frames)], axis=0)
        return out
    if not aligned:
        return a
    if len(set(a.reshape(-1).tolist())) == 1:
        return np.full(n_output, a.reshape(-1)[0], dtype=a.dtype)
    return a


def _zarray(a: np.ndarray) -> bytes:
    shape = list(a.shape)
    numeric = np.issubdtype(a.dtype, np.number) or a.dtype.kind == "b"
    meta = {"chunks": [max(d, 1) for d in shape] or [], "compressor":
            {"blocksize": 0, "clevel": 5, "cname": "lz4", "id": "blosc", "shuffle": 1},
            "dtype": a.dtype.str, "fill_value": 0 if numeric else None, "filters": None,
            "order": "C", "shape": shape, "zarr_format": 2}
    return json.dumps(meta, indent=4, sort_keys=True).encode()


def _blosc_stored(raw: bytes, typesize: int) -> bytes:
    """A Blosc1 chunk holding raw as is (flag 0x02, "memcpyed")."""
    n = len(raw)
    return struct.pack("<BBBBiii", 2, 1, 0x02, min(max(typesize, 1), 255), n, n, n + 16) + raw


class _Out:
    """Entries to add to the output store (name -> bytes)."""

    def __init__(self):
        self.entries: dict[str, bytes] = {}

    def group(self, path: str):
        self.entries.setdefault(f"{path}/.zgroup", b'{\n    "zarr_format": 2\n}')
        self.entries.setdefault(f"{path}/.zattrs", b"{}")

    def array(self, path: str, a: np.ndarray, attrs: dict | None = None):
        a = np.ascontiguousarray(a)
        if a.dtype.byteorder == ">":
            a = a.astype(a.dtype.newbyteorder("<"))
        self.entries[f"{path}/.zarray"] = _zarray(a)
        self.entries[f"{path}/.zattrs"] = json.dumps(attrs or {}, indent=4,
                                                     sort_keys=True).encode()
        if a.size:
            key = ".".join(["0"] * max(a.ndim, 1))
            self.entries[f"{path}/{key}"] = _blosc_stored(a.tobytes(), a.dtype.itemsize)


def _children(names: set, prefix: str):
    """Direct children of group `prefix` ("" = root): (name, is_group)."""
    pre = prefix + "/" if prefix else ""
    out = {}
    for n in names:
        if not n.startswith(pre):
            continue
        rest = n[len(pre):]
        parts = rest.split("/")
        if len(parts) == 2 and parts[1] in (".zgroup", ".zarray"):
            out[parts[0]] = parts[1] == ".zgroup"
    return sorted(out.items())


def _copy_group(zin: ZarrZip, names, src: str, dst: str, out: _Out, total, osf, exclude=()):
    """_copy_group_recursive (integrator.py:513-567)."""
    n_out = max(1, math.ceil(total / osf)) if total > 0 else 0
    out.group(dst)
    for key, is_group in _children(names, src):
        if key in exclude:
            continue
        s, d = f"{src}/{key}", f"{dst}/{key}"
        if is_group:
            _copy_group(zin, names, s, d, out, total, osf)
            continue
        meta = zin._meta(s)
        nbytes = np.dtype(meta["dtype"]).itemsize * int(np.prod(meta["shape"] or [1]))
        if nbytes > MAX_METADATA_BYTES:
            continue
        a = zin._read_array(s)
        if meta["shape"] == []:
            a = a.reshape(())
        out.array(d, coerce_metadata_array(a, total, osf, n_out))


def _rewrite(zip_path: Path, new: dict[str, bytes]):
    """Copies the archive with `new` entries replacing or adding to its own
    (a zip entry cannot be replaced in place)."""
    fd, tmp = tempfile.mkstemp(dir=zip_path.parent, suffix=".tmp")
    os.close(fd)
    try:
        with zipfile.ZipFile(zip_path) as zi, \
                zipfile.ZipFile(tmp, "w", zipfile.ZIP_STORED, allowZip64=True) as zo:
            for info in zi.infolist():
                if info.filename in new:
                    continue
                with zi.open(info) as src, zo.open(info.filename, "w", force_zip64=True) as dst:
                    shutil.copyfileobj(src, dst, 1 << 20)
            for name, data in new.items():
                zo.writestr(name, data)
        os.replace(tmp, zip_path)
    finally:
        if os.path.exists(tmp):
            os.remove(tmp)


def enrich_zarr(output_zip, input_zip, *, version: str | None = None) -> None:
    """Adds the input zip's metadata and the version attributes to a caked
    Zarr zip, as integrator.py:682-781 does (see the module text)."""
    output_zip = Path(output_zip)
    out = _Out()
    with ZarrZip(input_zip) as zin:
        names = zin._names
        total = zin.n_frames_total
        osf = 1
        try:
            v = zin.analysis.get("OmegaSumFrames")
            if v is not None:
                osf = int(np.asarray(v).reshape(-1)[0])
        except (TypeError, ValueError):
            osf = 1
        osf = max(osf, 1) if osf else 1
        if f"{SCAN}/.zgroup" in names or any(n.startswith(SCAN + "/") for n in names):
            for g in ("measurement", "measurement/process"):
                out.group(g)
            _copy_group(zin, names, SCAN, SCAN, out, total, osf,
                        exclude={"datatype", "start", "step"})
        n_out = max(1, math.ceil(total / osf)) if total > 0 else 0
        for key, is_group in _children(names, ""):
            if key in INTEGRATOR_OWNED:
                continue
            if is_group:
                _copy_group(zin, names, key, key, out, total, osf)
            else:
                meta = zin._meta(key)
                nbytes = np.dtype(meta["dtype"]).itemsize * int(np.prod(meta["shape"] or [1]))
                if nbytes <= MAX_METADATA_BYTES:
                    a = zin._read_array(key)
                    if meta["shape"] == []:
                        a = a.reshape(())
                    out.array(key, coerce_metadata_array(a, total, osf, n_out))
    with zipfile.ZipFile(output_zip) as z:
        root = json.loads(z.read(".zattrs")) if ".zattrs" in z.namelist() else {}
        have = set(z.namelist())
    from . import __version__
    root.update({"midas_version": version or __version__, "midas_git_commit": "",
                 "midas_git_date": "",
                 "creation_timestamp": datetime.datetime.now().isoformat()})
    out.entries[".zattrs"] = json.dumps(root, indent=4, sort_keys=True).encode()
    # groups the output already holds keep their own .zgroup/.zattrs
    new = {k: v for k, v in out.entries.items()
           if k == ".zattrs" or not (k.endswith((".zgroup", ".zattrs")) and k in have)}
    _rewrite(output_zip, new)
