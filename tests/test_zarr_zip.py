"""Zarr-zip input (midas_amd.zarr_zip) on zips written here in the layout
of the reference's zipper (midas_zipper/ff_zip.py: exchange/data|dark,
analysis_parameters scalars and strings, scan_parameters, one frame per
chunk, Blosc zstd + bitshuffle) -- the reference ships no zip, so the
layout is restated from its writer and its reader (IntegratorZarrOMP.c
:933-1454).  Checks: every frame comes back bit for bit for each pixel
type; SkipFrame drops frames and darks; the dark is the reference's
sum of dark/(nDarks - SkipFrame) in order; omegas start at
start + SkipFrame*step; parameters land in IntegrationParams."""
import json
import zipfile

import numpy as np
import pytest

import blosc_enc as be
from midas_amd.zarr_zip import ZarrZip

BLOSC = {"id": "blosc", "cname": "zstd", "clevel": 1, "shuffle": 2, "blocksize": 0}


def _add_array(zf, path, arr, chunks=None, compressor=BLOSC, blocksize=65536):
    arr = np.asarray(arr)
    chunks = list(chunks or arr.shape)
    meta = {"chunks": chunks, "compressor": compressor, "dtype": arr.dtype.str,
            "fill_value": 0 if arr.dtype.kind != "S" else "", "filters": None, "order": "C",
            "shape": list(arr.shape), "zarr_format": 2}
    zf.writestr(f"{path}/.zarray", json.dumps(meta))
    per = chunks[0]
    for c in range(0, arr.shape[0], per):
        block = np.zeros(chunks, arr.dtype)
        part = arr[c:c + per]
        block[:len(part)] = part
        raw = block.tobytes()
        if compressor is not None:
            raw = be.compress(raw, typesize=arr.dtype.itemsize, blocksize=blocksize,
                              codec=compressor["cname"], shuffle=compressor["shuffle"])
        key = ".".join([str(c // per)] + ["0"] * (arr.ndim - 1))
        zf.writestr(f"{path}/{key}", raw)


def write_zip(path, frames, darks, params, scan, datatype=None):
    with zipfile.ZipFile(path, "w", zipfile.ZIP_STORED) as zf:
        zf.writestr(".zgroup", json.dumps({"zarr_format": 2}))
        nz, ny = frames.shape[1:]
        _add_array(zf, "exchange/data", frames, chunks=[1, nz, ny])
        if darks is not None:
            _add_array(zf, "exchange/dark", darks, chunks=[1, nz, ny])
        for k, v in params.items():
            a = np.array([v.encode()], dtype="S4096") if isinstance(v, str) else np.array([v])
            _add_array(zf, "analysis/process/analysis_parameters/" + k, a)
        for k, v in scan.items():
            _add_array(zf, "measurement/process/scan_parameters/" + k, np.asarray(v))
        if datatype:
            _add_array(zf, "measurement/process/scan_parameters/datatype",
                       np.array([datatype.encode()], dtype="S32"))


@pytest.mark.parametrize("dtype", [np.uint16, np.int32, np.float32, np.float64])
def test_frames_params_dark_omegas(tmp_path, dtype):
    rng = np.random.default_rng(5)
    frames = (rng.poisson(100, (7, 40, 52)) + rng.integers(0, 3, (7, 40, 52))).astype(dtype)
    darks = rng.poisson(10, (4, 40, 52)).astype(dtype)
    params = {"Lsd": 654321.5, "RMin": 10.0, "RMax": 90.0, "RBinSize": 0.5, "EtaBinSize": 5.0,
              "PixelSize": 172.0, "Wavelength": 0.1729, "SkipFrame": np.int32(2),
              "ResultFolder": "/tmp/out", "OmegaSumFrames": np.int32(3)}
    scan = {"start": np.array([-180.0]), "step": np.array([0.25])}
    p = tmp_path / "scan_000001.MIDAS.zip"
    write_zip(p, frames, darks, params, scan)
    with ZarrZip(p, threads=3) as z:
        assert (z.n_frames_total, z.nz, z.ny) == (7, 40, 52) and z.n_frames == 5
        assert z.dtype == np.dtype(dtype)
        np.testing.assert_array_equal(z.frames(), frames[2:])
        np.testing.assert_array_equal(z.frame(3), frames[5])
        np.testing.assert_array_equal(z.frames(1, 3), frames[3:5])
        want = np.zeros((40, 52))
        for k in range(2, 4):
            want += darks[k].astype(np.float64) / 2
        np.testing.assert_array_equal(z.dark_mean(), want)
        np.testing.assert_allclose(z.omegas(), -180.0 + 0.25 * (2 + np.arange(5)), rtol=0,
                                   atol=1e-12)
        assert z.analysis["ResultFolder"] == "/tmp/out"
        assert z.analysis["OmegaSumFrames"] == 3
        P = z.integration_params()
        assert (P.Lsd, P.RMin, P.RMax, P.RBinSize, P.pxY, P.pxZ, P.Lam) == \
            (654321.5, 10.0, 90.0, 0.5, 172.0, 172.0, 0.1729)
        assert (P.NrPixelsY, P.NrPixelsZ) == (52, 40)
        with pytest.raises(IndexError):
            z.frames(0, 6)


def test_datatype_string_and_uncompressed(tmp_path):
    frames = np.arange(2 * 8 * 16, dtype=np.uint32).reshape(2, 8, 16)
    p = tmp_path / "u32.zip"
    with zipfile.ZipFile(p, "w") as zf:
        _add_array(zf, "exchange/data", frames, chunks=[1, 8, 16], compressor=None)
        _add_array(zf, "measurement/process/scan_parameters/datatype",
                   np.array([b"uint32"], dtype="S16"))
    with ZarrZip(p) as z:
        assert z.dtype == np.uint32 and z.n_frames == 2
        np.testing.assert_array_equal(z.frames(), frames)
        np.testing.assert_array_equal(z.dark_mean(), np.zeros((8, 16)))
