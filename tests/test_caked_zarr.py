"""The Zarr copy of the caked file (MC_OUT_CAKED_ZARR).

The reference converts <name>.caked.hdf to <name>.caked.hdf.zarr.zip with
Hdf5ToZarr (utils/converters/midas2zip.py:31-93): every HDF5 group becomes
a Zarr group, every dataset an array of one chunk with the dataset's shape,
dtype and fill value and a Blosc zstd/bitshuffle compressor entry, and each
object's attributes its .zattrs (bytes as str, size-1 values as scalars,
arrays as lists).  The check reads the HDF5 file with the test reader (itself
checked on a libhdf5 file in test_outputs.py) and the zip with zipfile, JSON
and a Blosc header parse of its own, and compares object by object.  CPU
only: the output set makes no GPU call."""
import json
import struct
import zipfile

import numpy as np
import pytest

from tests import h5lite


@pytest.fixture(scope="module")
def lib():
    from midas_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libmidas_caking.so is not built")
    return _lib.lib()


def _params(**kw):
    from midas_amd.params import IntegrationParams
    P = IntegrationParams(NrPixelsY=64, NrPixelsZ=64, pxY=200.0, pxZ=200.0, Lsd=1e6,
                          RMin=10.0, RMax=30.0, RBinSize=2.0, EtaMin=-180.0, EtaMax=180.0,
                          EtaBinSize=45.0, OmegaStart=-10.0, OmegaStep=0.25, Lam=0.17,
                          Polariz=0.95)
    for k, v in kw.items():
        setattr(P, k, v)
    return P


def _blosc_chunk(raw: bytes) -> bytes:
    """A Blosc1 chunk stored as is (flag 0x02): 16-byte header + data."""
    version, vlz, flags, typesize, nbytes, blocksize, cbytes = struct.unpack("<BBBBiii", raw[:16])
    assert version == 2 and flags & 0x02, (version, flags)
    assert cbytes == len(raw) == nbytes + 16
    return raw[16:]


def _h5_attrs_as_zarr(attrs: dict) -> dict:
    """transfer_attrs of midas2zip.py:38-62."""
    out = {}
    for k, v in attrs.items():
        if isinstance(v, bytes):
            v = v.decode()
        elif isinstance(v, (np.ndarray, np.number)):
            v = v.reshape(-1)[0].tolist() if v.size == 1 else v.tolist()
        out[k] = v
    return out


def _compare(h5_path, zip_path):
    f = h5lite.File(h5_path)
    z = zipfile.ZipFile(zip_path)
    assert z.testzip() is None  # every entry's CRC-32
    names = set(z.namelist())
    assert json.loads(z.read(".zgroup")) == {"zarr_format": 2}
    n_arrays = 0
    for path, is_group in [("/", True)] + list(f.walk("/")):
        key = path.strip("/")
        pre = key + "/" if key else ""
        want = _h5_attrs_as_zarr(f.attrs(path))
        assert json.loads(z.read(pre + ".zattrs")) == want, path
        if is_group:
            assert json.loads(z.read(pre + ".zgroup")) == {"zarr_format": 2}, path
            continue
        n_arrays += 1
        a = f.dataset(path).read()
        a = np.asarray(a)
        meta = json.loads(z.read(pre + ".zarray"))
        assert meta["shape"] == list(a.shape) and meta["chunks"] == [max(d, 1) for d in a.shape]
        assert meta["compressor"] == {"blocksize": 0, "clevel": 3, "cname": "zstd",
                                      "id": "blosc", "shuffle": 2}
        assert meta["filters"] is None and meta["order"] == "C" and meta["zarr_format"] == 2
        dt = np.dtype(meta["dtype"])
        assert dt == a.dtype.newbyteorder("<"), path
        assert meta["fill_value"] == 0
        ck = pre + ".".join(["0"] * max(a.ndim, 1))
        if a.size == 0:
            assert ck not in names
            continue
        got = np.frombuffer(_blosc_chunk(z.read(ck)), dt).reshape(a.shape)
        np.testing.assert_array_equal(got, a, err_msg=path)
    # nothing else in the store
    objs = {p.strip("/") for p, _ in f.walk("/")} | {""}
    for n in names:
        head = n.rsplit("/", 1)[0] if "/" in n else ""
        assert head in objs, n
    return n_arrays


def test_zarr_copy_matches_caked_hdf(lib, tmp_path):
    from midas_amd.outputs import ResultFiles
    P = _params()
    rng = np.random.default_rng(3)
    reta = rng.normal(size=(5, 10, 8))
    files = {"caked_hdf", "per_frame", "sum_frames", "omegas", "caked_zarr"}
    cakes = rng.gamma(2.0, 50.0, (5, 10, 8)).astype(np.float32)
    cakes[:, 2, 3] = np.nan
    with ResultFiles(tmp_path, "scan_0001.zip", P, files=files, omega_sum_frames=2,
                     reta=reta) as rf:
        for i, c in enumerate(cakes):
            rf.frame(i, c, meta=(300.0 + i, 1.0, 2.0 + i, 3.0))
    h5 = tmp_path / "scan_0001.zip.caked.hdf"
    zp = tmp_path / "scan_0001.zip.caked.hdf.zarr.zip"
    assert h5.exists() and zp.exists()
    # 5 frames + 3 omega sums + SumFrames + Omegas + REtaMap + 10 instrument values
    assert _compare(h5, zp) == 21
    z = zipfile.ZipFile(zp)
    a = json.loads(z.read("IntegrationResult/FrameNr_3/.zattrs"))
    assert a == {"Header": "Radius,Eta", "Units": "Pixels,Degrees", "omega": -9.25}
    b = json.loads(z.read("OmegaSumFrame/LastFrameNumber_4/.zattrs"))
    assert b["Number Of Frames Summed"] == 1 and b["Temperature"] == 304.0


def test_zarr_copy_is_optional_and_readable_by_the_blosc_decoder(lib, tmp_path):
    import ctypes as C
    from midas_amd.outputs import ResultFiles
    P = _params()
    c = np.arange(80, dtype=np.float32).reshape(10, 8)
    with ResultFiles(tmp_path, "a.zip", P, files={"caked_hdf", "per_frame"}) as rf:
        rf.frame(0, c)
    assert not (tmp_path / "a.zip.caked.hdf.zarr.zip").exists()
    with ResultFiles(tmp_path, "b.zip", P, files={"per_frame", "caked_zarr"}) as rf:
        rf.frame(7, c)
    raw = zipfile.ZipFile(tmp_path / "b.zip.caked.hdf.zarr.zip").read(
        "IntegrationResult/FrameNr_7/0.0")
    out = np.empty(80, np.float64)
    n = C.c_int64()
    from midas_amd._lib import check
    check(lib.mc_blosc_decompress(raw, len(raw), out.ctypes.data, out.nbytes, C.byref(n)))
    assert n.value == out.nbytes
    np.testing.assert_array_equal(out.reshape(10, 8), c.astype(np.float64))


def test_zip64_directory_past_65535_entries(lib, tmp_path):
    """22000 one-bin frames make 66000+ zip entries: the directory count no
    longer fits 16 bits and the ZIP64 end records carry it."""
    from midas_amd.outputs import ResultFiles
    P = _params(RMin=10.0, RMax=12.0, RBinSize=2.0, EtaBinSize=360.0)
    n = 22000
    with ResultFiles(tmp_path, "big.zip", P, files={"per_frame", "caked_zarr"}) as rf:
        one = np.ones((1, 1), np.float32)
        for i in range(n):
            rf.frame(i, one * i)
    zp = tmp_path / "big.zip.caked.hdf.zarr.zip"
    tail = zp.read_bytes()[-120:]
    assert b"PK\x06\x06" in tail and b"PK\x06\x07" in tail
    z = zipfile.ZipFile(zp)
    names = z.namelist()
    assert len(names) > 65535
    raw = z.read(f"IntegrationResult/FrameNr_{n - 1}/0.0")
    assert np.frombuffer(_blosc_chunk(raw), "<f8")[0] == n - 1


# ---- metadata from the input zip (integrator.py:428-781) ----------------------

def test_coerce_metadata_array_rules():
    """The six rules of _coerce_metadata_array, integrator.py:433-443."""
    from midas_amd.zarr_out import coerce_metadata_array as co
    t = np.arange(5.0)
    assert co(np.array([7.0]), 5, 2, 3).tolist() == [7.0]                 # size 1
    np.testing.assert_array_equal(co(t, 5, 2, 3), [0.5, 2.5, 4.0])        # averaged
    m = np.arange(10.0).reshape(5, 2)
    np.testing.assert_array_equal(co(m, 5, 2, 3), [[1, 2], [5, 6], [8, 9]])
    np.testing.assert_array_equal(co(np.arange(4.0), 5, 2, 3), np.arange(4.0))  # other length
    np.testing.assert_array_equal(co(t, 5, 1, 5), t)                      # no sums
    same = np.array([b"Al"] * 5)
    np.testing.assert_array_equal(co(same, 5, 2, 3), np.array([b"Al"] * 3))     # tiled
    diff = np.array([b"a", b"b", b"a", b"a", b"a"])
    np.testing.assert_array_equal(co(diff, 5, 2, 3), diff)                # verbatim
    assert co(np.array([b"x", b"y"]), 5, 2, 3).tolist() == [b"x", b"y"]
    ints = np.array([1, 2, 3, 4, 5], np.int32)
    out = co(ints, 5, 2, 3)
    assert out.dtype == np.int32 and out.tolist() == [1, 3, 5]             # mean, cast back


def test_enrich_copies_input_metadata(lib, tmp_path):
    from midas_amd.outputs import ResultFiles
    from midas_amd.zarr_out import enrich_zarr
    from tests.test_zarr_zip import _add_array, write_zip
    frames = np.zeros((5, 4, 6), np.uint16)
    inp = tmp_path / "scan_0002.MIDAS.zip"
    write_zip(inp, frames, None, {"OmegaSumFrames": np.int32(2), "Lsd": 1e6},
              {"start": np.array([-180.0]), "step": np.array([0.25]),
               "energy": np.array([71.676]), "temps": np.arange(5.0)},
              datatype="uint16")
    with zipfile.ZipFile(inp, "a") as zf:
        zf.writestr("misc/.zgroup", json.dumps({"zarr_format": 2}))
        _add_array(zf, "misc/ring_current", np.array([101.0, 102.0, 99.0, 100.0, 98.0]))
        _add_array(zf, "misc/sample", np.array([b"Ti64"] * 5))
        _add_array(zf, "misc/label", np.array([b"a", b"b", b"a", b"a", b"a"]))
        _add_array(zf, "misc/short", np.array([1.0, 2.0, 3.0]))
        zf.writestr("instrument/.zgroup", json.dumps({"zarr_format": 2}))
        zf.writestr("instrument/detector/.zgroup", json.dumps({"zarr_format": 2}))
        _add_array(zf, "instrument/detector/pixel", np.array([200.0]))
        _add_array(zf, "top_scalar", np.array([3.5]))
    P = _params()
    c = np.ones((10, 8), np.float32)
    with ResultFiles(tmp_path, inp.name, P, files={"per_frame", "caked_zarr"}) as rf:
        for i in range(3):
            rf.frame(i, c * i)
    zp = tmp_path / (inp.name + ".caked.hdf.zarr.zip")
    enrich_zarr(zp, inp, version="test-1")
    z = zipfile.ZipFile(zp)
    assert z.testzip() is None
    names = set(z.namelist())

    def arr(path):
        meta = json.loads(z.read(path + "/.zarray"))
        raw = _blosc_chunk(z.read(path + "/" + ".".join(["0"] * max(len(meta["shape"]), 1))))
        return np.frombuffer(raw, np.dtype(meta["dtype"])).reshape(meta["shape"])

    sp = "measurement/process/scan_parameters"
    assert arr(sp + "/energy").tolist() == [71.676]
    np.testing.assert_array_equal(arr(sp + "/temps"), [0.5, 2.5, 4.0])
    for k in ("datatype", "start", "step"):
        assert f"{sp}/{k}/.zarray" not in names
    np.testing.assert_array_equal(arr("misc/ring_current"), [101.5, 99.5, 98.0])
    assert arr("misc/sample").tolist() == [b"Ti64"] * 3
    assert arr("misc/label").tolist() == [b"a", b"b", b"a", b"a", b"a"]
    assert arr("misc/short").tolist() == [1.0, 2.0, 3.0]
    assert arr("instrument/detector/pixel").tolist() == [200.0]
    assert arr("top_scalar").tolist() == [3.5]
    for g in ("misc", "instrument", "instrument/detector", "measurement", sp):
        assert json.loads(z.read(g + "/.zgroup")) == {"zarr_format": 2}
    assert not any(n.startswith(("exchange/", "analysis/")) for n in names)
    root = json.loads(z.read(".zattrs"))
    assert root["midas_version"] == "test-1" and "creation_timestamp" in root
    # the caked results are untouched
    np.testing.assert_array_equal(arr("IntegrationResult/FrameNr_2"), np.full((10, 8), 2.0))
    assert json.loads(z.read("IntegrationResult/FrameNr_2/.zattrs"))["Header"] == "Radius,Eta"
    assert sum(1 for n in z.namelist() if n == ".zattrs") == 1
