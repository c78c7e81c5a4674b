"""The Blosc1 decoder (mc_blosc_decompress) against chunks written from the
container's and the shuffles' definitions (tests/blosc_enc.py): every codec
the image has (zstd, lz4, zlib), no / byte / bit shuffle, split and unsplit
blocks, leftover blocks under format versions 2 and 3, stored chunks and
streams, and the error paths.  Host code only: no GPU.

Real c-blosc output is not available in this container (no Blosc library,
and the reference ships no Zarr zip), so agreement with c-blosc itself rests
on the published header layout and bitshuffle definition restated in
blosc_enc.py."""
import ctypes as C

import numpy as np
import pytest

import blosc_enc as be
from midas_amd import _lib


@pytest.fixture(scope="module")
def L():
    return _lib.lib()


def _decode(L, buf, cap):
    out = np.zeros(cap, np.uint8)
    n = C.c_int64()
    rc = L.mc_blosc_decompress(buf, len(buf), out.ctypes.data, cap, C.byref(n))
    return rc, out[:n.value].tobytes()


def _data(seed, n, dtype):
    rng = np.random.default_rng(seed)
    # detector-like: smooth background + Poisson noise, compressible
    x = 100 + 20 * np.sin(np.arange(n) / 50.0) + rng.poisson(5, n)
    return x.astype(dtype).tobytes()


@pytest.mark.parametrize("codec", ["zstd", "lz4", "zlib"])
@pytest.mark.parametrize("shuffle", [0, 1, 2])
@pytest.mark.parametrize("split", [False, True])
def test_round_trip(L, codec, shuffle, split):
    data = _data(1, 5000, np.uint16)  # 10000 bytes: blocks of 4096 + a leftover
    buf = be.compress(data, typesize=2, blocksize=4096, codec=codec, shuffle=shuffle,
                      split=split)
    rc, got = _decode(L, buf, len(data))
    assert rc == 0, _lib.last_error()
    assert got == data


@pytest.mark.parametrize("ts,dtype", [(1, np.uint8), (4, np.int32), (8, np.float64)])
@pytest.mark.parametrize("version", [2, 3])
def test_bitshuffle_leftovers(L, ts, dtype, version):
    # 1003 elements: the last block's element count is not a multiple of 8
    data = _data(2, 1003, dtype)
    buf = be.compress(data, typesize=ts, blocksize=512 * ts, codec="zstd", shuffle=2,
                      version=version)
    rc, got = _decode(L, buf, len(data))
    assert rc == 0, _lib.last_error()
    assert got == data


def test_bitshuffle_layout_by_hand(L):
    """Two uint16 elements per bit row are too few; 8 elements: the
    shuffled block is 16 rows of one byte, row (j, k) holding bit k of byte
    j of elements 0..7 (element e at bit e)."""
    el = np.array([0x0001, 0x0100, 0x8000, 0x00FF, 0, 0xFFFF, 0x1234, 0x0002], np.uint16)
    rows = np.zeros(16, np.uint8)
    for j in range(2):
        for k in range(8):
            for e in range(8):
                rows[8 * j + k] |= ((int(el[e]) >> (8 * j + k)) & 1) << e
    assert be.bit_shuffle(el.tobytes(), 2, 2) == rows.tobytes()
    hdr = bytes([2, 1, 0x4 | 0x10 | (4 << 5), 2]) + np.array([16, 16, 16 + 4 + 4 + 16],
                                                              np.uint32).tobytes()
    buf = hdr + np.array([20], np.int32).tobytes() + np.array([16], np.int32).tobytes() + \
        rows.tobytes()  # one stored stream
    rc, got = _decode(L, buf, 16)
    assert rc == 0, _lib.last_error()
    assert got == el.tobytes()


def test_stored_chunk_and_errors(L):
    data = _data(3, 100, np.uint16)
    rc, got = _decode(L, be.compress(data, typesize=2, blocksize=256, memcpyed=True), 200)
    assert rc == 0 and got == data
    buf = be.compress(data, typesize=2, blocksize=256, codec="zstd")
    assert _decode(L, buf, 100)[0] != 0            # destination too small
    assert "holds" in _lib.last_error()
    assert _decode(L, buf[:-5], 200)[0] != 0       # truncated
    bad = bytearray(buf)
    bad[2] = (bad[2] & 0x1F) | (0 << 5)            # blosclz
    assert _decode(L, bytes(bad), 200)[0] != 0
    assert "blosclz" in _lib.last_error()


def test_many_chunks_on_threads(L):
    chunks = [_data(10 + i, 3000, np.uint16) for i in range(7)]
    bufs = [be.compress(c, typesize=2, blocksize=2048, codec="zstd") for c in chunks]
    srcs = (C.c_void_p * 7)(*[C.cast(C.c_char_p(b), C.c_void_p) for b in bufs])
    keep = [np.frombuffer(b, np.uint8).copy() for b in bufs]
    srcs = (C.c_void_p * 7)(*[k.ctypes.data for k in keep])
    sizes = (C.c_int64 * 7)(*[len(b) for b in bufs])
    outs = [np.zeros(6000, np.uint8) for _ in range(7)]
    dsts = (C.c_void_p * 7)(*[o.ctypes.data for o in outs])
    got = (C.c_int64 * 7)()
    assert L.mc_blosc_decompress_many(srcs, sizes, dsts, 6000, 7, 3, got) == 0
    for o, c, n in zip(outs, chunks, got):
        assert n == 6000 and o.tobytes() == c
    keep[4][20] ^= 0xFF  # corrupt one chunk's first block start
    keep[4][21] ^= 0xFF
    rc = L.mc_blosc_decompress_many(srcs, sizes, dsts, 6000, 7, 3, got)
    assert rc != 0 and got[4] == -1 and "chunk 4" in _lib.last_error()
