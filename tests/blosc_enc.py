"""Blosc1 chunk writer for the tests (test infrastructure, not product code).

Writes the container that c-blosc 1.x and numcodecs produce (c-blosc
README_HEADER.rst): a 16-byte header (version, versionlz, flags, typesize,
nbytes, blocksize, cbytes), int32 block starts, then per block one stream
(or `typesize` streams when split), each an int32 size and the codec's
bytes (the raw bytes when compression does not shrink them).  Shuffles are
written from their definitions: byte shuffle puts byte j of every element
in plane j; bitshuffle puts bit k of byte j of every element in bit row
(j, k), element e at bit e % 8 of byte e // 8.  Codecs: the system's
libzstd and liblz4 through ctypes, and Python's zlib.
"""
from __future__ import annotations

import ctypes as C
import struct
import zlib

import numpy as np

CODES = {"blosclz": 0, "lz4": 1, "snappy": 2, "zlib": 3, "zstd": 4}


def _zstd():
    L = C.CDLL("libzstd.so.1")
    L.ZSTD_compressBound.restype = C.c_size_t
    L.ZSTD_compressBound.argtypes = [C.c_size_t]
    L.ZSTD_compress.restype = C.c_size_t
    L.ZSTD_compress.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int]
    return L


def _lz4():
    L = C.CDLL("liblz4.so.1")
    L.LZ4_compressBound.restype = C.c_int
    L.LZ4_compressBound.argtypes = [C.c_int]
    L.LZ4_compress_default.restype = C.c_int
    L.LZ4_compress_default.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int]
    return L


def compress_stream(data: bytes, codec: str, level: int = 1) -> bytes:
    if codec == "zstd":
        L = _zstd()
        cap = L.ZSTD_compressBound(len(data))
        out = C.create_string_buffer(cap)
        n = L.ZSTD_compress(out, cap, data, len(data), level)
        return out.raw[:n]
    if codec == "lz4":
        L = _lz4()
        cap = L.LZ4_compressBound(len(data))
        out = C.create_string_buffer(cap)
        n = L.LZ4_compress_default(data, out, len(data), cap)
        return out.raw[:n]
    if codec == "zlib":
        return zlib.compress(data, level)
    raise ValueError(codec)


def byte_shuffle(block: bytes, ts: int) -> bytes:
    a = np.frombuffer(block, np.uint8)
    n_el = len(a) // ts
    body = a[:n_el * ts].reshape(n_el, ts).T.reshape(-1)
    return body.tobytes() + a[n_el * ts:].tobytes()


def bit_shuffle(block: bytes, ts: int, version: int) -> bytes:
    a = np.frombuffer(block, np.uint8)
    n_el = len(a) // ts
    if version <= 2:
        if n_el % 8:
            return block
    else:
        n_el -= n_el % 8
    el = a[:n_el * ts].reshape(n_el, ts)
    rows = [np.packbits((el[:, j] >> k) & 1, bitorder="little")
            for j in range(ts) for k in range(8)]
    return b"".join(r.tobytes() for r in rows) + a[n_el * ts:].tobytes()


def compress(data: bytes, *, typesize: int, blocksize: int, codec: str = "zstd",
             shuffle: int = 2, split: bool = False, version: int = 2,
             memcpyed: bool = False, level: int = 1) -> bytes:
    """shuffle: 0 none, 1 byte, 2 bit (numcodecs' Blosc.NOSHUFFLE/SHUFFLE/BITSHUFFLE).
    Type sizes above 255 are written as 1, as c-blosc does."""
    if typesize > 255:
        typesize = 1
    nbytes = len(data)
    flags = CODES[codec] << 5
    if shuffle == 1:
        flags |= 0x1
    elif shuffle == 2:
        flags |= 0x4
    if not split:
        flags |= 0x10
    if memcpyed:
        flags |= 0x2
        hdr = struct.pack("<BBBBIII", version, 1, flags, typesize, nbytes, blocksize,
                          16 + nbytes)
        return hdr + data
    nblocks = (nbytes + blocksize - 1) // blocksize
    body = b""
    starts = []
    base = 16 + 4 * nblocks
    for b in range(nblocks):
        blk = data[b * blocksize:(b + 1) * blocksize]
        leftover = len(blk) < blocksize
        if shuffle == 1:
            blk = byte_shuffle(blk, typesize)
        elif shuffle == 2:
            blk = bit_shuffle(blk, typesize, version)
        nstreams = typesize if (split and not leftover) else 1
        ne = len(blk) // nstreams
        starts.append(base + len(body))
        for s in range(nstreams):
            part = blk[s * ne:(s + 1) * ne]
            c = compress_stream(part, codec, level)
            if len(c) >= ne:  # incompressible: stored raw
                c = part
            body += struct.pack("<i", len(c)) + c
    hdr = struct.pack("<BBBBIII", version, 1, flags, typesize, nbytes, blocksize,
                      16 + 4 * nblocks + len(body))
    return hdr + struct.pack(f"<{nblocks}i", *starts) + body
