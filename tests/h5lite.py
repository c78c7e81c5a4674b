"""Minimal HDF5 reader for the tests (test infrastructure, not product code).

Reads the subset of the HDF5 file format (HDF Group, "HDF5 File Format
Specification Version 3.0") that libhdf5 writes by default and that the
caked outputs use:

* superblock version 0 (8-byte offsets and lengths),
* version-1 object headers with continuation blocks,
* groups as symbol tables: version-1 B-trees of group nodes (any depth),
  symbol-table nodes and local heaps,
* dataspaces (versions 1 and 2), datatypes of class 0 (integers),
  1 (IEEE floats), 3 (fixed strings) and 9 (variable-length strings held in
  global heap collections),
* contiguous and compact data layouts (layout message version 3),
* attribute messages versions 1 to 3.

It is checked against a file written by libhdf5 (the reference's
``FF_HEDM/Example/consolidated_Output.h5``, kept as a fixture), and the
native writer's files are read back with it.
"""
from __future__ import annotations

import struct

import numpy as np

SIG = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


class H5Error(ValueError):
    pass


class Dataset:
    def __init__(self, f, addr, msgs):
        self.f, self.addr = f, addr
        self.shape = ()
        self.dtype = None
        self._layout = None
        self.attrs = {}
        for mt, body in msgs:
            if mt == 0x1:
                self.shape = f._dataspace(body)
            elif mt == 0x3:
                self.dtype = f._datatype(body)
            elif mt == 0x8:
                self._layout = body
            elif mt == 0xC:
                name, val = f._attribute(body)
                self.attrs[name] = val

    def read(self):
        lay = self._layout
        if lay is None or lay[0] != 3:
            raise H5Error("layout message version 3 expected")
        cls = lay[1]
        n = int(np.prod(self.shape)) if self.shape else 1
        kind, size = self.dtype[0], self.dtype[1]
        nbytes = n * size
        if cls == 1:  # contiguous
            a, ln = struct.unpack_from("<QQ", lay, 2)
            if a == UNDEF:
                raw = b"\0" * nbytes
            else:
                if ln != nbytes:
                    raise H5Error(f"contiguous size {ln} != {nbytes}")
                raw = self.f.d[a:a + ln]
        elif cls == 0:  # compact
            ln = struct.unpack_from("<H", lay, 2)[0]
            raw = lay[4:4 + ln]
        else:
            raise H5Error(f"layout class {cls} not supported")
        return self.f._decode(raw, self.dtype, self.shape)


class File:
    def __init__(self, path):
        with open(path, "rb") as fh:
            self.d = fh.read()
        d = self.d
        if d[:8] != SIG:
            raise H5Error("not an HDF5 file")
        if d[8] != 0:
            raise H5Error(f"superblock version {d[8]} not supported")
        if d[13] != 8 or d[14] != 8:
            raise H5Error("8-byte offsets and lengths expected")
        self.leaf_k, self.internal_k = struct.unpack_from("<HH", d, 16)
        base, _free, self.eof, _drv = struct.unpack_from("<QQQQ", d, 24)
        if base != 0:
            raise H5Error("base address 0 expected")
        _lno, self.root_addr, _ct = struct.unpack_from("<QQI", d, 56)
        self._gcol = {}

    # ---- structure ----
    def _object_header(self, addr):
        d = self.d
        if d[addr] != 1:
            raise H5Error(f"object header version {d[addr]} at {addr}")
        nmsg, _refc, hsize = struct.unpack_from("<HIH", d, addr + 2)
        msgs = []
        blocks = [(addr + 16, hsize)]
        while blocks:
            p, size = blocks.pop(0)
            end = p + size
            while p + 8 <= end:
                mt, ms, _fl = struct.unpack_from("<HHB", d, p)
                body = d[p + 8:p + 8 + ms]
                if mt == 0x10:
                    blocks.append(struct.unpack_from("<QQ", body, 0))
                elif mt != 0:
                    msgs.append((mt, body))
                p += 8 + ms
        if len(msgs) > nmsg:
            raise H5Error("more messages than the header declares")
        return msgs

    def _heap_string(self, heap, off):
        d = self.d
        if d[heap:heap + 4] != b"HEAP":
            raise H5Error("local heap signature")
        _size, _free, data = struct.unpack_from("<QQQ", d, heap + 8)
        start = data + off
        return d[start:d.index(b"\0", start)].decode()

    def _group_entries(self, oh_addr):
        msgs = self._object_header(oh_addr)
        st = [b for mt, b in msgs if mt == 0x11]
        if not st:
            return None, msgs
        bt, heap = struct.unpack_from("<QQ", st[0], 0)
        out = {}
        self._btree(bt, heap, out)
        return out, msgs

    def _btree(self, addr, heap, out, level_links=None):
        """Walks a group B-tree node and checks what libhdf5's name lookup
        relies on: child i holds exactly the names in (key i, key i+1]
        (byte order), and the nodes of each level are linked left to right.
        Returns the node's names in order."""
        d = self.d
        if d[addr:addr + 4] != b"TREE" or d[addr + 4] != 0:
            raise H5Error("group B-tree node expected")
        level, n = d[addr + 5], struct.unpack_from("<H", d, addr + 6)[0]
        left, right = struct.unpack_from("<QQ", d, addr + 8)
        if level_links is None:
            level_links = {}
        level_links.setdefault(level, []).append((addr, left, right))
        p = addr + 24
        keys, kids = [], []
        for _ in range(n):
            keys.append(struct.unpack_from("<Q", d, p)[0])
            kids.append(struct.unpack_from("<Q", d, p + 8)[0])
            p += 16
        keys.append(struct.unpack_from("<Q", d, p)[0])
        kn = [self._heap_string(heap, k).encode() for k in keys]
        names = []
        for i, c in enumerate(kids):
            sub = (self._btree(c, heap, out, level_links) if level > 0
                   else self._snod(c, heap, out))
            if sub and not (kn[i] < sub[0].encode() and sub[-1].encode() <= kn[i + 1]):
                raise H5Error(f"B-tree keys do not bound child {i} at {addr}")
            names += sub
        if level_links is not None and level == max(level_links):
            for lv, nodes in level_links.items():
                for j, (a, lft, rgt) in enumerate(nodes):
                    want_l = nodes[j - 1][0] if j > 0 else UNDEF
                    want_r = nodes[j + 1][0] if j + 1 < len(nodes) else UNDEF
                    if (lft, rgt) != (want_l, want_r):
                        raise H5Error(f"B-tree level {lv} siblings are not linked")
        return names

    def _snod(self, addr, heap, out):
        d = self.d
        if d[addr:addr + 4] != b"SNOD":
            raise H5Error("symbol table node signature")
        n = struct.unpack_from("<H", d, addr + 6)[0]
        p = addr + 8
        prev = None
        names = []
        for _ in range(n):
            lno, oh, _ct = struct.unpack_from("<QQI", d, p)
            p += 40
            name = self._heap_string(heap, lno)
            if prev is not None and name.encode() <= prev.encode():
                raise H5Error("symbol table entries not in name order")
            prev = name
            out[name] = oh
            names.append(name)
        return names

    def _resolve(self, path):
        addr = self.root_addr
        for part in [p for p in path.split("/") if p]:
            ents, _ = self._group_entries(addr)
            if ents is None or part not in ents:
                raise KeyError(path)
            addr = ents[part]
        return addr

    # ---- messages ----
    def _dataspace(self, b):
        v, rank, flags = b[0], b[1], b[2]
        if v == 1:
            off = 8
        elif v == 2:
            if b[3] == 0:  # scalar
                return ()
            if b[3] == 2:
                return (0,)
            off = 4
        else:
            raise H5Error(f"dataspace version {v}")
        return tuple(struct.unpack_from("<" + "Q" * rank, b, off))

    def _datatype(self, b):
        cls, ver = b[0] & 15, b[0] >> 4
        bits = b[1] | (b[2] << 8) | (b[3] << 16)
        size = struct.unpack_from("<I", b, 4)[0]
        if ver not in (1, 2, 3):
            raise H5Error(f"datatype version {ver}")
        if cls == 0:
            if bits & 1:
                raise H5Error("big-endian integers not supported")
            signed = bool(bits & 8)
            return ("i" if signed else "u", size)
        if cls == 1:
            if bits & 1:
                raise H5Error("big-endian floats not supported")
            return ("f", size)
        if cls == 3:
            return ("S", size)
        if cls == 9:
            if (bits & 15) != 1:
                raise H5Error("variable-length sequences not supported")
            return ("vlen_str", size)
        raise H5Error(f"datatype class {cls} not supported")

    def _attribute(self, b):
        v = b[0]
        if v == 1:
            nsize, tsize, ssize = struct.unpack_from("<HHH", b, 2)
            p = 8
            name = b[p:p + nsize].split(b"\0")[0].decode()
            p += (nsize + 7) & ~7
            dt = self._datatype(b[p:p + tsize])
            p += (tsize + 7) & ~7
            shape = self._dataspace(b[p:p + ssize])
            p += (ssize + 7) & ~7
        elif v in (2, 3):
            nsize, tsize, ssize = struct.unpack_from("<HHH", b, 2)
            p = 8 + (1 if v == 3 else 0)
            name = b[p:p + nsize].split(b"\0")[0].decode()
            p += nsize
            dt = self._datatype(b[p:p + tsize])
            p += tsize
            shape = self._dataspace(b[p:p + ssize])
            p += ssize
        else:
            raise H5Error(f"attribute version {v}")
        n = int(np.prod(shape)) if shape else 1
        val = self._decode(b[p:p + n * dt[1]], dt, shape)
        return name, val

    def _gheap_object(self, coll, idx):
        if coll not in self._gcol:
            d = self.d
            if d[coll:coll + 4] != b"GCOL":
                raise H5Error("global heap signature")
            size = struct.unpack_from("<Q", d, coll + 8)[0]
            objs = {}
            p, end = coll + 16, coll + size
            while p + 16 <= end:
                oi, _rc = struct.unpack_from("<HH", d, p)
                osz = struct.unpack_from("<Q", d, p + 8)[0]
                if oi == 0:
                    break
                objs[oi] = d[p + 16:p + 16 + osz]
                p += 16 + ((osz + 7) & ~7)
            self._gcol[coll] = objs
        return self._gcol[coll][idx]

    def _decode(self, raw, dt, shape):
        kind, size = dt
        if kind in ("i", "u", "f"):
            a = np.frombuffer(raw, dtype=np.dtype(f"<{kind}{size}"))
        elif kind == "S":
            a = np.array([raw[i * size:(i + 1) * size].split(b"\0")[0].decode()
                          for i in range(len(raw) // size)], dtype=object)
        else:  # variable-length string: u32 length, u64 collection, u32 index
            out = []
            for i in range(len(raw) // 16):
                ln, coll, idx = struct.unpack_from("<IQI", raw, 16 * i)
                out.append(self._gheap_object(coll, idx)[:ln].decode() if ln else "")
            a = np.array(out, dtype=object)
        return a.reshape(shape) if shape else a[0]

    # ---- public ----
    def keys(self, path="/"):
        ents, _ = self._group_entries(self._resolve(path))
        if ents is None:
            raise KeyError(f"{path} is not a group")
        return list(ents)

    def is_group(self, path):
        ents, _ = self._group_entries(self._resolve(path))
        return ents is not None

    def attrs(self, path):
        _, msgs = self._group_entries(self._resolve(path))
        out = {}
        for mt, b in msgs:
            if mt == 0xC:
                k, v = self._attribute(b)
                out[k] = v
        return out

    def dataset(self, path):
        addr = self._resolve(path)
        ents, msgs = self._group_entries(addr)
        if ents is not None:
            raise KeyError(f"{path} is a group")
        return Dataset(self, addr, msgs)

    def walk(self, path="/"):
        """Yields (path, is_group) for every object below `path`."""
        for k in self.keys(path):
            p = path.rstrip("/") + "/" + k
            g = self.is_group(p)
            yield p, g
            if g:
                yield from self.walk(p)
