"""Continuous-batching frame server (mc_server_*) and its TCP front end.

``FrameServer`` mirrors the reference package's server
(packages/midas_integrate/midas_integrate/server.py:224-333) and the
reference GPU stream's socket protocol (IntegratorFitPeaksGPUStream.cu
:568-689): each message is a 4-byte header <u16 dataset_num, u16 dtype>
followed by one frame of the registered geometry (dtype 6, "hybrid", adds
a u32 overflow count after the header, then u16[N] + i32[C] + i64[C]).
"""
from __future__ import annotations

import ctypes as C
import socket
import struct
import threading
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib
from .caking import CSRGeometry, Context, _flags, default_context

PORT = 60439
WIRE_BYTES = {0: 1, 1: 2, 2: 4, 3: 8, 4: 4, 5: 8, 7: 4}
WIRE_NUMPY = {0: np.uint8, 1: np.uint16, 2: np.uint32, 3: np.int64, 4: np.float32,
              5: np.float64, 7: np.int32}


@dataclass
class Result:
    request_id: int
    geometry: int
    cake: np.ndarray | None
    lineouts: np.ndarray | None
    status: int


class FrameServer:
    """Frames of several registered geometries in, cakes out, batched on
    the GPU while earlier batches run."""

    def __init__(self, hbm_budget_bytes: int = 64 << 30, max_batch: int = 64,
                 queue_depth: int = 512, keep_results: bool = True,
                 on_result=None, ctx: Context | None = None, batch_window_ms: float = 0.0):
        self.ctx = ctx or default_context()
        self.keep = keep_results
        self.on_result = on_result
        self.results: dict[int, Result] = {}
        self._shapes: dict[int, tuple] = {}
        self._lock = threading.Lock()
        # Without a consumer of the results no Python callback is
        # registered: the server's completion thread then never waits for
        # the interpreter lock.
        self._cb = (_lib.RESULT_CB(self._callback) if (keep_results or on_result is not None)
                    else _lib.RESULT_CB())  # a NULL function pointer
        h = C.c_void_p()
        check(lib().mc_server_create(self.ctx.handle, hbm_budget_bytes, max_batch, queue_depth,
                                     self._cb, None, C.byref(h)))
        self.handle = h
        self._geoms: dict[int, tuple] = {}
        if batch_window_ms:
            self.set_batch_window(batch_window_ms)

    def set_batch_window(self, ms: float):
        """A free slot waits up to ``ms`` after the oldest queued request's
        arrival for a full batch of its geometry (0: launch at once)."""
        check(lib().mc_server_set_batch_window(self.handle, int(round(ms * 1000.0))))

    def _callback(self, user, rid, geo, cake, lo, status):
        shape = self._shapes[geo]
        c = l = None
        # Cakes are copied only when kept; an on_result consumer always gets
        # the lineouts (3 rows of n_r), e.g. for the stream's peak fits.
        if status == 0 and self.keep:
            c = np.ctypeslib.as_array(cake, shape=(shape[0] * shape[1],)).reshape(shape).copy()
        if status == 0 and (self.keep or self.on_result is not None):
            l = np.ctypeslib.as_array(lo, shape=(3 * shape[0],)).reshape(3, shape[0]).copy()
        r = Result(rid, geo, c, l, status)
        if self.on_result is not None:
            self.on_result(r)
        if self.keep or status != 0:
            with self._lock:
                self.results[rid] = r

    def add_geometry(self, geom: CSRGeometry, dtype=np.uint16, *, mode: str = "bilinear",
                     normalize: bool = True, mask_nan: bool = True) -> int:
        from .caking import NP_TO_CODE
        code = NP_TO_CODE[np.dtype(dtype)]
        gid = C.c_int32()
        check(lib().mc_server_add_geometry(self.handle, geom.handle, code, _lib.MODES[mode],
                                           _flags(normalize, mask_nan), C.byref(gid)))
        self._shapes[gid.value] = (geom.n_r, geom.n_eta)
        self._geoms[gid.value] = (geom, np.dtype(dtype))
        return gid.value

    def submit(self, geometry: int, frame, request_id: int, timeout_ms: int = 10000):
        geom, dt = self._geoms[geometry]
        a = np.ascontiguousarray(frame, dtype=dt)
        if a.size != geom.n_pixels_y * geom.n_pixels_z:
            raise ValueError("frame size does not match the geometry")
        check(lib().mc_server_submit(self.handle, geometry, a.ctypes.data, request_id,
                                     timeout_ms))

    def acquire(self, geometry: int) -> np.ndarray:
        """A pinned frame buffer of the geometry (shape (NZ, NY), its dtype)
        to fill in place and hand to :meth:`submit_buffer` (zero copy), or
        give back with :meth:`release`."""
        geom, dt = self._geoms[geometry]
        p = C.c_void_p()
        check(lib().mc_server_acquire_buffer(self.handle, geometry, C.byref(p)))
        n = geom.n_pixels_y * geom.n_pixels_z
        raw = (C.c_char * (n * dt.itemsize)).from_address(p.value)
        return np.frombuffer(raw, dtype=dt).reshape(geom.n_pixels_z, geom.n_pixels_y)

    def submit_buffer(self, geometry: int, buf: np.ndarray, request_id: int,
                      timeout_ms: int = 10000):
        """Queues an acquired buffer; the server owns it again."""
        check(lib().mc_server_submit_buffer(self.handle, geometry, buf.ctypes.data, request_id,
                                            timeout_ms))

    def release(self, geometry: int, buf: np.ndarray):
        check(lib().mc_server_release_buffer(self.handle, geometry, buf.ctypes.data))

    def drain(self, timeout_ms: int = 60000):
        check(lib().mc_server_drain(self.handle, timeout_ms))

    def stats(self) -> dict:
        s = _lib.ServerStats()
        check(lib().mc_server_stats_get(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def close(self):
        if self.handle:
            lib().mc_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---------------------------------------------------------------- TCP
    def serve_tcp(self, geometry_for_dataset, host: str = "127.0.0.1", port: int = PORT):
        """Accepts frame streams in the reference wire format; message i
        becomes request id (connection << 40) + i.  Returns the listening
        socket's (host, port) and runs the accept loop in a thread."""
        srv = socket.create_server((host, port))
        self._tcp = srv
        self._conn_id = 0
        self.wire_errors = []  # (connection, message index, reason) of refused streams

        def client(conn, cid):
            n = 0
            with conn:
                while True:
                    hdr = _recv_exact(conn, 4)
                    if hdr is None:
                        return
                    dataset, dtype = struct.unpack("<HH", hdr)
                    g = geometry_for_dataset(dataset)
                    geom, dt = self._geoms[g]
                    npx = geom.n_pixels_y * geom.n_pixels_z
                    why = wire_refusal(dtype, dt)
                    if why:
                        # a frame the geometry's pixel type cannot hold
                        # exactly: refuse the stream rather than wrap or
                        # truncate its values
                        self.wire_errors.append((cid, n, why))
                        return
                    if dtype == 6:
                        raw = _recv_exact(conn, 4)
                        if raw is None:
                            return
                        cnt = struct.unpack("<I", raw)[0]
                        body = _recv_exact(conn, npx * 2 + cnt * 12)
                        if body is None:
                            return
                        frame = decode_hybrid(body, npx, cnt)
                    elif dtype in WIRE_BYTES and np.dtype(WIRE_NUMPY[dtype]) == dt:
                        # the wire carries the geometry's own pixel type:
                        # receive straight into a pinned buffer (zero copy)
                        buf = self.acquire(g)
                        if not _recv_into(conn, memoryview(buf).cast("B")):
                            self.release(g, buf)
                            return  # disconnect mid-payload: the frame is dropped
                        self.submit_buffer(g, buf, (cid << 40) + n)
                        n += 1
                        continue
                    elif dtype in WIRE_BYTES:
                        body = _recv_exact(conn, npx * WIRE_BYTES[dtype])
                        if body is None:
                            return  # disconnect mid-payload: the frame is dropped
                        frame = np.frombuffer(body, dtype=WIRE_NUMPY[dtype])
                    else:
                        return  # unknown dtype: close the connection
                    self.submit(g, frame.astype(dt, copy=False), (cid << 40) + n)
                    n += 1

        def accept():
            while getattr(self, "_tcp", None) is not None:
                try:
                    conn, _ = srv.accept()
                except OSError:
                    return
                self._conn_id += 1
                threading.Thread(target=client, args=(conn, self._conn_id), daemon=True).start()

        threading.Thread(target=accept, daemon=True).start()
        return srv.getsockname()

    def stop_tcp(self):
        s = getattr(self, "_tcp", None)
        self._tcp = None
        if s is not None:
            s.close()


def _recv_into(conn, view) -> bool:
    got = 0
    while got < len(view):
        k = conn.recv_into(view[got:])
        if k == 0:
            return False
        got += k
    return True


def _recv_exact(conn, n):
    buf = bytearray()
    while len(buf) < n:
        chunk = conn.recv(n - len(buf))
        if not chunk:
            return None
        buf += chunk
    return bytes(buf)


def wire_refusal(code: int, dt: np.dtype) -> str | None:
    """Why a wire dtype cannot feed a geometry registered with pixel type
    dt, or None.  The frame is converted to dt on the host, so the
    conversion must keep every value: numpy's safe casting for the fixed
    types; for hybrid frames (uint16 base plus int64 overflow values) a
    signed or floating type of at least 32 bits."""
    if code == 6:
        if dt.kind in "if" and dt.itemsize >= 4:
            return None
        return (f"hybrid frames (dtype 6) carry overflow counts above 65535: register the "
                f"geometry as int32, int64, float32 or float64, not {dt}")
    if code not in WIRE_NUMPY:
        return f"unknown wire dtype {code}"
    src = np.dtype(WIRE_NUMPY[code])
    if src == dt or np.can_cast(src, dt, casting="safe"):
        return None
    return f"wire dtype {code} ({src}) does not convert exactly to the geometry's {dt}"


def decode_hybrid(body: bytes, npx: int, count: int) -> np.ndarray:
    """dtype 6: u16 base image + (i32 index, i64 value) overflow patches
    (IntegratorFitPeaksGPUStream.cu:1468-1494)."""
    base = np.frombuffer(body, dtype=np.uint16, count=npx).astype(np.int64)
    idx = np.frombuffer(body, dtype=np.int32, count=count, offset=npx * 2)
    val = np.frombuffer(body, dtype=np.int64, count=count, offset=npx * 2 + count * 4)
    base[idx] = val
    return base


def send_frame(sock, frame: np.ndarray, dataset: int = 0, dtype: int | None = None):
    """Client side of the wire format (server.py:339-375 of the reference)."""
    inv = {np.dtype(v): k for k, v in WIRE_NUMPY.items()}
    code = inv[frame.dtype] if dtype is None else dtype
    sock.sendall(struct.pack("<HH", dataset, code) + np.ascontiguousarray(frame).tobytes())
