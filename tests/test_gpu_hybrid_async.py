"""Hybrid frames (wire dtype 6: uint16 base plus int64 overflow patches)
decoded on the GPU through mc_integrate (IntegratorFitPeaksGPUStream.cu
:1468-1494), the asynchronous integrate with its event and host callback,
and build_csr(compute_variance=True)."""
import ctypes as C
import socket
import threading
import time

import numpy as np
import pytest

import midas_amd as ma
from helpers import rel_err
from midas_amd import _lib
from midas_amd.params import IntegrationParams
from midas_amd.server import FrameServer
from helpers import small_cases

pytestmark = pytest.mark.gpu


def _frames(P, n, seed):
    rng = np.random.default_rng(seed)
    f = rng.poisson(300.0, (n, P.NrPixelsZ, P.NrPixelsY)).astype(np.int32)
    hot = rng.choice(f[0].size, 40, replace=False)
    for i in range(n):  # counts past 16 bits on a few pixels
        f[i].reshape(-1)[hot] = rng.integers(70000, 5_000_000, hot.size)
    return f


def test_hybrid_messages_decode_on_the_gpu():
    P = small_cases()["tilt"]
    g = ma.build_map(P)
    dark = np.random.default_rng(2).poisson(20, (P.NrPixelsZ, P.NrPixelsY)).astype(np.int32)
    g.set_dark(dark)
    f = _frames(P, 5, 1)
    msgs = b"".join(ma.hybrid_message(x) for x in f)
    want_c, want_l = ma.integrate_batch(f, g, mode="bilinear", mask_nan=True)
    got_c, got_l = ma.integrate_hybrid(msgs, len(f), g, mode="bilinear", mask_nan=True)
    np.testing.assert_array_equal(got_c, want_c)
    np.testing.assert_array_equal(got_l, want_l)
    torch = pytest.importorskip("torch")
    dm = torch.from_numpy(np.frombuffer(msgs, dtype=np.uint8).copy()).cuda()
    dev_c, dev_l = ma.integrate_hybrid(dm, len(f), g, mode="bilinear", mask_nan=True)
    np.testing.assert_array_equal(dev_c, want_c)
    one_c, _ = ma.integrate_hybrid(ma.hybrid_message(f[3]), 1, g, mode="bilinear",
                                   mask_nan=True)
    alone, _ = ma.integrate_batch(f[3], g, mode="bilinear", mask_nan=True)
    np.testing.assert_array_equal(one_c[0], alone[0])


def test_hybrid_frames_over_tcp():
    """A dtype 6 stream to an int32 geometry integrates like the decoded
    frames; the same stream to a uint16 geometry is refused (its overflow
    values would wrap)."""
    P = small_cases()["flat"]
    g = ma.build_map(P)
    f = _frames(P, 3, 4)
    with FrameServer(hbm_budget_bytes=4 << 30, max_batch=4) as srv:
        gi = srv.add_geometry(g, np.int32)
        gu = srv.add_geometry(g, np.uint16)
        host, port = srv.serve_tcp(lambda ds: gi if ds == 0 else gu, port=0)
        with socket.create_connection((host, port)) as s:
            for x in f:
                s.sendall(np.array([0, 6], dtype="<u2").tobytes() + ma.hybrid_message(x))
        with socket.create_connection((host, port)) as s:
            s.sendall(np.array([1, 6], dtype="<u2").tobytes() + ma.hybrid_message(f[0]))
        t0 = time.time()
        while (len(srv.results) < 3 or not srv.wire_errors) and time.time() - t0 < 30:
            time.sleep(0.05)
        srv.drain()
        srv.stop_tcp()
        assert len(srv.results) == 3
        assert len(srv.wire_errors) == 1 and "int32" in srv.wire_errors[0][2]
        for i, rid in enumerate(sorted(srv.results)):
            alone, _ = ma.integrate_batch(f[i], g, mode="bilinear", mask_nan=True)
            assert rel_err(srv.results[rid].cake, alone[0], alone[0]).max() <= 1e-5


def test_async_integrate_event_and_callback():
    torch = pytest.importorskip("torch")
    P = small_cases()["tilt"]
    g = ma.build_map(P)
    rng = np.random.default_rng(5)
    g.set_dark(rng.poisson(20, (P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16))
    f = rng.poisson(300, (9, P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    want_c, want_l = ma.integrate_batch(f, g, mode="bilinear", mask_nan=True)
    df = torch.from_numpy(f).cuda()
    cakes = torch.full((9, g.n_r, g.n_eta), -1.0, device="cuda")
    lo = torch.full((9, 3, g.n_r), -1.0, device="cuda")
    fired = threading.Event()
    cb = _lib.HOST_FN(lambda user: fired.set())
    ev = ma.integrate_async(df, g, cakes, lo, mode="bilinear", mask_nan=True)
    _lib.check(_lib.lib().mc_ctx_callback(g.ctx.handle, cb, None))
    ev.wait()
    assert ev.done
    assert fired.wait(10)
    np.testing.assert_array_equal(cakes.cpu().numpy(), want_c)
    np.testing.assert_array_equal(lo.cpu().numpy(), want_l)
    ev.close()
    # one frame: the one-frame kernel, asynchronously
    c1 = torch.empty((1, g.n_r, g.n_eta), device="cuda")
    ev1 = ma.integrate_async(df[4:5], g, c1, mode="bilinear", mask_nan=True)
    ev1.wait()
    one, _ = ma.integrate_batch(f[4], g, mode="bilinear", mask_nan=True)
    np.testing.assert_array_equal(c1.cpu().numpy(), one)


def test_build_csr_with_variance():
    P = small_cases()["flat"]
    g0 = ma.build_map(P)
    pm = g0.download()
    g = ma.build_csr(pm, n_r=P.n_r_bins, n_eta=P.n_eta_bins, n_pixels_y=P.NrPixelsY,
                     n_pixels_z=P.NrPixelsZ, compute_variance=True)
    f = np.random.default_rng(6).poisson(100, (2, P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    c, v = ma.integrate_with_variance(f, g, mode="floor")
    c0, v0 = ma.integrate_with_variance(f, g0, mode="floor")
    np.testing.assert_array_equal(c, c0)
    np.testing.assert_array_equal(v, v0)
