"""Continuous-batching server: three geometries interleaved; every result
equals the same frame integrated alone (1e-5 relative, the batched and
one-frame kernels sum in different orders); admission by HBM budget;
the TCP wire format."""
import socket
import time

import numpy as np
import pytest

import midas_amd as ma
from midas_amd.params import IntegrationParams
from midas_amd.server import FrameServer, send_frame
from helpers import rel_err

pytestmark = pytest.mark.gpu


def _geoms():
    a = IntegrationParams(NrPixelsY=128, NrPixelsZ=128, pxY=200.0, pxZ=200.0, Lsd=1e6,
                          BC_y=64.3, BC_z=63.7, RMin=2.0, RMax=90.0, RBinSize=0.25,
                          EtaBinSize=5.0)
    b = IntegrationParams(NrPixelsY=256, NrPixelsZ=256, pxY=200.0, pxZ=200.0, Lsd=1e6,
                          BC_y=128.0, BC_z=128.0, RMin=2.0, RMax=180.0, RBinSize=0.5,
                          EtaBinSize=2.0, SolidAngleCorrection=1, PolarizationCorrection=1)
    c = IntegrationParams(NrPixelsY=150, NrPixelsZ=170, pxY=172.0, pxZ=172.0, Lsd=4e5,
                          BC_y=70.4, BC_z=84.2, RMin=2.0, RMax=100.0, RBinSize=0.5,
                          EtaBinSize=5.0, NPanelsY=2, NPanelsZ=2, PanelSizeY=72,
                          PanelSizeZ=80, PanelGapsY=[6], PanelGapsZ=[10])
    return [(a, np.uint16), (b, np.uint16), (c, np.int32)]


def test_mix_equals_alone():
    rng = np.random.default_rng(0)
    maps, frames = [], []
    for P, dt in _geoms():
        g = ma.build_map(P)
        g.set_dark(rng.poisson(10, (P.NrPixelsZ, P.NrPixelsY)).astype(dt))
        maps.append((g, dt))
        frames.append(rng.poisson(300, (40, P.NrPixelsZ, P.NrPixelsY)).astype(dt))
    with FrameServer(hbm_budget_bytes=8 << 30, max_batch=16, queue_depth=64) as srv:
        gids = [srv.add_geometry(g, dt) for g, dt in maps]
        order = [(k, i) for k in range(3) for i in range(40)]
        rng.shuffle(order)
        for rid, (k, i) in enumerate(order):
            srv.submit(gids[k], frames[k][i], rid)
        srv.drain()
        st = srv.stats()
        assert st["frames_done"] == len(order)
        assert 0 < st["p50_ms"] <= st["p99_ms"]
        assert st["batches"] < len(order)  # frames were batched
        for rid, (k, i) in enumerate(order):
            r = srv.results[rid]
            assert r.status == 0 and r.geometry == gids[k]
            alone, lo = ma.integrate_batch(frames[k][i], maps[k][0], mode="bilinear",
                                           mask_nan=True)
            assert rel_err(r.cake, alone[0], alone[0]).max() <= 1e-5
            assert rel_err(r.lineouts, lo[0], lo[0]).max() <= 1e-5


def test_on_result_gets_lineouts_without_kept_cakes():
    """keep_results=False with an on_result consumer (the stream's peak
    fits): each callback gets the frame's lineouts, no cake is copied."""
    P, dt = _geoms()[1]
    g = ma.build_map(P)
    rng = np.random.default_rng(3)
    frames = rng.poisson(200, (12, P.NrPixelsZ, P.NrPixelsY)).astype(dt)
    got = {}
    with FrameServer(hbm_budget_bytes=4 << 30, max_batch=8, keep_results=False,
                     on_result=lambda r: got.__setitem__(r.request_id, r)) as srv:
        gid = srv.add_geometry(g, dt)
        for i, f in enumerate(frames):
            srv.submit(gid, f, i)
        srv.drain()
    assert sorted(got) == list(range(len(frames)))
    for i, f in enumerate(frames):
        r = got[i]
        assert r.status == 0 and r.cake is None
        _, lo = ma.integrate_batch(f, g, mode="bilinear", mask_nan=True)
        assert rel_err(r.lineouts, lo[0], lo[0]).max() <= 1e-5


def test_budget_admission():
    P, dt = _geoms()[1]
    g = ma.build_map(P)
    with FrameServer(hbm_budget_bytes=1 << 20, max_batch=8) as srv:
        with pytest.raises(ma._lib.MidasError) as ei:
            srv.add_geometry(g, dt)
        assert ei.value.rc == -5
        assert srv.stats()["rejected"] == 1


def test_tcp_wire_protocol():
    P, dt = _geoms()[0]
    g = ma.build_map(P)
    rng = np.random.default_rng(1)
    frames = rng.poisson(100, (5, P.NrPixelsZ, P.NrPixelsY)).astype(dt)
    with FrameServer(hbm_budget_bytes=4 << 30, max_batch=4) as srv:
        gid = srv.add_geometry(g, dt)
        host, port = srv.serve_tcp(lambda ds: gid, port=0)
        with socket.create_connection((host, port)) as s:
            for f in frames:
                send_frame(s, f, dataset=0)
        t0 = time.time()
        while len(srv.results) < 5 and time.time() - t0 < 30:
            time.sleep(0.05)
        srv.drain()
        srv.stop_tcp()
        assert len(srv.results) == 5
        ids = sorted(srv.results)
        for i, rid in enumerate(ids):
            alone, _ = ma.integrate_batch(frames[i], g, mode="bilinear", mask_nan=True)
            assert rel_err(srv.results[rid].cake, alone[0], alone[0]).max() <= 1e-5


def test_zero_copy_buffers_and_large_frames():
    """Frames of 4.5 MiB (copied on 4 threads by submit) and frames written
    straight into acquired pinned buffers give the same cakes as
    integrating them alone; a released buffer is reused."""
    P = IntegrationParams(NrPixelsY=1536, NrPixelsZ=1536, pxY=200.0, pxZ=200.0, Lsd=1e6,
                          BC_y=768.2, BC_z=767.9, RMin=5.0, RMax=700.0, RBinSize=1.0,
                          EtaBinSize=5.0)
    g = ma.build_map(P)
    rng = np.random.default_rng(2)
    frames = rng.poisson(50, (6, P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    print("[server] zero-copy and 4.5 MiB frames", flush=True)
    with FrameServer(hbm_budget_bytes=16 << 30, max_batch=4) as srv:
        gid = srv.add_geometry(g, np.uint16)
        spare = srv.acquire(gid)
        assert spare.shape == (P.NrPixelsZ, P.NrPixelsY) and spare.dtype == np.uint16
        srv.release(gid, spare)
        for i in range(3):
            srv.submit(gid, frames[i], i)
        for i in range(3, 6):
            buf = srv.acquire(gid)
            buf[...] = frames[i]
            srv.submit_buffer(gid, buf, i)
        srv.drain()
        for i in range(6):
            alone, lo = ma.integrate_batch(frames[i], g, mode="bilinear", mask_nan=True)
            assert rel_err(srv.results[i].cake, alone[0], alone[0]).max() <= 1e-5


def test_concurrent_batches_of_one_geometry():
    """Batches of 7-16 frames of ONE geometry in flight on every slot at
    once: each batched pass writes lineout partials, and each slot owns its
    own, so every frame's cake and lineouts equal the frame integrated
    alone.  Frames differ in scale (frame i carries (1 + i/4) x the rings),
    so partials mixed between passes would show."""
    from midas_amd import synth
    P = synth.plain_1024()
    rng = np.random.default_rng(3)
    g = ma.build_map(P)
    dark = (100 + rng.poisson(5.0, (P.NrPixelsZ, P.NrPixelsY))).astype(np.uint16)
    g.set_dark(dark)
    base = synth.ring_image(P, synth.ring_radii(P, 8))
    n = 96
    frames = np.stack([dark + rng.poisson(50.0 + base * (1.0 + i / 4.0)) for i in range(n)])
    frames = frames.astype(np.uint16)
    print("[server] 96 frames of one 1024^2 geometry, batches of 16", flush=True)
    # A 200 ms batching window: free slots wait for full batches of 16
    with FrameServer(hbm_budget_bytes=16 << 30, max_batch=16, queue_depth=128,
                     batch_window_ms=200.0) as srv:
        gid = srv.add_geometry(g, np.uint16)
        for rep in range(2):
            for i in range(n):
                srv.submit(gid, frames[i], rep * n + i)
            srv.drain()
        st = srv.stats()
        assert st["frames_done"] == 2 * n
        assert st["batches"] <= 2 * n // 16 + 2  # batched passes of (nearly) 16 frames
        for rid in range(2 * n):
            i = rid % n
            r = srv.results[rid]
            assert r.status == 0
            alone, lo = ma.integrate_batch(frames[i], g, mode="bilinear", mask_nan=True)
            assert rel_err(r.cake, alone[0], alone[0]).max() <= 1e-5
            np.testing.assert_allclose(r.lineouts, lo[0], rtol=1e-5, atol=0)


def test_mix_equals_alone_real_geometries():
    """The serving workload's three geometries at size: 1024^2 uint16 with
    0.1 % hot pixels masked, 2048^2 uint16 (solid angle + polarisation),
    Pilatus 1679x1475 int32 with the module gaps masked; each with its own
    resident map and dark.  240 frames in a shuffled mix through the server
    (batching window 2 ms, batches of up to 64); every cake and lineout
    equals the same frame integrated alone by the one-frame kernel within
    1e-5 relative (no absolute floor), masked bins NaN in both."""
    from midas_amd import synth
    from helpers import rel_err
    print("[server] three real geometries: generating frames", flush=True)
    cases = []
    for P, dt, seed in ((synth.plain_1024(), np.uint16, 1), (synth.batch_2048(), np.uint16, 2),
                        (synth.pilatus(), np.int32, 3)):
        fr, dark, mask = synth.frames(P, 80, seed=seed, n_rings=8, dtype=dt)
        if P.NPanelsY:
            mask = np.maximum(mask, synth.pilatus_gap_mask(P))
        g = ma.build_map(P, mask=mask)
        g.set_dark(dark.astype(dt))
        cases.append((g, dt, fr))
    rng = np.random.default_rng(7)
    order = [(k, i) for k in range(3) for i in range(80)]
    rng.shuffle(order)
    print("[server] 240 frames through the server", flush=True)
    with FrameServer(hbm_budget_bytes=64 << 30, max_batch=64, queue_depth=512,
                     batch_window_ms=2.0) as srv:
        gids = [srv.add_geometry(g, dt) for g, dt, _ in cases]
        for rid, (k, i) in enumerate(order):
            srv.submit(gids[k], cases[k][2][i], rid)
        srv.drain()
        st = srv.stats()
        print(f"[server] {st['batches']} batches, p50 {st['p50_ms']:.2f} ms, "
              f"p99 {st['p99_ms']:.2f} ms", flush=True)
        assert st["frames_done"] == len(order) and st["batches"] < len(order)
        worst = [0.0, 0.0, 0.0]
        for rid, (k, i) in enumerate(order):
            r = srv.results[rid]
            assert r.status == 0 and r.geometry == gids[k]
            alone, lo = ma.integrate_batch(cases[k][2][i], cases[k][0], mode="bilinear",
                                           mask_nan=True)
            e = rel_err(r.cake, alone[0], alone[0])
            el = rel_err(r.lineouts, lo[0], lo[0])
            worst[k] = max(worst[k], float(e.max()), float(el.max()))
        print(f"[server] worst relative difference per geometry: {worst}", flush=True)
        assert max(worst) <= 1e-5
