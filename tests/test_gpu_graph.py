"""The batched passes captured in a HIP graph (mc_graph_*) and replayed.

Each replay runs the same kernels on the same buffers as an eager
integrate_batch call, so the graph's cakes and lineouts must equal the
eager ones bit for bit; the eager call itself is checked against the FP64
oracle (1e-5 relative, as in test_gpu_integrate).  Replays alternate with
eager integrations, refills of the frame buffer and synchronisations, and
every replay is compared.
"""
import numpy as np
import pytest

import midas_amd as ma
from midas_amd._lib import MidasError
from helpers import rel_err, small_cases, to_oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _case(n_frames, seed):
    P = small_cases()["tilt"]
    rng = np.random.default_rng(seed)
    mask = (rng.random((P.NrPixelsZ, P.NrPixelsY)) < 0.01).astype(np.float64)
    frames = rng.poisson(200.0, (n_frames, P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    dark = rng.poisson(20.0, (2, P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    return P, mask, frames, dark, rng


@pytest.mark.parametrize("n_frames", [5, 70])
def test_graph_replays_equal_eager_calls(n_frames, oracle_lib):
    P, mask, frames, dark, rng = _case(n_frames, seed=21)
    print(f"[graph] tilt, {n_frames} frames", flush=True)
    g = ma.build_map(P, mask=mask)
    g.set_dark(dark)
    buf = torch.from_numpy(frames).cuda()
    gr = ma.IntegrationGraph(g, buf, mode="bilinear", mask_nan=True)
    # The eager result on the same frames, pinned to the oracle on frame 0.
    ce, le = ma.integrate_batch(frames, g, mode="bilinear", mask_nan=True)
    om = oracle_lib.build_map(to_oracle(P), mask=mask)
    img = frames[0].astype(np.float64) - dark.astype(np.float64).mean(axis=0)
    ref, _ = oracle_lib.integrate(om, img, P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z),
                                  normalize=True, semantics=0)
    sc, _ = oracle_lib.integrate(om, np.abs(img), P.NrPixelsY, P.NrPixelsZ,
                                 bc=(P.BC_y, P.BC_z), normalize=True, semantics=0)
    assert rel_err(ce[0], ref, sc).max() <= 1e-5
    for rep in range(4):
        if rep:
            frames = rng.poisson(100.0 + 50 * rep, frames.shape).astype(np.uint16)
            buf.copy_(torch.from_numpy(frames))
        c, l = gr.launch()
        gc, gl = c.cpu().numpy(), l.cpu().numpy()
        # eager work on the same stream between replays
        ce, le = ma.integrate_batch(frames, g, mode="bilinear", mask_nan=True)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(gc, ce, err_msg=f"replay {rep}")
        np.testing.assert_array_equal(gl, le, err_msg=f"replay {rep} lineouts")
    gr.close()


def test_graph_follows_a_new_dark_and_refuses_a_moved_beam():
    P, mask, frames, dark, rng = _case(9, seed=22)
    print("[graph] dark change, tuning change, beam move", flush=True)
    g = ma.build_map(P, mask=mask)
    g.set_dark(dark)
    buf = torch.from_numpy(frames).cuda()
    gr = ma.IntegrationGraph(g, buf, mode="gradient", normalize=False, lineouts=False)
    assert gr.lineouts is None
    c0 = gr.launch()[0].cpu().numpy()
    e0, _ = ma.integrate_batch(frames, g, mode="gradient", normalize=False)
    np.testing.assert_array_equal(c0, e0)
    g.set_dark(dark[:1] + 7)
    c1 = gr.launch()[0].cpu().numpy()
    e1, _ = ma.integrate_batch(frames, g, mode="gradient", normalize=False)
    np.testing.assert_array_equal(c1, e1)
    assert not np.array_equal(c1, c0, equal_nan=True)
    g.set_dark(None)
    c2 = gr.launch()[0].cpu().numpy()
    e2, _ = ma.integrate_batch(frames, g, mode="gradient", normalize=False)
    np.testing.assert_array_equal(c2, e2)
    # a new tuning builds a second layout; the graph keeps its own
    g.set_tuning(batch=32, rows_per_group=32, lds_bytes=65536)
    ma.integrate_batch(frames, g, mode="gradient", normalize=False)
    np.testing.assert_array_equal(gr.launch()[0].cpu().numpy(), c2)
    # a new beam centre rebuilds the gradient map and its layouts
    g.set_beam_center(P.BC_y + 0.5, P.BC_z)
    ma.integrate_batch(frames, g, mode="gradient", normalize=False)
    with pytest.raises(MidasError, match="changed since the graph was captured"):
        gr.launch()
    gr.close()


def test_graph_timing_kind():
    P, mask, frames, dark, _ = _case(64, seed=23)
    print("[graph] timing", flush=True)
    g = ma.build_map(P, mask=mask)
    buf = torch.from_numpy(frames).cuda()
    per_launch, per_iter = ma.time_integrate(g, buf, kind="graph", iters=3, warmup=1)
    assert 0 < per_launch <= per_iter
