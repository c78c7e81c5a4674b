#!/usr/bin/env python3
"""Throughput of the caking path on one MI355X.

Workload (batch_2048): 64 frames of 2048x2048 uint16 (12 Gaussian rings up
to 1000 px, dark 100 + Poisson(5), Poisson background 50), R 10-1000 px by
0.25 px x eta 1 deg (3960 x 360 bins), solid-angle + polarisation (0.99)
weights in the map, bilinear mode, Normalize=1, mask NaN, fused lineouts.
The map is built on the GPU by the FP64 mapper, frames are generated on the
GPU and stay resident: the timed region is integration only.

A step is one pass of the batched kernel over the 64 frames (one launch at
batch 64).  `value` = frames / s of the batched kernel; the one-frame kernel
(64 launches per step) is reported beside it, with per-frame latency
percentiles (host clock around a synchronous one-frame call), the HBM
roofline of the batched kernel and the FP64 CPU oracle (all host cores) on
a bounded sample of the same frames.

`roofline.traffic` is the measured HBM traffic per batched launch (the
transpose, k_batched and the lineout finalisation, as timed; FETCH_SIZE +
WRITE_SIZE from rocprofv3 --pmc runs of this script with
--copy-calibration, scaled by the counters' reading for a 1 GiB
elementwise kernel), read from profiles/traffic_batch_2048.json when that
file was measured for the same tuning and tile layout; otherwise null.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes_per_frame(n_px, n_bins, n_r, nnz, batch, frame_bytes=2):
    """bytes/frame = S_frame + S_out + (S_map + S_static)/B with
    S_frame = N_px * sizeof(pixel), S_out = n_bins*4 + 2*n_r*4,
    S_map = nnz*8 + (n_bins+1)*4 + n_bins*4 + n_bins, S_static = N_px*4."""
    s_frame = n_px * frame_bytes
    s_out = n_bins * 4 + 2 * n_r * 4
    s_map = nnz * 8 + (n_bins + 1) * 4 + n_bins * 4 + n_bins
    s_static = n_px * 4
    return s_frame + s_out + (s_map + s_static) / batch


def serving_bench(rates, n_frames, log):
    """Poisson arrivals mixing three geometries (1024^2 u16, 2048^2 u16,
    Pilatus 1679x1475 int32 with module gaps masked), each with its own
    resident map and dark; the server batches in-flight frames by map.
    Returns per rate: offered rate, achieved frames/s, p50/p99 latency (ms,
    submit to result ready, host frames over PCIe included)."""
    import numpy as np
    import torch

    import midas_amd as ma
    from midas_amd import synth
    from midas_amd.server import FrameServer

    geos = []
    for P, dt, mask in ((synth.plain_1024(), np.uint16, None),
                        (synth.batch_2048(), np.uint16, None),
                        (synth.pilatus(), np.int32, "gaps")):
        m = synth.pilatus_gap_mask(P) if mask == "gaps" else None
        g = ma.build_map(P, mask=m)
        fr, dk = synth.torch_frames(P, 8, seed=7, n_rings=8, device="cuda",
                                    dtype=torch.int32 if dt == np.int32 else None)
        g.set_dark(dk)
        geos.append((g, dt, fr.cpu().numpy().astype(dt)))
    rng = np.random.default_rng(11)
    # Warm-up: first launches of each kernel for each geometry (code-object
    # loading) happen here, outside the measured runs.
    with FrameServer(hbm_budget_bytes=64 << 30, max_batch=32, keep_results=False) as srv:
        gids = [srv.add_geometry(g, dt) for g, dt, _ in geos]
        for k in range(3):
            srv.submit(gids[k], geos[k][2][0], k)
            srv.drain()
            for i in range(24):
                srv.submit(gids[k], geos[k][2][i % 8], 100 + i)
            srv.drain()
    out = []
    for rate in rates:
        with FrameServer(hbm_budget_bytes=64 << 30, max_batch=32, queue_depth=256,
                         keep_results=False) as srv:
            gids = [srv.add_geometry(g, dt) for g, dt, _ in geos]
            which = rng.integers(0, 3, n_frames)
            gaps = rng.exponential(1.0 / rate, n_frames)
            t0 = time.perf_counter()
            t_next = t0
            for i in range(n_frames):
                t_next += gaps[i]
                # sleep (releasing the interpreter lock) until shortly before
                # the arrival, then spin to it
                dt = t_next - time.perf_counter()
                if dt > 3e-4:
                    time.sleep(dt - 2e-4)
                while time.perf_counter() < t_next:
                    pass
                k = int(which[i])
                srv.submit(gids[k], geos[k][2][i % 8], i)
            srv.drain()
            wall = time.perf_counter() - t0
            st = srv.stats()
        out.append({"offered_fps": rate, "frames": n_frames, "fps": n_frames / wall,
                    "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "max_ms": st["max_ms"],
                    "batches": st["batches"]})
        log(f"serving {rate} fps offered: {n_frames / wall:,.0f} fps, p50 {st['p50_ms']:.2f} ms, "
            f"p99 {st['p99_ms']:.2f} ms, {st['batches']} batches")
    return out


def fit_bench(g, P, frames, log, n_frames=32):
    """Integration + batched LM fits: 360 eta sectors x 10 rings = 3600
    windows of 41 radial points per frame, pV and TCH, one launch for all
    frames of the batch."""
    import torch

    import midas_amd as ma
    from midas_amd import peakfit, synth

    cakes, _ = ma.integrate_batch(frames[:n_frames], g, mode="bilinear", mask_nan=False)
    RLo, RHi, _, _ = ma.bin_edges(P)
    rc = (RLo + RHi) / 2
    rings = synth.ring_radii(P, 12)[:10]  # the first 10 of the frames' 12 rings
    res = {}
    for model in ("pv", "tch"):
        peakfit.fit_cake_rings(cakes, g, rc, rings, pad=20, model=model)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = peakfit.fit_cake_rings(cakes, g, rc, rings, pad=20, model=model)
        dt = time.perf_counter() - t0
        n = out.shape[0] * out.shape[1] * out.shape[2]
        ok = float((out[..., 7] > 0).mean())
        res[model] = {"fits": n, "fits_per_s": n / dt, "ms": dt * 1e3, "converged": ok}
        log(f"fits {model}: {n} windows in {dt * 1e3:.1f} ms -> {n / dt:,.0f} fits/s "
            f"({ok:.3f} converged)")
    return res


def qmode_bench(log, steps=5, warmup=2, serving=True):
    """The Q-mode workload (2880x2880 float32, 11500 x 720 Q x eta bins,
    SubPixelLevel 2, tilts, parallax, gradient mode): map build on the GPU,
    64-frame batched passes and the one-frame kernel, with the roofline of
    the batched launch (frames 4 bytes a pixel)."""
    import numpy as np
    import torch

    import midas_amd as ma
    from midas_amd import synth

    P = synth.qmode_2880()
    log("qmode_2880: building the map on the GPU")
    t0 = time.perf_counter()
    g = ma.build_map(P)
    t_map = time.perf_counter() - t0
    nnz = g.info().nnz
    frames, dark = synth.torch_qmode_frames(P, 64, seed=3)
    g.set_dark(dark)
    t0 = time.perf_counter()
    layout = g.layout_info(mode="gradient", dtype=np.float32)
    t_layout = time.perf_counter() - t0
    cakes = torch.empty((64, P.n_r_bins, P.n_eta_bins), dtype=torch.float32, device="cuda")
    lo = torch.empty((64, 3, P.n_r_bins), dtype=torch.float32, device="cuda")
    ms_l, ms_i = ma.time_integrate(g, frames, kind="batched", mode="gradient", iters=steps,
                                   warmup=warmup, cakes_dev=cakes, lineouts_dev=lo)
    ms_p, _ = ma.time_integrate(g, frames, kind="plain", mode="gradient", iters=1, warmup=1,
                                cakes_dev=cakes, lineouts_dev=lo)
    npx, nb = P.NrPixelsY * P.NrPixelsZ, P.n_r_bins * P.n_eta_bins
    bpf = algorithmic_bytes_per_frame(npx, nb, P.n_r_bins, nnz, 64, frame_bytes=4)
    achieved = bpf * 64 / (ms_l / 1e3) / 1e9
    fps = 64 / (ms_i / 1e3)
    log(f"qmode_2880: {nnz} map records ({t_map:.1f} s), layout {t_layout:.1f} s, "
        f"{fps:,.0f} frames/s batched ({ms_l * 1e3:.0f} us/launch), "
        f"one-frame {ms_p * 1e3:.0f} us, roofline {achieved / HBM_PEAK_GBS:.3f}")
    out = {"workload": f"qmode_2880: 64 x 2880x2880 float32 frames, {P.n_r_bins}x"
                       f"{P.n_eta_bins} Q x eta bins (Q 0.5-12 by 0.001, eta 0.5 deg), "
                       f"SubPixelLevel 2, tilts, parallax, gradient mode, map nnz {nnz}",
           "frames_per_s": fps, "ms_per_step": ms_i, "plain_fps": 1e3 / ms_p,
           "batched_over_plain": fps * ms_p / 1e3,
           "kernel_us": {"batched_per_launch": ms_l * 1e3, "plain_per_launch": ms_p * 1e3},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "bytes_per_frame": bpf},
           "map_build_s": t_map, "layout_s": t_layout, "layout": layout}
    out["serving"] = (qmode_serving_bench(g, P, frames[:8].cpu().numpy(), log)
                      if serving else None)
    g.close()
    del frames, cakes, lo
    torch.cuda.empty_cache()
    return out


def qmode_serving_bench(g, P, frames_host, log, n_frames=2000, rate=400.0, n_rate=1000):
    """The Q-mode workload streamed through the serving loop with fitting:
    host frames (float32, 2880^2, 33 MB each, the given frames in turn)
    submitted to a FrameServer, each frame's area-weighted lineout handed
    to a StreamPeakFitter that fits the 16 rings (TCH, one launch per
    queued group of frames).  Two runs: n_frames submitted back to back
    (frames/s from the first submit to the last fit; the latency is then
    queueing), and n_rate frames at Poisson arrivals of `rate` per second
    (submit-to-result latency)."""
    import tempfile

    import numpy as np

    import midas_amd as ma
    from midas_amd import synth
    from midas_amd.server import FrameServer
    from midas_amd.stream_fits import StreamPeakFitter

    RLo, RHi, _, _ = ma.bin_edges(P)
    rc = (RLo + RHi) / 2
    q = synth.ring_q()
    rings = (P.Lsd / P.pxY) * np.tan(2.0 * np.arcsin(q * P.Wavelength / (4.0 * np.pi)))
    rng = np.random.default_rng(5)

    def run(n, pace):
        with tempfile.TemporaryDirectory() as td:
            fitter = StreamPeakFitter(rc, float(np.median(RHi - RLo)), rings,
                                      fit_roi_padding=60, out_dir=td, max_batch=128)
            with FrameServer(hbm_budget_bytes=128 << 30, max_batch=16, queue_depth=48,
                             keep_results=False,
                             on_result=lambda r: fitter.submit(r.request_id,
                                                               r.lineouts[1])) as srv:
                gid = srv.add_geometry(g, np.float32, mode="gradient")
                for k in range(4):  # first launches of each kernel
                    srv.submit(gid, frames_host[k % len(frames_host)], -1 - k)
                srv.drain()
                gaps = rng.exponential(1.0 / pace, n) if pace else np.zeros(n)
                t0 = time.perf_counter()
                t_next = t0
                for i in range(n):
                    t_next += gaps[i]
                    dt = t_next - time.perf_counter()
                    if dt > 3e-4:
                        time.sleep(dt - 2e-4)
                    while time.perf_counter() < t_next:
                        pass
                    srv.submit(gid, frames_host[i % len(frames_host)], i, timeout_ms=60000)
                srv.drain()
                t_int = time.perf_counter() - t0
                st = srv.stats()
            fitter.close()
            wall = time.perf_counter() - t0
            fitted = [v for k, v in fitter.fitted.items() if k >= 0]
            n_fits = int(sum(int((v[:, 0] != 0).sum()) for v in fitted))
        return {"frames": n, "offered_fps": pace or None, "fps": n / wall,
                "fps_integration": n / t_int, "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"],
                "batches": st["batches"], "frames_fitted": len(fitted), "fits": n_fits,
                "fits_per_s": n_fits / wall}

    log(f"qmode_2880 serving: {n_frames} frames back to back, 16 rings fitted per frame")
    full = run(n_frames, 0.0)
    log(f"qmode_2880 serving: {full['fps']:,.0f} frames/s with fits "
        f"({full['fps_integration']:,.0f} integrated), {full['fits']} fits of "
        f"{full['frames_fitted']} frames")
    paced = run(n_rate, rate)
    log(f"qmode_2880 serving at {rate:.0f} frames/s offered: {paced['fps']:,.0f} frames/s, "
        f"p50 {paced['p50_ms']:.2f} ms, p99 {paced['p99_ms']:.2f} ms")
    return {"back_to_back": full, "paced": paced,
            "what": "host frames over PCIe (33 MB up, a 33 MB cake and the lineouts down per "
                    "frame), submit -> result latency; TCH fits of the 16 rings of each "
                    "frame's area-weighted lineout, off the frame path"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frames", type=int, default=64)
    ap.add_argument("--batch", type=int, default=0, help="manual tuning (skips the plan)")
    ap.add_argument("--rows", type=int, default=0, help="manual tuning (skips the plan)")
    ap.add_argument("--lds", type=int, default=0, help="manual tuning (skips the plan)")
    ap.add_argument("--plan-dir", default=str(ROOT / "midas_amd" / "plans"),
                    help="tuned plans (mc_map_autotune_plan): applied if present, else "
                         "tuned and written")
    ap.add_argument("--no-plan", action="store_true", help="built-in default tuning")
    ap.add_argument("--mode", default="bilinear")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="finish with a 1 GiB device copy (counter calibration)")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-frames", type=int, default=4000,
                    help="frames per offered rate (5 rates x 4000 = 20,000 frames)")
    ap.add_argument("--serving-rates", default="100,500,1000,2000,5000")
    ap.add_argument("--no-fits", action="store_true")
    ap.add_argument("--no-qmode", action="store_true")
    ap.add_argument("--no-qmode-serving", action="store_true",
                    help="skip the Q-mode workload's streaming leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(frames would shard by frame across GPUs, with no exchange)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import midas_amd as ma
    from midas_amd import synth

    P = synth.batch_2048()
    dev = torch.device("cuda:0")
    log("building the 2048^2 map on the GPU (FP64 mapper)")
    t0 = time.perf_counter()
    g = ma.build_map(P)
    t_map = time.perf_counter() - t0
    info = g.info()
    verify = g.verify_info()
    log(f"map: {info.nnz} entries, {P.n_r_bins}x{P.n_eta_bins} bins, {t_map:.2f} s; host "
        f"verification of {verify['pixels']} pixels: {verify['membership_differs']} membership "
        f"and {verify['values_differ']} value differences")
    frames, dark = synth.torch_frames(P, args.frames, seed=0, n_rings=12, device=dev)
    g.set_dark(dark)
    plan = None
    if args.batch or args.rows or args.lds:
        g.set_tuning(batch=args.batch, rows_per_group=args.rows, lds_bytes=args.lds)
    elif not args.no_plan:
        t0 = time.perf_counter()
        plan = g.autotune_plan(args.plan_dir, np.uint16)
        log(f"plan {plan['key']} {'loaded' if plan['loaded'] else 'tuned'} in "
            f"{time.perf_counter() - t0:.1f} s: batch {plan['batch']}, rows "
            f"{plan['rows_per_group']}, LDS {plan['lds_bytes']}, tile {plan['tile_bins']}/"
            f"{plan['tile_entries']}, one-frame lanes {plan['plain_lanes']}")
    B, H, L = g.tuning()
    n_bins = P.n_r_bins * P.n_eta_bins
    npx = P.NrPixelsY * P.NrPixelsZ
    cakes = torch.empty((args.frames, P.n_r_bins, P.n_eta_bins), dtype=torch.float32,
                        device=dev)
    lo = torch.empty((args.frames, 3, P.n_r_bins), dtype=torch.float32, device=dev)
    log("first integration (layout build)")
    ma.integrate_batch(frames[:2], g, mode=args.mode, mask_nan=True)
    ma.integrate_batch(frames[:1], g, mode=args.mode, mask_nan=True)

    # ---- batched: the timed steps ----
    log(f"warmup {args.warmup} + {args.steps} steps of {args.frames} frames (batch {B})")
    ma.time_integrate(g, frames, kind="batched", mode=args.mode, iters=1, warmup=args.warmup,
                      cakes_dev=cakes, lineouts_dev=lo)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms_launch_b, ms_iter_b = ma.time_integrate(g, frames, kind="batched", mode=args.mode,
                                               iters=args.steps, warmup=0, cakes_dev=cakes,
                                               lineouts_dev=lo)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    fps_b = args.frames * args.steps / wall
    log(f"batched: {fps_b:,.0f} frames/s wall, kernel {ms_launch_b*1e3:.1f} us/launch")

    # ---- the same steps replayed as one HIP graph each ----
    log("batched passes replayed as a HIP graph")
    _, ms_iter_g = ma.time_integrate(g, frames, kind="graph", mode=args.mode, iters=args.steps,
                                     warmup=args.warmup, cakes_dev=cakes, lineouts_dev=lo)
    fps_g = args.frames / (ms_iter_g / 1e3)
    log(f"graph: {fps_g:,.0f} frames/s ({ms_iter_g:.3f} ms per step)")

    # ---- one-frame kernel: latency distribution ----
    log("one-frame kernel")
    ms_launch_p, ms_iter_p = ma.time_integrate(g, frames, kind="plain", mode=args.mode,
                                               iters=args.plain_steps, warmup=1,
                                               cakes_dev=cakes, lineouts_dev=lo)
    fps_p = args.frames / (ms_iter_p / 1e3)
    lat = []
    for i in range(min(args.frames, 64)):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        t_a = time.perf_counter()
        ma.integrate_batch(frames[i], g, mode=args.mode, mask_nan=True, lineouts=True)
        lat.append((time.perf_counter() - t_a) * 1e3)
    p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))

    # ---- roofline of the batched kernel ----
    bpf = algorithmic_bytes_per_frame(npx, n_bins, P.n_r_bins, info.nnz, B)
    nf_launch = min(B, args.frames)
    achieved = bpf * nf_launch / (ms_launch_b / 1e3) / 1e9
    bpf_plain = algorithmic_bytes_per_frame(npx, n_bins, P.n_r_bins, info.nnz, 1)
    achieved_plain = bpf_plain / (ms_launch_p / 1e3) / 1e9

    # ---- CPU baseline: FP64 C oracle, all host cores ----
    cpu = None
    if not args.no_cpu:
        try:
            from oracle import oracle as orc
            log("CPU baseline (FP64 oracle, all cores)")
            pm = g.download()
            om = orc.MapArrays(pm.pxList, pm.pairs(), pm.mask_flags, P.n_r_bins, P.n_eta_bins)
            cores = os.cpu_count() or 1
            fr_h = frames[:4].cpu().numpy()
            dk = dark.cpu().numpy().astype(np.float64)
            done, t_cpu = 0, 0.0
            while t_cpu < args.cpu_seconds and done < 64:
                img = fr_h[done % len(fr_h)].astype(np.float64) - dk
                t_a = time.perf_counter()
                orc.integrate(om, img, P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z),
                              threads=cores)
                t_cpu += time.perf_counter() - t_a
                done += 1
            cpu = {"value": done / t_cpu, "unit": "frames/s", "cores": cores, "kind": "port",
                   "sample": f"{done} frames of the same workload, FP64 C restatement of "
                             "IntegratorZarrOMP.c:1710-1778 (OpenMP dynamic,64), "
                             "dark-subtraction excluded like t_integration"}
        except Exception as exc:  # the checker is optional on the box
            log(f"CPU baseline skipped: {exc}")

    fits = None if args.no_fits else fit_bench(g, P, frames, log)
    serving = None
    if not args.no_serving:
        rates = tuple(float(r) for r in args.serving_rates.split(",") if r)
        serving = serving_bench(rates, args.serving_frames, log)
    qmode = None if args.no_qmode else qmode_bench(log, serving=not args.no_qmode_serving)

    layout = g.layout_info(args.mode)
    traffic = None
    tf = ROOT / "profiles" / "traffic_batch_2048.json"
    if tf.exists():
        try:
            tj = json.loads(tf.read_text())
            same_layout = all(layout.get(k) == v for k, v in tj.get("layout", {}).items())
            if (tj.get("batch"), tj.get("rows_per_group"), tj.get("lds_bytes")) == (B, H, L) \
                    and same_layout:
                traffic = tj["hbm_bytes_per_launch"]
        except (ValueError, KeyError):
            traffic = None
    if args.copy_calibration:
        # an elementwise kernel that reads and writes exactly 1 GiB
        x = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        y = torch.empty_like(x)
        torch.mul(x, 2.0, out=y)
        torch.cuda.synchronize()
        del x, y

    out = {
        "metric": "frames/s (2048^2 uint16 -> RxEta cake) batched vs plain; p50/p99 frame latency",
        "value": fps_b,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Gaussian rings + Poisson noise, generated on the GPU)",
        "config": {"workload": f"batch_2048: {args.frames} x 2048x2048 uint16 frames, "
                               f"{P.n_r_bins}x{P.n_eta_bins} bins (R 0.25 px, eta 1 deg), "
                               f"solid angle + polarisation, {args.mode}, map nnz {info.nnz}",
                   "batch": B, "rows_per_group": H, "lds_bytes": L,
                   "frames_per_step": args.frames, "layout": layout, "plan": plan},
        "graph": {"frames_per_s": fps_g, "ms_per_step": ms_iter_g,
                  "what": "the step's batched passes replayed as one HIP graph (mc_graph_launch)"},
        "plain_fps": fps_p,
        "batched_over_plain": fps_b / fps_p,
        "latency_ms": {"p50": p50, "p99": p99, "what": "one frame through mc_integrate "
                       "(device-resident frame, host-synchronous call)"},
        "kernel_us": {"batched_per_launch": ms_launch_b * 1e3, "plain_per_launch": ms_launch_p * 1e3},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "bytes_per_frame": bpf, "plain_achieved": achieved_plain},
        "cpu_baseline": cpu,
        "fits": fits,
        "serving": serving,
        "qmode_2880": qmode,
        "map_build_s": t_map,
        "map_verify": verify,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
