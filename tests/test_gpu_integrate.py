"""HIP integration kernels (one-frame and batched) vs the FP64 oracle.

Tolerance: |gpu - oracle| <= 1e-5 * S_b per bin, with S_b the oracle's cake
of |frame - dark| (the sum of absolute terms, normalised like the bin); for
the non-negative frames used here S_b = |oracle|, i.e. 1e-5 relative.
NaN (masked) bins must coincide exactly.  Lineouts: within 1e-5 of the same
lineout reductions applied to the |frame - dark| cake.
"""
import numpy as np
import pytest

import midas_amd as ma
from midas_amd.params import IntegrationParams
from midas_amd import synth
from helpers import panel_shifts_for, rel_err, small_cases, to_oracle

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _setup(name, n_frames=3, seed=1, dtype=np.uint16):
    P = small_cases()[name]
    rng = np.random.default_rng(seed)
    mask = (rng.random((P.NrPixelsZ, P.NrPixelsY)) < 0.01).astype(np.float64)
    frames = rng.poisson(200.0, (n_frames, P.NrPixelsZ, P.NrPixelsY)).astype(dtype)
    dark = rng.poisson(20.0, (2, P.NrPixelsZ, P.NrPixelsY)).astype(dtype)
    return P, mask, frames, dark


def _oracle(orc, om, P, frame, dark_mean, mode, normalize=True):
    img = frame.astype(np.float64) - dark_mean
    grad = mode == "gradient"
    cake, _ = orc.integrate(om, img, P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z),
                            gradient=grad, normalize=normalize, semantics=0)
    scale, _ = orc.integrate(om, np.abs(img), P.NrPixelsY, P.NrPixelsZ, bc=(P.BC_y, P.BC_z),
                             gradient=grad, normalize=normalize, semantics=0)
    return cake, scale


@pytest.mark.parametrize("name", ["flat", "tilt", "subpix", "qmode", "panels"])
@pytest.mark.parametrize("mode", ["bilinear", "gradient"])
def test_one_frame_and_batched_match_oracle(name, mode, oracle_lib):
    P, mask, frames, dark = _setup(name)
    print(f"[integrate] case={name} mode={mode}", flush=True)
    sh = panel_shifts_for(name)
    g = ma.build_map(P, mask=mask, panel_shifts=sh)
    g.set_dark(dark)
    om = oracle_lib.build_map(to_oracle(P), mask=mask, panel_shifts=sh)
    dmean = dark.astype(np.float64).mean(axis=0)
    for F in (1, len(frames)):
        cakes, lo = ma.integrate_batch(frames[:F], g, mode=mode, mask_nan=True)
        for i in range(F):
            ref, scale_b = _oracle(oracle_lib, om, P, frames[i], dmean, mode)
            e = rel_err(cakes[i], ref, scale_b)
            assert e.max() <= TOL, f"F={F} frame {i}: max rel err {e.max():.3g}"
            ainb = oracle_lib.area_inbounds(om, P.NrPixelsY, P.NrPixelsZ)
            want = oracle_lib.lineouts(ref, ainb)
            # the same reductions of the |frame - dark| cake scale the error
            scale = oracle_lib.lineouts(np.where(np.isnan(ref), np.nan, scale_b), ainb)
            for row in range(3):
                el = np.abs(lo[i, row] - want[row]) / np.maximum(np.abs(scale[row]), 1e-30)
                assert el.max() <= TOL, f"F={F} frame {i} lineout row {row}: {el.max():.3g}"


def test_floor_mode_matches_gpu_stream_semantics(oracle_lib):
    """floor: integer pixel via the float offset of .cu:916, no NaN."""
    P, mask, frames, dark = _setup("subpix")
    g = ma.build_map(P, mask=mask)
    g.set_dark(dark)
    om = oracle_lib.build_map(to_oracle(P), mask=mask)
    dmean = dark.astype(np.float64).mean(axis=0)
    cakes, _ = ma.integrate_batch(frames, g, mode="floor", mask_nan=False)
    for i in range(len(frames)):
        img = frames[i].astype(np.float64) - dmean
        ref, _ = oracle_lib.integrate(om, img, P.NrPixelsY, P.NrPixelsZ, semantics=1)
        sc, _ = oracle_lib.integrate(om, np.abs(img), P.NrPixelsY, P.NrPixelsZ, semantics=1)
        assert rel_err(cakes[i], ref, sc).max() <= TOL


@pytest.mark.parametrize("batch", [8, 16, 32, 64, 128])
@pytest.mark.parametrize("rows", [16, 32])
def test_batched_tilings_agree(batch, rows):
    """Every batch size and band height gives the one-frame result,
    including a partial last pass (F not a multiple of the batch)."""
    P, mask, frames, dark = _setup("tilt", n_frames=batch + 3, seed=5)
    g = ma.build_map(P, mask=mask)
    g.set_dark(dark)
    g.set_tuning(batch=batch, rows_per_group=rows, lds_bytes=49152 if batch <= 64 else 81920)
    assert g.layout_info()["tiles"] > 0
    cb, lb = ma.integrate_batch(frames, g, mode="bilinear", mask_nan=True)
    for i in (0, batch - 1, batch, len(frames) - 1):
        c1, l1 = ma.integrate_batch(frames[i], g, mode="bilinear", mask_nan=True)
        np.testing.assert_allclose(cb[i], c1[0], rtol=3e-6, atol=1e-4, equal_nan=True)
        np.testing.assert_allclose(lb[i], l1[0], rtol=3e-6, atol=1e-4)


# The budgets leave room for 2048 staged entries (16 KiB) and about 46
# pixel rows of the pass width.
@pytest.mark.parametrize("batch,tight_lds", [(8, 16384 + 40 * 48), (64, 16384 + 128 * 48)])
def test_small_lds_budget_uses_direct_bins(batch, tight_lds):
    """Coarse bins under a small LDS budget: bins whose own pixel box does
    not fit become direct (global gathers) and columns split into row
    ranges; results still agree with the roomy layout.  Passes of 16 frames
    stage 32-byte pixel rows, 64-frame passes 128-byte rows: each gets a
    budget that leaves bins direct."""
    P = IntegrationParams(NrPixelsY=160, NrPixelsZ=144, pxY=200.0, pxZ=200.0, Lsd=1e6,
                          BC_y=80.5, BC_z=70.25, RMin=2.0, RMax=100.0, RBinSize=1.5,
                          EtaBinSize=30.0)
    rng = np.random.default_rng(11)
    frames = rng.poisson(200.0, (9, P.NrPixelsZ, P.NrPixelsY)).astype(np.uint16)
    g = ma.build_map(P)
    g.set_tuning(batch=batch, rows_per_group=16, lds_bytes=tight_lds)
    tight = g.layout_info()
    assert tight["direct_bins"] > 0
    cb, lb = ma.integrate_batch(frames, g, mode="bilinear")
    g.set_tuning(batch=batch, rows_per_group=32, lds_bytes=163840)
    assert g.layout_info()["direct_bins"] < tight["direct_bins"]
    cr, lr = ma.integrate_batch(frames, g, mode="bilinear")
    np.testing.assert_allclose(cb, cr, rtol=3e-6, atol=1e-4, equal_nan=True)
    np.testing.assert_allclose(lb, lr, rtol=3e-6, atol=1e-4)
    c1, _ = ma.integrate_batch(frames[4], g, mode="bilinear")
    np.testing.assert_allclose(cb[4], c1[0], rtol=3e-6, atol=1e-4, equal_nan=True)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32, np.int32, np.int64,
                                   np.float32, np.float64])
def test_frame_dtypes(dtype, oracle_lib):
    P, mask, frames, dark = _setup("centred", n_frames=2, seed=2, dtype=dtype)
    g = ma.build_map(P, mask=mask)
    g.set_dark(dark)
    om = oracle_lib.build_map(to_oracle(P), mask=mask)
    dmean = dark.astype(np.float64).mean(axis=0)
    cakes, _ = ma.integrate_batch(frames, g, mode="bilinear", mask_nan=True)
    for i in range(2):
        ref, sc = _oracle(oracle_lib, om, P, frames[i], dmean, "bilinear")
        assert rel_err(cakes[i], ref, sc).max() <= TOL


def test_normalize_off_and_no_dark(oracle_lib):
    P, mask, frames, _ = _setup("flat", n_frames=2)
    g = ma.build_map(P, mask=mask)
    om = oracle_lib.build_map(to_oracle(P), mask=mask)
    cakes, _ = ma.integrate_batch(frames, g, mode="bilinear", normalize=False, mask_nan=False)
    for i in range(2):
        img = frames[i].astype(np.float64)
        ref, _ = oracle_lib.integrate(om, img, P.NrPixelsY, P.NrPixelsZ, normalize=False)
        ref[np.isnan(ref)] = 0  # oracle CPU semantics NaN-masks; compare values
        flags = om.flags.reshape(ref.shape).astype(bool)
        refv, _ = oracle_lib.integrate(om, img, P.NrPixelsY, P.NrPixelsZ, normalize=False,
                                       semantics=1)
        e = np.abs(cakes[i] - refv) / np.maximum(np.abs(refv), 1e-30)
        assert e.max() <= TOL
        assert not np.isnan(cakes[i][flags]).any()


def test_plain_1024_workload(oracle_lib):
    """The 1024^2 uint16 workload with hot-pixel mask, against FP64."""
    P = synth.plain_1024()
    frames, dark, mask = synth.frames(P, 2, seed=0)
    print("[integrate] plain_1024", flush=True)
    g = ma.build_map(P, mask=mask)
    g.set_dark(dark)
    om = oracle_lib.build_map(to_oracle(P), mask=mask)
    cakes, lo = ma.integrate_batch(frames, g, mode="bilinear", mask_nan=True)
    c1 = ma.integrate(frames[0], g, mode="bilinear", mask_nan=True)
    for i in range(2):
        ref, sc = _oracle(oracle_lib, om, P, frames[i], dark.astype(np.float64), "bilinear")
        e = rel_err(cakes[i], ref, sc)
        assert e.max() <= TOL, e.max()
    assert rel_err(c1, _oracle(oracle_lib, om, P, frames[0], dark.astype(np.float64),
                               "bilinear")[0], np.abs(c1) + 1e-30).max() <= TOL


def test_profile_1d_matches_fused(oracle_lib):
    P, mask, frames, dark = _setup("tilt", n_frames=4)
    g = ma.build_map(P, mask=mask)
    g.set_dark(dark)
    cakes, lo = ma.integrate_batch(frames, g, mode="bilinear", mask_nan=True)
    for mode, row in (("cpu_mean", 0), ("area_weighted", 1), ("simple_mean", 2)):
        p = ma.profile_1d(cakes, g, mode=mode)
        # the same reduction of |cake| bounds the rounding of either order
        scale = ma.profile_1d(np.abs(cakes), g, mode=mode)
        err = np.abs(np.asarray(p, np.float64) - lo[:, row]) / np.maximum(scale, 1e-30)
        print(f"[profile] {mode}: worst {err.max():.2e}", flush=True)
        assert err.max() <= 1e-5


def test_torch_device_tensors(oracle_lib):
    torch = pytest.importorskip("torch")
    P, mask, frames, dark = _setup("flat", n_frames=5)
    g = ma.build_map(P, mask=mask)
    g.set_dark(torch.from_numpy(dark).cuda())
    dc, dl = ma.integrate_batch(torch.from_numpy(frames).cuda(), g, mode="bilinear")
    hc, hl = ma.integrate_batch(frames, g, mode="bilinear")
    np.testing.assert_array_equal(dc.cpu().numpy(), hc)
    np.testing.assert_array_equal(dl.cpu().numpy(), hl)


@pytest.mark.parametrize("ny,nz", [(96, 80), (96, 66), (97, 66)])
def test_transpose_paths(ny, nz):
    """The pixel-major batch through each transpose path: 128-pixel blocks
    (pixel count a multiple of 128), 64-pixel blocks with 16-byte loads (a
    multiple of 64 only) and the element-wise path; the batched results
    match the one-frame kernel's, which reads the frames directly."""
    P = IntegrationParams(NrPixelsY=ny, NrPixelsZ=nz, pxY=200.0, pxZ=200.0, Lsd=1e6,
                          BC_y=ny / 2 + 0.3, BC_z=nz / 2 - 0.2, RMin=2.0, RMax=40.0,
                          RBinSize=0.5, EtaBinSize=10.0)
    rng = np.random.default_rng(12)
    frames = rng.poisson(300.0, (5, nz, ny)).astype(np.uint16)
    print(f"[integrate] transpose path {ny}x{nz}", flush=True)
    g = ma.build_map(P)
    cb, lb = ma.integrate_batch(frames, g, mode="bilinear", mask_nan=True)
    for i in range(len(frames)):
        c1, l1 = ma.integrate_batch(frames[i], g, mode="bilinear", mask_nan=True)
        np.testing.assert_allclose(cb[i], c1[0], rtol=3e-6, atol=1e-4, equal_nan=True)
        np.testing.assert_allclose(lb[i], l1[0], rtol=3e-6, atol=1e-4)


def test_passes_match_single_passes():
    """150 frames run as three passes of up to 64 frames over one
    pixel-major batch; the results are bit-identical to one call per pass,
    and a call on the frames in reverse order gives the same again."""
    P, mask, frames, dark = _setup("tilt", n_frames=150, seed=6)
    print("[integrate] three passes, 150 frames", flush=True)
    g = ma.build_map(P, mask=mask)
    g.set_dark(dark)
    g.set_tuning(batch=64, rows_per_group=16, lds_bytes=81920)
    cb, lb = ma.integrate_batch(frames, g, mode="bilinear", mask_nan=True)
    for a, b in ((0, 64), (64, 128), (128, 150)):
        c1, l1 = ma.integrate_batch(frames[a:b], g, mode="bilinear", mask_nan=True)
        np.testing.assert_array_equal(cb[a:b], c1)
        np.testing.assert_array_equal(lb[a:b], l1)
    c2, l2 = ma.integrate_batch(frames[::-1].copy(), g, mode="bilinear", mask_nan=True)
    np.testing.assert_array_equal(c2[::-1], cb)
    np.testing.assert_array_equal(l2[::-1], lb)


@pytest.mark.parametrize("ny,nz", [(256, 192), (250, 192)])
def test_unused_pixel_blocks_are_skipped(ny, nz):
    """Rings up to 20 px on a 256 x 192 (250 x 192) detector: the
    transposes rewrite only the pixel blocks the map reads (most blocks
    hold none), with the 128-pixel kernel and with the 64-pixel one; the
    batched cakes still equal the one-frame kernel's, which reads the
    frames directly."""
    P = IntegrationParams(NrPixelsY=ny, NrPixelsZ=nz, pxY=200.0, pxZ=200.0, Lsd=1e6,
                          BC_y=60.3, BC_z=130.6, RMin=2.0, RMax=20.0, RBinSize=0.5,
                          EtaBinSize=10.0)
    rng = np.random.default_rng(13)
    frames = rng.poisson(300.0, (7, nz, ny)).astype(np.uint16)
    print(f"[integrate] sparse blocks {ny}x{nz}", flush=True)
    g = ma.build_map(P)
    assert 0 < g.layout_info()["px_blocks"] < ny * nz // 64 // 2
    cb, lb = ma.integrate_batch(frames, g, mode="bilinear", mask_nan=True)
    for i in range(len(frames)):
        c1, l1 = ma.integrate_batch(frames[i], g, mode="bilinear", mask_nan=True)
        np.testing.assert_allclose(cb[i], c1[0], rtol=3e-6, atol=1e-4, equal_nan=True)
        np.testing.assert_allclose(lb[i], l1[0], rtol=3e-6, atol=1e-4)


@pytest.mark.parametrize("bins,ents", [(192, 2560), (128, 1536), (320, 3072)])
def test_tile_shapes_are_bit_identical(bins, ents):
    """Tiles of other bin/entry capacities group bins differently but pad
    each bin's entries the same way: cakes equal the default shape's bit for
    bit; lineouts sum a row's chunk partials, which other chunks reassociate
    (a few ulp)."""
    P, mask, frames, dark = _setup("tilt", n_frames=9, seed=8)
    g = ma.build_map(P, mask=mask)
    g.set_dark(dark)
    cd, ld = ma.integrate_batch(frames, g, mode="bilinear", mask_nan=True)
    g.set_tile_shape(bins, ents)
    assert g.tile_shape() == (bins, ents)
    cs, ls = ma.integrate_batch(frames, g, mode="bilinear", mask_nan=True)
    np.testing.assert_array_equal(cs, cd)
    np.testing.assert_allclose(ls, ld, rtol=1e-6, atol=0)


@pytest.mark.parametrize("dtype", [np.int32, np.float32, np.uint8])
@pytest.mark.parametrize("batch", [16, 32])
def test_narrow_passes_match_oracle(dtype, batch, oracle_lib):
    """Passes of 16 and 32 frames (the narrow kernels: one or two frames
    per lane) for the 32-bit and 8-bit batch paths, against the FP64
    oracle at 1e-5."""
    P, mask, frames, dark = _setup("panels", n_frames=batch, seed=9, dtype=dtype)
    sh = panel_shifts_for("panels")
    g = ma.build_map(P, mask=mask, panel_shifts=sh)
    g.set_dark(dark)
    g.set_tuning(batch=batch)
    om = oracle_lib.build_map(to_oracle(P), mask=mask, panel_shifts=sh)
    dmean = dark.astype(np.float64).mean(axis=0)
    cakes, lo = ma.integrate_batch(frames, g, mode="bilinear", mask_nan=True)
    for i in (0, batch // 2, batch - 1):
        ref, scale_b = _oracle(oracle_lib, om, P, frames[i], dmean, "bilinear")
        assert rel_err(cakes[i], ref, scale_b).max() <= TOL
