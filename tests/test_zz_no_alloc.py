"""mc_fit_windows_noalloc (no device allocation; scratch in a static device
array) against mc_fit_windows (the allocating entry point) through the C
ABI, on device buffers: two calls in a row, the second with fewer windows,
each bit-identical to the allocating call on the same windows (same kernel,
same inputs: the scratch holds only values each call recomputes).  Sizes
above the static array are refused with MC_ERR_ARG."""
import ctypes as C

import numpy as np
import pytest

from midas_amd import _lib, peakfit
from midas_amd._lib import lib
from midas_amd.caking import default_context

pytestmark = pytest.mark.gpu


def _windows(n_win, n_pts, seed):
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(50, 950, n_win)
    x = x0[:, None] + 0.25 * np.arange(n_pts)[None, :]
    c = x0 + 0.25 * (n_pts // 2) + rng.uniform(-0.6, 0.6, n_win)
    y = np.stack([peakfit.pseudo_voigt(x[i], c[i], rng.uniform(0, 1), rng.uniform(1, 3),
                                       rng.uniform(200, 2000), rng.uniform(10, 100))
                  for i in range(n_win)])
    y = y + rng.normal(0, 5.0, y.shape)
    return x.astype(np.float32), y.astype(np.float32), (c + 0.1).astype(np.float32)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def test_noalloc_fits_equal_allocating_fits():
    torch = pytest.importorskip("torch")
    ctx = default_context()
    for model, n_win, n_pts in ((_lib.MC_FIT_PV, 3000, 41), (_lib.MC_FIT_PV, 700, 33),
                                (_lib.MC_FIT_TCH, 500, 41)):
        print(f"[noalloc] model {model}, {n_win} windows of {n_pts}", flush=True)
        x, y, g = _windows(n_win, n_pts, seed=n_win)
        dx, dy, dg = (torch.from_numpy(a).cuda() for a in (x, y, g))
        out_a = torch.zeros((n_win, _lib.MC_FIT_OUT), dtype=torch.float64, device="cuda")
        out_b = torch.full_like(out_a, -7.0)
        torch.cuda.synchronize()
        _lib.check(lib().mc_fit_windows(ctx.handle, model, n_win, n_pts, _ptr(dx), _ptr(dy),
                                        _ptr(dg), 0.25, 200, _ptr(out_a), 1))
        _lib.check(lib().mc_fit_windows_noalloc(ctx.handle, model, n_win, n_pts, _ptr(dx),
                                                _ptr(dy), _ptr(dg), 0.25, 200, _ptr(out_b)))
        _lib.check(lib().mc_ctx_sync(ctx.handle))
        a, b = out_a.cpu().numpy(), out_b.cpu().numpy()
        assert (a[:, 7] > 0).mean() > 0.95
        np.testing.assert_array_equal(b, a)


def test_noalloc_refuses_sizes_beyond_the_array():
    ctx = default_context()
    x = np.zeros(4, dtype=np.float32)
    rc = lib().mc_fit_windows_noalloc(ctx.handle, _lib.MC_FIT_PV, 1 << 20, 41,
                                      x.ctypes.data, x.ctypes.data, x.ctypes.data, 0.25, 10,
                                      x.ctypes.data)
    assert rc == -1  # MC_ERR_ARG
    assert "allocation-free" in lib().mc_last_error().decode()
