// mc_post.hip -- post-process filters on cakes along eta
// (postprocess.py:36-142 of the reference's midas_integrate package).
//
// Both filters work on [n_cakes][n_r][n_eta] float32 cakes, one output bin
// per thread, with the eta axis periodic (wrap) or reflected at its ends.
// Arithmetic follows the reference's numpy expressions operation by
// operation (FP64 products and sums in the same order, no contraction), so
// results match it to the rounding of the weights' exp():
//   * gauss_smooth_eta: out = sum_k w_k*IA[j+k] / sum_k w_k*A[j+k] with
//     IA = I*A (A = 1 without area weights), or I where the area sum is 0;
//   * median_filter_eta: the window's median (NaN if the window holds a
//     NaN), then per row the factor sum(I)/sum(out) in float32 pairwise
//     summation, numpy's order (loops_utils.h pairwise_sum).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "mc_internal.h"

#pragma clang fp contract(off)

namespace mc {
namespace {

constexpr int kMaxSmoothHalf = 1024;  // taps on each side
constexpr int kMaxMedianWindow = 63;

// Index of padded position i (may be < 0 or >= n) on an eta axis of n bins:
// periodic wrap, or numpy's "reflect" (mirror without repeating the edge,
// repeated for pads longer than the axis).
__device__ __forceinline__ int eta_index(int i, int n, int periodic) {
  if (periodic) {
    i %= n;
    return i < 0 ? i + n : i;
  }
  if (n == 1) return 0;
  const int p = 2 * (n - 1);
  i %= p;
  if (i < 0) i += p;
  return i < n ? i : p - i;
}

__global__ __launch_bounds__(256) void k_smooth_eta(const float *__restrict__ cakes,
                                                    const float *__restrict__ area,
                                                    const double *__restrict__ w, int half,
                                                    int n_r, int n_eta, int periodic,
                                                    float *__restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y, c = blockIdx.z;
  if (j >= n_eta) return;
  const float *row = cakes + ((int64_t)c * n_r + r) * n_eta;
  const float *arow = area ? area + (int64_t)r * n_eta : nullptr;
  double sia = 0.0, sa = 0.0;
  for (int k = -half; k <= half; ++k) {
    const int q = eta_index(j + k, n_eta, periodic);
    const double a = arow ? (double)arow[q] : 1.0;
    const double ia = arow ? (double)row[q] * a : (double)row[q];
    const double wk = w[k + half];
    sia = __dadd_rn(sia, __dmul_rn(wk, ia));
    sa = __dadd_rn(sa, __dmul_rn(wk, a));
  }
  out[((int64_t)c * n_r + r) * n_eta + j] =
      sa > 0.0 ? (float)__ddiv_rn(sia, sa) : row[j];
}

__global__ __launch_bounds__(256) void k_median_eta(const float *__restrict__ cakes, int window,
                                                    int n_r, int n_eta, int periodic,
                                                    float *__restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y, c = blockIdx.z;
  if (j >= n_eta) return;
  const float *row = cakes + ((int64_t)c * n_r + r) * n_eta;
  const int half = window / 2;
  float v[kMaxMedianWindow];
  bool nan = false;
  int n = 0;
  for (int k = -half; k <= half; ++k) {
    const float x = row[eta_index(j + k, n_eta, periodic)];
    nan |= isnan(x);
    // insertion sort, ascending
    int i = n++;
    while (i > 0 && v[i - 1] > x) {
      v[i] = v[i - 1];
      --i;
    }
    v[i] = x;
  }
  out[((int64_t)c * n_r + r) * n_eta + j] = nan ? NAN : v[half];
}

// Sum of a block of at most 128 floats, numpy's unrolled order.
__device__ float block_sum(const float *a, int m) {
  if (m < 8) {
    float s = 0.f;
    for (int i = 0; i < m; ++i) s += a[i];
    return s;
  }
  float r0 = a[0], r1 = a[1], r2 = a[2], r3 = a[3], r4 = a[4], r5 = a[5], r6 = a[6], r7 = a[7];
  int i;
  for (i = 8; i < m - (m % 8); i += 8) {
    r0 += a[i];
    r1 += a[i + 1];
    r2 += a[i + 2];
    r3 += a[i + 3];
    r4 += a[i + 4];
    r5 += a[i + 5];
    r6 += a[i + 6];
    r7 += a[i + 7];
  }
  float s = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
  for (; i < m; ++i) s += a[i];
  return s;
}

// numpy's float32 pairwise sum of a[0..n) (loops_utils.h.src
// pairwise_sum): blocks of <= 128 summed as above, longer runs split in two
// at n/2 rounded down to a multiple of 8.  The recursion runs on an
// explicit stack (post-order).
__device__ float pairwise_sum(const float *a, int n) {
  struct Frame {
    int lo, n, phase;
    float left;
  };
  Frame st[32];
  int sp = 0;
  st[0] = {0, n, 0, 0.f};
  float ret = 0.f;  // value of the call that completed last
  while (sp >= 0) {
    Frame &f = st[sp];
    if (f.n <= 128) {
      ret = block_sum(a + f.lo, f.n);
      --sp;
      continue;
    }
    int n2 = f.n / 2;
    n2 -= n2 % 8;
    if (f.phase == 0) {
      f.phase = 1;
      st[sp + 1] = {f.lo, n2, 0, 0.f};
      ++sp;
    } else if (f.phase == 1) {
      f.left = ret;
      f.phase = 2;
      st[sp + 1] = {f.lo + n2, f.n - n2, 0, 0.f};
      ++sp;
    } else {
      ret = f.left + ret;
      --sp;
    }
  }
  return ret;
}

// Per row: out *= where(sum(out) != 0, sum(in) / sum(out), 1), both sums
// in float32 pairwise order (a row with a NaN becomes all NaN).
__global__ __launch_bounds__(256) void k_flux_scale(const float *__restrict__ cakes, int64_t rows,
                                                    int n_eta, float *__restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const float *in = cakes + r * n_eta;
  float *o = out + r * n_eta;
  const float s_in = pairwise_sum(in, n_eta), s_out = pairwise_sum(o, n_eta);
  const float scale = s_out != 0.f ? s_in / s_out : 1.f;
  for (int j = 0; j < n_eta; ++j) o[j] = o[j] * scale;
}

// numpy's float64 pairwise sum for the (short) weight vector, host side.
double pairwise_sum_host(const double *a, int n) {
  if (n < 8) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i];
    return s;
  }
  if (n <= 128) {
    double r[8];
    for (int j = 0; j < 8; ++j) r[j] = a[j];
    int i;
    for (i = 8; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; ++j) r[j] += a[i + j];
    double s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) s += a[i];
    return s;
  }
  int n2 = n / 2;
  n2 -= n2 % 8;
  return pairwise_sum_host(a, n2) + pairwise_sum_host(a + n2, n - n2);
}

// Gaussian weights of gauss_smooth_eta (postprocess.py:71-74): half =
// max(1, round-half-even(4 sigma)), w_k = exp(-k^2 / (2 sigma^2)) / sum.
std::vector<double> gauss_weights(double sigma, int *half_out) {
  const int half = std::max(1, (int)std::nearbyint(4.0 * sigma));
  std::vector<double> w(2 * half + 1);
  for (int k = -half; k <= half; ++k) {
    const double kk = (double)k;
    w[k + half] = std::exp(-(kk * kk) / (2.0 * sigma * sigma));
  }
  const double s = pairwise_sum_host(w.data(), (int)w.size());
  for (auto &x : w) x /= s;
  *half_out = half;
  return w;
}

// Stages host inputs on the device when needed, runs `body` with device
// pointers, copies the result back.
template <class F>
int with_device_cakes(Ctx *ctx, const float *cakes, size_t n_cake, const float *area,
                      size_t n_area, int on_device, float *out, F body) {
  hipStream_t s = ctx->stream;
  DBuf dc, da, dout;
  int rc;
  const float *c = cakes, *a = area;
  float *o = out;
  if (!on_device) {
    if ((rc = dc.alloc(n_cake * 4)) || (rc = dout.alloc(n_cake * 4))) return rc;
    MC_HIP(hipMemcpyAsync(dc.p, cakes, n_cake * 4, hipMemcpyHostToDevice, s));
    c = dc.as<float>();
    o = dout.as<float>();
    if (area) {
      if ((rc = da.alloc(n_area * 4))) return rc;
      MC_HIP(hipMemcpyAsync(da.p, area, n_area * 4, hipMemcpyHostToDevice, s));
      a = da.as<float>();
    }
  }
  if ((rc = body(s, c, a, o))) return rc;
  if (!on_device) MC_HIP(hipMemcpyAsync(out, o, n_cake * 4, hipMemcpyDeviceToHost, s));
  MC_HIP(hipStreamSynchronize(s));
  return MC_OK;
}

}  // namespace
}  // namespace mc

using namespace mc;

extern "C" {

int mc_smooth_eta(mc_ctx *ctx, const float *cakes, int32_t n_cakes, int32_t n_r, int32_t n_eta,
                  double sigma_bins, const float *area, int32_t periodic, int32_t on_device,
                  float *out) {
  MC_CHECK(ctx && cakes && out, MC_ERR_ARG, "null argument");
  MC_CHECK(n_cakes >= 1 && n_r >= 1 && n_eta >= 1, MC_ERR_ARG, "bad cake shape %d x %d x %d",
           n_cakes, n_r, n_eta);
  MC_CHECK(std::isfinite(sigma_bins) && sigma_bins > 0, MC_ERR_ARG,
           "sigma_bins must be positive, got %g", sigma_bins);
  MC_CHECK(cakes != out, MC_ERR_ARG, "out must not alias cakes");
  int half = 0;
  std::vector<double> w = gauss_weights(sigma_bins, &half);
  MC_CHECK(half <= kMaxSmoothHalf, MC_ERR_ARG, "sigma_bins %g needs %d taps per side (max %d)",
           sigma_bins, half, kMaxSmoothHalf);
  // The reference pads with slices of the cake itself: a periodic pad
  // longer than the axis is not defined there.
  MC_CHECK(!periodic || half <= n_eta, MC_ERR_ARG,
           "periodic smoothing with %d taps per side over %d eta bins", half, n_eta);
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  MC_HIP(hipSetDevice(ctx->device));
  DBuf dw;
  int rc;
  if ((rc = dw.alloc(w.size() * 8))) return rc;
  MC_HIP(hipMemcpyAsync(dw.p, w.data(), w.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  const size_t n_cake = (size_t)n_cakes * n_r * n_eta, n_area = (size_t)n_r * n_eta;
  return with_device_cakes(ctx, cakes, n_cake, area, n_area, on_device, out,
                           [&](hipStream_t s, const float *c, const float *a, float *o) -> int {
                             dim3 grid((unsigned)((n_eta + 255) / 256), (unsigned)n_r,
                                       (unsigned)n_cakes);
                             hipLaunchKernelGGL(k_smooth_eta, grid, dim3(256), 0, s, c, a,
                                                dw.as<double>(), half, n_r, n_eta, periodic, o);
                             MC_HIP(hipGetLastError());
                             return MC_OK;
                           });
}

int mc_median_eta(mc_ctx *ctx, const float *cakes, int32_t n_cakes, int32_t n_r, int32_t n_eta,
                  int32_t window, int32_t periodic, int32_t preserve_ring_flux,
                  int32_t on_device, float *out) {
  MC_CHECK(ctx && cakes && out, MC_ERR_ARG, "null argument");
  MC_CHECK(n_cakes >= 1 && n_r >= 1 && n_eta >= 1, MC_ERR_ARG, "bad cake shape %d x %d x %d",
           n_cakes, n_r, n_eta);
  MC_CHECK(window >= 1 && window % 2 == 1, MC_ERR_ARG,
           "window must be a positive odd integer, got %d", window);
  MC_CHECK(window <= kMaxMedianWindow, MC_ERR_ARG, "window %d over the maximum %d", window,
           kMaxMedianWindow);
  MC_CHECK(!periodic || window / 2 <= n_eta, MC_ERR_ARG,
           "periodic window of %d over %d eta bins", window, n_eta);
  MC_CHECK(cakes != out, MC_ERR_ARG, "out must not alias cakes");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  MC_HIP(hipSetDevice(ctx->device));
  const size_t n_cake = (size_t)n_cakes * n_r * n_eta;
  return with_device_cakes(ctx, cakes, n_cake, nullptr, 0, on_device, out,
                           [&](hipStream_t s, const float *c, const float *, float *o) -> int {
                             dim3 grid((unsigned)((n_eta + 255) / 256), (unsigned)n_r,
                                       (unsigned)n_cakes);
                             hipLaunchKernelGGL(k_median_eta, grid, dim3(256), 0, s, c, window,
                                                n_r, n_eta, periodic, o);
                             MC_HIP(hipGetLastError());
                             if (preserve_ring_flux) {
                               const int64_t rows = (int64_t)n_cakes * n_r;
                               hipLaunchKernelGGL(k_flux_scale, dim3((unsigned)((rows + 255) / 256)),
                                                  dim3(256), 0, s, c, rows, n_eta, o);
                               MC_HIP(hipGetLastError());
                             }
                             return MC_OK;
                           });
}

}  // extern "C"
