// mc_kernels.hip -- the gather-reduce kernels of the caking path (gfx950).
//
// Per (R, eta) bin b and frame f the cake value is
//     I[f][b] = ( sum_k frame_f[pix_k] * w_k  -  D_b ) * s_b
// where (pix_k, w_k) are the bin's CSR entries (w = Map.bin frac, or frac
// times a bilinear corner weight), D_b = sum_k dark[pix_k] * w_k is the dark
// term of the bin (computed once per dark in FP64), and s_b = 1 / sum
// areaWeight (Normalize=1) or 1.  This is the per-bin loop of
// IntegratorZarrOMP.c:1710-1778 with "Image = raw - dark" (:1646-1669)
// distributed over the sum: subtracting the dark per bin instead of per
// pixel is exact in real arithmetic and removes the dark from the
// per-frame traffic.  Bins flagged in maskMap.bin carry a negative s_b and
// come out NaN when MC_F_MASK_NAN is set (IntegratorZarrOMP.c:1766-1767).
//
// Accumulation is FP32: uint16/int32 counts convert exactly (int32 up to
// 2^24), weights are FP32, sums use Kahan compensation per lane and a
// pairwise tree across lanes.
//
// The 1-D lineouts of the three reference definitions are fused:
//   cpu  : mean over eta of values != 0 and not NaN (IntegratorZarrOMP.c
//          :1829-1843)
//   area : sum(I*A)/sum(A) over bins with A > 1e-9, A = in-bounds area
//          (IntegratorFitPeaksGPUStream.cu:1162-1254)
//   mean : mean of I over bins with A > 1e-9 (.cu:1256-1334)
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "mc_internal.h"

namespace mc {

namespace {

template <typename T> __device__ __forceinline__ float to_f(T v) {
  return static_cast<float>(v);
}

// Sum over groups of 16 lanes with DPP row operations: quad xor 1, quad
// xor 2, half-row mirror, row mirror.  Every lane of the group ends with the
// group's total.
__device__ __forceinline__ float dpp_sum16(float v) {
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xf, 0xf, false));
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xf, 0xf, false));
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xf, 0xf, false));
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xf, 0xf, false));
  return v;
}

struct Kahan {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float x) {
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ float value() const { return s - c; }
};

// lo = (sum of nonzero values, their count, sum I*A, sum I) of one row;
// rs = (sum A, count) over valid bins, then over valid unflagged bins.
__device__ __forceinline__ void write_lineouts(float4 lo, float4 rs, int mask_nan, float *L,
                                               int n_r, int r) {
  const float sa = mask_nan ? rs.z : rs.x, na = mask_nan ? rs.w : rs.y;
  L[0 * n_r + r] = lo.y > 0.f ? lo.x / lo.y : 0.f;
  L[1 * n_r + r] = sa > 1e-9f ? lo.z / sa : 0.f;
  L[2 * n_r + r] = na > 0.f ? lo.w / na : 0.f;
}

__device__ __forceinline__ bool flagged(float scale) {
  return __float_as_int(scale) < 0;
}

// Block-wide sum of 4 floats (256 threads), result valid in thread 0.
__device__ __forceinline__ void block_sum4(float4 &v, float4 *red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_down(v.x, off);
    v.y += __shfl_down(v.y, off);
    v.z += __shfl_down(v.z, off);
    v.w += __shfl_down(v.w, off);
  }
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    float4 t = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); i++) {
      t.x += red[i].x;
      t.y += red[i].y;
      t.z += red[i].z;
      t.w += red[i].w;
    }
    v = t;
  }
}

// ---------------------------------------------------------------------
// One-frame kernel: workgroup = one R row of one frame, 16 lanes per bin.
// grid = (n_r, n_frames).
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_plain(
    const int32_t *__restrict__ row, const int32_t *__restrict__ pix,
    const float *__restrict__ w, const float2 *__restrict__ binfo,
    const float *__restrict__ a_inb, const float4 *__restrict__ row_static,
    const T *__restrict__ frames, int64_t frame_stride, int n_eta, int n_r,
    int64_t nb, int mask_nan, float *__restrict__ cake, float *__restrict__ lineouts) {
  constexpr int LPB = 16, SLOTS = 256 / LPB;
  const int r = blockIdx.x, f = blockIdx.y;
  const T *__restrict__ img = frames + (int64_t)f * frame_stride;
  float *__restrict__ out = cake + (int64_t)f * nb;
  const int slot = threadIdx.x / LPB, sub = threadIdx.x % LPB;
  float4 lo = make_float4(0.f, 0.f, 0.f, 0.f);  // S_cpu, N_cpu, S_area, S_mean
  for (int e = slot; e < n_eta; e += SLOTS) {
    const int64_t b = (int64_t)r * n_eta + e;
    const int beg = row[b], end = row[b + 1];
    Kahan acc;
    for (int k = beg + sub; k < end; k += LPB) acc.add(to_f(img[pix[k]]) * w[k]);
    float v = dpp_sum16(acc.value());
    if (sub == 0) {
      const float2 bi = binfo[b];
      const float val = (v - bi.y) * fabsf(bi.x);
      const float o = flagged(bi.x) ? NAN : val;
      out[b] = o;
      if (lineouts) {
        if (o != 0.f && !isnan(o)) {
          lo.x += o;
          lo.y += 1.f;
        }
        const float a = a_inb[b];
        if (a > 1e-9f && !isnan(o)) {
          lo.z += o * a;
          lo.w += o;
        }
      }
    }
  }
  if (lineouts) {
    __shared__ float4 red[4];
    block_sum4(lo, red);
    if (threadIdx.x == 0) write_lineouts(lo, row_static[r], mask_nan, lineouts + (int64_t)f * 3 * n_r, n_r, r);
  }
}

// ---------------------------------------------------------------------
// Lineouts of existing cakes: workgroup = one R row of one frame.
// ---------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_profile(
    const float *__restrict__ cakes, const float *__restrict__ a_inb,
    const float4 *__restrict__ row_static, int n_eta, int n_r, int64_t nb,
    int mask_nan, float *__restrict__ lineouts) {
  const int r = blockIdx.x, f = blockIdx.y;
  const float *c = cakes + (int64_t)f * nb + (int64_t)r * n_eta;
  const float *a = a_inb + (int64_t)r * n_eta;
  float4 lo = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int e = threadIdx.x; e < n_eta; e += 256) {
    const float o = c[e];
    if (o != 0.f && !isnan(o)) {
      lo.x += o;
      lo.y += 1.f;
    }
    if (a[e] > 1e-9f && !isnan(o)) {
      lo.z += o * a[e];
      lo.w += o;
    }
  }
  __shared__ float4 red[4];
  block_sum4(lo, red);
  if (threadIdx.x == 0) write_lineouts(lo, row_static[r], mask_nan, lineouts + (int64_t)f * 3 * n_r, n_r, r);
}

// ---------------------------------------------------------------------
// Batched kernel.  Workgroup = one row block (rows r0..r0+H-1, all eta).
// The block's bins are cut into tiles along eta; a tile's pixel bounding
// box is staged into LDS for all B frames as [pixel][frame] (pitch P), so
// one map entry read by the wave serves every frame: lane = frame.
//   tiles: y0 z0 bw bh bin_begin bin_end
//   tbins: cake bin, ent_begin, ent_end, (row in block) | (direct << 30)
//   ents : (LDS pixel index, or global pixel for direct bins), weight
// ---------------------------------------------------------------------
template <typename TIn, typename TL, int B>
__global__ __launch_bounds__(256) void k_batched(
    const int32_t *__restrict__ group_tile, const int32_t *__restrict__ group_row0,
    const int32_t *__restrict__ tiles, const int4 *__restrict__ tbins,
    const uint2 *__restrict__ ents, const float2 *__restrict__ binfo,
    const float *__restrict__ a_inb, const float4 *__restrict__ row_static,
    const TIn *__restrict__ frames, int64_t frame_stride,
    const int32_t *__restrict__ slots, int n_frames, int ny, int n_eta,
    int n_r, int rows_per_group, int pitch, int64_t nb, int mask_nan,
    float *__restrict__ cakes, float *__restrict__ lineouts) {
  extern __shared__ __attribute__((aligned(16))) char lds_raw[];
  constexpr int SPW = B >= 64 ? 1 : 64 / B;   // bins per wave at once
  constexpr int FPL = B > 64 ? B / 64 : 1;     // frames per lane
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int slot = SPW > 1 ? lane / B : 0;
  const int f0 = SPW > 1 ? lane % B : lane * FPL;
  const int g = blockIdx.x;
  const int r0 = group_row0[g];
  const int t_beg = group_tile[g], t_end = group_tile[g + 1];
  float *lo = reinterpret_cast<float *>(lds_raw);  // [H][4][B]
  TL *tile = reinterpret_cast<TL *>(lds_raw + (size_t)rows_per_group * 4 * B * sizeof(float));
  const bool want_lo = lineouts != nullptr;
  for (int i = threadIdx.x; i < rows_per_group * 4 * B; i += 256) lo[i] = 0.f;

  for (int t = t_beg; t < t_end; ++t) {
    const int y0 = tiles[6 * t + 0], z0 = tiles[6 * t + 1];
    const int bw = tiles[6 * t + 2], bh = tiles[6 * t + 3];
    const int b_beg = tiles[6 * t + 4], b_end = tiles[6 * t + 5];
    __syncthreads();
    // ---- stage the bounding box of every frame into LDS ----
    if (bw > 0) {
      const int cols = bw <= 16 ? 16 : (bw <= 32 ? 32 : 64);
      const int rpi = 64 / cols;
      const int lc = lane % cols, lr = lane / cols;
      const int nrows = n_frames * bh;
      for (int pr = wave * rpi; pr < nrows; pr += 4 * rpi) {
        const int q = pr + lr;
        if (q < nrows) {
          const int f = q / bh, rr = q - f * bh;
          const TIn *src = frames + (int64_t)slots[f] * frame_stride +
                           (int64_t)(z0 + rr) * ny + y0;
          TL *dst = tile + (size_t)(rr * bw) * pitch + f;
          for (int c = lc; c < bw; c += cols) dst[(size_t)c * pitch] = static_cast<TL>(src[c]);
        }
      }
    }
    __syncthreads();
    // ---- every wave reduces its share of the tile's bins ----
    for (int bi = b_beg + wave * SPW + slot; bi < b_end; bi += 4 * SPW) {
      const int4 tb = tbins[bi];
      const bool direct = (tb.w >> 30) & 1;
      const int rib = tb.w & 0x3fffffff;
      Kahan acc[FPL];
      if (!direct) {
        for (int k = tb.y; k < tb.z; ++k) {
          const uint2 e = ents[k];
          const float wk = __uint_as_float(e.y);
          const TL *p = tile + (size_t)e.x * pitch + f0;
#pragma unroll
          for (int j = 0; j < FPL; ++j) acc[j].add(to_f(p[j]) * wk);
        }
      } else {
        for (int k = tb.y; k < tb.z; ++k) {
          const uint2 e = ents[k];
          const float wk = __uint_as_float(e.y);
#pragma unroll
          for (int j = 0; j < FPL; ++j) {
            const int f = f0 + j;
            if (f < n_frames)
              acc[j].add(to_f(frames[(int64_t)slots[f] * frame_stride + e.x]) * wk);
          }
        }
      }
      const float2 inf = binfo[tb.x];
      const float a = want_lo ? a_inb[tb.x] : 0.f;
#pragma unroll
      for (int j = 0; j < FPL; ++j) {
        const int f = f0 + j;
        if (f < n_frames) {
          const float val = (acc[j].value() - inf.y) * fabsf(inf.x);
          const float o = flagged(inf.x) ? NAN : val;
          cakes[(int64_t)f * nb + tb.x] = o;
          if (want_lo) {
            float *L = lo + (size_t)rib * 4 * B;
            if (o != 0.f && !isnan(o)) {
              atomicAdd(&L[0 * B + f], o);
              atomicAdd(&L[1 * B + f], 1.f);
            }
            if (a > 1e-9f && !isnan(o)) {
              atomicAdd(&L[2 * B + f], o * a);
              atomicAdd(&L[3 * B + f], o);
            }
          }
        }
      }
    }
  }
  if (want_lo) {
    __syncthreads();
    for (int i = threadIdx.x; i < rows_per_group * B; i += 256) {
      const int rib = i / B, f = i - rib * B;
      const int r = r0 + rib;
      if (r < n_r && f < n_frames) {
        const float *L = lo + (size_t)rib * 4 * B;
        write_lineouts(make_float4(L[f], L[B + f], L[2 * B + f], L[3 * B + f]), row_static[r],
                       mask_nan, lineouts + (int64_t)f * 3 * n_r, n_r, r);
      }
    }
  }
}

// ---------------------------------------------------------------------
// Setup kernels.
// ---------------------------------------------------------------------
__global__ void k_fill(float *p, int64_t n, float v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
__global__ void k_dark_mean(const T *__restrict__ frames, int n_frames,
                            int64_t npx, double scale, double *__restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npx) return;
  double s = 0;
  for (int f = 0; f < n_frames; ++f) s += (double)frames[(int64_t)f * npx + i];
  out[i] = s / n_frames * scale;
}

__global__ void k_dark_term(const int32_t *__restrict__ row, const int32_t *__restrict__ pix,
                            const float *__restrict__ w, const double *__restrict__ dark,
                            int64_t nb, float *__restrict__ dterm) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double s = 0;
  if (dark)
    for (int k = row[b]; k < row[b + 1]; ++k) s += dark[pix[k]] * (double)w[k];
  dterm[b] = (float)s;
}

__global__ void k_binfo(const float *__restrict__ inv_area, const int32_t *__restrict__ flags,
                        const float *__restrict__ dterm, int64_t nb, int normalize,
                        int mask_nan, float2 *__restrict__ binfo) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  float s = normalize ? inv_area[b] : 1.f;
  if (mask_nan && flags && flags[b]) s = -s;
  binfo[b] = make_float2(s, dterm ? dterm[b] : 0.f);
}

template <typename F> int dispatch_dtype(int dtype, F &&f) {
  switch (dtype) {
    case MC_DTYPE_U8: return f((const uint8_t *)nullptr);
    case MC_DTYPE_U16: return f((const uint16_t *)nullptr);
    case MC_DTYPE_U32: return f((const uint32_t *)nullptr);
    case MC_DTYPE_I64: return f((const int64_t *)nullptr);
    case MC_DTYPE_F32: return f((const float *)nullptr);
    case MC_DTYPE_F64: return f((const double *)nullptr);
    case MC_DTYPE_I32: return f((const int32_t *)nullptr);
    default:
      set_error("unsupported frame dtype %d", dtype);
      return MC_ERR_ARG;
  }
}

inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

size_t dtype_size(int dtype) {
  switch (dtype) {
    case MC_DTYPE_U8: return 1;
    case MC_DTYPE_U16: return 2;
    case MC_DTYPE_U32: return 4;
    case MC_DTYPE_I64: return 8;
    case MC_DTYPE_F32: return 4;
    case MC_DTYPE_F64: return 8;
    case MC_DTYPE_I32: return 4;
    default: return 0;
  }
}

int batched_elem_size(int dtype) {
  return (dtype == MC_DTYPE_U16 || dtype == MC_DTYPE_U8) ? 2 : 4;
}

// LDS pitch (elements per pixel): the frames of one pixel are contiguous,
// and the odd dword stride between pixels keeps the staging writes of
// consecutive pixels on different banks.
int batched_lds_pitch(int batch, int dtype) {
  if (batched_elem_size(dtype) == 2) return batch + ((batch / 2) % 2 == 1 ? 0 : 2);
  return batch + 1;
}

int launch_area_fill(hipStream_t s, float *p, int64_t n, float v) {
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n, v);
  MC_HIP(hipGetLastError());
  return MC_OK;
}

int launch_dark_mean(hipStream_t s, const void *frames, int dtype, int n_frames,
                     int64_t npx, double scale, double *out) {
  return dispatch_dtype(dtype, [&](auto tag) -> int {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(tag)>>;
    hipLaunchKernelGGL(k_dark_mean<T>, dim3(blocks_for(npx, 256)), dim3(256), 0, s,
                       (const T *)frames, n_frames, npx, scale, out);
    MC_HIP(hipGetLastError());
    return MC_OK;
  });
}

int launch_dark_term(hipStream_t s, const int32_t *row, const int32_t *pix,
                     const float *w, const double *dark, int64_t nb, float *dterm) {
  hipLaunchKernelGGL(k_dark_term, dim3(blocks_for(nb, 256)), dim3(256), 0, s, row, pix,
                     w, dark, nb, dterm);
  MC_HIP(hipGetLastError());
  return MC_OK;
}

int launch_binfo(hipStream_t s, const float *inv_area, const int32_t *flags,
                 const float *dterm, int64_t nb, int normalize, int mask_nan,
                 float2 *binfo) {
  hipLaunchKernelGGL(k_binfo, dim3(blocks_for(nb, 256)), dim3(256), 0, s, inv_area,
                     flags, dterm, nb, normalize, mask_nan, binfo);
  MC_HIP(hipGetLastError());
  return MC_OK;
}

int launch_plain(hipStream_t s, const Map &m, const Csr &c, const void *frames,
                 int dtype, float *cake, float *lineouts, int n_frames,
                 int64_t frame_stride) {
  return dispatch_dtype(dtype, [&](auto tag) -> int {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(tag)>>;
    hipLaunchKernelGGL(k_plain<T>, dim3(m.n_r, n_frames), dim3(256), 0, s,
                       c.row.as<int32_t>(), c.pix.as<int32_t>(), c.w.as<float>(),
                       m.d_binfo.as<float2>(), m.d_area_inb.as<float>(),
                       m.d_row_static.as<float4>(), (const T *)frames, frame_stride,
                       m.n_eta, m.n_r, m.nb(), (m.binfo_flags & MC_F_MASK_NAN) ? 1 : 0, cake,
                       lineouts);
    MC_HIP(hipGetLastError());
    return MC_OK;
  });
}

int launch_profile(hipStream_t s, const Map &m, const float *cakes, int n_frames,
                   int mask_nan, float *lineouts) {
  hipLaunchKernelGGL(k_profile, dim3(m.n_r, n_frames), dim3(256), 0, s, cakes,
                     m.d_area_inb.as<float>(), m.d_row_static.as<float4>(), m.n_eta, m.n_r,
                     m.nb(), mask_nan, lineouts);
  MC_HIP(hipGetLastError());
  return MC_OK;
}

template <typename TIn, int B>
static int launch_batched_t(hipStream_t s, const Map &m, const Tiles &t,
                            const void *frames, int n_frames, const int32_t *slots,
                            float *cakes, float *lineouts) {
  using TL = typename std::conditional<sizeof(TIn) <= 2, uint16_t, float>::type;
  auto kern = k_batched<TIn, TL, B>;
  const size_t lds = t.lds_needed;
  if (lds > 65536) MC_HIP(hipFuncSetAttribute((const void *)kern,
                                              hipFuncAttributeMaxDynamicSharedMemorySize,
                                              (int)lds));
  hipLaunchKernelGGL(kern, dim3(t.n_groups), dim3(256), lds, s,
                     t.group_tile.as<int32_t>(), t.group_row0.as<int32_t>(),
                     t.tiles.as<int32_t>(), t.tbins.as<int4>(), t.ents.as<uint2>(),
                     m.d_binfo.as<float2>(), m.d_area_inb.as<float>(),
                     m.d_row_static.as<float4>(), (const TIn *)frames, m.npx(), slots,
                     n_frames, m.ny, m.n_eta, m.n_r, t.rows, t.pitch, m.nb(),
                     (m.binfo_flags & MC_F_MASK_NAN) ? 1 : 0, cakes, lineouts);
  MC_HIP(hipGetLastError());
  return MC_OK;
}

int launch_batched(hipStream_t s, const Map &m, const Tiles &t, const void *frames,
                   int dtype, int n_frames, const int32_t *slots, float *cakes,
                   float *lineouts) {
  return dispatch_dtype(dtype, [&](auto tag) -> int {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(tag)>>;
    switch (t.batch) {
      case 8: return launch_batched_t<T, 8>(s, m, t, frames, n_frames, slots, cakes, lineouts);
      case 16: return launch_batched_t<T, 16>(s, m, t, frames, n_frames, slots, cakes, lineouts);
      case 32: return launch_batched_t<T, 32>(s, m, t, frames, n_frames, slots, cakes, lineouts);
      case 64: return launch_batched_t<T, 64>(s, m, t, frames, n_frames, slots, cakes, lineouts);
      case 128: return launch_batched_t<T, 128>(s, m, t, frames, n_frames, slots, cakes, lineouts);
      default:
        set_error("batch %d not in {8,16,32,64,128}", t.batch);
        return MC_ERR_ARG;
    }
  });
}

}  // namespace mc
