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
// Batched kernel.  A workgroup takes one chunk: a run of consecutive tiles
// of one row block (rows r0..r0+H-1).  For each tile it stages into LDS
//   * the tile's bin records (cake bin, entry range, per-bin scale/dark,
//     in-bounds area),
//   * the tile's map entries (pixel index inside the tile box, weight),
//   * the pixel box of every frame of the batch, one plane per frame
//     (plane stride S elements, chosen so lanes reading the same pixel of
//     consecutive frames hit different LDS banks),
// then reduces every bin with lane = frame: one entry read (an LDS
// broadcast) serves all frames of the batch.  Bins whose pixel box does not
// fit the LDS budget ("direct" bins) gather from global memory instead.
// Lineout sums per (row, frame) are kept in LDS and written as one partial
// per chunk; k_lineout_finalize adds the partials of each row block.
//   chunk: tile_begin tile_end row0 n_rows
//   tile : y0 z0 bwp bh kshift S b_beg b_end e_beg e_end
//          (bwp = padded box width = nv * V, nv = 1 << kshift vector loads)
//   tbin : cake bin, ent_begin, ent_end, (row in block) | (direct << 30)
//   ent  : (pixel index in the tile box, or global pixel if direct), weight
// ---------------------------------------------------------------------
constexpr int kTileBins = 128;    // bin records per tile (LDS)
constexpr int kTileEnts = 1024;   // map entries per tile (LDS)

struct BinRec {
  int4 tb;
  float2 info;
  float a;
  float pad;
};

template <typename TIn, int V> struct alignas(sizeof(TIn) * V) VecLoad {
  TIn v[V];
};

template <typename TIn, typename TL, int B, int V>
__global__ __launch_bounds__(256) void k_batched(
    const int32_t *__restrict__ chunks, const int32_t *__restrict__ tiles,
    const int4 *__restrict__ tbins, const uint2 *__restrict__ ents,
    const float2 *__restrict__ binfo, const float *__restrict__ a_inb,
    const TIn *__restrict__ frames, int64_t frame_stride,
    const int32_t *__restrict__ slots, int n_frames, int ny, int rows_per_group,
    int64_t nb, float *__restrict__ cakes, float *__restrict__ part) {
  extern __shared__ __attribute__((aligned(16))) char lds_raw[];
  constexpr int SPW = B >= 64 ? 1 : 64 / B;  // bins per wave at once
  constexpr int FPL = B > 64 ? B / 64 : 1;    // frames per lane
  constexpr int U = 8;                        // staging loads in flight per lane
  BinRec *lbin = reinterpret_cast<BinRec *>(lds_raw);
  uint2 *lent = reinterpret_cast<uint2 *>(lds_raw + kTileBins * sizeof(BinRec));
  float *lo = reinterpret_cast<float *>(lds_raw + kTileBins * sizeof(BinRec) +
                                        kTileEnts * sizeof(uint2));
  TL *plane = reinterpret_cast<TL *>(reinterpret_cast<char *>(lo) +
                                     (size_t)rows_per_group * 4 * B * sizeof(float));
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int slot = SPW > 1 ? lane / B : 0;
  const int f0 = SPW > 1 ? lane % B : lane * FPL;
  const int c = blockIdx.x;
  const int t_beg = chunks[4 * c], t_end = chunks[4 * c + 1];
  const bool want_lo = part != nullptr;
  if (want_lo)
    for (int i = tid; i < rows_per_group * 4 * B; i += 256) lo[i] = 0.f;

  for (int t = t_beg; t < t_end; ++t) {
    const int32_t *T = tiles + 10 * t;
    const int y0 = T[0], z0 = T[1], bwp = T[2], bh = T[3], ks = T[4], S = T[5];
    const int b_beg = T[6], nbt = T[7] - T[6], e_beg = T[8], net = T[9] - T[8];
    __syncthreads();
    // ---- bin records and entries ----
    for (int i = tid; i < nbt; i += 256) {
      BinRec r;
      r.tb = tbins[b_beg + i];
      r.info = binfo[r.tb.x];
      r.a = a_inb[r.tb.x];
      r.pad = 0.f;
      lbin[i] = r;
    }
    for (int k = tid; k < net; k += 256) lent[k] = ents[e_beg + k];
    // ---- pixel boxes of all frames: plane f holds the box row-major ----
    if (bwp > 0) {
      const int nv = 1 << ks;
      const int rp = 64 >> ks;  // box rows per wave instruction
      const int cv = lane & (nv - 1), lr = lane >> ks;
      const int passes = (bh + rp - 1) / rp;
      const int jobs = n_frames * passes;
      for (int j0 = wave; j0 < jobs; j0 += 4 * U) {
        VecLoad<TIn, V> v[U];
        int dst[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int j = j0 + 4 * u;
          dst[u] = -1;
          if (j < jobs) {
            const int f = j / passes, rr = (j - f * passes) * rp + lr;
            if (rr < bh) {
              const TIn *src = frames + (int64_t)slots[f] * frame_stride +
                               (int64_t)(z0 + rr) * ny + y0 + cv * V;
              v[u] = *reinterpret_cast<const VecLoad<TIn, V> *>(src);
              dst[u] = f * S + rr * bwp + cv * V;
            }
          }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (dst[u] >= 0) {
#pragma unroll
            for (int q = 0; q < V; ++q) plane[dst[u] + q] = static_cast<TL>(v[u].v[q]);
          }
      }
    }
    __syncthreads();
    // ---- every wave reduces its share of the tile's bins ----
    for (int bi = wave * SPW + slot; bi < nbt; bi += 4 * SPW) {
      const BinRec r = lbin[bi];
      const bool direct = (r.tb.w >> 30) & 1;
      const int rib = r.tb.w & 0x3fffffff;
      Kahan acc[FPL];
      if (!direct) {
        for (int k = r.tb.y - e_beg; k < r.tb.z - e_beg; ++k) {
          const uint2 e = lent[k];
          const float wk = __uint_as_float(e.y);
#pragma unroll
          for (int j = 0; j < FPL; ++j)
            acc[j].add(to_f(plane[(f0 + j) * S + (int)e.x]) * wk);
        }
      } else {
        for (int k = r.tb.y; k < r.tb.z; ++k) {  // entries stay in global memory
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
#pragma unroll
      for (int j = 0; j < FPL; ++j) {
        const int f = f0 + j;
        if (f < n_frames) {
          const float val = (acc[j].value() - r.info.y) * fabsf(r.info.x);
          const float o = flagged(r.info.x) ? NAN : val;
          cakes[(int64_t)f * nb + r.tb.x] = o;
          if (want_lo) {
            float *L = lo + (size_t)rib * 4 * B;
            if (o != 0.f && !isnan(o)) {
              atomicAdd(&L[0 * B + f], o);
              atomicAdd(&L[1 * B + f], 1.f);
            }
            if (r.a > 1e-9f && !isnan(o)) {
              atomicAdd(&L[2 * B + f], o * r.a);
              atomicAdd(&L[3 * B + f], o);
            }
          }
        }
      }
    }
  }
  if (want_lo) {
    __syncthreads();
    float *P = part + (int64_t)c * rows_per_group * 4 * B;
    for (int i = tid; i < rows_per_group * 4 * B; i += 256) P[i] = lo[i];
  }
}

// Lineouts from the per-chunk partial sums: one thread per (frame, row);
// the chunks of a row block are [block_chunk[b], block_chunk[b+1]).
__global__ void k_lineout_finalize(const float *__restrict__ part,
                                   const int32_t *__restrict__ block_chunk,
                                   const float4 *__restrict__ row_static, int n_frames, int n_r,
                                   int rows_per_group, int B, int mask_nan,
                                   float *__restrict__ lineouts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n_frames * n_r) return;
  const int f = (int)(i / n_r), r = (int)(i % n_r);
  const int blk = r / rows_per_group, rib = r - blk * rows_per_group;
  float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int c = block_chunk[blk]; c < block_chunk[blk + 1]; ++c) {
    const float *L = part + ((int64_t)c * rows_per_group + rib) * 4 * B;
    s.x += L[f];
    s.y += L[B + f];
    s.z += L[2 * B + f];
    s.w += L[3 * B + f];
  }
  write_lineouts(s, row_static[r], mask_nan, lineouts + (int64_t)f * 3 * n_r, n_r, r);
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

int launch_plain(hipStream_t s, const Map &m, const Csr &c, const float2 *binfo, int mask_nan,
                 const void *frames, int dtype, float *cake, float *lineouts, int n_frames,
                 int64_t frame_stride) {
  return dispatch_dtype(dtype, [&](auto tag) -> int {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(tag)>>;
    hipLaunchKernelGGL(k_plain<T>, dim3(m.n_r, n_frames), dim3(256), 0, s,
                       c.row.as<int32_t>(), c.pix.as<int32_t>(), c.w.as<float>(), binfo,
                       m.d_area_inb.as<float>(), m.d_row_static.as<float4>(),
                       (const T *)frames, frame_stride, m.n_eta, m.n_r, m.nb(), mask_nan,
                       cake, lineouts);
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

template <typename TIn, int B, int V>
static int launch_batched_v(hipStream_t s, const Map &m, const Tiles &t, const float2 *binfo,
                            int mask_nan, const void *frames, int n_frames,
                            const int32_t *slots, float *cakes, float *lineouts) {
  using TL = typename std::conditional<sizeof(TIn) <= 2, uint16_t, float>::type;
  auto kern = k_batched<TIn, TL, B, V>;
  const size_t lds = t.lds_needed;
  if (lds > 65536)
    MC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
  float *part = lineouts ? t.part.as<float>() : nullptr;
  hipLaunchKernelGGL(kern, dim3(t.n_chunks), dim3(256), lds, s, t.chunks.as<int32_t>(),
                     t.tiles.as<int32_t>(), t.tbins.as<int4>(), t.ents.as<uint2>(), binfo,
                     m.d_area_inb.as<float>(), (const TIn *)frames, m.npx(), slots, n_frames,
                     m.ny, t.rows, m.nb(), cakes, part);
  MC_HIP(hipGetLastError());
  if (lineouts) {
    const int64_t n = (int64_t)n_frames * m.n_r;
    hipLaunchKernelGGL(k_lineout_finalize, dim3(blocks_for(n, 256)), dim3(256), 0, s, part,
                       t.block_chunk.as<int32_t>(), m.d_row_static.as<float4>(), n_frames,
                       m.n_r, t.rows, B, mask_nan, lineouts);
    MC_HIP(hipGetLastError());
  }
  return MC_OK;
}

template <typename TIn, int B>
static int launch_batched_t(hipStream_t s, const Map &m, const Tiles &t, const float2 *binfo,
                            int mask_nan, const void *frames, int n_frames,
                            const int32_t *slots, float *cakes, float *lineouts) {
  constexpr int VV = sizeof(TIn) < 8 ? 8 / (int)sizeof(TIn) : 1;
  if (t.vec == VV)
    return launch_batched_v<TIn, B, VV>(s, m, t, binfo, mask_nan, frames, n_frames, slots,
                                        cakes, lineouts);
  return launch_batched_v<TIn, B, 1>(s, m, t, binfo, mask_nan, frames, n_frames, slots, cakes,
                                     lineouts);
}

int launch_batched(hipStream_t s, const Map &m, const Tiles &t, const float2 *binfo,
                   int mask_nan, const void *frames, int dtype, int n_frames,
                   const int32_t *slots, float *cakes, float *lineouts) {
  return dispatch_dtype(dtype, [&](auto tag) -> int {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(tag)>>;
#define MC_LB(Bv) \
  launch_batched_t<T, Bv>(s, m, t, binfo, mask_nan, frames, n_frames, slots, cakes, lineouts)
    switch (t.batch) {
      case 8: return MC_LB(8);
      case 16: return MC_LB(16);
      case 32: return MC_LB(32);
      case 64: return MC_LB(64);
      case 128: return MC_LB(128);
#undef MC_LB
      default:
        set_error("batch %d not in {8,16,32,64,128}", t.batch);
        return MC_ERR_ARG;
    }
  });
}

}  // namespace mc
