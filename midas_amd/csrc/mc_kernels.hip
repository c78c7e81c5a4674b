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
#include <type_traits>

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
// Batched kernel.  The cake rows are cut into bands of 16*G rows and a band
// into tiles: rectangles of rows x eta columns whose pixel footprint (the
// bounding box of their map entries, padded to whole staging vectors) fits
// the LDS plane budget.  A workgroup of 16 waves takes one chunk -- a run of
// consecutive tiles of one band -- and for each tile stages into LDS
//   * the tile's bin records (cake bin, entry range, scale/dark, area),
//   * its map entries (pixel index inside the box, weight),
//   * the box of every frame of the batch, one plane per frame (plane
//     stride S elements, so that lanes reading the same pixel of
//     consecutive frames hit different LDS banks),
// then reduces the bins with lane = frame: one entry read (an LDS
// broadcast) serves every frame of the batch.  Wave w owns the band rows
// j = w (mod 16), so it keeps the lineout sums of its G rows in registers
// across the whole chunk and writes them once at the end;
// k_lineout_finalize adds the chunks of each band.  Bins whose own
// footprint exceeds the budget ("direct" bins) gather from global memory.
//   chunk: tile_begin tile_end band_row0 band_rows
//   tile : y0 z0 bwp bh kshift S b_beg b_end e_beg e_end ra nc pv 0
//          (box of bwp x bh pixels at (y0, z0), bwp = (1 << kshift) * V;
//          plane stride S elements with a zero pad element at S-1;
//          band rows ra .. ra + (b_end - b_beg) / nc - 1, nc columns,
//          records row-major; pv = staging vectors per frame plane)
//   rec  : cake bin, ent_begin, ent_end, direct flag, scale, dark term,
//          in-bounds area
//   ent  : (byte offset of the pixel in its box plane, or global pixel if
//          direct), weight
// The loads of tile t+1 (this thread's share of the boxes, records and
// entries, kStageVecs staging vectors at most) are issued into registers
// before the waves reduce tile t, so global latency hides behind the
// reduction.
// Workgroup b takes chunk (b % 8) * per + b / 8: hardware hands
// consecutive workgroups to the 8 XCDs in turn, so each XCD works through
// one contiguous range of bands and neighbouring bands share its L2.
// ---------------------------------------------------------------------
constexpr int kWaves = 16;
constexpr int kTileBins = 256;    // bin records per tile (LDS)
constexpr int kTileEnts = 2048;   // staged map entries per tile (LDS)
constexpr int kTileInts = 14;
constexpr int kStageVecs = 8;     // staging vectors per thread and tile

struct BinRec {
  int4 tb;
  float2 info;
  float a;
  float pad;
};

template <typename TIn, int V> struct alignas(sizeof(TIn) * V) VecLoad {
  TIn v[V];
};

// Cake values are written per lane (frame) in whole granules of LG
// consecutive bins: each row slot of a wave keeps the granule it is
// filling in a register vector (the position inside the granule is a
// run-time index, which the hardware's register indexing serves) and
// stores it when its next bin falls in another granule.  Storing a frame's
// bins one at a time would leave partially written lines for the L2 to
// evict (a 16-byte write per 4-byte value).
template <int LG> struct GranVec;
template <> struct GranVec<16> { typedef float type __attribute__((ext_vector_type(16))); };
template <> struct GranVec<8> { typedef float type __attribute__((ext_vector_type(8))); };
typedef float f4v __attribute__((ext_vector_type(4)));

template <int LG, typename VT>
__device__ __forceinline__ void flush_granule(const VT &v, int idx, uint32_t mask, int f,
                                              float *__restrict__ cakes, int64_t nb,
                                              bool vec_ok) {
  float *dst = cakes + (int64_t)f * nb + (int64_t)idx * LG;
  if (mask == ((1u << LG) - 1u) && vec_ok) {
#pragma unroll
    for (int q = 0; q < LG; q += 4) {
      f4v w;
      w.x = v[q];
      w.y = v[q + 1];
      w.z = v[q + 2];
      w.w = v[q + 3];
      *reinterpret_cast<f4v *>(dst + q) = w;
    }
  } else {
#pragma unroll
    for (int q = 0; q < LG; ++q)
      if ((mask >> q) & 1u) dst[q] = v[q];
  }
}

// Registers of one tile's prefetch: this thread's share of the frame
// boxes (kStageVecs staging vectors), of the bin records and entries.
typedef int i4v __attribute__((ext_vector_type(4)));
typedef unsigned u2v __attribute__((ext_vector_type(2)));

template <typename TIn, int V> struct TilePrefetch {
  VecLoad<TIn, V> v[kStageVecs];
  i4v r0, r1;  // one 32-byte bin record
  u2v e[kTileEnts / (64 * kWaves)];
};

struct TileHdr {
  int y0, z0, bwp, bh, ks, S, b_beg, nbt, e_beg, net, ra, nc, pv;
};

__device__ __forceinline__ TileHdr tile_hdr(const int32_t *__restrict__ tiles, int t) {
  const int32_t *T = tiles + kTileInts * t;
  TileHdr h;
  h.y0 = T[0];
  h.z0 = T[1];
  h.bwp = T[2];
  h.bh = T[3];
  h.ks = T[4];
  h.S = T[5];
  h.b_beg = T[6];
  h.nbt = T[7] - T[6];
  h.e_beg = T[8];
  h.net = T[9] - T[8];
  h.ra = T[10];
  h.nc = T[11];
  h.pv = T[12];
  return h;
}

// Staging slot u of a thread: frame wave + kWaves * (u >> spf_log) (the
// wave's frames), vector q = lane + 64 * (u & (spf - 1)) of that frame's
// box.  Boxes are at most 64 vectors wide, so a lane keeps its column.
template <typename TIn, int V>
__device__ __forceinline__ void tile_issue(TilePrefetch<TIn, V> &P, const TileHdr &h, int tid,
                                           int lane, int wave, int spf_log,
                                           const BinRec *__restrict__ recs,
                                           const uint2 *__restrict__ ents,
                                           const TIn *__restrict__ frames, int64_t frame_stride,
                                           int n_frames, int ny) {
  if (tid < h.nbt) {
    const i4v *rp = reinterpret_cast<const i4v *>(recs + h.b_beg + tid);
    P.r0 = rp[0];
    P.r1 = rp[1];
  }
#pragma unroll
  for (int q = 0; q < kTileEnts / (64 * kWaves); ++q) {
    const int k = tid + 64 * kWaves * q;
    if (k < h.net) P.e[q] = reinterpret_cast<const u2v *>(ents)[h.e_beg + k];
  }
  if (h.bwp > 0) {
    const int cv = lane & ((1 << h.ks) - 1), r0 = lane >> h.ks, rstep = 64 >> h.ks;
    const TIn *col = frames + (int64_t)(h.z0 + r0) * ny + h.y0 + cv * V;
#pragma unroll
    for (int u = 0; u < kStageVecs; ++u) {
      const int f = wave + kWaves * (u >> spf_log);
      const int s = u & ((1 << spf_log) - 1);
      if (f < n_frames && lane + 64 * s < h.pv)
        P.v[u] = *reinterpret_cast<const VecLoad<TIn, V> *>(
            col + (int64_t)f * frame_stride + (int64_t)(s * rstep) * ny);
    }
  }
}

template <typename TIn, typename TL, int V>
__device__ __forceinline__ void tile_commit(const TilePrefetch<TIn, V> &P, const TileHdr &h,
                                            int tid, int lane, int wave, int spf_log,
                                            int n_frames, BinRec *lbin, uint2 *lent, TL *plane) {
  if (tid < h.nbt) {
    i4v *rp = reinterpret_cast<i4v *>(lbin + tid);
    rp[0] = P.r0;
    rp[1] = P.r1;
  }
#pragma unroll
  for (int q = 0; q < kTileEnts / (64 * kWaves); ++q) {
    const int k = tid + 64 * kWaves * q;
    if (k < h.net) reinterpret_cast<u2v *>(lent)[k] = P.e[q];
  }
  // Zero pad element S-1 of every plane: the padding entries of each bin
  // (weight 0) point at it.
  if (tid < n_frames) plane[tid * h.S + h.S - 1] = static_cast<TL>(0);
  if (h.bwp > 0) {
    const int cv = lane & ((1 << h.ks) - 1), r0 = lane >> h.ks, rstep = 64 >> h.ks;
    const int colo = r0 * h.bwp + cv * V;
#pragma unroll
    for (int u = 0; u < kStageVecs; ++u) {
      const int f = wave + kWaves * (u >> spf_log);
      const int s = u & ((1 << spf_log) - 1);
      if (f < n_frames && lane + 64 * s < h.pv) {
        const int dst = f * h.S + colo + s * rstep * h.bwp;
        if constexpr (sizeof(TL) == 2 && V % 2 == 0) {
          // pairs of 16-bit values as 32-bit words (dst is even)
#pragma unroll
          for (int q = 0; q < V; q += 2)
            *reinterpret_cast<uint32_t *>(plane + dst + q) =
                (uint32_t)(uint16_t)P.v[u].v[q] | ((uint32_t)(uint16_t)P.v[u].v[q + 1] << 16);
        } else {
#pragma unroll
          for (int q = 0; q < V; ++q) plane[dst + q] = static_cast<TL>(P.v[u].v[q]);
        }
      }
    }
  }
}

// Compensated sum: products are added in blocks of 8 with plain FMAs and
// each block sum enters a Kahan accumulator (relative error about 8 ulp of
// a block plus 2 ulp, independent of the entry count).
__device__ __forceinline__ void kahan_add(Kahan &k, float x) {
  const float y = x - k.c;
  const float t = k.s + y;
  k.c = (t - k.s) - y;
  k.s = t;
}

template <typename TIn, typename TL, int FPL, int V, int G>
__global__ __launch_bounds__(64 * kWaves) void k_batched(
    const int32_t *__restrict__ chunks, int n_chunks, const int32_t *__restrict__ tiles,
    const BinRec *__restrict__ recs, const uint2 *__restrict__ ents,
    const TIn *__restrict__ frames, int64_t frame_stride, int n_frames, int spf_log, int ny,
    int64_t nb, float *__restrict__ cakes, float *__restrict__ part) {
  extern __shared__ __attribute__((aligned(16))) char lds_raw[];
  constexpr int NP = 64 * FPL;  // planes (frames) per launch
  constexpr int RB = kWaves * G;
  BinRec *lbin = reinterpret_cast<BinRec *>(lds_raw);
  uint2 *lent = reinterpret_cast<uint2 *>(lds_raw + kTileBins * sizeof(BinRec));
  TL *plane = reinterpret_cast<TL *>(lds_raw + kTileBins * sizeof(BinRec) +
                                     kTileEnts * sizeof(uint2));
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int per = (n_chunks + 7) / 8;
  const int c = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (c >= n_chunks) return;
  const int t_beg = chunks[4 * c], t_end = chunks[4 * c + 1];
  const bool want_lo = part != nullptr;
  constexpr int LG = 16 / FPL;
  typedef typename GranVec<LG>::type VT;
  float lo[G][4][FPL];
  VT ob[G][FPL];
  int og[G];
  uint32_t om[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    og[g] = -1;
    om[g] = 0u;
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int j = 0; j < FPL; ++j) lo[g][k][j] = 0.f;
#pragma unroll
    for (int j = 0; j < FPL; ++j) ob[g][j] = 0.f;
  }
  const bool vec_ok = ((nb & 3) == 0) && (((uintptr_t)cakes & 15) == 0);

  // Software pipeline: the loads of tile t+1 are in flight while the waves
  // reduce tile t.
  TilePrefetch<TIn, V> P;
  TileHdr hn = tile_hdr(tiles, t_beg);
  tile_issue(P, hn, tid, lane, wave, spf_log, recs, ents, frames, frame_stride, n_frames, ny);
  for (int t = t_beg; t < t_end; ++t) {
    const TileHdr h = hn;
    __syncthreads();  // every wave is done with the previous tile's LDS
    tile_commit<TIn, TL, V>(P, h, tid, lane, wave, spf_log, n_frames, lbin, lent, plane);
    __syncthreads();
    if (t + 1 < t_end) {
      hn = tile_hdr(tiles, t + 1);
      tile_issue(P, hn, tid, lane, wave, spf_log, recs, ents, frames, frame_stride, n_frames,
                 ny);
    }
    const int nrows = h.nbt / h.nc;
    // Byte address of this lane's plane(s); entries hold byte offsets.
    // Lanes past the batch read the last staged plane; their sums are
    // dropped.
    const char *pb[FPL];
#pragma unroll
    for (int j = 0; j < FPL; ++j)
      pb[j] = reinterpret_cast<const char *>(plane) +
              (size_t)min(lane + 64 * j, n_frames - 1) * h.S * sizeof(TL);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int jr = kWaves * g + wave - h.ra;  // record row inside the tile
      if (jr < 0 || jr >= nrows) continue;
      const BinRec *rrow = lbin + jr * h.nc;
      BinRec rn = rrow[0];
      for (int e = 0; e < h.nc; ++e) {
        const BinRec r = rn;
        if (e + 1 < h.nc) rn = rrow[e + 1];  // next record early: hides its LDS latency
        Kahan acc[FPL];
        if (!r.tb.w) {
          // Staged entries come in whole groups of 8 (zero-weight padding on
          // the plane's zero pad element), 16-byte aligned in LDS.
          const int kb = r.tb.y - h.e_beg, ke = r.tb.z - h.e_beg;
          for (int k = kb; k < ke; k += 8) {
            uint2 en[8];
            const uint4 *ep = reinterpret_cast<const uint4 *>(lent + k);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const uint4 two = ep[q];
              en[2 * q] = make_uint2(two.x, two.y);
              en[2 * q + 1] = make_uint2(two.z, two.w);
            }
            TL pv[8][FPL];
#pragma unroll
            for (int q = 0; q < 8; ++q)
#pragma unroll
              for (int j = 0; j < FPL; ++j)
                pv[q][j] = *reinterpret_cast<const TL *>(pb[j] + en[q].x);
            float gs[FPL];
#pragma unroll
            for (int j = 0; j < FPL; ++j) gs[j] = to_f(pv[0][j]) * __uint_as_float(en[0].y);
#pragma unroll
            for (int q = 1; q < 8; ++q)
#pragma unroll
              for (int j = 0; j < FPL; ++j)
                gs[j] = fmaf(to_f(pv[q][j]), __uint_as_float(en[q].y), gs[j]);
#pragma unroll
            for (int j = 0; j < FPL; ++j) kahan_add(acc[j], gs[j]);
          }
        } else {
          for (int k = r.tb.y; k < r.tb.z; ++k) {  // entries stay in global memory
            const uint2 en = ents[k];
#pragma unroll
            for (int j = 0; j < FPL; ++j) {
              const int f = lane + 64 * j;
              if (f < n_frames)
                kahan_add(acc[j],
                          to_f(frames[(int64_t)f * frame_stride + en.x]) * __uint_as_float(en.y));
            }
          }
        }
        const int gr = r.tb.x / LG, gk = r.tb.x - gr * LG;
        if (gr != og[g]) {
          if (om[g])
#pragma unroll
            for (int j = 0; j < FPL; ++j)
              if (lane + 64 * j < n_frames)
                flush_granule<LG>(ob[g][j], og[g], om[g], lane + 64 * j, cakes, nb, vec_ok);
          og[g] = gr;
          om[g] = 0u;
        }
        om[g] |= 1u << gk;
#pragma unroll
        for (int j = 0; j < FPL; ++j) {
          const float val = (acc[j].value() - r.info.y) * fabsf(r.info.x);
          const float o = flagged(r.info.x) ? NAN : val;
          ob[g][j][gk] = o;
          if (o != 0.f && !isnan(o)) {
            lo[g][0][j] += o;
            lo[g][1][j] += 1.f;
          }
          if (r.a > 1e-9f && !isnan(o)) {
            lo[g][2][j] += o * r.a;
            lo[g][3][j] += o;
          }
        }
      }
    }
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    if (om[g])
#pragma unroll
      for (int j = 0; j < FPL; ++j)
        if (lane + 64 * j < n_frames)
          flush_granule<LG>(ob[g][j], og[g], om[g], lane + 64 * j, cakes, nb, vec_ok);
    if (want_lo) {
      // part[c][row][k][NP] for the band row this slot holds.
      float *Pp = part + ((int64_t)c * RB + kWaves * g + wave) * 4 * NP;
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int j = 0; j < FPL; ++j) Pp[k * NP + lane + 64 * j] = lo[g][k][j];
    }
  }
}

// Tile bin records with the per-bin scale/dark and in-bounds area of the
// current (mode, flags, dark), so the kernel reads each record once.
__global__ void k_tile_recs(const int4 *__restrict__ tbins, const float2 *__restrict__ binfo,
                            const float *__restrict__ a_inb, int64_t n, BinRec *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  BinRec r;
  r.tb = tbins[i];
  r.info = binfo[r.tb.x];
  r.a = a_inb[r.tb.x];
  r.pad = 0.f;
  out[i] = r;
}

// Lineouts from the per-chunk partial sums: one thread per (frame, row);
// the chunks of band b are [band_chunk[b], band_chunk[b+1]).
__global__ void k_lineout_finalize(const float *__restrict__ part,
                                   const int32_t *__restrict__ band_chunk,
                                   const float4 *__restrict__ row_static, int n_frames, int n_r,
                                   int band_rows, int np, int mask_nan,
                                   float *__restrict__ lineouts) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n_frames * n_r) return;
  // Consecutive threads take consecutive frames of one row: the partials
  // of a (chunk, row) are frame-contiguous.
  const int r = (int)(i / n_frames), f = (int)(i % n_frames);
  const int band = r / band_rows, rib = r - band * band_rows;
  float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int c = band_chunk[band]; c < band_chunk[band + 1]; ++c) {
    const float *L = part + ((int64_t)c * band_rows + rib) * 4 * np;
    s.x += L[f];
    s.y += L[np + f];
    s.z += L[2 * np + f];
    s.w += L[3 * np + f];
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

int launch_tile_recs(hipStream_t s, const Map &m, Tiles &t, const float2 *binfo) {
  if (!t.recs.p) {
    int rc = t.recs.alloc((size_t)t.n_tbins * sizeof(BinRec));
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_tile_recs, dim3(blocks_for(t.n_tbins, 256)), dim3(256), 0, s,
                     t.tbins.as<int4>(), binfo, m.d_area_inb.as<float>(), t.n_tbins,
                     t.recs.as<BinRec>());
  MC_HIP(hipGetLastError());
  return MC_OK;
}

template <typename TIn, int FPL, int V, int G>
static int launch_batched_v(hipStream_t s, const Map &m, const Tiles &t, const float2 *binfo,
                            int mask_nan, const void *frames, int n_frames, float *cakes,
                            float *lineouts) {
  using TL = typename std::conditional<sizeof(TIn) <= 2, uint16_t, float>::type;
  auto kern = k_batched<TIn, TL, FPL, V, G>;
  if (V > 1 && ((uintptr_t)frames % (sizeof(TIn) * V)) != 0) {
    set_error("frame buffer not %d-byte aligned for vector staging", (int)(sizeof(TIn) * V));
    return MC_ERR_ARG;
  }
  if (n_frames < 1 || n_frames > t.batch) {
    set_error("batch of %d frames outside 1..%d", n_frames, t.batch);
    return MC_ERR_ARG;
  }
  const size_t lds = t.lds_needed;
  if (lds > 65536)
    MC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
  float *part = lineouts ? t.part.as<float>() : nullptr;
  const unsigned grid = 8u * (unsigned)((t.n_chunks + 7) / 8);
  (void)binfo;  // folded into the tile records (launch_tile_recs)
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * kWaves), lds, s, t.chunks.as<int32_t>(),
                     t.n_chunks, t.tiles.as<int32_t>(), t.recs.as<BinRec>(), t.ents.as<uint2>(),
                     (const TIn *)frames, m.npx(), n_frames, t.spf_log, m.ny, m.nb(), cakes, part);
  MC_HIP(hipGetLastError());
  if (lineouts) {
    const int64_t n = (int64_t)n_frames * m.n_r;
    hipLaunchKernelGGL(k_lineout_finalize, dim3(blocks_for(n, 256)), dim3(256), 0, s, part,
                       t.band_chunk.as<int32_t>(), m.d_row_static.as<float4>(), n_frames,
                       m.n_r, t.rows, 64 * FPL, mask_nan, lineouts);
    MC_HIP(hipGetLastError());
  }
  return MC_OK;
}

template <typename TIn, int FPL, int V>
static int launch_batched_g(hipStream_t s, const Map &m, const Tiles &t, const float2 *binfo,
                            int mask_nan, const void *frames, int n_frames, float *cakes,
                            float *lineouts) {
  switch (t.rows / kWaves) {
    case 1:
      return launch_batched_v<TIn, FPL, V, 1>(s, m, t, binfo, mask_nan, frames, n_frames, cakes,
                                              lineouts);
    case 2:
      return launch_batched_v<TIn, FPL, V, 2>(s, m, t, binfo, mask_nan, frames, n_frames, cakes,
                                              lineouts);
    default:
      set_error("rows per band %d not in {16, 32}", t.rows);
      return MC_ERR_ARG;
  }
}

template <typename TIn, int FPL>
static int launch_batched_t(hipStream_t s, const Map &m, const Tiles &t, const float2 *binfo,
                            int mask_nan, const void *frames, int n_frames, float *cakes,
                            float *lineouts) {
  constexpr int VV = sizeof(TIn) < 8 ? 8 / (int)sizeof(TIn) : 1;
  if (t.vec == VV)
    return launch_batched_g<TIn, FPL, VV>(s, m, t, binfo, mask_nan, frames, n_frames, cakes,
                                          lineouts);
  if constexpr (VV != 1) {
    if (t.vec != 1) {
      set_error("staging vector %d not supported", t.vec);
      return MC_ERR_STATE;
    }
    return launch_batched_g<TIn, FPL, 1>(s, m, t, binfo, mask_nan, frames, n_frames, cakes,
                                         lineouts);
  }
  set_error("staging vector %d not supported", t.vec);
  return MC_ERR_STATE;
}

int launch_batched(hipStream_t s, const Map &m, const Tiles &t, const float2 *binfo,
                   int mask_nan, const void *frames, int dtype, int n_frames, float *cakes,
                   float *lineouts) {
  return dispatch_dtype(dtype, [&](auto tag) -> int {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(tag)>>;
    if (t.fpl == 1)
      return launch_batched_t<T, 1>(s, m, t, binfo, mask_nan, frames, n_frames, cakes,
                                    lineouts);
    if (t.fpl == 2)
      return launch_batched_t<T, 2>(s, m, t, binfo, mask_nan, frames, n_frames, cakes,
                                    lineouts);
    set_error("frames per lane %d not in {1, 2}", t.fpl);
    return MC_ERR_STATE;
  });
}

}  // namespace mc
