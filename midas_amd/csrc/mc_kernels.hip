// mc_kernels.hip -- the gather-reduce kernels of the caking path (gfx950).
//
// Per (R, eta) bin b and frame f the cake value is
//     I[f][b] = ( sum_k (frame_f[pix_k] - dhi[pix_k]) * w_k  -  D_b ) * s_b
// where (pix_k, w_k) are the bin's CSR entries (w = Map.bin frac, or frac
// times a bilinear corner weight), dhi is the dark rounded to an integer
// (clamped to [0, 65535]), D_b = sum_k (dark - dhi)[pix_k] * w_k is the
// bin's dark remainder (computed once per dark in FP64), and s_b =
// 1 / sum areaWeight (Normalize=1) or 1.  This is the per-bin loop of
// IntegratorZarrOMP.c:1710-1778 with "Image = raw - dark" (:1646-1669):
// the integer part of the dark is subtracted from each pixel before the
// product, exactly (integer counts and integer darks are exact in FP32), so
// the FP32 sums see dark-subtracted values even under a dark pedestal of
// thousands of counts; only the fractional remainder, below 0.5 a pixel,
// is distributed over the sum.  Since |dark - dhi| <= |frame - dark| on a
// pixel whose count equals dhi, and |frame - dark| >= 1/2 on any other,
// the remainder is at most twice sum_k w_k |frame - dark|[pix_k]: the FP32
// rounding stays relative to the dark-subtracted signal of the bin.  Bins
// flagged in maskMap.bin carry a NaN
// s_b when MC_F_MASK_NAN is set and come out NaN (IntegratorZarrOMP.c
// :1766-1767).
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
// Sum over aligned groups of N lanes (N = 4, 8 or 16), every lane of a
// group receiving the total: quad swaps, then the half-row and row mirrors.
template <int N> __device__ __forceinline__ float dpp_sum(float v) {
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xf, 0xf, false));
  v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xf, 0xf, false));
  if (N >= 8)
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xf, 0xf, false));
  if (N >= 16)
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
// One-frame kernel: workgroup = one R row of one frame, LPB lanes per bin
// (4, 8 or 16, a tuning of the map; 8 by default: a bin of about 15
// entries leaves half of 16 lanes idle, while 4 lanes lengthen each lane's
// chain of dependent loads).  grid = (n_r, n_frames).
// ---------------------------------------------------------------------
//
// With a dark the kernel reads frames already dark-subtracted by
// k_frame_dark (FP32, count minus the pixel's integer dark, exact): a second
// gather per entry for the dark costs more than that elementwise pass.
template <typename T, int LPB>
__global__ __launch_bounds__(256) void k_plain(
    const int32_t *__restrict__ row, const int32_t *__restrict__ pix,
    const float *__restrict__ w, const float2 *__restrict__ binfo,
    const float *__restrict__ a_inb, const float4 *__restrict__ row_static,
    const T *__restrict__ frames, int64_t frame_stride, int n_eta, int n_r, int64_t nb,
    int mask_nan, float *__restrict__ cake, float *__restrict__ lineouts) {
  constexpr int SLOTS = 256 / LPB;
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
    float v = dpp_sum<LPB>(acc.value());
    if (sub == 0) {
      const float2 bi = binfo[b];
      const float o = (v - bi.y) * bi.x;  // NaN scale: a masked bin
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
// One-frame kernel with variance propagation (integrate_with_variance,
// kernels.py:476-561): beside each cake value, the sum of w^2 * var over
// the same merged (bin, pixel) weights, scaled by the square of the value's
// own scale.  var is the dark-subtracted frame (Poisson: Var(I) = I; the
// integer dark per pixel, the remainder's sum w^2 (dark - dint) per bin)
// or a variance image.  With a dark the frames are k_frame_dark's output.
// grid = (n_r, n_frames).
// ---------------------------------------------------------------------
template <typename T, bool POISSON, int LPB>
__global__ __launch_bounds__(256) void k_plain_var(
    const int32_t *__restrict__ row, const int32_t *__restrict__ pix,
    const float *__restrict__ w, const float2 *__restrict__ binfo,
    const float *__restrict__ d2term, const T *__restrict__ frames, const float *__restrict__ vimg,
    int64_t frame_stride, int n_eta, int64_t nb, float *__restrict__ cake,
    float *__restrict__ var) {
  constexpr int SLOTS = 256 / LPB;  // the one-frame kernel's order for the same lanes
  const int r = blockIdx.x, f = blockIdx.y;
  const T *__restrict__ img = frames + (int64_t)f * frame_stride;
  const float *__restrict__ vi = POISSON ? nullptr : vimg + (int64_t)f * frame_stride;
  const int slot = threadIdx.x / LPB, sub = threadIdx.x % LPB;
  for (int e = slot; e < n_eta; e += SLOTS) {
    const int64_t b = (int64_t)r * n_eta + e;
    const int beg = row[b], end = row[b + 1];
    Kahan acc, acc2;
    for (int k = beg + sub; k < end; k += LPB) {
      const int p = pix[k];
      const float x = to_f(img[p]);
      const float ww = w[k];
      acc.add(x * ww);
      acc2.add((POISSON ? x : vi[p]) * ww * ww);
    }
    const float v = dpp_sum<LPB>(acc.value());
    const float v2 = dpp_sum<LPB>(acc2.value());
    if (sub == 0) {
      const float2 bi = binfo[b];
      const float sc = bi.x;  // NaN: a masked bin
      cake[(int64_t)f * nb + b] = (v - bi.y) * sc;
      const float raw = POISSON && d2term ? v2 - d2term[b] : v2;
      var[(int64_t)f * nb + b] = raw * sc * sc;
    }
  }
}

// ---------------------------------------------------------------------
// Frames minus the integer dark per pixel, as FP32 (exact for integer
// counts below 2^24): the one-frame kernels' input when a dark is set.
// Four pixels per thread; grid = (pixels / 1024, frames).
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_frame_dark(const T *__restrict__ frames, int64_t npx,
                                                    const uint16_t *__restrict__ dint,
                                                    float *__restrict__ out) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (i >= npx) return;
  const T *__restrict__ src = frames + (int64_t)blockIdx.y * npx;
  float *__restrict__ dst = out + (int64_t)blockIdx.y * npx;
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (i + q < npx) dst[i + q] = to_f(src[i + q]) - (float)dint[i + q];
}

// ---------------------------------------------------------------------
// Hybrid frames (wire dtype 6, IntegratorFitPeaksGPUStream.cu:1468-1494):
// a uint16 base image, then count (int32 pixel, int64 value) patches for
// the pixels whose count overflowed 16 bits.  Decoded to int32 (values
// clamped to the int32 range).  The patch arrays follow the base image
// unaligned, so they are read byte by byte.
// ---------------------------------------------------------------------
__global__ void k_hybrid_base(const uint16_t *__restrict__ base, int64_t npx,
                              int32_t *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < npx) out[i] = base[i];
}

__global__ void k_hybrid_patch(const uint8_t *__restrict__ idx, const uint8_t *__restrict__ val,
                               int64_t count, int64_t npx, int32_t *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint32_t p = 0;
  uint64_t v = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) p |= (uint32_t)idx[4 * i + b] << (8 * b);
#pragma unroll
  for (int b = 0; b < 8; ++b) v |= (uint64_t)val[8 * i + b] << (8 * b);
  const int64_t px = (int32_t)p, x = (int64_t)v;
  if (px >= 0 && px < npx)
    out[px] = (int32_t)(x > INT32_MAX ? INT32_MAX : (x < INT32_MIN ? INT32_MIN : x));
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

// Lineout of a cake with propagated variance (profile_1d_with_variance,
// packages/midas_integrate/midas_integrate/kernels.py:604-644): per (row
// r, frame f), over the bins whose in-bounds area is above 1e-9 and whose
// value is not NaN, with weight w = A (area_weighted) or 1 (simple_mean):
// I = sum w I / sum w and V = sum w^2 Var / (sum w)^2, 0 where no bin
// counts.  Sums in FP64.
__global__ __launch_bounds__(256) void k_profile_var(
    const float *__restrict__ cakes, const float *__restrict__ vars,
    const float *__restrict__ a_inb, int n_eta, int n_r, int64_t nb, int simple,
    float *__restrict__ prof, float *__restrict__ pvar) {
  const int r = blockIdx.x, f = blockIdx.y;
  const float *c = cakes + (int64_t)f * nb + (int64_t)r * n_eta;
  const float *v = vars + (int64_t)f * nb + (int64_t)r * n_eta;
  const float *a = a_inb + (int64_t)r * n_eta;
  double s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int e = threadIdx.x; e < n_eta; e += 256) {
    const float o = c[e];
    if (a[e] > 1e-9f && !isnan(o)) {
      const double w = simple ? 1.0 : (double)a[e];
      s1 += w * (double)o;
      s2 += w * w * (double)v[e];
      s3 += w;
    }
  }
  __shared__ double red[3][256];
  red[0][threadIdx.x] = s1;
  red[1][threadIdx.x] = s2;
  red[2][threadIdx.x] = s3;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double sw = red[2][0];
    const bool ok = simple ? sw > 0.0 : sw > 1e-9;
    prof[(int64_t)f * n_r + r] = ok ? (float)(red[0][0] / sw) : 0.f;
    pvar[(int64_t)f * n_r + r] = ok ? (float)(red[1][0] / (sw * sw)) : 0.f;
  }
}

// ---------------------------------------------------------------------
// Batched path: two kernels per pass of up to 64 frames.
//
// k_transpose turns the pass's frames [F][npx] into the pixel-major batch
// Xt[npx][W] (W = 16, 32 or 64 frames, frame fastest, frames past F zero):
// 2W bytes per pixel for 8/16-bit frames (kept as uint16), 4W bytes for
// the other types (FP32, exact for integer counts below 2^24).
//
// k_batched cuts the cake rows into bands of 16*G rows and a band into
// tiles: rectangles of rows x eta columns whose pixel footprint fits the
// LDS budget.  A workgroup of 16 waves takes one chunk -- every k-th tile
// of one band, the band's k chunks running side by side on one XCD -- and
// for each tile stages into LDS
//   * its map entries (byte offset of the pixel row in the footprint,
//     weight),
//   * the footprint of the batch: each pixel's W frames contiguous --
//     whole lines of Xt, so staging reads only what it keeps.
// Staging uses global->LDS copies (no registers), so a workgroup needs at
// most 64 VGPRs per lane and two workgroups share a CU: one stages its
// next tile while the other reduces.
//
// The bins of a band row are reduced in quads: four consecutive cake bins,
// the first a multiple of 4 (slots outside the tile's columns of the row
// hold no bin).  The four 16-lane groups of a wave take the four bins, one
// each, and lane l of a group holds frames l*W/16 .. (l+1)*W/16 - 1: one
// LDS read serves W/16 frames of one entry, and no lane exchange is needed
// to finish a bin.  A quad's staged slots (those with staged entries, P of
// them) are interleaved: pair m of the r-th staged slot at entries
// e_off + 2 (m P + r) and the next (pairs past a bin's end are
// zero-weight entries on a zero pixel row), so each iteration reads one
// 16-byte pair per group; the groups of the other slots read the tile's
// zero pair.  The staged entries of a row are one run, quad after quad,
// and the next pair is read an iteration ahead across full quads.
// The quad records (bins, flags, entry run, scale/dark/area of the four
// bins) are uniform over the wave and come in through scalar loads.
// After a quad, two lane-half swaps and two 16-lane swaps transpose the
// groups' 4 x 4 blocks (64-frame passes): lane (g, l) then holds frame
// 4l + g of the four bins, so every lane applies the bins' scale and dark
// term from scalar registers, and the four values fill half of an 8-bin
// granule of its frame in registers, stored whole (32 bytes).  Narrower
// passes store the values where they are (8 or 4 bytes).
// Wave w owns the band rows j = w (mod 16), so it keeps the lineout sums
// of its G rows in registers across the chunk; k_lineout_finalize adds the
// chunks of each band.  Bins whose own footprint exceeds the budget
// ("direct" bins) gather from Xt in global memory.
//   chunk: tile_begin tile_end band_row0 band_rows
//   tile : px_beg npx zp 0 q_beg nq e_beg e_end ra nc 0 0 nrows 0 0 0
//          (footprint of npx pixels listed at px_beg, band rows ra .. ra +
//          nrows - 1, columns of nc eta bins, nq quads per row: the quads
//          of row j at q_beg + j * nq, rows padded with empty quads; the
//          zero pair at entry zp of the tile)
//   quad : QuadRec (below); qdir[4 quad + s] = the direct entry range of
//          slot s
//   ent  : (byte offset of the pixel row in the footprint, or global pixel
//          if direct), weight
// Workgroup b takes chunk xcd[b % 8] + b / 8: hardware hands consecutive
// workgroups to the 8 XCDs in turn, so each XCD works through one
// contiguous range of bands of equal estimated cost, and neighbouring bands
// share its L2.
// ---------------------------------------------------------------------
constexpr int kWaves = 16;
constexpr int kTileInts = 16;
// Cache policy of a global->LDS copy read once (non-temporal).
constexpr int kNonTemporal = 2;
constexpr int kXtFrames = 64;     // frames per pixel row of the batch

// Four consecutive cake bins of one band row, the first a multiple of 4.
// Slots that hold no bin of the row's columns in the tile get scale, dark
// term and areas 0 (they add nothing and are not stored).
struct QuadRec {
  int32_t b0;     // cake bin of slot 0
  int32_t flags;  // bits 0-3: slot s holds a bin; bits 4-7: its entries are direct;
                  // bits 8-11: it has staged entries (a staged slot)
  int32_t e_off;  // first staged entry of the quad, from the tile's first
  int32_t n_it;   // staged entry pairs of each staged slot (the staged slots'
                  // pairs are interleaved: pair m of the r-th staged slot is
                  // entries e_off + 2 (m P + r), P staged slots)
  float scale[4];
  float dterm[4];
  float a[4];   // in-bounds area where above 1e-9, else 0
  float af[4];  // 1 where the in-bounds area is above 1e-9, else 0
};
static_assert(sizeof(QuadRec) == kQuadRecBytes, "quad record size");

typedef int i4v __attribute__((ext_vector_type(4)));
typedef unsigned u2v __attribute__((ext_vector_type(2)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

// ---- pixel-major batch ----
// FP32 batches (32/64-bit frames) carry the dark-subtracted value: the
// integer dark dhi of the pixel is subtracted here, exactly for integer
// counts below 2^24 (uint16 batches keep the raw count and subtract the
// dark in k_batched, from the staged entries).
template <typename TIn, typename TL, int XF>
__global__ __launch_bounds__(256) void k_transpose(const TIn *__restrict__ frames,
                                                   int64_t frame_stride, int n_frames,
                                                   int64_t npx, int vec_ok,
                                                   const int32_t *__restrict__ blocks,
                                                   const uint16_t *__restrict__ dint,
                                                   TL *__restrict__ xt) {
  constexpr int PX = 64;                       // pixels per block
  constexpr int EV = 16 / (int)sizeof(TL);     // elements per 16-byte store
  __shared__ TL tile[XF][PX + 4 / (int)sizeof(TL)];
  const int64_t p0 = (int64_t)(blocks ? blocks[blockIdx.x] : (int32_t)blockIdx.x) * PX;
  const int t = threadIdx.x;
  if (vec_ok && sizeof(TIn) == 2 && sizeof(TL) == 2) {
    // 16-byte loads: thread -> (frame t/4, 16 pixels as two vectors)
    const int f = t >> 2, q0 = (t & 3) * 16;
    if (f < XF) {
      u4v a = 0u, b = 0u;
      if (f < n_frames) {
        const u4v *src =
            reinterpret_cast<const u4v *>(frames + (int64_t)f * frame_stride + p0 + q0);
        a = __builtin_nontemporal_load(src);
        b = __builtin_nontemporal_load(src + 1);
      }
      uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        tile[f][q0 + 2 * q] = (TL)(w[q] & 0xffffu);
        tile[f][q0 + 2 * q + 1] = (TL)(w[q] >> 16);
      }
    }
  } else if (vec_ok && sizeof(TL) == 4) {
    // Vector loads: thread -> (frames t/16 + 16k, 4 consecutive pixels), so
    // 16 lanes read 64 pixels of one frame in one piece; the pixels' dark
    // is read once for the four passes.
    typedef TIn V4 __attribute__((ext_vector_type(4)));
    const int fs = t >> 4, q = (t & 15) * 4;
    float dk[4] = {0.f, 0.f, 0.f, 0.f};
    if (dint) {
      const u2v d2 = *reinterpret_cast<const u2v *>(dint + p0 + q);
      dk[0] = (float)(d2.x & 0xffffu);
      dk[1] = (float)(d2.x >> 16);
      dk[2] = (float)(d2.y & 0xffffu);
      dk[3] = (float)(d2.y >> 16);
    }
#pragma unroll
    for (int k = 0; k < XF / 16; ++k) {
      const int f = fs + 16 * k;
      float x[4] = {0.f, 0.f, 0.f, 0.f};  // frames past the batch stay zero
      if (f < n_frames) {
        const V4 v = __builtin_nontemporal_load(
            reinterpret_cast<const V4 *>(frames + (int64_t)f * frame_stride + p0 + q));
        x[0] = static_cast<float>(v.x) - dk[0];
        x[1] = static_cast<float>(v.y) - dk[1];
        x[2] = static_cast<float>(v.z) - dk[2];
        x[3] = static_cast<float>(v.w) - dk[3];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) tile[f][q + j] = (TL)x[j];
    }
  } else {
    const int f = t >> 2, q0 = (t & 3) * 16;
    const TIn *src = frames + (int64_t)f * frame_stride + p0 + q0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      if (f >= XF) break;
      TL v = static_cast<TL>(0);
      if (f < n_frames && p0 + q0 + q < npx) {
        if constexpr (sizeof(TL) == 4) {
          float x = static_cast<float>(src[q]);
          if (dint) x -= (float)dint[p0 + q0 + q];
          v = x;
        } else {
          v = static_cast<TL>(src[q]);
        }
      }
      tile[f][q0 + q] = v;
    }
  }
  __syncthreads();
  // 16-byte stores: thread -> (pixel, EV frames); XF / EV stores per pixel.
  constexpr int SPP = XF / EV;
  for (int job = t; job < PX * SPP; job += 256) {
    const int p = job / SPP, c = job - p * SPP;
    if (p0 + p >= npx) continue;
    u4v out;
    if constexpr (sizeof(TL) == 2) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        out[k] = (uint32_t)(uint16_t)tile[c * EV + 2 * k][p] |
                 ((uint32_t)(uint16_t)tile[c * EV + 2 * k + 1][p] << 16);
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) out[k] = __float_as_uint((float)tile[c * EV + k][p]);
    }
    __builtin_nontemporal_store(out, reinterpret_cast<u4v *>(xt + (p0 + p) * XF + c * EV));
  }
}

// The transposes read each frame once and write each pixel row of the
// batch once per pass: non-temporal loads and stores keep both streams
// from displacing the L2 lines the batched kernel reuses (footprint rows
// shared by neighbouring tiles).  On 2048^2 (64 frames) this took the
// launch from 655 to 624 us.
//
// uint16 frames, frame base 16-byte aligned, npx a multiple of 128: blocks
// of 128 pixels x 64 frames with 16-byte loads and LDS writes.  LDS holds
// [frame][16 chunks of 8 pixels]; chunk c of frame f sits at slot
// c ^ (f / 8), so the 2-byte gathers of the store phase (8 frame groups x
// 8 pixels per wave) fall on 32 different banks.
template <int XF>
__global__ __launch_bounds__(256) void k_transpose16(const uint16_t *__restrict__ frames,
                                                     int64_t frame_stride, int n_frames,
                                                     int64_t npx, const int32_t *__restrict__ blocks,
                                                     uint16_t *__restrict__ xt) {
  __shared__ __attribute__((aligned(16))) uint16_t tile[XF * 128];
  const int64_t p0 = (int64_t)(blocks ? blocks[blockIdx.x] : (int32_t)blockIdx.x) * 128;
  const int t = threadIdx.x;
#pragma unroll
  for (int u = 0; u < XF / 16; ++u) {
    const int j = t + 256 * u;
    const int f = j >> 4, c = j & 15;  // 16 consecutive threads read 256 bytes of a frame
    u4v v = 0u;
    if (f < n_frames)
      v = __builtin_nontemporal_load(
          reinterpret_cast<const u4v *>(frames + (int64_t)f * frame_stride + p0 + c * 8));
    *reinterpret_cast<u4v *>(tile + f * 128 + (c ^ (f >> 3)) * 8) = v;
  }
  __syncthreads();
  constexpr int GPP = XF / 8;  // 16-byte stores (8 frames) per pixel
#pragma unroll
  for (int u = 0; u < XF / 16; ++u) {
    const int k = t + 256 * u;
    const int p = k / GPP, g = k % GPP;  // pixel, group of 8 frames
    if (p0 + p >= npx) continue;
    const int ch = p >> 3, w = p & 7;
    u4v out;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int fa = 8 * g + 2 * q, fb = fa + 1;  // fa / 8 = fb / 8 = g
      const uint32_t a = tile[fa * 128 + (ch ^ g) * 8 + w];
      const uint32_t b = tile[fb * 128 + (ch ^ g) * 8 + w];
      out[q] = a | (b << 16);
    }
    __builtin_nontemporal_store(out, reinterpret_cast<u4v *>(xt + (p0 + p) * XF + 8 * g));
  }
}

// Four frames of one pixel row: uint16 x4 in 8 bytes or float x4 in 16.
template <typename TL> struct Quad;
template <> struct Quad<uint16_t> {
  typedef u2v type;
  __device__ __forceinline__ static void unpack(const u2v v, float o[4]) {
    o[0] = (float)(v.x & 0xffffu);
    o[1] = (float)(v.x >> 16);
    o[2] = (float)(v.y & 0xffffu);
    o[3] = (float)(v.y >> 16);
  }
};
template <> struct Quad<float> {
  typedef f4v type;
  __device__ __forceinline__ static void unpack(const f4v v, float o[4]) {
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
  }
};

// FPL frames of one pixel row for one lane (FPL = 4, 2 or 1: passes of 64,
// 32 or 16 frames).
template <typename TL, int FPL> struct Lanes;
template <typename TL> struct Lanes<TL, 4> {
  typedef typename Quad<TL>::type type;
  __device__ __forceinline__ static void unpack(const type v, float o[4]) {
    Quad<TL>::unpack(v, o);
  }
};
template <> struct Lanes<uint16_t, 2> {
  typedef uint32_t type;
  __device__ __forceinline__ static void unpack(const uint32_t v, float o[2]) {
    o[0] = (float)(v & 0xffffu);
    o[1] = (float)(v >> 16);
  }
};
template <> struct Lanes<uint16_t, 1> {
  typedef uint16_t type;
  __device__ __forceinline__ static void unpack(const uint16_t v, float o[1]) { o[0] = (float)v; }
};
template <> struct Lanes<float, 2> {
  typedef float type __attribute__((ext_vector_type(2)));
  __device__ __forceinline__ static void unpack(const type v, float o[2]) {
    o[0] = v.x;
    o[1] = v.y;
  }
};
template <> struct Lanes<float, 1> {
  typedef float type;
  __device__ __forceinline__ static void unpack(const float v, float o[1]) { o[0] = v; }
};

template <int LG> struct GranVec;
template <> struct GranVec<8> { typedef float type __attribute__((ext_vector_type(8))); };

// Stores a row slot's granule for this lane's frame f (8 consecutive bins;
// mask: the bins this workgroup holds).
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
  } else if ((mask == 0x0fu || mask == 0xf0u) && vec_ok) {
    // half a granule (a tile of 4 columns): one 16-byte store
    const int q = mask == 0x0fu ? 0 : 4;
    f4v w;
    w.x = v[q];
    w.y = v[q + 1];
    w.z = v[q + 2];
    w.w = v[q + 3];
    *reinterpret_cast<f4v *>(dst + q) = w;
  } else {
#pragma unroll
    for (int q = 0; q < LG; ++q)
      if ((mask >> q) & 1u) dst[q] = v[q];
  }
}

struct TileHdr {
  int px_beg, npx, zp, q_beg, nq, e_beg, net, ra, nrows;
};

// Tile record: {first listed pixel, pixels, zero pair (entry offset), 0,
// first quad, quads per row, entry range, first row, columns, 0, 0, rows,
// 0, 0, 0}.
__device__ __forceinline__ TileHdr tile_hdr(const int32_t *__restrict__ tiles, int t) {
  const int32_t *T = tiles + kTileInts * t;
  TileHdr h;
  h.px_beg = T[0];
  h.npx = T[1];
  h.zp = T[2];
  h.q_beg = T[4];
  h.nq = T[5];
  h.e_beg = T[6];
  h.net = T[7] - T[6];
  h.ra = T[8];
  h.nrows = T[12];
  return h;
}

typedef __attribute__((address_space(3))) void lds_void;

// Copies a tile's pixel list (npx int32 at a multiple of 4) to LDS with
// global->LDS copies (16 bytes per lane).
__device__ __forceinline__ void stage_pixels(const int32_t *__restrict__ tpix, int px_beg,
                                             int npx, int wave, int lane, char *dst) {
  const int nv = (npx + 3) >> 2;
  for (int ch = kWaves - 1 - wave; ch * 64 < nv; ch += kWaves) {
    const int j = ch * 64 + lane;
    if (j < nv)
      __builtin_amdgcn_global_load_lds(reinterpret_cast<const u4v *>(tpix + px_beg) + j,
                                       (lds_void *)(dst + ch * 1024), 16, 0, 0);
  }
}

// Stages one tile into LDS with gfx950 global->LDS copies (16 bytes per
// lane; a wave fills 1 KB of consecutive LDS, so the copies need no
// registers): the rows of the tile's pixels (its footprint, listed in LDS
// at ltp), the entries (at lent; a tile's entries start at a multiple of
// 8), and the zero pixel row after the staged rows (plain stores; the
// padding entries point at it).  The caller waits for the copies
// (s_waitcnt) and then for every wave (barrier).
template <typename TL, int XF>
__device__ __forceinline__ void tile_stage(const TileHdr &h, int wave, int lane, char *lent,
                                           char *plane, const uint2 *__restrict__ ents,
                                           const TL *__restrict__ xt,
                                           const int32_t *__restrict__ ltp) {
  constexpr int RB = XF * (int)sizeof(TL);  // bytes per pixel row
  constexpr int VPP = RB / 16;                     // 16-byte vectors per pixel row
  const int jobs = h.npx * VPP;
  for (int ch = wave; ch * 64 < jobs; ch += kWaves) {
    const int j = ch * 64 + lane;
    if (j < jobs) {
      const int p = ltp[j / VPP];
      const char *src = reinterpret_cast<const char *>(xt + (int64_t)p * XF) +
                        (j % VPP) * 16;
      __builtin_amdgcn_global_load_lds(src, (lds_void *)(plane + ch * 1024), 16, 0, 0);
    }
  }
  const int nev = h.net >> 1;  // 16-byte entry vectors
  for (int ch = kWaves - 1 - wave; ch * 64 < nev; ch += kWaves) {
    const int j = ch * 64 + lane;
    if (j < nev)
      __builtin_amdgcn_global_load_lds(reinterpret_cast<const u4v *>(ents + h.e_beg) + j,
                                       (lds_void *)(lent + ch * 1024), 16, 0, kNonTemporal);
  }
  if (wave == kWaves - 1 && lane < VPP)
    reinterpret_cast<u4v *>(plane + (size_t)h.npx * RB)[lane] = 0u;
}

// Sum over the four 16-lane groups, every group receiving the total (lane l
// of each group holds the same frames): the two lane-half swaps return each
// value's partner, and the sums are the same bits in every group.
__device__ __forceinline__ float group_allreduce(float v) {
  const auto a = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false,
                                                  false);
  v = __uint_as_float(a[0]) + __uint_as_float(a[1]);
  const auto b = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false,
                                                  false);
  return __uint_as_float(b[0]) + __uint_as_float(b[1]);
}

// Sum of lanes l and l + 32 (groups g and g ^ 2), both receiving it.
__device__ __forceinline__ float half_sum(float v) {
  const auto a = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false,
                                                  false);
  return __uint_as_float(a[0]) + __uint_as_float(a[1]);
}

// Compensated sum: products are added with plain FMAs over blocks of up to
// 32 entries; a bin with more entries adds its blocks through a Kahan
// accumulator (relative error about 32 ulp of a block plus 2 ulp,
// independent of the entry count).
__device__ __forceinline__ void kahan_add(Kahan &k, float x) {
  const float y = x - k.c;
  const float t = k.s + y;
  k.c = (t - k.s) - y;
  k.s = t;
}

// Row sums of the three lineouts: masked (NaN) values add nothing, zeros
// add nothing to the sums (only the count skips them); a and af are 0
// where the in-bounds area is not above 1e-9, and for empty slots.
__device__ __forceinline__ void lineout_add(float lo[4], float o, float a, float af) {
  const float on = isnan(o) ? 0.f : o;
  lo[0] += on;
  lo[1] += on != 0.f ? 1.f : 0.f;
  lo[2] = fmaf(on, a, lo[2]);
  lo[3] = fmaf(on, af, lo[3]);
}

// Two workgroups share a CU (8 waves per SIMD, at most 64 VGPRs and half
// the LDS each): while one waits for its tile's copies, the other reduces.
//
// DARK (uint16 batches only): the first word of a staged entry holds the
// byte offset of its pixel row in the low half and the pixel's integer dark
// in the high half (k_ents_dark); each value is converted and the dark
// subtracted before the product, so the sums are of dark-subtracted counts.
// Direct bins read the dark from dint.
template <typename TL, int G, bool DARK, int FPL, bool VAR>
__global__ __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_batched(
    const int32_t *__restrict__ chunks, const int32_t *__restrict__ xcd,
    const int32_t *__restrict__ tiles, const QuadRec *__restrict__ quads,
    const int2 *__restrict__ qdir, const uint2 *__restrict__ ents,
    const int32_t *__restrict__ tpix, const TL *__restrict__ xt,
    const uint16_t *__restrict__ dint, int n_frames, int plane_off, int tp_off, int tp_len,
    int64_t nb, float *__restrict__ cakes, float *__restrict__ part,
    float *__restrict__ vcakes, const float *__restrict__ d2term) {
  static_assert(!VAR || FPL == 2, "variance sums run in 32-frame passes");
  extern __shared__ __attribute__((aligned(16))) char lds_raw[];
  typedef typename Lanes<TL, FPL>::type QT;
  constexpr int XF = 16 * FPL;                // frames per pass (pixel row)
  constexpr int QB = FPL * (int)sizeof(TL);   // bytes of one lane's frames
  constexpr int RBAND = kWaves * G;
  constexpr int LG = 8;  // cake bins per granule: 32-byte stores
  typedef typename GranVec<LG>::type VT;
  // LDS: the tile's entries at 0, its pixel rows at plane_off, the two
  // pixel lists at tp_off.
  const uint2 *lent = reinterpret_cast<const uint2 *>(lds_raw);
  char *plane = lds_raw + plane_off;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int grp = lane >> 4, l16 = lane & 15;
  // The frame of this lane's values after a quad: 64-frame passes spread
  // the 64 frames over the 64 lanes; 32-frame passes give lanes (g, l)
  // frame 2l + (g & 1), 16-frame passes frame l.
  const int frame = FPL == 4 ? 4 * l16 + grp : FPL == 2 ? 2 * l16 + (grp & 1) : l16;
  // Lanes whose lineout sums are written (the others hold partial sums of
  // the same frames, added in first).
  const bool keeper = FPL == 4 || grp < 4 / FPL;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int x = blockIdx.x & 7;
  const int c = xcd[x] + (int)(blockIdx.x >> 3);
  if (c >= xcd[x + 1]) return;
  const int t_beg = chunks[4 * c], t_end = chunks[4 * c + 1];
  const bool want_lo = part != nullptr;
  float lo[G][4];
#pragma unroll
  for (int g = 0; g < G; ++g) {
#pragma unroll
    for (int k = 0; k < 4; ++k) lo[g][k] = 0.f;
  }
  const bool vec_ok = ((nb & 3) == 0) && (((uintptr_t)cakes & 15) == 0);
  const bool vec2_ok = ((nb & 1) == 0) && (((uintptr_t)cakes & 7) == 0);
  const bool live = frame < n_frames;
  const char *lpl = plane + l16 * QB;  // this lane's frames in a pixel row

  // Pixel lists of tiles t (buffer (t - t_beg) & 1) and t + 1 (the other):
  // a tile's list is in LDS before its staging starts.
  {
    const TileHdr h0 = tile_hdr(tiles, t_beg);
    stage_pixels(tpix, h0.px_beg, h0.npx, wave, lane, lds_raw + tp_off);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  // The next tile's header is read while the current one is staged, and
  // carried to the next iteration.
  TileHdr hn = tile_hdr(tiles, t_beg);
  for (int t = t_beg; t < t_end; ++t) {
    const TileHdr h = hn;
    const int cur = (t - t_beg) & 1;
    if (t > t_beg) __syncthreads();  // every wave is done with the previous tile's LDS
    tile_stage<TL, XF>(h, wave, lane, lds_raw, plane, ents, xt,
                       reinterpret_cast<const int32_t *>(lds_raw + tp_off + cur * tp_len));
    if (t + 1 < t_end) {
      hn = tile_hdr(tiles, t + 1);
      stage_pixels(tpix, hn.px_beg, hn.npx, wave, lane, lds_raw + tp_off + (cur ^ 1) * tp_len);
    }
    __builtin_amdgcn_s_waitcnt(0);  // this wave's copies have landed
    __syncthreads();
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int jr = kWaves * g + wave - h.ra;  // row inside the tile
      if (jr < 0 || jr >= h.nrows) continue;
      const int q_row = h.q_beg + jr * h.nq;
      // 64-frame passes: the granule being filled (this lane's frame, 8
      // bins of the row) and its bins so far.  A chunk's tiles are every
      // k-th tile of its band, so a granule never continues into the
      // chunk's next tile: it is stored when the row's quads of this tile
      // are done.
      VT ob = 0.f;
      int og = -1;
      uint32_t om = 0u;
      // The pair this group reads next and its entry index: a quad whose
      // slots all hold staged entries reads on where the previous full quad
      // ended (the row's entries are one run), others read their first pair.
      u4v tw;
      int ra = -1;
      for (int qi = 0; qi < h.nq; ++qi) {
        const QuadRec &Q = quads[q_row + qi];
        const int flags = Q.flags;
        if ((flags & 15) == 0) continue;  // row padding
        const int nit = Q.n_it;
        // This group's first pair and the step between its pairs; a group
        // whose slot has no staged entries reads the zero pair throughout.
        const uint32_t sm = ((uint32_t)flags >> 8) & 15u;
        const bool mine = (sm >> grp) & 1u;
        const int rank = __builtin_popcount(sm & ((1u << grp) - 1u));
        int ka = mine ? Q.e_off + 2 * rank : h.zp;  // in entries
        const int kstep = mine ? 2 * __builtin_popcount(sm) : 0;
        if (ka != ra) tw = *reinterpret_cast<const u4v *>(lent + ka);
        float tot[FPL];  // this group's bin, this lane's frames
        constexpr int NV = VAR ? FPL : 1;
        float vtot[NV];  // VAR: the same bins' sums of w^2 x
        {
          float gs[FPL], vs[NV];
#pragma unroll
          for (int j = 0; j < FPL; ++j) gs[j] = 0.f;
#pragma unroll
          for (int j = 0; j < NV; ++j) vs[j] = 0.f;
          // Blocks of 16 pairs (32 entries) of plain FMAs; a bin with more
          // adds its blocks through a Kahan accumulator.
          Kahan acc[FPL], vacc[NV];
          int m = 0;
          for (;;) {
            const int m_end = min(nit, m + 16);
            for (; m < m_end; ++m) {
              const u4v two = tw;
              ka += kstep;
              // the next pair (past the run's end this reads LDS words
              // inside the allocation that are not used)
              tw = *reinterpret_cast<const u4v *>(lent + ka);
              const uint32_t oa = DARK ? (two.x & 0xffffu) : two.x;
              const uint32_t obb = DARK ? (two.z & 0xffffu) : two.z;
              const QT da = *reinterpret_cast<const QT *>(lpl + oa);
              const QT db = *reinterpret_cast<const QT *>(lpl + obb);
              float fa[FPL], fb[FPL];
              Lanes<TL, FPL>::unpack(da, fa);
              Lanes<TL, FPL>::unpack(db, fb);
              if constexpr (DARK) {
                const float dka = (float)(two.x >> 16), dkb = (float)(two.z >> 16);
#pragma unroll
                for (int j = 0; j < FPL; ++j) {
                  fa[j] -= dka;
                  fb[j] -= dkb;
                }
              }
              const float wa = __uint_as_float(two.y), wb = __uint_as_float(two.w);
#pragma unroll
              for (int j = 0; j < FPL; ++j) gs[j] = fmaf(fb[j], wb, fmaf(fa[j], wa, gs[j]));
              if constexpr (VAR) {
                // Poisson: the dark-subtracted count is its own variance
                const float wa2 = wa * wa, wb2 = wb * wb;
#pragma unroll
                for (int j = 0; j < NV; ++j) vs[j] = fmaf(fb[j], wb2, fmaf(fa[j], wa2, vs[j]));
              }
            }
            if (m >= nit) break;
#pragma unroll
            for (int j = 0; j < FPL; ++j) {
              kahan_add(acc[j], gs[j]);
              gs[j] = 0.f;
            }
            if constexpr (VAR) {
#pragma unroll
              for (int j = 0; j < NV; ++j) {
                kahan_add(vacc[j], vs[j]);
                vs[j] = 0.f;
              }
            }
          }
          if (nit > 16) {
#pragma unroll
            for (int j = 0; j < FPL; ++j) {
              kahan_add(acc[j], gs[j]);
              tot[j] = acc[j].value();
            }
            if constexpr (VAR) {
#pragma unroll
              for (int j = 0; j < NV; ++j) {
                kahan_add(vacc[j], vs[j]);
                vtot[j] = vacc[j].value();
              }
            }
          } else {
#pragma unroll
            for (int j = 0; j < FPL; ++j) tot[j] = gs[j];
#pragma unroll
            for (int j = 0; j < NV; ++j) vtot[j] = vs[j];
          }
        }
        ra = ka;  // tw holds the pair at ka
        if (flags & 0xf0) {
          // Direct slots: the slot's group gathers its entries from the
          // batch in global memory.
          if ((flags >> (4 + grp)) & 1) {
            const int2 dr = qdir[4 * (q_row + qi) + grp];
            Kahan acc[FPL], vacc[NV];
#pragma unroll
            for (int j = 0; j < FPL; ++j) acc[j].s = tot[j];
#pragma unroll
            for (int j = 0; j < NV; ++j) vacc[j].s = vtot[j];
            for (int e = dr.x; e < dr.y; ++e) {
              const uint2 en = ents[e];
              const QT d = *reinterpret_cast<const QT *>(
                  reinterpret_cast<const char *>(xt + (int64_t)en.x * XF) + l16 * QB);
              float fv[FPL];
              Lanes<TL, FPL>::unpack(d, fv);
              if constexpr (DARK) {
                const float dk = (float)dint[en.x];
#pragma unroll
                for (int j = 0; j < FPL; ++j) fv[j] -= dk;
              }
              const float w = __uint_as_float(en.y);
#pragma unroll
              for (int j = 0; j < FPL; ++j) kahan_add(acc[j], fv[j] * w);
              if constexpr (VAR) {
#pragma unroll
                for (int j = 0; j < NV; ++j) kahan_add(vacc[j], fv[j] * (w * w));
              }
            }
#pragma unroll
            for (int j = 0; j < FPL; ++j) tot[j] = acc[j].value();
#pragma unroll
            for (int j = 0; j < NV; ++j) vtot[j] = vacc[j].value();
          }
        }
        const uint32_t pm = (uint32_t)flags & 15u;
        if constexpr (FPL == 4) {
          // 4 x 4 transpose over the groups: lane (g, l) gets frame 4l + g
          // of slots 0..3.
          const auto p0 = __builtin_amdgcn_permlane32_swap(__float_as_uint(tot[0]),
                                                           __float_as_uint(tot[2]), false, false);
          const auto p1 = __builtin_amdgcn_permlane32_swap(__float_as_uint(tot[1]),
                                                           __float_as_uint(tot[3]), false, false);
          const auto c0 = __builtin_amdgcn_permlane16_swap(p0[0], p1[0], false, false);
          const auto c1 = __builtin_amdgcn_permlane16_swap(p0[1], p1[1], false, false);
          float o[4] = {__uint_as_float(c0[0]), __uint_as_float(c0[1]), __uint_as_float(c1[0]),
                        __uint_as_float(c1[1])};
#pragma unroll
          for (int s = 0; s < 4; ++s) o[s] = (o[s] - Q.dterm[s]) * Q.scale[s];  // NaN: masked
          const int gr = Q.b0 / LG;
          const bool hi = (Q.b0 & 4) != 0;
          if (gr != og) {
            if (om && live) flush_granule<LG>(ob, og, om, frame, cakes, nb, vec_ok);
            og = gr;
            om = 0u;
          }
          om |= hi ? pm << 4 : pm;
          if (pm == 15u) {
            if (hi) {
              ob[4] = o[0];
              ob[5] = o[1];
              ob[6] = o[2];
              ob[7] = o[3];
            } else {
              ob[0] = o[0];
              ob[1] = o[1];
              ob[2] = o[2];
              ob[3] = o[3];
            }
          } else {
#pragma unroll
            for (int s = 0; s < 4; ++s)
              if ((pm >> s) & 1u) {
                if (hi) ob[4 + s] = o[s];
                else ob[s] = o[s];
              }
          }
#pragma unroll
          for (int s = 0; s < 4; ++s) lineout_add(lo[g], o[s], Q.a[s], Q.af[s]);
        } else if constexpr (FPL == 2) {
          // One 16-lane swap: lane (g, l) gets frame 2l + (g & 1) of slots
          // 2 (g >> 1) and 2 (g >> 1) + 1.
          const auto p = __builtin_amdgcn_permlane16_swap(__float_as_uint(tot[0]),
                                                          __float_as_uint(tot[1]), false, false);
          const bool up = grp >= 2;
          const int s0 = up ? 2 : 0;
          const float o0 = (__uint_as_float(p[0]) - (up ? Q.dterm[2] : Q.dterm[0])) *
                           (up ? Q.scale[2] : Q.scale[0]);
          const float o1 = (__uint_as_float(p[1]) - (up ? Q.dterm[3] : Q.dterm[1])) *
                           (up ? Q.scale[3] : Q.scale[1]);
          if (live) {
            const uint32_t pb = (pm >> s0) & 3u;
            float *dst = cakes + (int64_t)frame * nb + Q.b0 + s0;
            if (pb == 3u && vec2_ok) {
              *reinterpret_cast<float2 *>(dst) = make_float2(o0, o1);
            } else {
              if (pb & 1u) dst[0] = o0;
              if (pb & 2u) dst[1] = o1;
            }
          }
          if constexpr (VAR) {
            // The variance: the same swap of the w^2 sums, less the dark
            // remainder's sum w^2 (dark - dint), times the value's scale
            // squared (NaN where the value is).
            const auto pv = __builtin_amdgcn_permlane16_swap(
                __float_as_uint(vtot[0]), __float_as_uint(vtot[1]), false, false);
            const uint32_t pb = (pm >> s0) & 3u;
            float r0 = 0.f, r1 = 0.f;
            if (d2term) {
              if (pb & 1u) r0 = d2term[Q.b0 + s0];
              if (pb & 2u) r1 = d2term[Q.b0 + s0 + 1];
            }
            const float sc0 = up ? Q.scale[2] : Q.scale[0], sc1 = up ? Q.scale[3] : Q.scale[1];
            const float v0 = (__uint_as_float(pv[0]) - r0) * sc0 * sc0;
            const float v1 = (__uint_as_float(pv[1]) - r1) * sc1 * sc1;
            if (live) {
              float *dst = vcakes + (int64_t)frame * nb + Q.b0 + s0;
              if (pb == 3u && vec2_ok) {
                *reinterpret_cast<float2 *>(dst) = make_float2(v0, v1);
              } else {
                if (pb & 1u) dst[0] = v0;
                if (pb & 2u) dst[1] = v1;
              }
            }
          }
          lineout_add(lo[g], o0, up ? Q.a[2] : Q.a[0], up ? Q.af[2] : Q.af[0]);
          lineout_add(lo[g], o1, up ? Q.a[3] : Q.a[1], up ? Q.af[3] : Q.af[1]);
        } else {
          // Lane (g, l) holds frame l of slot g.
          const float sc = grp == 0 ? Q.scale[0] : grp == 1 ? Q.scale[1]
                         : grp == 2 ? Q.scale[2] : Q.scale[3];
          const float dt = grp == 0 ? Q.dterm[0] : grp == 1 ? Q.dterm[1]
                         : grp == 2 ? Q.dterm[2] : Q.dterm[3];
          const float aa = grp == 0 ? Q.a[0] : grp == 1 ? Q.a[1] : grp == 2 ? Q.a[2] : Q.a[3];
          const float af = grp == 0 ? Q.af[0] : grp == 1 ? Q.af[1]
                         : grp == 2 ? Q.af[2] : Q.af[3];
          const float o = (tot[0] - dt) * sc;
          if (live && ((pm >> grp) & 1u)) cakes[(int64_t)frame * nb + Q.b0 + grp] = o;
          lineout_add(lo[g], o, aa, af);
        }
      }
      if constexpr (FPL == 4) {
        if (om && live) flush_granule<LG>(ob, og, om, frame, cakes, nb, vec_ok);
      }
    }
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    if (want_lo) {
      float v[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        v[k] = lo[g][k];
        if constexpr (FPL == 2) v[k] = half_sum(v[k]);
        if constexpr (FPL == 1) v[k] = group_allreduce(v[k]);
      }
      if (keeper) {
        // part[c][row][k][XF] for the band row this slot holds (read once,
        // by k_lineout_finalize: non-temporal).
        float *Pp = part + ((int64_t)c * RBAND + kWaves * g + wave) * 4 * XF;
#pragma unroll
        for (int k = 0; k < 4; ++k) __builtin_nontemporal_store(v[k], Pp + k * XF + frame);
      }
    }
  }
}

// Staged entries of the uint16 layout: the low half of the first word is
// the byte offset of the pixel row in the tile's staged footprint; the high
// half receives the pixel's integer dark (dhi, 0 without a dark).  One
// workgroup per tile; the padding entries point at the zero row (offset
// npx * 128) and get 0.
__global__ __launch_bounds__(256) void k_ents_dark(const int32_t *__restrict__ tiles,
                                                   const int32_t *__restrict__ tpix,
                                                   const uint16_t *__restrict__ dint,
                                                   int row_bytes, uint2 *__restrict__ ents) {
  const int32_t *T = tiles + kTileInts * blockIdx.x;
  const int px_beg = T[0], npx = T[1], e_beg = T[6], e_end = T[7];
  for (int i = e_beg + (int)threadIdx.x; i < e_end; i += 256) {
    const uint32_t off = ents[i].x & 0xffffu;
    const int slot = (int)(off / row_bytes);
    uint32_t d = 0u;
    if (dint && slot < npx) d = dint[tpix[px_beg + slot]];
    ents[i].x = off | (d << 16);
  }
}

// Quad records of the current (mode, flags, dark): the scale, dark term and
// in-bounds area of each slot's bin (0 for empty slots), so the batched
// kernel reads each quad once.  qs: cake bin of slot 0, flags, entry
// offset, pairs.
__global__ void k_quad_recs(const int4 *__restrict__ qs, const float2 *__restrict__ binfo,
                            const float *__restrict__ a_inb, int64_t n,
                            QuadRec *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int4 q = qs[i];
  QuadRec r;
  r.b0 = q.x;
  r.flags = q.y;
  r.e_off = q.z;
  r.n_it = q.w;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    r.scale[s] = r.dterm[s] = r.a[s] = r.af[s] = 0.f;
    if ((q.y >> s) & 1) {
      const float2 bi = binfo[q.x + s];
      const float a = a_inb[q.x + s];
      r.scale[s] = bi.x;
      r.dterm[s] = bi.y;
      r.a[s] = a > 1e-9f ? a : 0.f;
      r.af[s] = a > 1e-9f ? 1.f : 0.f;
    }
  }
  out[i] = r;
}

// Lineouts from the per-chunk partial sums; the chunks of band b are
// [band_chunk[b], band_chunk[b+1]).  A block takes band blockIdx.x and
// 256 / band_rows frames from blockIdx.y on: one thread per (row, frame),
// consecutive threads on consecutive frames (the partials of a (chunk, row)
// are frame-contiguous).  Up to 8 chunks' loads are in flight before the
// first add (the sums keep chunk order); the three values of each thread
// are staged in LDS and stored as runs of the band's rows of one frame's
// lineout.
__global__ __launch_bounds__(256) void k_lineout_finalize(
    const float *__restrict__ part, const int32_t *__restrict__ band_chunk,
    const float4 *__restrict__ row_static, int n_frames, int n_r, int band_rows, int np,
    int mask_nan, float *__restrict__ lineouts) {
  __shared__ float st[3 * 256];  // [frame in block][lineout][row in band]
  const int fb = 256 / band_rows;  // frames per block
  const int band = blockIdx.x, r0 = band * band_rows;
  const int rows = min(band_rows, n_r - r0);
  const int rib = threadIdx.x / fb, fl = threadIdx.x - rib * fb;
  const int f0 = blockIdx.y * fb, f = f0 + fl;
  const int nf = min(fb, n_frames - f0);
  if (rib < rows && fl < nf) {
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    const int c1 = band_chunk[band + 1];
    const int64_t cs = (int64_t)band_rows * 4 * np;  // floats from one chunk's partials to the next
    for (int c = band_chunk[band]; c < c1; c += 8) {
      const float *L = part + ((int64_t)c * band_rows + rib) * 4 * np + f;
      const int n = c1 - c;
      float4 v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < n)
          v[k] = make_float4(__builtin_nontemporal_load(L + k * cs),
                             __builtin_nontemporal_load(L + k * cs + np),
                             __builtin_nontemporal_load(L + k * cs + 2 * np),
                             __builtin_nontemporal_load(L + k * cs + 3 * np));
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < n) {
          s.x += v[k].x;
          s.y += v[k].y;
          s.z += v[k].z;
          s.w += v[k].w;
        }
    }
    write_lineouts(s, row_static[r0 + rib], mask_nan, st + fl * 3 * band_rows + rib, band_rows,
                   0);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * band_rows * nf; i += 256) {
    const int fk = i / band_rows, j = i - fk * band_rows;  // (frame, lineout), row
    if (j < rows) {
      const int g = fk / 3, k = fk - 3 * g;
      lineouts[((int64_t)(f0 + g) * 3 + k) * n_r + r0 + j] = st[i];
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

// The integer part of the dark that the kernels subtract per pixel: the
// nearest integer, clamped to the uint16 range the batched entries carry.
__global__ void k_dark_int(const double *__restrict__ dark, int64_t npx,
                           uint16_t *__restrict__ dint) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npx) return;
  dint[i] = (uint16_t)fmin(fmax(rint(dark[i]), 0.0), 65535.0);
}

// The dark remainder of each bin, sum w * (dark - dhi), in FP64.
__global__ void k_dark_term(const int32_t *__restrict__ row, const int32_t *__restrict__ pix,
                            const float *__restrict__ w, const double *__restrict__ dark,
                            const uint16_t *__restrict__ dint, int64_t nb,
                            float *__restrict__ dterm) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double s = 0;
  if (dark)
    for (int k = row[b]; k < row[b + 1]; ++k) {
      const int p = pix[k];
      s += (dark[p] - (double)dint[p]) * (double)w[k];
    }
  dterm[b] = (float)s;
}

__global__ void k_dark_term2(const int32_t *__restrict__ row, const int32_t *__restrict__ pix,
                             const float *__restrict__ w, const double *__restrict__ dark,
                             const uint16_t *__restrict__ dint, int64_t nb,
                             float *__restrict__ d2term) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double s = 0;
  for (int k = row[b]; k < row[b + 1]; ++k) {
    const int p = pix[k];
    s += (dark[p] - (double)dint[p]) * (double)w[k] * (double)w[k];
  }
  d2term[b] = (float)s;
}

// Per-bin {scale, dark remainder}; the scale of a bin masked to NaN is NaN.
__global__ void k_binfo(const float *__restrict__ inv_area, const int32_t *__restrict__ flags,
                        const float *__restrict__ dterm, int64_t nb, int normalize,
                        int mask_nan, float2 *__restrict__ binfo) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  float s = normalize ? inv_area[b] : 1.f;
  if (mask_nan && flags && flags[b]) s = NAN;
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

int launch_dark_int(hipStream_t s, const double *dark, int64_t npx, uint16_t *dint) {
  hipLaunchKernelGGL(k_dark_int, dim3(blocks_for(npx, 256)), dim3(256), 0, s, dark, npx,
                     dint);
  MC_HIP(hipGetLastError());
  return MC_OK;
}

int launch_dark_term(hipStream_t s, const int32_t *row, const int32_t *pix,
                     const float *w, const double *dark, const uint16_t *dint, int64_t nb,
                     float *dterm) {
  hipLaunchKernelGGL(k_dark_term, dim3(blocks_for(nb, 256)), dim3(256), 0, s, row, pix,
                     w, dark, dint, nb, dterm);
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

template <int LPB>
static int launch_plain_l(hipStream_t s, const Map &m, const Csr &c, const float2 *binfo,
                          int mask_nan, const void *frames, int dtype, float *cake,
                          float *lineouts, int n_frames, int64_t frame_stride, float *sub,
                          int64_t sub_frames) {
  const int64_t npx = m.npx();
  if (m.has_dark) {
    if (!sub || sub_frames < 1) {
      set_error("the one-frame kernel needs scratch for dark-subtracted frames");
      return MC_ERR_STATE;
    }
    const size_t es = dtype_size(dtype);
    for (int f0 = 0; f0 < n_frames; f0 += (int)sub_frames) {
      const int nf = (int)std::min<int64_t>(sub_frames, n_frames - f0);
      const char *fb = (const char *)frames + (size_t)f0 * frame_stride * es;
      int rc = dispatch_dtype(dtype, [&](auto tag) -> int {
        using T = std::remove_const_t<std::remove_pointer_t<decltype(tag)>>;
        hipLaunchKernelGGL(k_frame_dark<T>, dim3(blocks_for((npx + 3) / 4, 256), nf),
                           dim3(256), 0, s, (const T *)fb, npx, m.d_dint.as<uint16_t>(), sub);
        MC_HIP(hipGetLastError());
        return MC_OK;
      });
      if (rc) return rc;
      hipLaunchKernelGGL((k_plain<float, LPB>), dim3(m.n_r, nf), dim3(256), 0, s,
                         c.row.as<int32_t>(), c.pix.as<int32_t>(), c.w.as<float>(), binfo,
                         m.d_area_inb.as<float>(), m.d_row_static.as<float4>(), sub, npx,
                         m.n_eta, m.n_r, m.nb(), mask_nan, cake + (size_t)f0 * m.nb(),
                         lineouts ? lineouts + (size_t)f0 * 3 * m.n_r : nullptr);
      MC_HIP(hipGetLastError());
    }
    return MC_OK;
  }
  return dispatch_dtype(dtype, [&](auto tag) -> int {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(tag)>>;
    hipLaunchKernelGGL((k_plain<T, LPB>), dim3(m.n_r, n_frames), dim3(256), 0, s,
                       c.row.as<int32_t>(), c.pix.as<int32_t>(), c.w.as<float>(), binfo,
                       m.d_area_inb.as<float>(), m.d_row_static.as<float4>(), (const T *)frames,
                       frame_stride, m.n_eta, m.n_r, m.nb(), mask_nan, cake, lineouts);
    MC_HIP(hipGetLastError());
    return MC_OK;
  });
}

// Calls f(std::integral_constant<int, L>) for the map's lanes per bin.
template <typename F> int with_plain_lanes(int lanes, F &&f) {
  switch (lanes) {
    case 4: return f(std::integral_constant<int, 4>());
    case 8: return f(std::integral_constant<int, 8>());
    case 16: return f(std::integral_constant<int, 16>());
    default: set_error("lanes per bin %d not in {4, 8, 16}", lanes); return MC_ERR_ARG;
  }
}

int launch_plain(hipStream_t s, const Map &m, const Csr &c, const float2 *binfo, int mask_nan,
                 const void *frames, int dtype, float *cake, float *lineouts, int n_frames,
                 int64_t frame_stride, float *sub, int64_t sub_frames) {
  return with_plain_lanes(m.plain_lanes, [&](auto lanes) -> int {
    constexpr int LPB = decltype(lanes)::value;
    return launch_plain_l<LPB>(s, m, c, binfo, mask_nan, frames, dtype, cake, lineouts,
                               n_frames, frame_stride, sub, sub_frames);
  });
}

int launch_dark_term2(hipStream_t s, const int32_t *row, const int32_t *pix, const float *w,
                      const double *dark, const uint16_t *dint, int64_t nb, float *d2term) {
  hipLaunchKernelGGL(k_dark_term2, dim3(blocks_for(nb, 256)), dim3(256), 0, s, row, pix, w,
                     dark, dint, nb, d2term);
  MC_HIP(hipGetLastError());
  return MC_OK;
}

int launch_plain_var(hipStream_t s, const Map &m, const Csr &c, const float2 *binfo,
                     const float *d2term, const void *frames, int dtype, const float *vimg,
                     float *cake, float *var, int n_frames, float *sub, int64_t sub_frames) {
  const int64_t npx = m.npx(), nb = m.nb();
  auto run = [&](auto tag, const void *fr, const float *vi, float *ck, float *vr, int nf) -> int {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(tag)>>;
    const dim3 grid(m.n_r, nf);
    return with_plain_lanes(m.plain_lanes, [&](auto lanes) -> int {
      constexpr int LPB = decltype(lanes)::value;
      auto kern = vi ? k_plain_var<T, false, LPB> : k_plain_var<T, true, LPB>;
      hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, c.row.as<int32_t>(), c.pix.as<int32_t>(),
                         c.w.as<float>(), binfo, d2term, (const T *)fr, vi, npx, m.n_eta, nb,
                         ck, vr);
      MC_HIP(hipGetLastError());
      return MC_OK;
    });
  };
  if (!m.has_dark)
    return dispatch_dtype(dtype, [&](auto tag) -> int {
      return run(tag, frames, vimg, cake, var, n_frames);
    });
  if (!sub || sub_frames < 1) {
    set_error("the variance kernel needs scratch for dark-subtracted frames");
    return MC_ERR_STATE;
  }
  const size_t es = dtype_size(dtype);
  for (int f0 = 0; f0 < n_frames; f0 += (int)sub_frames) {
    const int nf = (int)std::min<int64_t>(sub_frames, n_frames - f0);
    const char *fb = (const char *)frames + (size_t)f0 * npx * es;
    int rc = dispatch_dtype(dtype, [&](auto tag) -> int {
      using T = std::remove_const_t<std::remove_pointer_t<decltype(tag)>>;
      hipLaunchKernelGGL(k_frame_dark<T>, dim3(blocks_for((npx + 3) / 4, 256), nf),
                         dim3(256), 0, s, (const T *)fb, npx, m.d_dint.as<uint16_t>(), sub);
      MC_HIP(hipGetLastError());
      return MC_OK;
    });
    if (rc) return rc;
    if ((rc = run((const float *)nullptr, sub, vimg ? vimg + (size_t)f0 * npx : nullptr,
                  cake + (size_t)f0 * nb, var + (size_t)f0 * nb, nf)))
      return rc;
  }
  return MC_OK;
}

int launch_profile(hipStream_t s, const Map &m, const float *cakes, int n_frames,
                   int mask_nan, float *lineouts) {
  hipLaunchKernelGGL(k_profile, dim3(m.n_r, n_frames), dim3(256), 0, s, cakes,
                     m.d_area_inb.as<float>(), m.d_row_static.as<float4>(), m.n_eta, m.n_r,
                     m.nb(), mask_nan, lineouts);
  MC_HIP(hipGetLastError());
  return MC_OK;
}

int launch_profile_var(hipStream_t s, const Map &m, const float *cakes, const float *vars,
                       int n_frames, int simple, float *prof, float *pvar) {
  hipLaunchKernelGGL(k_profile_var, dim3(m.n_r, n_frames), dim3(256), 0, s, cakes, vars,
                     m.d_area_inb.as<float>(), m.n_eta, m.n_r, m.nb(), simple, prof, pvar);
  MC_HIP(hipGetLastError());
  return MC_OK;
}

int launch_hybrid_decode(hipStream_t s, const void *msg, int64_t npx, int64_t count,
                         int32_t *out) {
  const uint8_t *b = static_cast<const uint8_t *>(msg) + 4;  // after the u32 count
  hipLaunchKernelGGL(k_hybrid_base, dim3(blocks_for(npx, 256)), dim3(256), 0, s,
                     reinterpret_cast<const uint16_t *>(b), npx, out);
  MC_HIP(hipGetLastError());
  if (count > 0) {
    const uint8_t *idx = b + 2 * npx, *val = idx + 4 * count;
    hipLaunchKernelGGL(k_hybrid_patch, dim3(blocks_for(count, 256)), dim3(256), 0, s, idx, val,
                       count, npx, out);
    MC_HIP(hipGetLastError());
  }
  return MC_OK;
}

int launch_quad_recs(hipStream_t s, const Map &m, Tiles &t, const float2 *binfo) {
  if (!t.qrecs.p) {
    int rc = t.qrecs.alloc((size_t)std::max<int64_t>(t.n_quads, 1) * sizeof(QuadRec));
    if (rc) return rc;
  }
  if (t.n_quads == 0) return MC_OK;
  hipLaunchKernelGGL(k_quad_recs, dim3(blocks_for(t.n_quads, 256)), dim3(256), 0, s,
                     t.qs.as<int4>(), binfo, m.d_area_inb.as<float>(), t.n_quads,
                     t.qrecs.as<QuadRec>());
  MC_HIP(hipGetLastError());
  return MC_OK;
}

int launch_ents_dark(hipStream_t s, const Map &m, Tiles &t) {
  if (t.in_size != 2 || t.n_tiles == 0) return MC_OK;
  hipLaunchKernelGGL(k_ents_dark, dim3((unsigned)t.n_tiles), dim3(256), 0, s,
                     t.tiles.as<int32_t>(), t.tpix.as<int32_t>(),
                     m.has_dark ? m.d_dint.as<uint16_t>() : nullptr, t.width * 2,
                     t.ents.as<uint2>());
  MC_HIP(hipGetLastError());
  return MC_OK;
}

size_t batch_xt_bytes(const Map &m, int dtype) {
  return (size_t)m.npx() * kXtFrames * (dtype_size(dtype) <= 2 ? 2 : 4);
}

template <typename TIn, typename TL, int G, int FPL>
static int launch_batched_g(hipStream_t s, const Map &m, const Tiles &t, int mask_nan,
                            const void *frames, int n_frames, float *cakes, float *lineouts,
                            void *xt, float *part, float *vcakes, const float *d2term) {
  constexpr int XF = 16 * FPL;
  const uint16_t *dhi = m.has_dark ? m.d_dint.as<uint16_t>() : nullptr;
  // 1. pixel-major batch
  const int64_t npx = m.npx();
  const int vec_ok = (npx % 64 == 0) && (((uintptr_t)frames & 15) == 0) ? 1 : 0;
  // Only the blocks holding pixels the map reads are rewritten.
  if (sizeof(TIn) == 2 && sizeof(TL) == 2 && vec_ok && npx % 128 == 0) {
    if (t.n_blocks128)
      hipLaunchKernelGGL(k_transpose16<XF>, dim3((unsigned)t.n_blocks128), dim3(256), 0, s,
                         (const uint16_t *)frames, npx, n_frames, npx,
                         t.blocks128.as<int32_t>(), (uint16_t *)xt);
  } else if (t.n_blocks64) {
    hipLaunchKernelGGL((k_transpose<TIn, TL, XF>), dim3((unsigned)t.n_blocks64), dim3(256), 0, s,
                       (const TIn *)frames, npx, n_frames, npx, vec_ok, t.blocks64.as<int32_t>(),
                       dhi, (TL *)xt);
  }
  MC_HIP(hipGetLastError());
  // 2. tiles (uint16 batches subtract the dark carried by the entries;
  // FP32 batches were dark-subtracted by the transpose)
  constexpr bool kU16 = sizeof(TL) == 2;
  auto kern = (kU16 && dhi) ? k_batched<TL, G, kU16, FPL, false>
                            : k_batched<TL, G, false, FPL, false>;
  if constexpr (FPL == 2 && G == 1) {
    if (vcakes)
      kern = (kU16 && dhi) ? k_batched<TL, G, kU16, FPL, true> : k_batched<TL, G, false, FPL, true>;
  }
  if (vcakes && !(FPL == 2 && G == 1)) {
    set_error("variance sums need 32-frame passes of 16-row bands");
    return MC_ERR_STATE;
  }
  const size_t lds = t.lds_needed;
  if (lds > 65536)
    MC_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds));
  if (!lineouts) part = nullptr;
  hipLaunchKernelGGL(kern, dim3((unsigned)t.grid), dim3(64 * kWaves), lds, s,
                     t.chunks.as<int32_t>(), t.xcd.as<int32_t>(), t.tiles.as<int32_t>(),
                     t.qrecs.as<QuadRec>(), t.qdir.as<int2>(), t.ents.as<uint2>(),
                     t.tpix.as<int32_t>(), (const TL *)xt, dhi, n_frames, t.plane_off, t.tp_off,
                     t.tp_len, m.nb(), cakes, part, vcakes, d2term);
  MC_HIP(hipGetLastError());
  // 3. lineouts
  if (lineouts) {
    const int fb = 256 / t.rows;
    hipLaunchKernelGGL(k_lineout_finalize, dim3((unsigned)t.n_bands, (n_frames + fb - 1) / fb),
                       dim3(256), 0, s, part,
                       t.band_chunk.as<int32_t>(), m.d_row_static.as<float4>(), n_frames,
                       m.n_r, t.rows, XF, mask_nan, lineouts);
    MC_HIP(hipGetLastError());
  }
  return MC_OK;
}

int launch_batched(hipStream_t s, const Map &m, const Tiles &t, const float2 *binfo,
                   int mask_nan, const void *frames, int dtype, int n_frames, float *cakes,
                   float *lineouts, void *xt, float *part, float *vcakes,
                   const float *d2term) {
  (void)binfo;  // folded into the quad records (launch_quad_recs)
  if (n_frames < 1 || n_frames > t.width) {
    set_error("batched pass of %d frames outside 1..%d", n_frames, t.width);
    return MC_ERR_ARG;
  }
  if ((dtype_size(dtype) <= 2 ? 2 : 4) != t.in_size) {
    set_error("tile layout built for %d-byte batch elements", t.in_size);
    return MC_ERR_STATE;
  }
  if (lineouts && !part) {
    set_error("batched lineouts need a partials buffer");
    return MC_ERR_ARG;
  }
  return dispatch_dtype(dtype, [&](auto tag) -> int {
    using T = std::remove_const_t<std::remove_pointer_t<decltype(tag)>>;
    using TL = typename std::conditional<sizeof(T) <= 2, uint16_t, float>::type;
    auto go = [&](auto fpl) -> int {
      constexpr int FPL = decltype(fpl)::value;
      switch (t.rows / kWaves) {
        case 1:
          return launch_batched_g<T, TL, 1, FPL>(s, m, t, mask_nan, frames, n_frames, cakes,
                                                 lineouts, xt, part, vcakes, d2term);
        case 2:
          return launch_batched_g<T, TL, 2, FPL>(s, m, t, mask_nan, frames, n_frames, cakes,
                                                 lineouts, xt, part, vcakes, d2term);
        default:
          set_error("rows per band %d not in {16, 32}", t.rows);
          return MC_ERR_ARG;
      }
    };
    switch (t.width) {
      case 16: return go(std::integral_constant<int, 1>());
      case 32: return go(std::integral_constant<int, 2>());
      case 64: return go(std::integral_constant<int, 4>());
      default: set_error("pass width %d not in {16, 32, 64}", t.width); return MC_ERR_ARG;
    }
  });
}

}  // namespace mc
