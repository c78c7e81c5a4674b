// mc_internal.h -- shared types of the MI355X caking library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/midas_caking.h"

namespace mc {

void set_error(const char *fmt, ...);

#define MC_HIP(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      ::mc::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,            \
                      hipGetErrorString(e_));                               \
      return MC_ERR_HIP;                                                    \
    }                                                                       \
  } while (0)

#define MC_CHECK(cond, code, ...)                                           \
  do {                                                                      \
    if (!(cond)) {                                                          \
      ::mc::set_error(__VA_ARGS__);                                         \
      return (code);                                                        \
    }                                                                       \
  } while (0)

// One Map.bin record, MapperCore.h:19-25 (24 bytes, 4-byte packed so the
// device copy has no padding either).
#pragma pack(push, 4)
struct Rec {
  float y, z;
  double frac;
  float deltaR, areaWeight;
};
#pragma pack(pop)
static_assert(sizeof(Rec) == 24, "Map.bin record must be 24 bytes");

// Device buffer with RAII free.
struct DBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  int alloc(size_t n) {
    reset();
    if (n == 0) n = 16;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      set_error("hipMalloc(%zu): %s", n, hipGetErrorString(e));
      p = nullptr;
      return MC_ERR_NOMEM;
    }
    bytes = n;
    return MC_OK;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::recursive_mutex mu;  // one submitter on the stream at a time
  int64_t map_verify = 4096;  // pixels of each GPU map build recomputed on the host
};

// Bin-major CSR of one integration mode: the one-frame kernel's layout.
struct Csr {
  bool built = false;
  double bc_y = 0, bc_z = 0;  // beam centre it was built for (gradient)
  int64_t nnz = 0;
  DBuf row;   // int32 [nb+1]
  DBuf pix;   // int32 [nnz]
  DBuf w;     // float [nnz]
  std::vector<int32_t> h_row, h_pix;  // host copies for the tile builder
  std::vector<float> h_w;
  DBuf dterm;                         // float [nb], dark remainder sum w (dark - dint)
  bool dterm_valid = false;
  DBuf d2term;                        // float [nb], sum w^2 (dark - dint) (variance)
  bool d2_valid = false;
};

// Bytes of a quad record of the batched kernel (4 ints, 16 floats).
constexpr int kQuadRecBytes = 80;

// Row-block/tile layout of the batched kernel for one (mode, batch, rows,
// LDS budget).
struct Tiles {
  bool built = false;
  uint64_t id = 0;   // unique per layout built (captured graphs check it)
  int batch = 0, rows = 0, lds_bytes = 0;  // tuned batch, band rows (16 * G), LDS budget
  int width = 64;    // frames per pass: 16, 32 or 64 (pixel row of the pixel-major batch)
  int bin_cap = 0, ent_cap = 0;  // bins and staged entries a tile holds (entries in LDS)
  int in_size = 0;   // bytes per element of the pixel-major batch (2 or 4)
  int n_chunks = 0, n_bands = 0;
  size_t lds_needed = 0;
  int plane_off = 0;  // LDS bytes: the staged pixel rows (after the entries)
  int tp_off = 0, tp_len = 0;  // LDS bytes: the two pixel-list buffers (offset, length of one)
  DBuf chunks;       // int32 x4 per chunk: tile_begin tile_end band_row0 band_rows
  DBuf tiles;        // int32 x16 per tile (see mc_kernels.hip)
  DBuf qs;           // int32 x4 per quad: cake bin of slot 0, flags, entry offset, pairs
  DBuf qrecs;        // QuadRec per quad: qs + scale, dark term, areas of one binfo table
  DBuf qdir;         // int32 x2 per quad slot: entry range of a direct slot
  const void *recs_binfo = nullptr;
  uint64_t recs_version = 0;
  DBuf ents;         // uint32 x2: byte offset of the staged pixel row (or global pixel), weight bits
  DBuf band_chunk;   // int32 [n_bands+1]: chunk range of each band
  DBuf tpix;         // int32: staged pixels of every tile (tile t at px_beg .. + npx)
  DBuf blocks64, blocks128;  // int32: 64- / 128-pixel blocks holding pixels of the map
  int n_blocks64 = 0, n_blocks128 = 0;
  DBuf xcd;          // int32 [9]: chunk range of each XCD (equal cost)
  int grid = 0;      // workgroups of the launch: 8 x the largest XCD range
  uint64_t ents_dark = 0;  // dark version whose pixel darks the staged entries carry
  int64_t n_tiles = 0, n_quads = 0, n_ents = 0, n_direct = 0, staged_px = 0;
  // Bytes of lineout partials a pass over this layout writes:
  // float [n_chunks][rows][4][64].
  size_t part_bytes() const { return (size_t)n_chunks * rows * 4 * width * sizeof(float); }
};

// Scratch of the batched passes of one caller: the pixel-major batch and
// the lineout partials.  Each stream that runs passes owns one (the map's
// context, each server slot, each captured graph): two passes on different
// streams never share a buffer that a pass writes.
struct Work {
  DBuf xt;
  DBuf part;
};

struct Map {
  Ctx *ctx = nullptr;
  std::mutex mu;
  // A fused map (mc_map_fuse) holds one combined matrix as its bilinear and
  // gradient CSR, unit areas and no Map.bin records.
  bool fused = false;
  int ny = 0, nz = 0, n_r = 0, n_eta = 0;
  int64_t nnz = 0;
  int q_mode = 0;
  int64_t skipped[3] = {0, 0, 0};
  // host verification of the GPU build: pixels checked, records compared,
  // pixels whose bin membership differs, records whose values differ
  int64_t verify[4] = {0, 0, 0, 0};
  std::vector<Rec> h_rec;           // canonical records (bin-major)
  std::vector<int32_t> h_pairs;     // nMap pairs
  std::vector<int32_t> h_flags;     // maskMap flags
  std::vector<double> h_area_all;   // sum areaWeight per bin (all entries)
  std::vector<double> h_area_inb;   // in-bounds entries only (.cu:817-899)
  DBuf d_rec, d_pairs;
  DBuf d_flags;                     // int32 [nb]
  DBuf d_area_inb;                  // float [nb]
  DBuf d_row_static;                // float4 [n_r]: sum A_inb, n over valid bins;
                                    // same over valid unflagged bins
  DBuf inv_area;                    // float [nb]: 1 / sum areaWeight (0 if empty)
  double bc_y = 0, bc_z = 0;
  // dark
  bool has_dark = false;
  uint64_t dark_version = 1;        // bumped by every mc_map_set_dark
  DBuf d_dark;                      // double [npx]: mean dark x scale
  DBuf d_dint;                      // uint16 [npx]: the dark rounded to an integer in
                                    // [0, 65535], subtracted from each pixel exactly;
                                    // the remainder is folded per bin (dark term)
  // per mode layouts
  Csr csr[3];
  std::vector<std::unique_ptr<Tiles>> tiles[3];
  // per-bin float2 {signed scale, dark term}, one table per (mode, flags):
  // index mode * 4 + flags
  DBuf d_binfo[12];
  bool binfo_ok[12] = {};
  uint64_t binfo_version[12] = {};  // bumped whenever a table is rebuilt
  // tuning
  int batch = 64, rows_per_group = 16, lds_bytes = 81920;
  int tile_bins = 256, tile_ents = 2048;  // LDS slots of a tile: bin records, entries
  int plain_lanes = 8;                    // lanes per bin of the one-frame kernels
  Work work;  // scratch of the calls made through the map's context stream
  int64_t nb() const { return (int64_t)n_r * n_eta; }
  int64_t npx() const { return (int64_t)ny * nz; }
};

size_t dtype_size(int dtype);

// ---- kernel launchers (mc_kernels.hip) ----
int launch_area_fill(hipStream_t s, float *p, int64_t n, float v);
int launch_dark_mean(hipStream_t s, const void *frames, int dtype,
                     int n_frames, int64_t npx, double scale, double *out);
// dint[p] = the mean dark rounded to an integer and clamped to [0, 65535].
int launch_dark_int(hipStream_t s, const double *dark, int64_t npx, uint16_t *dint);
// dterm[b] = sum_k w_k * (dark - dint)[pix_k] in FP64 (0 without a dark).
int launch_dark_term(hipStream_t s, const int32_t *row, const int32_t *pix,
                     const float *w, const double *dark, const uint16_t *dint, int64_t nb,
                     float *dterm);
int launch_binfo(hipStream_t s, const float *inv_area, const int32_t *flags,
                 const float *dterm, int64_t nb, int normalize, int mask_nan,
                 float2 *binfo);
// d2term[b] = sum_k w_k^2 * (dark - dint)[pix_k] (Poisson variance).
int launch_dark_term2(hipStream_t s, const int32_t *row, const int32_t *pix, const float *w,
                      const double *dark, const uint16_t *dint, int64_t nb, float *d2term);
// One-frame kernel with variance propagation; vimg NULL = Poisson (the
// frame itself, minus the dark's d2term when that is given).
int launch_plain_var(hipStream_t s, const Map &m, const Csr &c, const float2 *binfo,
                     const float *d2term, const void *frames, int dtype, const float *vimg,
                     float *cake, float *var, int n_frames, float *sub, int64_t sub_frames);
// One-frame kernel.  With a dark, each frame is first written to sub as
// FP32 minus the pixel's integer dark (sub holds sub_frames frames; the
// frames go through in chunks of that many).
int launch_plain(hipStream_t s, const Map &m, const Csr &c, const float2 *binfo,
                 int mask_nan, const void *frame, int dtype, float *cake,
                 float *lineouts, int n_frames, int64_t frame_stride, float *sub,
                 int64_t sub_frames);
// Tile records of t for one binfo table (scale/dark/area folded in).
int launch_quad_recs(hipStream_t s, const Map &m, Tiles &t, const float2 *binfo);
// Writes each staged entry's pixel dark (dint, 0 without a dark) into the
// high half of its first word (2-byte batch elements).
int launch_ents_dark(hipStream_t s, const Map &m, Tiles &t);
// One batched pass of up to 64 frames; xt is scratch of batch_xt_bytes(),
// part of t.part_bytes() (needed when lineouts is not NULL).  With vcakes
// (32-frame layouts of 16-row bands) the pass also writes the Poisson
// variance of each bin, sum w^2 (x - dint) less d2term (sum w^2 (dark -
// dint), NULL for 0), times the value's scale squared.
int launch_batched(hipStream_t s, const Map &m, const Tiles &t, const float2 *binfo,
                   int mask_nan, const void *frames, int dtype, int n_frames, float *cakes,
                   float *lineouts, void *xt, float *part, float *vcakes = nullptr,
                   const float *d2term = nullptr);
size_t batch_xt_bytes(const Map &m, int dtype);
int launch_profile(hipStream_t s, const Map &m, const float *cakes, int n_frames,
                   int mask_nan, float *lineouts);
// Lineouts and their variances of cakes with propagated variance
// ([n_frames][n_r] each); simple: simple_mean instead of area_weighted.
int launch_profile_var(hipStream_t s, const Map &m, const float *cakes, const float *vars,
                       int n_frames, int simple, float *prof, float *pvar);
// One hybrid message (u32 count, u16 base[npx], i32 index[count], i64
// value[count]) in device memory at msg, decoded to int32 out[npx].
int launch_hybrid_decode(hipStream_t s, const void *msg, int64_t npx, int64_t count,
                         int32_t *out);

// ---- mapper (mc_mapper.hip) ----
int map_build_gpu(Ctx *ctx, const mc_geometry *g, const double *mask,
                  const double *distY, const double *distZ,
                  const double *residual, const mc_panel *panels,
                  int n_panels, Map *m);

// ---- host helpers (mc_core.cpp) ----
int map_finish_host(Map *m);  // derived arrays + upload
int ensure_csr(Map *m, int mode);
// Pass width of a batch size: the narrowest of 16, 32, 64 frames holding it.
inline int width_of(int batch) { return batch <= 16 ? 16 : batch <= 32 ? 32 : 64; }
// Tile layout of (mode, element size, pass width; 0: the width of the
// map's tuned batch) under the map's tuning.
int ensure_tiles(Map *m, int mode, int esize, Tiles **out, int width = 0);
// Launches the integration of n_frames device frames on stream s (no sync).
// work: the caller's scratch for batched passes on s (NULL: the map's own,
// for the map's context stream only).
int run_integrate(Map *m, hipStream_t s, const void *frames_dev, int dtype, int n_frames,
                  int mode, uint32_t flags, float *cakes_dev, float *lineouts_dev,
                  bool force_plain, bool force_batched, Work *work = nullptr,
                  int width = 0);
// Sizes work for batched passes of (mode, dtype) over m's current layout.
int ensure_work(Map *m, int mode, int dtype, Work *work, int width = 0);
// Everything the batched kernel reads for (mode, flags, dtype), built ahead.
int prepare_batched(Map *m, int mode, uint32_t flags, int dtype, int width = 0);
// The per-bin table of (mode, flags); builds it (and the dark term) once.
int ensure_binfo(Map *m, int mode, uint32_t flags, const float2 **out);

}  // namespace mc

struct mc_ctx : mc::Ctx {};
struct mc_map : mc::Map {};
