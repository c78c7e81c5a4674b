/*
 * midas_caking.h -- C ABI of the MI355X caking library (libmidas_caking.so).
 *
 * Plain C types only: pointers, sizes, status codes.  Every function returns
 * an int status (MC_OK = 0, negative on error) and never exits the process;
 * mc_last_error() returns a thread-local message for the last failure.
 *
 * Ownership: the caller owns host arrays; the library owns device memory and
 * handles, which are freed explicitly (mc_*_destroy).  A context serialises
 * its GPU work on one HIP stream; map handles are immutable after they are
 * built, except for the dark frame and the cached kernel layouts, which are
 * guarded by the map's own lock.
 *
 * Reference interfaces replaced (paths relative to the MIDAS repository):
 *   mc_map_build        mapper_build_map, FF_HEDM/src/MapperCore.h:52-84, and
 *                       the DetectorMapper executable, DetectorMapper.c:53-888
 *   mc_map_write        Map.bin/nMap.bin/maskMap.bin writer,
 *                       DetectorMapper.c:813-881 + MapHeader.h:200-240
 *   mc_map_load         ReadBins, IntegratorZarrOMP.c:97-191 (and the GPU
 *                       stream's mmap, IntegratorFitPeaksGPUStream.cu:742-797)
 *   mc_map_from_host    load_map/build_csr, packages/midas_integrate/
 *                       midas_integrate/bin_io.py:206, kernels.py:242-375
 *   mc_map_set_dark     dark averaging, IntegratorZarrOMP.c:1416-1454 and
 *                       IntegratorFitPeaksGPUStream.cu:2192-2238
 *   mc_integrate        per-bin loop IntegratorZarrOMP.c:1710-1778 (==
 *                       integration_apply_map, IntegrationCore.h:28-36),
 *                       integrate_noMapMask .cu:975-1050, the 1-D lineouts
 *                       IntegratorZarrOMP.c:1829-1863 / .cu:1162-1334, and
 *                       kernels.py integrate/profile_1d (:416-601)
 *   mc_fit_windows      per-eta peak fits, IntegratorZarrOMP.c:1903-2070
 *                       (pf_fit_single_peak PeakFit.h:158-160, fitPeaks
 *                       PeakFit.h:114-116)
 *   mc_server_*         the frame server of IntegratorFitPeaksGPUStream.cu
 *                       :567-734, 2567-2883 and server.py FrameServer
 *   mc_out_*            the result files: <data>.caked.hdf
 *                       (IntegratorZarrOMP.c:1671-1687, 2084-2168),
 *                       <prefix>_lineout.bin/.xy (:1514-1527, 1829-1863),
 *                       <prefix>_fit.bin/_fit_per_eta.csv (:1530-1549,
 *                       1879-1883, 2038-2058), <prefix>_caked_peaks.h5
 *                       (PeakFitIO.c:52-267, IntegratorZarrOMP.c:2176-2195),
 *                       and the GPU stream's lineout.bin,
 *                       lineout_simple_mean.bin, Int2D.bin and its writer
 *                       thread (IntegratorFitPeaksGPUStream.cu:506-565);
 *                       the caked file's Zarr zip copy of
 *                       FF_HEDM/workflows/integrator.py:379-426
 *   mc_blosc_*          blosc1_decompress of the Zarr-zip chunks,
 *                       IntegratorZarrOMP.c:1416-1454, 1646-1669 and
 *                       ZarrReader.c:7-62
 *   mc_map_fuse         build_fused_geometry, packages/midas_integrate/
 *                       midas_integrate/fused_csr.py:126-260
 *   mc_smooth_eta       gauss_smooth_eta, packages/midas_integrate/
 *   mc_median_eta       midas_integrate/postprocess.py:36-142
 */
#ifndef MIDAS_CAKING_H
#define MIDAS_CAKING_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define MC_OK 0
#define MC_ERR_ARG (-1)
#define MC_ERR_HIP (-2)
#define MC_ERR_IO (-3)
#define MC_ERR_FORMAT (-4)
#define MC_ERR_BUDGET (-5)
#define MC_ERR_STATE (-6)
#define MC_ERR_NOMEM (-7)
#define MC_ERR_TIMEOUT (-8)

/* ---- frame dtype codes: the wire codes of the reference frame server
 * (IntegratorFitPeaksGPUStream.cu:568-585) plus an int32 extension ---- */
#define MC_DTYPE_U8 0
#define MC_DTYPE_U16 1
#define MC_DTYPE_U32 2
#define MC_DTYPE_I64 3
#define MC_DTYPE_F32 4
#define MC_DTYPE_F64 5
#define MC_DTYPE_HYBRID 6 /* u16 base + (i32 index, i64 value) patches */
#define MC_DTYPE_I32 7

/* ---- integration modes (kernels.py:1-30 vocabulary) ---- */
#define MC_MODE_FLOOR 0    /* integer pixel (float offset, .cu:916) */
#define MC_MODE_BILINEAR 1 /* bilinear at stored y,z (IntegratorZarrOMP) */
#define MC_MODE_GRADIENT 2 /* bilinear at y,z - deltaR*(dy,dz)/R */

/* ---- integration flags ---- */
#define MC_F_NORMALIZE 1u /* divide by sum areaWeight (Normalize=1) */
#define MC_F_MASK_NAN 2u  /* NaN for bins flagged in maskMap.bin */

/* ---- lineout rows written per frame when requested ---- */
#define MC_LINEOUT_CPU_MEAN 0 /* mean of values != 0 and not NaN */
#define MC_LINEOUT_AREA 1     /* sum(I*A)/sum(A) over A > 1e-9 */
#define MC_LINEOUT_MEAN 2     /* mean of I over A > 1e-9 */
#define MC_N_LINEOUTS 3

/* ---- library / context ---- */
int mc_version(int *major, int *minor, int *patch);
const char *mc_last_error(void);
int mc_device_count(int *n);

typedef struct mc_ctx mc_ctx;
int mc_ctx_create(int device, mc_ctx **out);
int mc_ctx_destroy(mc_ctx *ctx);
int mc_ctx_sync(mc_ctx *ctx);
/* The context's HIP stream, as an opaque pointer (hipStream_t). */
void *mc_ctx_stream(mc_ctx *ctx);
/* Device buffers owned by the library (for callers without an allocator). */
int mc_dev_alloc(mc_ctx *ctx, uint64_t bytes, void **dptr);
int mc_dev_free(mc_ctx *ctx, void *dptr);
int mc_memcpy_h2d(mc_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int mc_memcpy_d2h(mc_ctx *ctx, void *dst, const void *src, uint64_t bytes);

/* ---- geometry: the DetectorMapper parameter set (DetectorMapper.c:84-114) */
typedef struct mc_geometry {
  int32_t NrPixelsY, NrPixelsZ;
  double pxY, pxZ, yCen, zCen, Lsd, RhoD, tx, ty, tz;
  double p[15];
  double RMin, RMax, RBinSize, EtaMin, EtaMax, EtaBinSize;
  double QMin, QMax, QBinSize, Wavelength;
  int32_t NrTransOpt;
  int32_t TransOpt[10];
  int32_t SubPixelLevel;
  double SubPixelCardinalWidth;
  double Parallax;
  int32_t SolidAngleCorrection, PolarizationCorrection;
  double PolarizationFraction;
} mc_geometry;

/* Panel.h:9-22. */
typedef struct mc_panel {
  int32_t yMin, yMax, zMin, zMax;
  double dY, dZ, dTheta, dLsd, dP2, centerY, centerZ;
} mc_panel;

typedef struct mc_map mc_map;

typedef struct mc_map_info {
  int64_t nnz;          /* Map.bin records */
  int32_t n_r, n_eta;   /* bins */
  int32_t ny, nz;       /* detector pixels */
  int32_t n_flagged;    /* bins flagged by masked pixels */
  int32_t q_mode;
  int64_t hbm_bytes;    /* device memory held by this map */
  int64_t skipped[3];   /* mapper rejections: <3 vertices, <3 unique, area */
} mc_map_info;

/* GPU FP64 DetectorMapper.  mask/distY/distZ are NrPixelsZ x NrPixelsY
 * doubles in raw detector orientation (ImTransOpt is applied inside, as in
 * DetectorMapper.c:640-676); residual is not transformed.  Any of them may
 * be NULL.  Records come out bin-major, within a bin in the reference's
 * single-thread loop order. */
int mc_map_build(mc_ctx *ctx, const mc_geometry *g, const double *mask,
                 const double *distY, const double *distZ,
                 const double *residual, const mc_panel *panels,
                 int32_t n_panels, mc_map **out);
/* Map.bin / nMap.bin (+ maskMap.bin when any bin is flagged) into dir,
 * each with the given 64-byte header (NULL: no header, legacy layout). */
int mc_map_write(mc_map *m, const char *dir, const uint8_t *header64);
/* Reads Map.bin/nMap.bin (header auto-detected) and optional maskMap.bin. */
int mc_map_load(mc_ctx *ctx, const char *map_path, const char *nmap_path,
                const char *maskmap_path, int32_t ny, int32_t nz, int32_t n_r,
                int32_t n_eta, mc_map **out);
/* From host arrays: records are 24-byte MapPixelData, pairs are nMap
 * (count, offset) int32 pairs, flags may be NULL. */
int mc_map_from_host(mc_ctx *ctx, const void *records, int64_t nnz,
                     const int32_t *pairs, const int32_t *flags, int32_t ny,
                     int32_t nz, int32_t n_r, int32_t n_eta, mc_map **out);
int mc_map_info_get(mc_map *m, mc_map_info *info);
/* Host verification of GPU map builds (the GPU mapper's membership is
 * pinned to the host's libm, MapperCore.c:45-424 run on the CPU): each
 * later mc_map_build on this context recomputes n_pixels pixels, spread over
 * the detector, with the same per-pixel code built for the host, and
 * compares them with the GPU's records (default 4096; 0 turns it off;
 * n_pixels >= NY*NZ checks every pixel).  mc_map_verify_info: out[4] =
 * pixels checked, records compared, pixels whose bin membership differs,
 * records whose values differ beyond 1e-6 px^2 (areas) or 1e-5 px
 * (deltaR).  Differences are reported, not refused: where a vertex lies
 * exactly on a bin boundary the two libms' last bits can flip a case.
 * Maps not built by mc_map_build report zeros. */
int mc_ctx_set_map_verify(mc_ctx *ctx, int64_t n_pixels);
int mc_map_verify_info(mc_map *m, int64_t *out);
/* Copies the map back: records (nnz*24 B), pairs (2*n_bins), flags. */
int mc_map_download(mc_map *m, void *records, int32_t *pairs,
                    int32_t *flags);
/* Multi-map averaging fused into one map (replaces build_fused_geometry,
 * packages/midas_integrate/midas_integrate/fused_csr.py:126-260): maps[k]
 * are built for the same detector and eta bins with R edges shifted by
 * k * dR / n_maps (their R bin counts may differ by the shift); r_axes
 * holds each map's R bin centres, map after map (n_r of map 0, then of
 * map 1, ...).  The fused map has map 0's bins.
 * The result holds M = (1/N) sum_k P_k D_k M_k, with M_k the map's
 * bilinear or gradient matrix (mode), D_k = diag(1/A) over its in-bounds
 * bin areas A > 1e-9 (0 elsewhere) and P_k the linear interpolation from
 * its R axis onto map 0's, clamped at the ends.  FP64 products and sums in
 * the reference's order, stored as FP32 weights.  The fused map integrates
 * in bilinear or gradient mode (both use M) with unit areas, so normalise
 * is a no-op, as in the reference; it has no Map.bin records (mc_map_write
 * and mc_map_download refuse it).  Every map needs n_r >= 2. */
int mc_map_fuse(mc_map *const *maps, int32_t n_maps, const double *r_axes,
                int32_t mode, mc_map **out);

/* Static per-bin axes (REtaMap, IntegratorZarrOMP.c:1754-1763): 5 rows
 * R, 2theta, eta, sum areaWeight, Q; needs the bin edges. */
int mc_map_reta(mc_map *m, const double *r_lo, const double *r_hi,
                const double *eta_lo, const double *eta_hi, double px,
                double lsd, double lambda, double *out5);
/* The same rows as the GPU stream computes them for RTthEtaAreaMap.bin
 * (initialize_PerFrameArr_Area_kernel, IntegratorFitPeaksGPUStream.cu
 * :817-899, written at :2443-2462): the area row sums areaWeight over the
 * entries whose pixel lies on the detector only. */
int mc_map_reta_stream(mc_map *m, const double *r_lo, const double *r_hi,
                       const double *eta_lo, const double *eta_hi, double px,
                       double lsd, double lambda, double *out5);
int mc_map_destroy(mc_map *m);

/* Dark: mean of n_frames frames (FP64 on the GPU) times scale (the
 * reference scales by framesToSum, IntegratorZarrOMP.c:887-892).  NULL
 * clears the dark.  bc_y/bc_z are only used by MC_MODE_GRADIENT. */
int mc_map_set_dark(mc_map *m, const void *dark, int32_t dtype,
                    int32_t n_frames, double scale, int32_t on_device);
int mc_map_set_beam_center(mc_map *m, double bc_y, double bc_z);

/* Tuning of the batched kernel: batch (8, 16, 32, 64 or 128) sets the
 * pass width, the narrowest of 16, 32 or 64 frames holding it (a pass of
 * 16 or 32 frames runs a kernel with one or two frames per lane; 128 runs
 * passes of 64), R rows per band (16 or 32), LDS budget per workgroup in
 * bytes (8192..163840).  0 keeps the current value. */
int mc_map_set_tuning(mc_map *m, int32_t batch, int32_t rows_per_group,
                      int32_t lds_bytes);
int mc_map_get_tuning(mc_map *m, int32_t *batch, int32_t *rows_per_group,
                      int32_t *lds_bytes);
/* Tile shape of the batched kernel: how many bin records (multiple of 16,
 * 32..1024) and staged map entries (multiple of 64, 256..16384) a tile
 * holds in LDS; the rest of the LDS budget holds the tile's pixel rows.
 * 0 keeps a value.  Defaults 256 / 2048.  No counterpart in the reference
 * (its GPU kernel is one thread per bin, .cu:975-1050). */
int mc_map_set_tile_shape(mc_map *m, int32_t bins, int32_t entries);
int mc_map_get_tile_shape(mc_map *m, int32_t *bins, int32_t *entries);
/* Layout of the batched kernel for (mode, dtype) under the current tuning
 * (built if needed): out[9] = tiles, chunks (workgroups), direct bins,
 * staged pixels per frame, entries, LDS bytes per workgroup, bands,
 * bytes per batch element (2 or 4), and 64-pixel blocks holding pixels of
 * the map (the blocks the transposes rewrite). */
int mc_map_layout_info(mc_map *m, int32_t mode, int32_t dtype, int64_t *out);
/* Lanes per bin of the one-frame kernels (4, 8 or 16; default 8). */
int mc_map_set_plain_lanes(mc_map *m, int32_t lanes);
int mc_map_get_plain_lanes(mc_map *m, int32_t *lanes);
/* Times candidate plans on synthetic frames of the map's shape and keeps
 * the fastest: batch 16/32/64, rows per band 16/32, LDS budget 80/160 KiB and
 * three tile shapes for the batched kernel (128 frames a timing), then 4/8/
 * 16 lanes per bin for the one-frame kernel; writes the winner's batched
 * frames/s.  No counterpart in the reference (it has no tuner). */
int mc_map_autotune(mc_map *m, int32_t dtype, double *best_fps);
/* Key of a tuned plan: 16 hex digits of a 64-bit FNV-1a hash of the map's
 * size and bins, its per-bin record counts and offsets and mask flags, the
 * pixel type, the device (architecture name, CU count) and key_extra (may
 * be NULL; e.g. the map header's parameter hash, MapHeader.h:44-55).
 * out holds at least 17 bytes. */
int mc_map_plan_key(mc_map *m, int32_t dtype, const char *key_extra, char *out, int32_t cap);
/* mc_map_autotune with a persistent plan: plan_dir/<key>.plan, when it
 * exists, is applied without timing (*loaded = 1); otherwise the candidates
 * are timed, the winner applied and written there with the table of every
 * candidate (*loaded = 0; plan_dir is created if missing).  *best_fps: the
 * plan's batched frames/s as measured when it was tuned. */
int mc_map_autotune_plan(mc_map *m, int32_t dtype, const char *plan_dir, const char *key_extra,
                         double *best_fps, int32_t *loaded);

/* Integrates n_frames frames (each NrPixelsZ x NrPixelsY, row-major) into
 * cakes [n_frames][n_r][n_eta] float32 and, when lineouts != NULL,
 * [n_frames][3][n_r] float32 (rows MC_LINEOUT_*).  *_on_device selects
 * whether each pointer is a device or host address.  n_frames == 1 uses
 * the one-frame kernel; larger counts use the batched kernel in passes of
 * the tuned batch size.  Synchronous.
 * MC_DTYPE_HYBRID: frames holds n_frames messages back to back, each a u32
 * overflow count C, the uint16 base image and C (i32 pixel, i64 value)
 * patches (the wire format of IntegratorFitPeaksGPUStream.cu:568-689); they
 * are decoded on the GPU (.cu:1468-1494) to int32 and integrated as such
 * (exact in FP32 up to 2^24 counts, as the reference's FP32 conversion). */
int mc_integrate(mc_map *m, const void *frames, int32_t dtype,
                 int32_t n_frames, int32_t frames_on_device, int32_t mode,
                 uint32_t flags, float *cakes, int32_t cakes_on_device,
                 float *lineouts, int32_t lineouts_on_device);
/* mc_integrate on device buffers, queued on the map's context stream
 * without waiting (the stream/event form of IntegratorFitPeaksGPUStream.cu
 * :2511-2555): the cakes are complete after mc_ctx_sync, after an event
 * recorded behind the call is done (mc_event_*), or when a host callback
 * queued behind it runs (mc_ctx_callback).  Calls on one context run in
 * order.  The first call for a (mode, dtype) builds the map's layout and
 * may allocate; later calls allocate only when a buffer must grow.  The
 * hybrid dtype is refused (it reads the messages' counts on the host). */
int mc_integrate_async(mc_map *m, const void *frames_dev, int32_t dtype,
                       int32_t n_frames, int32_t mode, uint32_t flags,
                       float *cakes_dev, float *lineouts_dev);
/* Completion of work queued on a context's stream: an event recorded at
 * the current end of the stream, polled (done = 1 when everything queued
 * before it has finished), waited for, destroyed; or a host function the
 * runtime calls once everything queued before it has finished (it must not
 * call into this library). */
typedef struct mc_event mc_event;
int mc_event_record(mc_ctx *ctx, mc_event **out);
int mc_event_query(mc_event *ev, int32_t *done);
int mc_event_wait(mc_event *ev);
int mc_event_destroy(mc_event *ev);
typedef void (*mc_host_fn)(void *user);
int mc_ctx_callback(mc_ctx *ctx, mc_host_fn fn, void *user);

/* ---- the batched passes captured in a HIP graph ----
 * For a caller that integrates the same device buffers over and over (the
 * per-frame loop of IntegratorFitPeaksGPUStream.cu:2567-2883 refills one
 * frame buffer per batch): mc_graph_create runs mc_integrate's batched
 * passes once on device frames[n_frames] (n_frames >= 1, batched kernel
 * also for one frame), captures the same passes into a graph and
 * instantiates it; mc_graph_launch replays it on the context's stream
 * (asynchronous: mc_ctx_sync waits), reading whatever frames_dev holds at
 * that time and writing cakes_dev / lineouts_dev (device pointers, NULL
 * lineouts for none).  A new dark (mc_map_set_dark) is folded into the
 * graph's tables before each replay; a new tuning leaves the graph on the
 * layout it captured; a new beam centre in MC_MODE_GRADIENT, or a frame
 * type needing a larger batch buffer, makes mc_graph_launch fail with
 * MC_ERR_STATE (create a new graph).  Launch the graph only while its map
 * lives; mc_graph_destroy needs only the map's context.  The reference has no counterpart:
 * the CUDA path launches kernel by kernel. */
typedef struct mc_graph mc_graph;
int mc_graph_create(mc_map *m, const void *frames_dev, int32_t dtype, int32_t n_frames,
                    int32_t mode, uint32_t flags, float *cakes_dev, float *lineouts_dev,
                    mc_graph **out);
int mc_graph_launch(mc_graph *g);
int mc_graph_destroy(mc_graph *g);

/* Cakes with propagated variance (integrate_with_variance,
 * packages/midas_integrate/midas_integrate/kernels.py:476-561): per bin,
 * besides the value mc_integrate returns for (mode, flags), the variance
 * sum(w^2 * Var(pixel)) over the same merged (bin, pixel) weights, scaled by
 * the square of the value's own scale (1/sum A when MC_F_NORMALIZE), NaN
 * where the value is NaN.  variance = [n_frames][NZ][NY] float32 variance
 * images, or NULL for Poisson statistics: Var = the frame minus the dark
 * (the dark's sum w^2 * dark is subtracted).  cakes and variances are
 * [n_frames][n_r][n_eta] float32, host or device (out_on_device).  The
 * hybrid dtype is not accepted here.  Poisson statistics over 3 frames or
 * more run batched passes of 32 frames (values and variances from one
 * gather); given variance images run the one-frame kernel per frame. */
int mc_integrate_variance(mc_map *m, const void *frames, int32_t dtype, int32_t n_frames,
                          int32_t frames_on_device, const float *variance,
                          int32_t variance_on_device, int32_t mode, uint32_t flags,
                          float *cakes, float *variances, int32_t out_on_device);

/* 1-D lineouts of cakes already computed ([n_frames][n_r][n_eta] float32)
 * into [n_frames][3][n_r] (rows MC_LINEOUT_*); the profile kernels of
 * IntegratorFitPeaksGPUStream.cu:1162-1334 and the CPU lineout of
 * IntegratorZarrOMP.c:1829-1843.  NaN bins (mask) count as missing in all
 * three definitions, as in the lineouts mc_integrate fuses. */
int mc_profile(mc_map *m, const float *cakes, int32_t n_frames,
               int32_t cakes_on_device, float *lineouts,
               int32_t lineouts_on_device);

/* 1-D lineouts and their variances of cakes with propagated variance
 * (profile_1d_with_variance, packages/midas_integrate/midas_integrate/
 * kernels.py:604-644): mode MC_LINEOUT_AREA gives sum(A*I)/sum(A) and
 * sum(A^2*Var)/sum(A)^2 over the bins with A > 1e-9, MC_LINEOUT_MEAN the
 * mean of I and sum(Var)/N^2 over the same bins; NaN bins count as
 * missing, a row without bins gives 0.  cakes and variances are
 * [n_frames][n_r][n_eta], profile and profile_var [n_frames][n_r], float32,
 * host or device (*_on_device). */
int mc_profile_variance(mc_map *m, const float *cakes, const float *variances,
                        int32_t n_frames, int32_t in_on_device, int32_t mode,
                        float *profile, float *profile_var, int32_t out_on_device);

/* ---- Zarr-zip input: Blosc1 chunks ---- */

/* Decompresses one Blosc1 chunk (the chunk format of the reference's
 * Zarr zips, read there with blosc1_decompress: IntegratorZarrOMP.c:1440,
 * 1658; ZarrReader.c:7-62).  Handles stored chunks, split and unsplit
 * blocks, byte shuffle and bitshuffle (format versions 2 and 3), and the
 * zstd, lz4/lz4hc and zlib codecs through the system's libzstd.so.1,
 * liblz4.so.1 and libz.so.1 (loaded on first use); blosclz and snappy
 * chunks return MC_ERR_FORMAT.  *out_size = decompressed bytes; an error
 * if they exceed dst_capacity.  Host memory only; no GPU is used. */
int mc_blosc_decompress(const void *src, int64_t src_size, void *dst,
                        int64_t dst_capacity, int64_t *out_size);

/* n chunks at once on `threads` host threads (0 = all cores); out_sizes[i]
 * is -1 for a chunk that failed, and the error names the first one. */
int mc_blosc_decompress_many(const void *const *srcs, const int64_t *src_sizes,
                             void *const *dsts, int64_t dst_capacity, int32_t n,
                             int32_t threads, int64_t *out_sizes);

/* ---- post-process filters on cakes along eta ---- */

/* Gaussian smoothing along eta of n_cakes cakes [n_cakes][n_r][n_eta]
 * (replaces gauss_smooth_eta, midas_integrate/postprocess.py:36-103).
 * Weights w_k = exp(-k^2 / (2 sigma^2)) normalised to sum 1 over
 * k = -h..h, h = max(1, round-half-even(4 sigma)).  area ([n_r][n_eta],
 * optional) makes the smoothing area-weighted: out = sum w I A / sum w A,
 * which keeps each ring's sum of I*A when the axis is periodic; with
 * area NULL every bin has area 1.  periodic 1 wraps the eta axis, 0
 * reflects it at the ends (numpy "reflect").  FP64 arithmetic in the
 * reference's order.  cakes, area and out are device pointers when
 * on_device is 1, host pointers otherwise; out must not alias cakes.
 * Errors: sigma <= 0, h > 1024, or h > n_eta when periodic. */
int mc_smooth_eta(mc_ctx *ctx, const float *cakes, int32_t n_cakes, int32_t n_r,
                  int32_t n_eta, double sigma_bins, const float *area,
                  int32_t periodic, int32_t on_device, float *out);

/* Median filter along eta over an odd window (replaces median_filter_eta,
 * postprocess.py:106-142): the window's median, NaN where the window holds
 * a NaN; with preserve_ring_flux each row is then scaled by
 * sum(cake row) / sum(filtered row) (float32 sums in numpy's pairwise
 * order; factor 1 where the filtered sum is 0).  window 1..63.  Pointers
 * as in mc_smooth_eta. */
int mc_median_eta(mc_ctx *ctx, const float *cakes, int32_t n_cakes, int32_t n_r,
                  int32_t n_eta, int32_t window, int32_t periodic,
                  int32_t preserve_ring_flux, int32_t on_device, float *out);

/* Launch timing of the kernels on device-resident frames, with HIP events
 * on the context's stream.  kind 0 = one-frame kernel (n_frames launches),
 * 1 = batched kernel (n_frames/batch launches per iteration), 2 = the
 * batched passes of one iteration replayed as one HIP graph (mc_graph_*).
 * ms_per_launch = mean kernel time (kind 2: ms_per_iter / passes),
 * ms_per_iter = wall time per pass over all n_frames. */
int mc_time_integrate(mc_map *m, const void *frames_dev, int32_t dtype,
                      int32_t n_frames, int32_t mode, uint32_t flags,
                      int32_t kind, int32_t iters, int32_t warmup,
                      float *cakes_dev, float *lineouts_dev,
                      double *ms_per_launch, double *ms_per_iter);

/* ---- batched peak fits (per eta sector and ring window) ---- */
#define MC_FIT_PV 0  /* height-normalised pV, 5 params [R, mu, G, Imax, BG] */
#define MC_FIT_TCH 1 /* area-normalised TCH pV, 6 free params
                        [area, center, sig, gam, bg0, bg1] */
#define MC_FIT_PV2 2 /* pV doublet, 8 params [R1, R2, mu, G1, I1, G2, I2, BG]
                        (pf_fit_doublet_pv, PeakFit.c:984-1101); windows
                        only; peak_guess holds 2 values per window (R1, R2);
                        out: R1, R2, SNR1, SNR2, Gamma1, Gamma2, mu, status */
/* TCH with N = 2..MC_FIT_MAX_PEAKS peaks on one window: a merged fitPeaks
 * job (PeakFit.c:321-351, 356-544), 4 params per peak + the 2 background
 * terms; N = 2 is the TCH doublet of the per-eta fits (:1138-1170). */
#define MC_FIT_MAX_PEAKS 8
#define MC_FIT_TCH_PEAKS(n) ((n) == 1 ? MC_FIT_TCH : (n) + 1)
#define MC_FIT_TCH2 MC_FIT_TCH_PEAKS(2)
#define MC_FIT_OUT 8 /* per window: 7 PF_PARAMS_PER_PEAK values + status */

/* Fits n_windows windows of n_points samples each.  x, y are
 * [n_windows][n_points] float32 and peak_guess [n_windows][k] (the window's
 * k expected centres: k = 1 for PV and TCH, 2 for PV2, N for
 * MC_FIT_TCH_PEAKS(N)); rstep is the sample spacing.  out is
 * [n_windows][MC_FIT_OUT] double (PV, TCH, PV2) or [n_windows][N]
 * [MC_FIT_OUT] (TCH with N peaks, one row per peak in guess order): area, center,
 * sig, gam, FWHM, eta, chi2 (chi2/dof for TCH, SNR for pV, as the
 * reference reports them), status (iterations, negative on failure).
 * Pointers are host or device. */
int mc_fit_windows(mc_ctx *ctx, int32_t model, int32_t n_windows,
                   int32_t n_points, const float *x, const float *y,
                   const float *peak_guess, double rstep, int32_t max_iter,
                   double *out, int32_t on_device);
/* mc_fit_windows without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  x, y,
 * peak_guess and out are device pointers laid out as for mc_fit_windows.
 * The kernel's window-relative x values live in a static device array of
 * 8 Mi floats (32 MiB) per GPU: n_windows * n_points above that is
 * refused (MC_ERR_ARG).  The fits are queued on the context's stream and
 * the call returns without waiting (mc_ctx_sync waits).  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (from
 * contexts with different streams) race; calls on one context's stream
 * run in order. */
int mc_fit_windows_noalloc(mc_ctx *ctx, int32_t model, int32_t n_windows,
                           int32_t n_points, const float *x, const float *y,
                           const float *peak_guess, double rstep, int32_t max_iter,
                           double *out);
/* Cuts windows out of device-resident cakes and fits them in one launch:
 * for every frame, every eta bin and every ring radius ring_r[k] (same
 * units as r_centers, the n_r R-bin centres), the ROI of R bins whose
 * centre lies within pad bin widths of the ring (IntegratorZarrOMP.c
 * :1966-1988); eta profiles with fewer than 5 nonzero bins are skipped
 * (:1943-1951, status -3).  out is [n_frames][n_eta][n_rings][MC_FIT_OUT]. */
int mc_fit_cake_rings(mc_map *m, const float *cakes_dev, int32_t n_frames,
                      const double *r_centers, const double *ring_r,
                      int32_t n_rings, int32_t pad, int32_t model,
                      int32_t max_iter, double *out, int32_t out_on_device);

/* ---- result files ---- */
typedef struct mc_out mc_out;

#define MC_OUT_CAKED_HDF 0x01u  /* <dir>/<name>.caked.hdf: /REtaMap, /InstrumentParameters */
#define MC_OUT_PER_FRAME 0x02u  /* .caked.hdf /IntegrationResult/FrameNr_<i> (IndividualSave) */
#define MC_OUT_SUM_FRAMES 0x04u /* .caked.hdf /SumFrames (SumImages) */
#define MC_OUT_OMEGAS 0x08u     /* .caked.hdf /Omegas */
#define MC_OUT_LINEOUT 0x10u    /* <dir>/<prefix>_lineout.bin and _lineout.xy */
#define MC_OUT_STREAM 0x20u     /* <dir>/lineout.bin and lineout_simple_mean.bin */
#define MC_OUT_INT2D 0x40u      /* <dir>/Int2D.bin */
#define MC_OUT_PEAKS 0x80u      /* <prefix>_fit.bin, _fit_per_eta.csv, _caked_peaks.h5 */
/* With any of the .caked.hdf parts: also <dir>/<name>.caked.hdf.zarr.zip, the
 * same groups, datasets and attributes as a Zarr (format 2) zip store, as the
 * reference's conversion writes it (FF_HEDM/workflows/integrator.py:379-426,
 * utils/converters/midas2zip.py:31-93): one chunk per array, Blosc entry
 * (zstd, level 3, bit shuffle), chunks stored uncompressed. */
#define MC_OUT_CAKED_ZARR 0x100u

/* An output set.  <name> is the data file's name: the .caked.hdf file is
 * <dir>/<name>.caked.hdf and <prefix> is <name> without its extension.
 * n_r x n_eta cakes; omega_sum_frames > 0 adds /OmegaSumFrame sums of that
 * many frames.  px, lsd (um) and wavelength (A) give 2-theta, Q and
 * d-spacing.  q_mode selects Q (q_min + q_bin_size * (r + 0.5)) instead of
 * 2-theta as the abscissa of _lineout.xy.  reta is the [5][n_r][n_eta]
 * REtaMap of mc_map_reta (NULL: no /REtaMap); r_centers [n_r] (pixels) and
 * eta_lo/eta_hi [n_eta] (degrees) are the bin axes; instrument [10] holds
 * Polariz, Lam, SH_L, U, V, W, X, Y, Z, Distance (NULL: no
 * /InstrumentParameters); source is the zarr_source attribute of
 * _caked_peaks.h5.  The arrays are copied. */
typedef struct mc_out_config {
  int32_t n_r, n_eta;
  uint32_t files;
  int32_t omega_sum_frames;
  double px, lsd, wavelength;
  int32_t q_mode;
  double q_min, q_bin_size;
  const double *reta;
  const double *r_centers;
  const double *eta_lo, *eta_hi;
  const double *instrument;
  const char *source;
} mc_out_config;

/* One per-eta peak fit (IntegratorZarrOMP.c:2038-2058): the 7 PeakFit
 * parameters (PeakFit.h:40-48) [area, center (px), sig, gam, FWHM, eta,
 * chi2 or SNR] of peak `peak` in eta bin `eta_bin` of frame `frame`. */
typedef struct mc_peak_row {
  int32_t frame, eta_bin, peak, reserved;
  double p[7];
} mc_peak_row;

int mc_out_open(const char *dir, const char *name, const mc_out_config *cfg, mc_out **out);
/* Queues one frame: cake [n_r][n_eta] float32 (host), lineouts
 * [3][n_r] float32 (rows MC_LINEOUT_*; needed by MC_OUT_STREAM, else may be
 * NULL), omega in degrees, meta [4] = Temperature, Pressure, I, I0 (or
 * NULL).  The arrays are copied and the files are written by the output
 * set's writer thread; frames are written in call order.  Blocks while 64
 * frames are queued. */
int mc_out_frame(mc_out *o, int32_t frame, double omega, const float *cake,
                 const float *lineouts, const double *meta);
/* Fits of one frame's lineout: n_peaks x 7 PeakFit parameters
 * (<prefix>_fit.bin, IntegratorZarrOMP.c:1879-1883). */
int mc_out_lineout_fits(mc_out *o, int32_t frame, int32_t n_peaks, const double *params);
/* Per-eta fits: rows with center == 0 are skipped, as the reference does. */
int mc_out_peaks(mc_out *o, int32_t n_rows, const mc_peak_row *rows);
/* Drains the queue, writes the closing datasets and closes every file.
 * Returns the first error of the output set; the handle is freed. */
int mc_out_close(mc_out *o);

/* ---- continuous-batching frame server ---- */
typedef struct mc_server mc_server;
typedef struct mc_server_stats {
  int64_t frames_done;
  int64_t batches;
  double p50_ms, p99_ms, mean_ms, max_ms;
  double frames_per_s;
  int64_t admitted_bytes, peak_bytes;
  int64_t rejected;
} mc_server_stats;
/* Result callback: called from the server's completion thread with the
 * request id, its geometry id and host pointers to the cake / lineouts
 * (valid during the call). */
typedef void (*mc_result_cb)(void *user, int64_t request_id, int32_t geometry,
                             const float *cake, const float *lineouts,
                             int32_t status);
int mc_server_create(mc_ctx *ctx, int64_t hbm_budget_bytes,
                     int32_t max_batch, int32_t queue_depth, mc_result_cb cb,
                     void *user, mc_server **out);
/* Registers a geometry (map, dtype, mode, flags); returns its id.  Fails
 * with MC_ERR_BUDGET when its resident footprint does not fit. */
int mc_server_add_geometry(mc_server *s, mc_map *m, int32_t dtype,
                           int32_t mode, uint32_t flags, int32_t *geometry);
/* Submits one frame (host memory, copied before return).  Thread-safe.
 * Blocks while the queue is full (backpressure) up to timeout_ms. */
int mc_server_submit(mc_server *s, int32_t geometry, const void *frame,
                     int64_t request_id, int32_t timeout_ms);
/* Zero-copy submission: a pinned host buffer of the geometry's frame size
 * (from the server's pool, allocated when the pool is empty).  The caller
 * fills it (for instance with a socket receive) and either submits it --
 * the server owns it again and returns it to the pool after the frame's
 * batch -- or releases it unused.  Same queueing and backpressure as
 * mc_server_submit. */
int mc_server_acquire_buffer(mc_server *s, int32_t geometry, void **buf);
int mc_server_submit_buffer(mc_server *s, int32_t geometry, void *buf,
                            int64_t request_id, int32_t timeout_ms);
int mc_server_release_buffer(mc_server *s, int32_t geometry, void *buf);
int mc_server_drain(mc_server *s, int32_t timeout_ms);
/* Batching window (0 = none, the default): a free slot waits up to
 * window_us after the oldest queued request arrived for max_batch frames of
 * its geometry before launching what is queued; mc_server_drain closes the
 * window.  No counterpart in the reference, which launches one frame per
 * request (IntegratorFitPeaksGPUStream.cu:2567-2883): it trades latency for
 * larger batches under load. */
int mc_server_set_batch_window(mc_server *s, int64_t window_us);
int mc_server_stats_get(mc_server *s, mc_server_stats *st);
int mc_server_destroy(mc_server *s);

#ifdef __cplusplus
}
#endif
#endif
