// mc_core.cpp -- host runtime of the MI355X caking library: contexts, map
// residency, layouts for the kernels, dark handling and the integrate
// entry points of the C ABI (include/midas_caking.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <iterator>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include <sys/stat.h>

#include "mc_internal.h"

namespace mc {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

// ------------------------------------------------------------------
// Derived per-bin arrays + upload of the canonical map.
// ------------------------------------------------------------------
int map_finish_host(Map *m) {
  const int64_t nb = m->nb();
  MC_CHECK((int64_t)m->h_pairs.size() == 2 * nb, MC_ERR_FORMAT,
           "nMap has %lld pairs, expected %lld", (long long)m->h_pairs.size() / 2,
           (long long)nb);
  if (m->h_flags.empty()) m->h_flags.assign(nb, 0);
  for (int64_t b = 0; b < nb; b++) {
    int64_t c = m->h_pairs[2 * b], o = m->h_pairs[2 * b + 1];
    MC_CHECK(c >= 0 && o >= 0 && o + c <= m->nnz, MC_ERR_FORMAT,
             "bin %lld references records [%lld, %lld) past %lld", (long long)b,
             (long long)o, (long long)(o + c), (long long)m->nnz);
  }
  m->h_area_all.assign(nb, 0.0);
  m->h_area_inb.assign(nb, 0.0);
  std::vector<float> inv(nb), ainb(nb);
  std::vector<float4> rs(m->n_r, make_float4(0.f, 0.f, 0.f, 0.f));
  for (int64_t b = 0; b < nb; b++) {
    double A = 0, Ai = 0;
    const int64_t o = m->h_pairs[2 * b + 1], c = m->h_pairs[2 * b];
    for (int64_t k = o; k < o + c; k++) {
      const Rec &r = m->h_rec[k];
      A += r.areaWeight;
      if (!(r.y < 0 || r.y >= m->ny || r.z < 0 || r.z >= m->nz)) Ai += r.areaWeight;
    }
    m->h_area_all[b] = A;
    m->h_area_inb[b] = Ai;
    inv[b] = A > 0 ? (float)(1.0 / A) : 0.f;
    ainb[b] = (float)Ai;
    if (Ai > 1e-9) {
      float4 &q = rs[b / m->n_eta];
      q.x += (float)Ai;
      q.y += 1.f;
      if (!m->h_flags[b]) {
        q.z += (float)Ai;
        q.w += 1.f;
      }
    }
  }
  int rc;
  if ((rc = m->d_rec.alloc(m->nnz * sizeof(Rec))) || (rc = m->d_pairs.alloc(2 * nb * 4)) ||
      (rc = m->d_flags.alloc(nb * 4)) || (rc = m->d_area_inb.alloc(nb * 4)) ||
      (rc = m->d_row_static.alloc(m->n_r * sizeof(float4))))
    return rc;
  static_assert(sizeof(float2) == 8, "float2");
  MC_HIP(hipMemcpy(m->d_rec.p, m->h_rec.data(), m->nnz * sizeof(Rec), hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(m->d_pairs.p, m->h_pairs.data(), 2 * nb * 4, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(m->d_flags.p, m->h_flags.data(), nb * 4, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(m->d_area_inb.p, ainb.data(), nb * 4, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(m->d_row_static.p, rs.data(), m->n_r * sizeof(float4), hipMemcpyHostToDevice));
  for (bool &ok : m->binfo_ok) ok = false;
  if ((rc = m->inv_area.alloc(nb * 4))) return rc;
  MC_HIP(hipMemcpy(m->inv_area.p, inv.data(), nb * 4, hipMemcpyHostToDevice));
  return MC_OK;
}

// ------------------------------------------------------------------
// CSR of one integration mode (host build, then upload).
// ------------------------------------------------------------------
namespace {

struct PW {
  int32_t pix;
  double w;
};

// Appends the (pixel, weight) contributions of record r for mode.
inline void expand(const Map &m, const Rec &r, int mode, double bcy, double bcz,
                   std::vector<PW> &out) {
  const int NY = m.ny, NZ = m.nz;
  const int64_t npx = m.npx();
  if (mode == MC_MODE_FLOOR) {
    // (long long)z * NY + y in float arithmetic, IntegratorFitPeaksGPUStream.cu:916.
    long long zz = (long long)r.z * NY;
    float o = (float)zz + r.y;
    long long p = (long long)o;
    if (p < 0) p = 0;
    if (p >= npx) p = npx - 1;
    out.push_back({(int32_t)p, r.frac});
    return;
  }
  double ry = r.y, rz = r.z;
  if (mode == MC_MODE_GRADIENT && r.deltaR != 0.0f) {
    double dy = r.y - bcy, dz = r.z - bcz;
    double R = sqrt(dy * dy + dz * dz);
    if (R > 1.0) {
      ry -= r.deltaR * dy / R;
      rz -= r.deltaR * dz / R;
    }
  }
  // Bilinear sampler of IntegratorZarrOMP.c:1733-1744 (floorf on the
  // coordinate converted to float, edge clamps).
  int iy = (int)floorf((float)ry), iz = (int)floorf((float)rz);
  double fy = ry - iy, fz = rz - iz;
  if (iy < 0) { iy = 0; fy = 0; }
  if (iy >= NY - 1) { iy = NY - 2; fy = 1; }
  if (iz < 0) { iz = 0; fz = 0; }
  if (iz >= NZ - 1) { iz = NZ - 2; fz = 1; }
  const int64_t b00 = (int64_t)iz * NY + iy;
  const double w00 = (1 - fy) * (1 - fz), w10 = fy * (1 - fz), w01 = (1 - fy) * fz,
               w11 = fy * fz;
  if (w00 != 0) out.push_back({(int32_t)b00, w00 * r.frac});
  if (w10 != 0) out.push_back({(int32_t)(b00 + 1), w10 * r.frac});
  if (w01 != 0) out.push_back({(int32_t)(b00 + NY), w01 * r.frac});
  if (w11 != 0) out.push_back({(int32_t)(b00 + NY + 1), w11 * r.frac});
}

template <class F> void parallel_for(int64_t n, F &&f) {
  int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      int64_t a = n * t / nt, b = n * (t + 1) / nt;
      f(a, b);
    });
  for (auto &x : th) x.join();
}

}  // namespace

int ensure_csr(Map *m, int mode) {
  MC_CHECK(mode >= 0 && mode <= 2, MC_ERR_ARG, "mode %d", mode);
  Csr &c = m->csr[mode];
  if (m->fused) {
    MC_CHECK(c.built, MC_ERR_ARG, "a fused map integrates in bilinear or gradient mode only");
    return MC_OK;
  }
  if (c.built && (mode != MC_MODE_GRADIENT || (c.bc_y == m->bc_y && c.bc_z == m->bc_z)))
    return MC_OK;
  const int64_t nb = m->nb();
  // Pass 1: per-bin coalesced entry counts, built per thread range.
  std::vector<std::vector<PW>> per_bin_tmp;  // not kept: two passes instead
  std::vector<int32_t> cnt(nb, 0);
  auto bin_entries = [&](int64_t b, std::vector<PW> &v) {
    v.clear();
    const int64_t o = m->h_pairs[2 * b + 1], n = m->h_pairs[2 * b];
    for (int64_t k = o; k < o + n; k++) expand(*m, m->h_rec[k], mode, m->bc_y, m->bc_z, v);
    std::sort(v.begin(), v.end(), [](const PW &a, const PW &b) { return a.pix < b.pix; });
    size_t w = 0;
    for (size_t i = 0; i < v.size(); i++) {
      if (w > 0 && v[w - 1].pix == v[i].pix) v[w - 1].w += v[i].w;
      else v[w++] = v[i];
    }
    v.resize(w);
  };
  parallel_for(nb, [&](int64_t a, int64_t b) {
    std::vector<PW> v;
    for (int64_t q = a; q < b; q++) {
      bin_entries(q, v);
      cnt[q] = (int32_t)v.size();
    }
  });
  c.h_row.assign(nb + 1, 0);
  for (int64_t b = 0; b < nb; b++) c.h_row[b + 1] = c.h_row[b] + cnt[b];
  const int64_t nnz = c.h_row[nb];
  c.h_pix.assign(nnz, 0);
  c.h_w.assign(nnz, 0.f);
  parallel_for(nb, [&](int64_t a, int64_t b) {
    std::vector<PW> v;
    for (int64_t q = a; q < b; q++) {
      bin_entries(q, v);
      int64_t o = c.h_row[q];
      for (size_t i = 0; i < v.size(); i++) {
        c.h_pix[o + i] = v[i].pix;
        c.h_w[o + i] = (float)v[i].w;
      }
    }
  });
  c.nnz = nnz;
  int rc;
  if ((rc = c.row.alloc((nb + 1) * 4)) || (rc = c.pix.alloc(nnz * 4)) ||
      (rc = c.w.alloc(nnz * 4)) || (rc = c.dterm.alloc(nb * 4)))
    return rc;
  MC_HIP(hipMemcpy(c.row.p, c.h_row.data(), (nb + 1) * 4, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(c.pix.p, c.h_pix.data(), nnz * 4, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(c.w.p, c.h_w.data(), nnz * 4, hipMemcpyHostToDevice));
  c.bc_y = m->bc_y;
  c.bc_z = m->bc_z;
  c.built = true;
  c.dterm_valid = false;
  c.d2_valid = false;
  m->tiles[mode].clear();
  for (int f = 0; f < 4; f++) m->binfo_ok[mode * 4 + f] = false;
  return MC_OK;
}

// ------------------------------------------------------------------
// Tile planner of the batched kernel.
// ------------------------------------------------------------------
// LDS of the batched kernel: staged entries, the staged pixel rows of the
// pixel-major batch, a zero pixel row, and two pixel-list buffers (the
// current tile's and the next one's).  Quad records are read with scalar
// loads and take no LDS.
constexpr int kWaves = 16, kTileInts = 16;

// Tiles of the batched kernel.  Bands of H cake rows; inside a band, tiles
// are rectangles of rows x eta columns grown while the footprint of their
// entries (the distinct pixels) fits the LDS budget -- four columns at a
// time where the rows hold a multiple of 4 bins, so that tiles cut the rows
// between quads; a tile stages exactly its footprint, one pixel row of the
// batch per pixel.  A column that does not fit whole is split into row
// ranges, and a bin that does not fit alone is "direct" (gathered from
// global memory).  A band gets one chunk (the work of one workgroup) per
// `target` entries of its tiles, and its tiles are dealt to its chunks in
// turn.  in_size is the
// element size of the pixel-major batch (2 or 4 bytes).
int ensure_tiles(Map *m, int mode, int in_size, Tiles **out, int width) {
  int rc = ensure_csr(m, mode);
  if (rc) return rc;
  const int B = m->batch, H = m->rows_per_group;
  const int W = width ? width : width_of(B);
  MC_CHECK(W == 16 || W == 32 || W == 64, MC_ERR_ARG, "pass width %d not in {16, 32, 64}", W);
  const int kTileBins = m->tile_bins, kTileEnts = m->tile_ents;
  for (auto &t : m->tiles[mode])
    if (t->width == W && t->rows == H && t->lds_bytes == m->lds_bytes && t->in_size == in_size &&
        t->bin_cap == kTileBins && t->ent_cap == kTileEnts) {
      *out = t.get();
      return MC_OK;
    }
  MC_CHECK(H == 16 || H == 32, MC_ERR_ARG, "rows per band %d not in {16, 32}", H);
  MC_CHECK(B >= 1 && B <= 128, MC_ERR_ARG, "batch %d not in 1..128", B);
  MC_CHECK(in_size == 2 || in_size == 4, MC_ERR_ARG, "batch element size %d", in_size);
  const Csr &c = m->csr[mode];
  const int nE = m->n_eta, nR = m->n_r;
  const int64_t nb = m->nb();
  const int RB = W * in_size;  // bytes per pixel row
  const int64_t fixed = (int64_t)kTileEnts * 8;
  // Pixels a tile can stage: what the LDS budget holds after the entries,
  // the zero row and the two pixel lists (4 bytes a pixel, rounded up to
  // 16 bytes).
  int64_t cap_px = (m->lds_bytes - fixed - RB - 24) / (RB + 8);
  // uint16 entries carry the row's byte offset in 16 bits (the high half
  // holds the pixel's dark): the zero row, at npx * RB, must stay below 2^16.
  if (in_size == 2) cap_px = std::min<int64_t>(cap_px, 65535 / RB - 1);
  MC_CHECK(cap_px >= 16, MC_ERR_ARG, "LDS budget %d too small", m->lds_bytes);
  auto cnt = [&](int64_t b) { return (int64_t)(c.h_row[b + 1] - c.h_row[b]); };
  // Staged entries of a quad: the slots holding staged entries ("staged
  // slots") are interleaved a pair at a time, each padded to the pairs of
  // the quad's largest bin; the other slots read the tile's zero pair.
  auto pairs = [](int64_t n) { return (n + 1) / 2; };
  auto qents = [&](int64_t n, int64_t p) { return 2 * pairs(n) * p; };
  // Per-bin footprints: the sorted distinct pixels of its entries; a bin
  // whose footprint or entries do not fit a tile alone is direct.
  std::vector<std::vector<int32_t>> fp(nb);
  std::vector<char> big(nb, 0);
  parallel_for(nb, [&](int64_t a, int64_t b) {
    for (int64_t q = a; q < b; q++) {
      auto &v = fp[q];
      v.assign(c.h_pix.begin() + c.h_row[q], c.h_pix.begin() + c.h_row[q + 1]);
      std::sort(v.begin(), v.end());
      v.erase(std::unique(v.begin(), v.end()), v.end());
      big[q] = (int64_t)v.size() > cap_px || qents(cnt(q), 1) + 2 > kTileEnts;
      if (big[q]) std::vector<int32_t>().swap(v);
    }
  });
  auto staged_cnt = [&](int64_t b) { return big[b] ? (int64_t)0 : cnt(b); };
  // Size of the union of two sorted pixel sets.
  auto union_size = [](const std::vector<int32_t> &x, const std::vector<int32_t> &y) {
    size_t i = 0, j = 0, n = 0;
    while (i < x.size() && j < y.size()) {
      if (x[i] < y[j]) i++;
      else if (y[j] < x[i]) j++;
      else { i++; j++; }
      n++;
    }
    return (int64_t)(n + (x.size() - i) + (y.size() - j));
  };
  auto merge_into = [](std::vector<int32_t> &x, const std::vector<int32_t> &y) {
    std::vector<int32_t> u;
    u.reserve(x.size() + y.size());
    std::set_union(x.begin(), x.end(), y.begin(), y.end(), std::back_inserter(u));
    x.swap(u);
  };
  // Last quad of each row of a tile being grown: its largest staged bin and
  // its staged slots.
  struct RowQ {
    int64_t q = -1, n = 0, p = 0;
  };
  struct TileH {
    int ra = 0, nrows = 0, e0 = 0, nc = 0;
    std::vector<int32_t> px;  // footprint, sorted
    std::vector<RowQ> rq;     // per row (whole-band tiles)
    int64_t ents = 0;         // staged entries
    bool chunk_start = false; // first tile of a workgroup's run
  };
  // Chunks (one workgroup each): the band planner cuts a band's tiles into
  // runs of about `target` cost (a few thousand for a 2048^2 map), each
  // ending before an eta column that is a multiple of 8; the number of runs
  // is the number of workgroups the band gets.  The cost of a tile counts
  // its entries plus a fixed charge per bin and per tile: a bin of 2
  // entries costs a wave about as much as one of 30.  The band's tiles are
  // then dealt to its chunks in turn (below).
  constexpr int64_t kBinCost = 24, kTileCost = 256;
  const int64_t target =
      std::max<int64_t>(4096, (c.nnz + kBinCost * nb + kTileCost * (nb / kTileBins + 1)) / 4096);
  const int n_bands = (nR + H - 1) / H;
  // Columns added at a time: 4 where every row starts at a multiple of 4
  // bins (tiles then hold whole quads), else 1.
  const int U = nE % 4 == 0 ? 4 : 1;
  std::vector<std::vector<TileH>> bands(n_bands);
  parallel_for(n_bands, [&](int64_t a, int64_t bq) {
    for (int64_t gi = a; gi < bq; gi++) {
      auto &tv = bands[gi];
      const int row0 = (int)gi * H, nrows = std::min(H, nR - row0);
      // (a tile's staged entries end with its zero pair: 2 entries)
      auto room = [&](const TileH &T, const std::vector<int32_t> &add, int64_t ents,
                      int nbins) {
        return (int64_t)T.nrows * T.nc + nbins <= kTileBins && T.ents + ents + 2 <= kTileEnts &&
               union_size(T.px, add) <= cap_px;
      };
      // Staged entries that columns e .. e + n - 1 of all rows add to T
      // (rq empty: a fresh tile); with `commit` T's row quads advance.
      auto add_cols = [&](TileH &T, int e, int n, bool commit) {
        std::vector<RowQ> rq = T.rq;
        rq.resize(nrows);
        int64_t add = 0;
        for (int j = 0; j < nrows; j++)
          for (int u = 0; u < n; u++) {
            const int64_t b = (int64_t)(row0 + j) * nE + e + u, q = b / 4, k = staged_cnt(b);
            const int64_t st = k > 0 ? 1 : 0;
            if (rq[j].q == q) {
              const int64_t was = qents(rq[j].n, rq[j].p);
              rq[j].n = std::max(rq[j].n, k);
              rq[j].p += st;
              add += qents(rq[j].n, rq[j].p) - was;
            } else {
              add += qents(k, st);
              rq[j].q = q;
              rq[j].n = k;
              rq[j].p = st;
            }
          }
        if (commit) T.rq.swap(rq);
        return add;
      };
      TileH cur;
      bool next_starts = true;  // the next tile begins a chunk
      int64_t acc = 0;          // cost since the current chunk began
      auto push = [&](TileH &&T) {
        T.chunk_start = next_starts;
        next_starts = false;
        acc += kTileCost;
        std::vector<RowQ>().swap(T.rq);
        tv.push_back(std::move(T));
      };
      auto flush = [&]() {
        if (cur.nc > 0) push(std::move(cur));
        cur = TileH();
      };
      // Adds columns e .. e + n - 1 (all rows) to the current tile or a
      // fresh one; false when they do not fit a fresh tile.
      std::vector<int32_t> col;
      auto place = [&](int e, int n) {
        col.clear();
        for (int j = 0; j < nrows; j++)
          for (int u = 0; u < n; u++) {
            const int64_t b = (int64_t)(row0 + j) * nE + e + u;
            if (!big[b]) merge_into(col, fp[b]);
          }
        if (cur.nc > 0 && cur.nrows == nrows) {
          const int64_t ce = add_cols(cur, e, n, false);
          if (room(cur, col, ce, nrows * n)) {
            merge_into(cur.px, col);
            cur.ents += add_cols(cur, e, n, true);
            cur.nc += n;
            return true;
          }
        }
        flush();
        TileH fresh;
        fresh.nrows = nrows;
        const int64_t ce = add_cols(fresh, e, n, false);
        if (!room(fresh, col, ce, nrows * n)) return false;
        cur = std::move(fresh);
        cur.e0 = e;
        cur.nc = n;
        cur.px = col;
        cur.ents = add_cols(cur, e, n, true);
        return true;
      };
      for (int e = 0; e < nE; e += U) {
        const int n = std::min(U, nE - e);
        if (e % 8 == 0 && acc >= target) {  // end the chunk here
          flush();
          next_starts = true;
          acc = 0;
        }
        for (int j = 0; j < nrows; j++)
          for (int u = 0; u < n; u++) acc += cnt((int64_t)(row0 + j) * nE + e + u) + kBinCost;
        if (place(e, n)) continue;
        // Four columns (whole quads) that do not fit a fresh tile: row
        // ranges of the four columns, a quad per row, while a single row of
        // them fits.
        if (n == 4) {
          std::vector<std::vector<int32_t>> radd(nrows);
          std::vector<int64_t> rents(nrows);
          bool fits = true;
          for (int j = 0; j < nrows && fits; j++) {
            int64_t kmax = 0, st = 0;
            for (int u = 0; u < 4; u++) {
              const int64_t b = (int64_t)(row0 + j) * nE + e + u;
              if (!big[b]) merge_into(radd[j], fp[b]);
              kmax = std::max(kmax, staged_cnt(b));
              st += staged_cnt(b) > 0 ? 1 : 0;
            }
            rents[j] = qents(kmax, st);
            fits = room(TileH(), radd[j], rents[j], 4);
          }
          if (fits) {
            flush();
            TileH piece;
            for (int j = 0; j < nrows; j++) {
              if (piece.nrows > 0 && !room(piece, radd[j], rents[j], 4)) {
                push(std::move(piece));
                piece = TileH();
              }
              if (piece.nrows == 0) {
                piece.e0 = e;
                piece.nc = 4;
                piece.ra = j;
              }
              merge_into(piece.px, radd[j]);
              piece.ents += rents[j];
              piece.nrows++;
            }
            push(std::move(piece));
            continue;
          }
        }
        // The columns do not fit a fresh tile together: one at a time.
        for (int e1 = e; e1 < e + n; e1++) {
          if (n > 1 && place(e1, 1)) continue;
          // The whole column does not fit: row ranges of one column each,
          // a quad per row.
          flush();
          TileH piece;
          piece.e0 = e1;
          piece.nc = 1;
          static const std::vector<int32_t> none;
          for (int j = 0; j < nrows; j++) {
            const int64_t b = (int64_t)(row0 + j) * nE + e1;
            const std::vector<int32_t> &add = big[b] ? none : fp[b];
            const int64_t ae = qents(staged_cnt(b), staged_cnt(b) > 0 ? 1 : 0);
            if (piece.nrows > 0 && !room(piece, add, ae, 1)) {
              push(std::move(piece));
              piece = TileH();
              piece.e0 = e1;
              piece.nc = 1;
              piece.ra = j;
            }
            merge_into(piece.px, add);
            piece.ents += ae;
            piece.nrows++;
          }
          push(std::move(piece));
        }
      }
      flush();
    }
  });
  // Deal each band's tiles to its k chunks in turn: tile i (in eta order)
  // goes to chunk i mod k.  Consecutive chunks run on one XCD at the same
  // time (see k_batched), so the band's workgroups work on neighbouring
  // tiles together: the 8-bin granules of a 128-byte cake line are stored
  // about a tile's time apart and merge in that XCD's L2 before they are
  // written back, and neighbouring footprints share their border pixels
  // there.  (A workgroup that walks its own run of tiles leaves a line
  // open for four tiles, and every granule goes to memory on its own.)
  parallel_for(n_bands, [&](int64_t a, int64_t bq) {
    for (int64_t gi = a; gi < bq; gi++) {
      std::vector<TileH> &v = bands[gi];
      size_t k = 0;
      for (const TileH &T : v) k += T.chunk_start ? 1 : 0;
      if (k < 2) continue;
      std::vector<TileH> dealt;
      dealt.reserve(v.size());
      for (size_t j = 0; j < k; j++)
        for (size_t i = j; i < v.size(); i += k) {
          v[i].chunk_start = i == j;
          dealt.push_back(std::move(v[i]));
        }
      v.swap(dealt);
    }
  });
  // Flatten, starting a chunk at each tile marked above.
  std::vector<int64_t> chunk_cost;
  std::vector<int32_t> chunks, band_chunk(n_bands + 1, 0), tiles;
  std::vector<int4> qs;
  std::vector<int2> qdir;
  std::vector<uint2> ents;
  std::vector<int32_t> tpix;  // staged pixels of every tile, tile after tile
  ents.reserve(c.nnz + c.nnz / 4);
  int64_t n_direct = 0, max_npx = 0, staged = 0, real_ents = 0, n_slots = 0;
  for (int gi = 0; gi < n_bands; gi++) {
    const int row0 = gi * H, nrows_band = std::min(H, nR - row0);
    for (auto &T : bands[gi]) {
      if (T.chunk_start) {
        chunks.insert(chunks.end(), {(int32_t)(tiles.size() / kTileInts),
                                     (int32_t)(tiles.size() / kTileInts), row0, nrows_band});
        chunk_cost.push_back(0);
      }
      const int32_t npx = (int32_t)T.px.size();
      const int32_t px_beg = (int32_t)tpix.size();
      tpix.insert(tpix.end(), T.px.begin(), T.px.end());
      while (tpix.size() % 4) tpix.push_back(0);  // lists start at 16-byte boundaries
      max_npx = std::max<int64_t>(max_npx, npx);
      staged += npx;
      const int32_t e_beg = (int32_t)ents.size(), q_beg = (int32_t)qs.size();
      const uint32_t zero_row = (uint32_t)((int64_t)npx * RB);
      // Quads of each row: the aligned groups of 4 bins that cover the
      // row's columns; rows are padded to the tile's largest count.
      int nq = 0;
      for (int j = 0; j < T.nrows; j++) {
        const int64_t b_lo = (int64_t)(row0 + T.ra + j) * nE + T.e0, b_hi = b_lo + T.nc;
        nq = std::max<int>(nq, (int)((b_hi - 1) / 4 - b_lo / 4 + 1));
      }
      int64_t all_ents = 0;
      for (int j = 0; j < T.nrows; j++) {
        const int64_t b_lo = (int64_t)(row0 + T.ra + j) * nE + T.e0, b_hi = b_lo + T.nc;
        const int64_t q_lo = b_lo / 4, nq_j = (b_hi - 1) / 4 - q_lo + 1;
        for (int qi = 0; qi < nq; qi++) {
          const int32_t e_off = (int32_t)(ents.size() - (size_t)e_beg);
          qdir.insert(qdir.end(), 4, make_int2(0, 0));
          if (qi >= nq_j) {
            qs.push_back(make_int4(0, 0, e_off, 0));
            continue;
          }
          const int64_t b0 = 4 * (q_lo + qi);
          int32_t flags = 0;
          int64_t np = 0;
          for (int s = 0; s < 4; s++) {
            const int64_t b = b0 + s;
            if (b < b_lo || b >= b_hi) continue;
            flags |= 1 << s;
            n_slots++;
            all_ents += cnt(b);
            if (big[b]) flags |= 16 << s;
            if (staged_cnt(b) > 0) flags |= 256 << s;  // a staged slot
            np = std::max(np, pairs(staged_cnt(b)));
          }
          for (int64_t mi = 0; mi < np; mi++)
            for (int s = 0; s < 4; s++)
              for (int u = 0; u < 2; u++) {
                if (!((flags >> (8 + s)) & 1)) continue;
                const int64_t b = b0 + s, idx = 2 * mi + u;
                if (idx < cnt(b)) {
                  const int32_t k = c.h_row[b] + (int32_t)idx;
                  const int p = c.h_pix[k];
                  // byte offset of the pixel's row in the tile's staged list
                  const int64_t slot =
                      std::lower_bound(T.px.begin(), T.px.end(), p) - T.px.begin();
                  uint32_t wb;
                  memcpy(&wb, &c.h_w[k], 4);
                  ents.push_back(make_uint2((uint32_t)(slot * RB), wb));
                  real_ents++;
                } else {
                  ents.push_back(make_uint2(zero_row, 0u));
                }
              }
          qs.push_back(make_int4((int32_t)b0, flags, e_off, (int32_t)np));
        }
      }
      // The zero pair: the lanes of a quad's slots without staged entries
      // read it (the zero row, weight 0).
      const int32_t zp = (int32_t)(ents.size() - (size_t)e_beg);
      ents.push_back(make_uint2(zero_row, 0u));
      ents.push_back(make_uint2(zero_row, 0u));
      const int32_t e_end = (int32_t)ents.size();
      MC_CHECK(e_end - e_beg <= kTileEnts, MC_ERR_STATE, "tile stages %d entries over %d",
               e_end - e_beg, kTileEnts);
      // Direct slots: their entries (global pixel, weight) after the staged ones.
      for (int32_t qq = q_beg; qq < (int32_t)qs.size(); qq++)
        for (int s = 0; s < 4; s++) {
          if (!((qs[qq].y >> (4 + s)) & 1)) continue;
          const int64_t b = qs[qq].x + s;
          const int32_t eb = (int32_t)ents.size();
          for (int32_t k = c.h_row[b]; k < c.h_row[b + 1]; k++) {
            uint32_t wb;
            memcpy(&wb, &c.h_w[k], 4);
            ents.push_back(make_uint2((uint32_t)c.h_pix[k], wb));
            real_ents++;
          }
          qdir[4 * (size_t)qq + s] = make_int2(eb, (int32_t)ents.size());
          n_direct++;
        }
      // Every tile's entries start at a multiple of 8 (16-byte staging).
      while (ents.size() % 8) ents.push_back(make_uint2(0u, 0u));
      tiles.insert(tiles.end(), {px_beg, npx, zp, 0, q_beg, nq, e_beg, e_end, T.ra, T.nc, 0, 0,
                                 T.nrows, 0, 0, 0});
      chunks[chunks.size() - 3] = (int32_t)(tiles.size() / kTileInts);
      const int64_t tcost = all_ents + kTileCost + kBinCost * T.nrows * T.nc;
      chunk_cost.back() += tcost;
    }
    band_chunk[gi + 1] = (int32_t)(chunks.size() / 4);
  }
  MC_CHECK(n_slots == nb, MC_ERR_STATE, "tile planner covered %lld of %lld bins",
           (long long)n_slots, (long long)nb);
  MC_CHECK(real_ents == c.nnz, MC_ERR_STATE, "tile planner placed %lld of %lld entries",
           (long long)real_ents, (long long)c.nnz);
  // The 8 XCDs take contiguous chunk ranges of equal cost (workgroups are
  // handed to the XCDs in turn; see k_batched).
  std::vector<int32_t> xcd(9, 0);
  {
    const int64_t n = (int64_t)chunk_cost.size();
    int64_t sum = 0, all = 0;
    for (int64_t v : chunk_cost) all += v;
    int x = 1;
    for (int64_t i = 0; i < n && x < 8; i++) {
      sum += chunk_cost[i];
      while (x < 8 && sum * 8 >= all * x) xcd[x++] = (int32_t)(i + 1);
    }
    while (x < 9) xcd[x++] = (int32_t)n;
  }
  int32_t per_xcd = 0;
  for (int x = 0; x < 8; x++) per_xcd = std::max(per_xcd, xcd[x + 1] - xcd[x]);
  auto t = std::make_unique<Tiles>();
  t->batch = B;
  t->width = W;
  t->bin_cap = kTileBins;
  t->ent_cap = kTileEnts;
  t->grid = 8 * per_xcd;
  t->rows = H;
  t->lds_bytes = m->lds_bytes;
  t->in_size = in_size;
  t->n_chunks = (int)(chunks.size() / 4);
  t->n_bands = n_bands;
  t->plane_off = (int)fixed;
  t->tp_off = (int)(fixed + (max_npx + 1) * RB);
  t->tp_len = (int)(((max_npx + 3) & ~(int64_t)3) * 4);
  t->lds_needed = (size_t)t->tp_off + 2 * (size_t)t->tp_len;
  MC_CHECK((int64_t)t->lds_needed <= m->lds_bytes, MC_ERR_STATE, "tile LDS %zu over budget %d",
           t->lds_needed, m->lds_bytes);
  t->n_tiles = (int64_t)tiles.size() / kTileInts;
  t->n_quads = (int64_t)qs.size();
  t->n_ents = (int64_t)ents.size();
  t->n_direct = n_direct;
  t->staged_px = staged;
  if ((rc = t->chunks.alloc(chunks.size() * 4)) || (rc = t->tiles.alloc(tiles.size() * 4)) ||
      (rc = t->qs.alloc(std::max<size_t>(qs.size(), 1) * 16)) ||
      (rc = t->qdir.alloc(std::max<size_t>(qdir.size(), 1) * 8)) ||
      (rc = t->ents.alloc(ents.size() * 8)) ||
      (rc = t->band_chunk.alloc(band_chunk.size() * 4)) || (rc = t->xcd.alloc(9 * 4)) ||
      (rc = t->tpix.alloc(std::max<size_t>(tpix.size(), 1) * 4)))
    return rc;
  if (!tpix.empty())
    MC_HIP(hipMemcpy(t->tpix.p, tpix.data(), tpix.size() * 4, hipMemcpyHostToDevice));
  // Pixel blocks the transposes rewrite: those holding a pixel of the map.
  {
    const int64_t npx = m->npx();
    std::vector<char> used((size_t)((npx + 63) / 64), 0);
    for (int32_t p : c.h_pix) used[(size_t)p / 64] = 1;
    std::vector<int32_t> b64, b128;
    for (size_t i = 0; i < used.size(); i++)
      if (used[i]) b64.push_back((int32_t)i);
    for (size_t i = 0; i < used.size(); i += 2)
      if (used[i] || (i + 1 < used.size() && used[i + 1])) b128.push_back((int32_t)(i / 2));
    t->n_blocks64 = (int)b64.size();
    t->n_blocks128 = (int)b128.size();
    if ((rc = t->blocks64.alloc(std::max<size_t>(b64.size(), 1) * 4)) ||
        (rc = t->blocks128.alloc(std::max<size_t>(b128.size(), 1) * 4)))
      return rc;
    if (!b64.empty())
      MC_HIP(hipMemcpy(t->blocks64.p, b64.data(), b64.size() * 4, hipMemcpyHostToDevice));
    if (!b128.empty())
      MC_HIP(hipMemcpy(t->blocks128.p, b128.data(), b128.size() * 4, hipMemcpyHostToDevice));
  }
  MC_HIP(hipMemcpy(t->chunks.p, chunks.data(), chunks.size() * 4, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(t->tiles.p, tiles.data(), tiles.size() * 4, hipMemcpyHostToDevice));
  if (!qs.empty()) {
    MC_HIP(hipMemcpy(t->qs.p, qs.data(), qs.size() * 16, hipMemcpyHostToDevice));
    MC_HIP(hipMemcpy(t->qdir.p, qdir.data(), qdir.size() * 8, hipMemcpyHostToDevice));
  }
  MC_HIP(hipMemcpy(t->ents.p, ents.data(), ents.size() * 8, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(t->band_chunk.p, band_chunk.data(), band_chunk.size() * 4,
                   hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(t->xcd.p, xcd.data(), 9 * 4, hipMemcpyHostToDevice));
  t->built = true;
  static std::atomic<uint64_t> next_id{1};
  t->id = next_id++;
  *out = t.get();
  m->tiles[mode].push_back(std::move(t));
  return MC_OK;
}

int ensure_binfo(Map *m, int mode, uint32_t flags, const float2 **out) {
  MC_CHECK(flags < 4, MC_ERR_ARG, "unknown integration flags 0x%x", flags);
  int rc = ensure_csr(m, mode);
  if (rc) return rc;
  Csr &c = m->csr[mode];
  hipStream_t s = m->ctx->stream;
  const int key = mode * 4 + (int)flags;
  if (!c.dterm_valid) {
    if ((rc = launch_dark_term(s, c.row.as<int32_t>(), c.pix.as<int32_t>(), c.w.as<float>(),
                               m->has_dark ? m->d_dark.as<double>() : nullptr,
                               m->has_dark ? m->d_dint.as<uint16_t>() : nullptr, m->nb(),
                               c.dterm.as<float>())))
      return rc;
    c.dterm_valid = true;
    for (int f = 0; f < 4; f++) m->binfo_ok[mode * 4 + f] = false;
  }
  if (!m->binfo_ok[key]) {
    if (!m->d_binfo[key].p && (rc = m->d_binfo[key].alloc(m->nb() * sizeof(float2)))) return rc;
    if ((rc = launch_binfo(s, m->inv_area.as<float>(), m->d_flags.as<int32_t>(),
                           c.dterm.as<float>(), m->nb(), (flags & MC_F_NORMALIZE) ? 1 : 0,
                           (flags & MC_F_MASK_NAN) ? 1 : 0, m->d_binfo[key].as<float2>())))
      return rc;
    // Tables are read by kernels on other streams (the frame server).
    MC_HIP(hipStreamSynchronize(s));
    m->binfo_ok[key] = true;
    m->binfo_version[key]++;
  }
  *out = m->d_binfo[key].as<float2>();
  return MC_OK;
}

// Tile records of t for the binfo table of (mode, flags); rebuilt when the
// table was (a new dark, other flags).  The uint16 layout's staged entries
// carry the pixel darks of the current dark.  Synchronous like the table
// itself: records and entries are read by kernels on other streams (the
// frame server).
static int ensure_recs(Map *m, hipStream_t s, Tiles *t, int mode, uint32_t flags,
                       const float2 *binfo) {
  const int key = mode * 4 + (int)flags;
  int rc;
  if (t->in_size == 2 && t->ents_dark != m->dark_version) {
    if ((rc = launch_ents_dark(s, *m, *t))) return rc;
    MC_HIP(hipStreamSynchronize(s));
    t->ents_dark = m->dark_version;
  }
  if (t->qrecs.p && t->recs_binfo == binfo && t->recs_version == m->binfo_version[key])
    return MC_OK;
  if ((rc = launch_quad_recs(s, *m, *t, binfo))) return rc;
  MC_HIP(hipStreamSynchronize(s));
  t->recs_binfo = binfo;
  t->recs_version = m->binfo_version[key];
  return MC_OK;
}

// Element size of the pixel-major batch for a frame dtype.
static int batch_elem(int dtype) { return dtype_size(dtype) <= 2 ? 2 : 4; }

// Frames per batched pass: the layout's pass width.
static int pass_frames(const Tiles &t) { return t.width; }

static int ensure_xt(DBuf &xt, const Map &m, int dtype) {
  const size_t need = batch_xt_bytes(m, dtype);
  return xt.bytes >= need ? MC_OK : xt.alloc(need);
}

// Scratch of the one-frame kernels with a dark: up to 8 frames of FP32
// dark-subtracted pixels (a larger buffer, e.g. a batched pass's, holds
// more).  Frames per chunk in *frames_out.
static int ensure_sub(DBuf &xt, const Map &m, int n_frames, int64_t *frames_out) {
  const size_t per = (size_t)m.npx() * sizeof(float);
  *frames_out = 0;
  if (!m.has_dark) return MC_OK;
  const size_t want = per * (size_t)std::min(n_frames, 8);
  if (xt.bytes < want) {
    int rc = xt.alloc(want);
    if (rc) return rc;
  }
  *frames_out = (int64_t)(xt.bytes / per);
  return MC_OK;
}

static int ensure_part(DBuf &part, const Tiles &t) {
  const size_t need = t.part_bytes();
  return part.bytes >= need ? MC_OK : part.alloc(need);
}

int ensure_work(Map *m, int mode, int dtype, Work *work, int width) {
  Tiles *t = nullptr;
  int rc;
  if ((rc = ensure_tiles(m, mode, batch_elem(dtype), &t, width)) ||
      (rc = ensure_xt(work->xt, *m, dtype)) || (rc = ensure_part(work->part, *t)))
    return rc;
  return MC_OK;
}

int prepare_batched(Map *m, int mode, uint32_t flags, int dtype, int width) {
  const float2 *binfo = nullptr;
  Tiles *t = nullptr;
  int rc;
  if ((rc = ensure_binfo(m, mode, flags, &binfo)) ||
      (rc = ensure_tiles(m, mode, batch_elem(dtype), &t, width)))
    return rc;
  return ensure_recs(m, m->ctx->stream, t, mode, flags, binfo);
}

int run_integrate(Map *m, hipStream_t s, const void *frames_dev, int dtype, int n_frames,
                  int mode, uint32_t flags, float *cakes_dev, float *lineouts_dev,
                  bool force_plain, bool force_batched, Work *work, int width) {
  int rc;
  const float2 *binfo = nullptr;
  if ((rc = ensure_binfo(m, mode, flags, &binfo))) return rc;
  const int mask_nan = (flags & MC_F_MASK_NAN) ? 1 : 0;
  const size_t es = dtype_size(dtype);
  MC_CHECK(es > 0, MC_ERR_ARG, "unsupported dtype %d", dtype);
  Work &ws = work ? *work : m->work;
  if ((n_frames == 1 && !force_batched) || force_plain) {
    int64_t sub_frames = 0;
    if ((rc = ensure_sub(ws.xt, *m, n_frames, &sub_frames))) return rc;
    return launch_plain(s, *m, m->csr[mode], binfo, mask_nan, frames_dev, dtype, cakes_dev,
                        lineouts_dev, n_frames, m->npx(), ws.xt.as<float>(), sub_frames);
  }
  Tiles *t = nullptr;
  if ((rc = ensure_tiles(m, mode, batch_elem(dtype), &t, width))) return rc;
  if ((rc = ensure_recs(m, s, t, mode, flags, binfo))) return rc;
  if ((rc = ensure_xt(ws.xt, *m, dtype)) || (rc = ensure_part(ws.part, *t))) return rc;
  const int step = pass_frames(*t);
  for (int f0 = 0; f0 < n_frames; f0 += step) {
    const int nf = std::min(step, n_frames - f0);
    const char *fb = (const char *)frames_dev + (size_t)f0 * m->npx() * es;
    float *cb = cakes_dev + (size_t)f0 * m->nb();
    float *lb = lineouts_dev ? lineouts_dev + (size_t)f0 * 3 * m->n_r : nullptr;
    if ((rc = launch_batched(s, *m, *t, binfo, mask_nan, fb, dtype, nf, cb, lb, ws.xt.p,
                             ws.part.as<float>())))
      return rc;
  }
  return MC_OK;
}

}  // namespace mc

using namespace mc;

// ====================================================================
// C ABI
// ====================================================================
extern "C" {

int mc_version(int *major, int *minor, int *patch) {
  if (major) *major = 0;
  if (minor) *minor = 1;
  if (patch) *patch = 0;
  return MC_OK;
}

const char *mc_last_error(void) { return g_err.c_str(); }

int mc_device_count(int *n) {
  MC_CHECK(n, MC_ERR_ARG, "null");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) c = 0;
  *n = c;
  return MC_OK;
}

int mc_ctx_create(int device, mc_ctx **out) {
  MC_CHECK(out, MC_ERR_ARG, "null out");
  int n = 0;
  MC_HIP(hipGetDeviceCount(&n));
  MC_CHECK(device >= 0 && device < n, MC_ERR_ARG, "device %d not present (%d devices)",
           device, n);
  MC_HIP(hipSetDevice(device));
  auto *c = new mc_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    set_error("hipStreamCreate: %s", hipGetErrorString(e));
    return MC_ERR_HIP;
  }
  *out = c;
  return MC_OK;
}

int mc_ctx_destroy(mc_ctx *ctx) {
  if (!ctx) return MC_OK;
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return MC_OK;
}

int mc_ctx_sync(mc_ctx *ctx) {
  MC_CHECK(ctx, MC_ERR_ARG, "null ctx");
  MC_HIP(hipStreamSynchronize(ctx->stream));
  return MC_OK;
}

void *mc_ctx_stream(mc_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

struct mc_event {
  hipEvent_t e = nullptr;
};

int mc_event_record(mc_ctx *ctx, mc_event **out) {
  MC_CHECK(ctx && out, MC_ERR_ARG, "null argument");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  MC_HIP(hipSetDevice(ctx->device));
  auto *ev = new mc_event();
  hipError_t e = hipEventCreateWithFlags(&ev->e, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(ev->e, ctx->stream);
  if (e != hipSuccess) {
    if (ev->e) (void)hipEventDestroy(ev->e);
    delete ev;
    set_error("event record: %s", hipGetErrorString(e));
    return MC_ERR_HIP;
  }
  *out = ev;
  return MC_OK;
}

int mc_event_query(mc_event *ev, int32_t *done) {
  MC_CHECK(ev && done, MC_ERR_ARG, "null argument");
  const hipError_t e = hipEventQuery(ev->e);
  if (e == hipErrorNotReady) {
    *done = 0;
    return MC_OK;
  }
  MC_HIP(e);
  *done = 1;
  return MC_OK;
}

int mc_event_wait(mc_event *ev) {
  MC_CHECK(ev, MC_ERR_ARG, "null event");
  MC_HIP(hipEventSynchronize(ev->e));
  return MC_OK;
}

int mc_event_destroy(mc_event *ev) {
  if (!ev) return MC_OK;
  (void)hipEventDestroy(ev->e);
  delete ev;
  return MC_OK;
}

namespace {
struct HostCb {
  mc_host_fn fn;
  void *user;
};
void run_host_cb(void *p) {
  std::unique_ptr<HostCb> cb(static_cast<HostCb *>(p));
  cb->fn(cb->user);
}
}  // namespace

int mc_ctx_callback(mc_ctx *ctx, mc_host_fn fn, void *user) {
  MC_CHECK(ctx && fn, MC_ERR_ARG, "null argument");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  MC_HIP(hipSetDevice(ctx->device));
  auto *cb = new HostCb{fn, user};
  const hipError_t e = hipLaunchHostFunc(ctx->stream, run_host_cb, cb);
  if (e != hipSuccess) {
    delete cb;
    set_error("hipLaunchHostFunc: %s", hipGetErrorString(e));
    return MC_ERR_HIP;
  }
  return MC_OK;
}

int mc_dev_alloc(mc_ctx *ctx, uint64_t bytes, void **dptr) {
  MC_CHECK(ctx && dptr, MC_ERR_ARG, "null");
  MC_HIP(hipSetDevice(ctx->device));
  hipError_t e = hipMalloc(dptr, bytes ? bytes : 16);
  if (e != hipSuccess) {
    set_error("hipMalloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e));
    return MC_ERR_NOMEM;
  }
  return MC_OK;
}

int mc_dev_free(mc_ctx *ctx, void *dptr) {
  (void)ctx;
  if (dptr) MC_HIP(hipFree(dptr));
  return MC_OK;
}

int mc_memcpy_h2d(mc_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  MC_CHECK(ctx, MC_ERR_ARG, "null ctx");
  MC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  MC_HIP(hipStreamSynchronize(ctx->stream));
  return MC_OK;
}

int mc_memcpy_d2h(mc_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
  MC_CHECK(ctx, MC_ERR_ARG, "null ctx");
  MC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  MC_HIP(hipStreamSynchronize(ctx->stream));
  return MC_OK;
}

int mc_map_build(mc_ctx *ctx, const mc_geometry *g, const double *mask, const double *distY,
                 const double *distZ, const double *residual, const mc_panel *panels,
                 int32_t n_panels, mc_map **out) {
  MC_CHECK(ctx && g && out, MC_ERR_ARG, "null argument");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  MC_HIP(hipSetDevice(ctx->device));
  auto *m = new mc_map();
  m->ctx = ctx;
  m->bc_y = g->yCen;
  m->bc_z = g->zCen;
  int rc = map_build_gpu(ctx, g, mask, distY, distZ, residual, panels, n_panels, m);
  if (!rc) rc = map_finish_host(m);
  if (rc) {
    delete m;
    return rc;
  }
  *out = m;
  return MC_OK;
}

static int read_file(const char *path, std::vector<char> &buf) {
  FILE *f = fopen(path, "rb");
  MC_CHECK(f, MC_ERR_IO, "cannot open %s", path);
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  buf.resize(n);
  size_t got = n > 0 ? fread(buf.data(), 1, n, f) : 0;
  fclose(f);
  MC_CHECK((long)got == n, MC_ERR_IO, "short read of %s", path);
  return MC_OK;
}

static size_t header_offset(const std::vector<char> &b) {
  if (b.size() >= 64) {
    uint32_t magic;
    memcpy(&magic, b.data(), 4);
    if (magic == 0x3050414Du) return 64;  // "MAP0", MapHeader.h:40
  }
  return 0;
}

int mc_map_from_host(mc_ctx *ctx, const void *records, int64_t nnz, const int32_t *pairs,
                     const int32_t *flags, int32_t ny, int32_t nz, int32_t n_r, int32_t n_eta,
                     mc_map **out) {
  MC_CHECK(ctx && pairs && out && (records || nnz == 0), MC_ERR_ARG, "null argument");
  MC_CHECK(ny > 1 && nz > 1 && n_r > 0 && n_eta > 0 && nnz >= 0, MC_ERR_ARG,
           "bad sizes ny=%d nz=%d n_r=%d n_eta=%d", ny, nz, n_r, n_eta);
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  MC_HIP(hipSetDevice(ctx->device));
  auto *m = new mc_map();
  m->ctx = ctx;
  m->ny = ny;
  m->nz = nz;
  m->n_r = n_r;
  m->n_eta = n_eta;
  m->nnz = nnz;
  m->bc_y = 0;
  m->bc_z = 0;
  const int64_t nb = (int64_t)n_r * n_eta;
  m->h_rec.resize(nnz);
  if (nnz) memcpy(m->h_rec.data(), records, nnz * sizeof(Rec));
  m->h_pairs.assign(pairs, pairs + 2 * nb);
  if (flags) m->h_flags.assign(flags, flags + nb);
  int rc = map_finish_host(m);
  if (rc) {
    delete m;
    return rc;
  }
  *out = m;
  return MC_OK;
}

int mc_map_load(mc_ctx *ctx, const char *map_path, const char *nmap_path,
                const char *maskmap_path, int32_t ny, int32_t nz, int32_t n_r, int32_t n_eta,
                mc_map **out) {
  MC_CHECK(ctx && map_path && nmap_path && out, MC_ERR_ARG, "null argument");
  std::vector<char> a, b, c;
  int rc;
  if ((rc = read_file(map_path, a)) || (rc = read_file(nmap_path, b))) return rc;
  size_t oa = header_offset(a), ob = header_offset(b);
  MC_CHECK((a.size() - oa) % 24 == 0, MC_ERR_FORMAT,
           "%s: %zu data bytes is not a multiple of the 24-byte record (legacy map?)",
           map_path, a.size() - oa);
  const int64_t nb = (int64_t)n_r * n_eta;
  MC_CHECK((int64_t)(b.size() - ob) == nb * 8, MC_ERR_FORMAT,
           "%s holds %zu pairs, expected %lld bins", nmap_path, (b.size() - ob) / 8,
           (long long)nb);
  const int32_t *flags = nullptr;
  if (maskmap_path && maskmap_path[0]) {
    FILE *f = fopen(maskmap_path, "rb");
    if (f) {
      fclose(f);
      if ((rc = read_file(maskmap_path, c))) return rc;
      size_t oc = header_offset(c);
      MC_CHECK((int64_t)(c.size() - oc) == nb * 4, MC_ERR_FORMAT,
               "%s holds %zu flags, expected %lld", maskmap_path, (c.size() - oc) / 4,
               (long long)nb);
      flags = (const int32_t *)(c.data() + oc);
    }
  }
  rc = mc_map_from_host(ctx, a.data() + oa, (int64_t)((a.size() - oa) / 24),
                        (const int32_t *)(b.data() + ob), flags, ny, nz, n_r, n_eta, out);
  if (!rc && oa == 64) (*out)->q_mode = (unsigned char)a[40];
  return rc;
}

int mc_map_write(mc_map *m, const char *dir, const uint8_t *header64) {
  MC_CHECK(m && dir, MC_ERR_ARG, "null argument");
  MC_CHECK(!m->fused, MC_ERR_ARG, "a fused map has no Map.bin records to write");
  std::string d(dir);
  auto wr = [&](const std::string &name, const void *p, size_t n) -> int {
    std::string path = d + "/" + name;
    FILE *f = fopen(path.c_str(), "wb");
    MC_CHECK(f, MC_ERR_IO, "cannot create %s", path.c_str());
    if (header64) fwrite(header64, 1, 64, f);
    size_t w = n ? fwrite(p, 1, n, f) : 0;
    fclose(f);
    MC_CHECK(w == n, MC_ERR_IO, "short write to %s", path.c_str());
    return MC_OK;
  };
  int rc;
  if ((rc = wr("Map.bin", m->h_rec.data(), m->nnz * sizeof(Rec)))) return rc;
  if ((rc = wr("nMap.bin", m->h_pairs.data(), m->h_pairs.size() * 4))) return rc;
  bool any = false;
  for (int32_t f : m->h_flags) any |= f != 0;
  if (any && (rc = wr("maskMap.bin", m->h_flags.data(), m->h_flags.size() * 4))) return rc;
  return MC_OK;
}

int mc_ctx_set_map_verify(mc_ctx *ctx, int64_t n_pixels) {
  MC_CHECK(ctx && n_pixels >= 0, MC_ERR_ARG, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  ctx->map_verify = n_pixels;
  return MC_OK;
}

int mc_map_verify_info(mc_map *m, int64_t *out) {
  MC_CHECK(m && out, MC_ERR_ARG, "null argument");
  for (int k = 0; k < 4; k++) out[k] = m->verify[k];
  return MC_OK;
}

int mc_map_info_get(mc_map *m, mc_map_info *info) {
  MC_CHECK(m && info, MC_ERR_ARG, "null argument");
  memset(info, 0, sizeof(*info));
  info->nnz = m->nnz;
  info->n_r = m->n_r;
  info->n_eta = m->n_eta;
  info->ny = m->ny;
  info->nz = m->nz;
  int nf = 0;
  for (int32_t f : m->h_flags) nf += f != 0;
  info->n_flagged = nf;
  info->q_mode = m->q_mode;
  int64_t hb = m->d_rec.bytes + m->d_pairs.bytes + m->d_flags.bytes + m->d_area_inb.bytes +
               m->d_row_static.bytes + m->inv_area.bytes + m->d_dark.bytes +
               m->d_dint.bytes + m->work.xt.bytes + m->work.part.bytes;
  for (const DBuf &b : m->d_binfo) hb += b.bytes;
  for (int q = 0; q < 3; q++) {
    const Csr &c = m->csr[q];
    hb += c.row.bytes + c.pix.bytes + c.w.bytes + c.dterm.bytes + c.d2term.bytes;
    for (auto &t : m->tiles[q])
      hb += t->chunks.bytes + t->tiles.bytes + t->qs.bytes + t->qdir.bytes + t->ents.bytes +
            t->band_chunk.bytes + t->qrecs.bytes + t->tpix.bytes + t->blocks64.bytes +
            t->blocks128.bytes;
  }
  info->hbm_bytes = hb;
  for (int q = 0; q < 3; q++) info->skipped[q] = m->skipped[q];
  return MC_OK;
}

int mc_map_download(mc_map *m, void *records, int32_t *pairs, int32_t *flags) {
  MC_CHECK(m, MC_ERR_ARG, "null map");
  MC_CHECK(!m->fused, MC_ERR_ARG, "a fused map has no Map.bin records");
  if (records) memcpy(records, m->h_rec.data(), m->nnz * sizeof(Rec));
  if (pairs) memcpy(pairs, m->h_pairs.data(), m->h_pairs.size() * 4);
  if (flags) memcpy(flags, m->h_flags.data(), m->h_flags.size() * 4);
  return MC_OK;
}

static int reta_rows(mc_map *m, const double *r_lo, const double *r_hi, const double *eta_lo,
                     const double *eta_hi, double px, double lsd, double lambda,
                     const std::vector<double> &area, double *out5) {
  MC_CHECK(m && r_lo && r_hi && eta_lo && eta_hi && out5, MC_ERR_ARG, "null argument");
  MC_CHECK((int64_t)area.size() == m->nb(), MC_ERR_STATE, "map has no per-bin areas");
  const int64_t nb = m->nb();
  for (int r = 0; r < m->n_r; r++) {
    const double Rm = (r_lo[r] + r_hi[r]) / 2;
    const double tth = atan(Rm * px / lsd);
    for (int e = 0; e < m->n_eta; e++) {
      const int64_t b = (int64_t)r * m->n_eta + e;
      out5[b] = Rm;
      out5[nb + b] = (180.0 / M_PI) * tth;
      out5[2 * nb + b] = (eta_lo[e] + eta_hi[e]) / 2;
      out5[3 * nb + b] = area[b];
      out5[4 * nb + b] = lambda > 0 ? (4.0 * M_PI / lambda) * sin(tth / 2.0) : 0.0;
    }
  }
  return MC_OK;
}

int mc_map_reta(mc_map *m, const double *r_lo, const double *r_hi, const double *eta_lo,
                const double *eta_hi, double px, double lsd, double lambda, double *out5) {
  MC_CHECK(m, MC_ERR_ARG, "null map");
  return reta_rows(m, r_lo, r_hi, eta_lo, eta_hi, px, lsd, lambda, m->h_area_all, out5);
}

int mc_map_reta_stream(mc_map *m, const double *r_lo, const double *r_hi,
                       const double *eta_lo, const double *eta_hi, double px, double lsd,
                       double lambda, double *out5) {
  MC_CHECK(m, MC_ERR_ARG, "null map");
  return reta_rows(m, r_lo, r_hi, eta_lo, eta_hi, px, lsd, lambda, m->h_area_inb, out5);
}

// Linear interpolation of the source R axis at r (fused_csr.py:64-99):
// (j, w_lo, w_hi) with r = w_lo * src[j] + w_hi * src[j + 1], clamped to the
// first / last pair.
static void r_interp(const double *src, int n, double r, int *j, double *wl, double *wh) {
  if (r <= src[0]) {
    *j = 0;
    *wl = 1.0;
    *wh = 0.0;
  } else if (r >= src[n - 1]) {
    *j = n - 2;
    *wl = 0.0;
    *wh = 1.0;
  } else {
    int k = (int)(std::upper_bound(src, src + n, r) - src) - 1;
    k = std::max(0, std::min(k, n - 2));
    const double den = src[k + 1] - src[k];
    const double t = den > 0 ? (r - src[k]) / den : 0.0;
    *j = k;
    *wl = 1.0 - t;
    *wh = t;
  }
}

int mc_map_fuse(mc_map *const *maps, int32_t n_maps, const double *r_axes, int32_t mode,
                mc_map **out) {
  MC_CHECK(maps && r_axes && out && n_maps >= 1, MC_ERR_ARG, "bad arguments");
  MC_CHECK(mode == MC_MODE_BILINEAR || mode == MC_MODE_GRADIENT, MC_ERR_ARG,
           "mode must be bilinear or gradient, got %d", mode);
  Map *m0 = maps[0];
  MC_CHECK(m0, MC_ERR_ARG, "null map");
  for (int k = 0; k < n_maps; k++) {
    const Map *mk = maps[k];
    MC_CHECK(mk && mk->ctx == m0->ctx && mk->ny == m0->ny && mk->nz == m0->nz &&
                 mk->n_eta == m0->n_eta && !mk->fused,
             MC_ERR_ARG, "map %d does not match map 0 (context, detector, eta bins) or is fused",
             k);
    MC_CHECK(mk->n_r >= 2, MC_ERR_ARG, "R interpolation needs at least 2 R bins (map %d)", k);
  }
  const int nR = m0->n_r, nE = m0->n_eta;
  std::vector<size_t> roff(n_maps + 1, 0);  // each map's R axis in r_axes
  for (int k = 0; k < n_maps; k++) roff[k + 1] = roff[k] + maps[k]->n_r;
  Ctx *ctx = m0->ctx;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  MC_HIP(hipSetDevice(ctx->device));
  int rc;
  std::vector<std::vector<double>> inv_a(n_maps);
  for (int k = 0; k < n_maps; k++) {
    std::lock_guard<std::mutex> lk2(maps[k]->mu);
    if ((rc = ensure_csr(maps[k], mode))) return rc;
    const auto &A = maps[k]->h_area_inb;
    inv_a[k].resize(A.size());
    for (size_t b = 0; b < A.size(); b++)
      inv_a[k][b] = A[b] > 1e-9 ? 1.0 / std::max(A[b], 1e-9) : 0.0;
  }
  // Per output R row: the source rows and weights of every map.
  std::vector<int> jj((size_t)n_maps * nR);
  std::vector<double> wlo((size_t)n_maps * nR), whi((size_t)n_maps * nR);
  for (int k = 0; k < n_maps; k++)
    for (int i = 0; i < nR; i++)
      r_interp(r_axes + roff[k], maps[k]->n_r, r_axes[i], &jj[(size_t)k * nR + i],
               &wlo[(size_t)k * nR + i], &whi[(size_t)k * nR + i]);
  // Row (i, e) of M = (1/N) sum_k (P_k D_k) M_k, in the order the
  // reference's sparse products take: per map, a pixel's contribution is
  // (w_lo / A_j) * w(j) + (w_hi / A_j+1) * w(j+1); maps add in order; the
  // sum is divided by N; exact zeros drop.
  const int64_t nb = (int64_t)nR * nE;
  std::vector<int32_t> cnt(nb, 0);
  std::vector<std::vector<int32_t>> rp(nb);
  std::vector<std::vector<float>> rw(nb);
  parallel_for(nb, [&](int64_t a, int64_t bq) {
    struct T {
      int32_t pix, k, side;
      double v;
    };
    std::vector<T> v;
    for (int64_t ob = a; ob < bq; ob++) {
      const int i = (int)(ob / nE), e = (int)(ob % nE);
      v.clear();
      for (int k = 0; k < n_maps; k++) {
        const Csr &c = maps[k]->csr[mode];
        const int j = jj[(size_t)k * nR + i];
        for (int side = 0; side < 2; side++) {
          const int64_t sb = (int64_t)(j + side) * nE + e;
          const double pw = side ? whi[(size_t)k * nR + i] : wlo[(size_t)k * nR + i];
          const double ia = inv_a[k][sb];
          for (int32_t q = c.h_row[sb]; q < c.h_row[sb + 1]; q++)
            v.push_back({c.h_pix[q], k, side, (pw * ia) * (double)c.h_w[q]});
        }
      }
      std::sort(v.begin(), v.end(), [](const T &x, const T &y) {
        return x.pix != y.pix ? x.pix < y.pix : (x.k != y.k ? x.k < y.k : x.side < y.side);
      });
      auto &P = rp[ob];
      auto &W = rw[ob];
      size_t q = 0;
      while (q < v.size()) {
        const int32_t pix = v[q].pix;
        double tot = 0.0;
        bool first = true;
        while (q < v.size() && v[q].pix == pix) {
          const int k = v[q].k;
          double ck = 0.0;
          while (q < v.size() && v[q].pix == pix && v[q].k == k) ck += v[q++].v;
          tot = first ? ck : tot + ck;
          first = false;
        }
        const double val = tot / n_maps;
        if (val != 0.0) {
          P.push_back(pix);
          W.push_back((float)val);
        }
      }
      cnt[ob] = (int32_t)P.size();
    }
  });
  auto *m = new mc_map();
  std::unique_ptr<mc_map> guard(m);
  m->ctx = ctx;
  m->fused = true;
  m->ny = m0->ny;
  m->nz = m0->nz;
  m->n_r = nR;
  m->n_eta = nE;
  m->nnz = 0;
  m->q_mode = m0->q_mode;
  m->bc_y = m0->bc_y;
  m->bc_z = m0->bc_z;
  m->h_flags.assign(nb, 0);
  m->h_area_all.assign(nb, 1.0);
  m->h_area_inb.assign(nb, 1.0);
  Csr &c = m->csr[MC_MODE_BILINEAR];
  c.h_row.assign(nb + 1, 0);
  for (int64_t b = 0; b < nb; b++) c.h_row[b + 1] = c.h_row[b] + cnt[b];
  const int64_t nnz = c.h_row[nb];
  c.h_pix.resize(nnz);
  c.h_w.resize(nnz);
  for (int64_t b = 0; b < nb; b++) {
    std::copy(rp[b].begin(), rp[b].end(), c.h_pix.begin() + c.h_row[b]);
    std::copy(rw[b].begin(), rw[b].end(), c.h_w.begin() + c.h_row[b]);
  }
  c.nnz = nnz;
  c.bc_y = m->bc_y;
  c.bc_z = m->bc_z;
  if ((rc = c.row.alloc((nb + 1) * 4)) || (rc = c.pix.alloc(nnz * 4)) ||
      (rc = c.w.alloc(nnz * 4)) || (rc = c.dterm.alloc(nb * 4)))
    return rc;
  MC_HIP(hipMemcpy(c.row.p, c.h_row.data(), (nb + 1) * 4, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(c.pix.p, c.h_pix.data(), nnz * 4, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(c.w.p, c.h_w.data(), nnz * 4, hipMemcpyHostToDevice));
  c.built = true;
  // The gradient slot holds the same matrix (the reference exposes the
  // fused matrix under both names).
  Csr &g = m->csr[MC_MODE_GRADIENT];
  g.h_row = c.h_row;
  g.h_pix = c.h_pix;
  g.h_w = c.h_w;
  g.nnz = nnz;
  g.bc_y = c.bc_y;
  g.bc_z = c.bc_z;
  if ((rc = g.row.alloc((nb + 1) * 4)) || (rc = g.pix.alloc(nnz * 4)) ||
      (rc = g.w.alloc(nnz * 4)) || (rc = g.dterm.alloc(nb * 4)))
    return rc;
  MC_HIP(hipMemcpy(g.row.p, c.row.p, (nb + 1) * 4, hipMemcpyDeviceToDevice));
  MC_HIP(hipMemcpy(g.pix.p, c.pix.p, nnz * 4, hipMemcpyDeviceToDevice));
  MC_HIP(hipMemcpy(g.w.p, c.w.p, nnz * 4, hipMemcpyDeviceToDevice));
  g.built = true;
  // Unit areas: scale 1, every bin valid for the lineouts.
  std::vector<float> ones(nb, 1.f);
  std::vector<float4> rs(nR, make_float4((float)nE, (float)nE, (float)nE, (float)nE));
  std::vector<int32_t> pairs0(2 * nb, 0);
  if ((rc = m->d_rec.alloc(16)) || (rc = m->d_pairs.alloc(2 * nb * 4)) ||
      (rc = m->d_flags.alloc(nb * 4)) || (rc = m->d_area_inb.alloc(nb * 4)) ||
      (rc = m->d_row_static.alloc(nR * sizeof(float4))) || (rc = m->inv_area.alloc(nb * 4)))
    return rc;
  MC_HIP(hipMemcpy(m->d_pairs.p, pairs0.data(), 2 * nb * 4, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(m->d_flags.p, m->h_flags.data(), nb * 4, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(m->d_area_inb.p, ones.data(), nb * 4, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(m->inv_area.p, ones.data(), nb * 4, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(m->d_row_static.p, rs.data(), nR * sizeof(float4), hipMemcpyHostToDevice));
  m->h_pairs = std::move(pairs0);
  *out = guard.release();
  return MC_OK;
}

int mc_map_destroy(mc_map *m) {
  if (!m) return MC_OK;
  {
    std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
    (void)hipStreamSynchronize(m->ctx->stream);
  }
  delete m;
  return MC_OK;
}

int mc_map_set_dark(mc_map *m, const void *dark, int32_t dtype, int32_t n_frames, double scale,
                    int32_t on_device) {
  MC_CHECK(m, MC_ERR_ARG, "null map");
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  std::lock_guard<std::mutex> lk2(m->mu);
  MC_HIP(hipSetDevice(m->ctx->device));
  for (auto &c : m->csr) c.dterm_valid = c.d2_valid = false;
  for (bool &ok : m->binfo_ok) ok = false;
  m->dark_version++;
  if (!dark) {
    m->has_dark = false;
    return MC_OK;
  }
  MC_CHECK(n_frames >= 1, MC_ERR_ARG, "n_frames %d", n_frames);
  const size_t es = dtype_size(dtype);
  MC_CHECK(es > 0, MC_ERR_ARG, "unsupported dark dtype %d", dtype);
  int rc;
  m->has_dark = false;
  if ((rc = m->d_dark.alloc(m->npx() * 8)) || (rc = m->d_dint.alloc(m->npx() * 2))) return rc;
  const void *src = dark;
  DBuf tmp;
  const size_t bytes = (size_t)n_frames * m->npx() * es;
  if (!on_device) {
    if ((rc = tmp.alloc(bytes))) return rc;
    MC_HIP(hipMemcpyAsync(tmp.p, dark, bytes, hipMemcpyHostToDevice, m->ctx->stream));
    src = tmp.p;
  }
  if ((rc = launch_dark_mean(m->ctx->stream, src, dtype, n_frames, m->npx(), scale,
                             m->d_dark.as<double>())) ||
      (rc = launch_dark_int(m->ctx->stream, m->d_dark.as<double>(), m->npx(),
                            m->d_dint.as<uint16_t>())))
    return rc;
  MC_HIP(hipStreamSynchronize(m->ctx->stream));
  m->has_dark = true;
  return MC_OK;
}

int mc_map_set_beam_center(mc_map *m, double bc_y, double bc_z) {
  MC_CHECK(m, MC_ERR_ARG, "null map");
  std::lock_guard<std::mutex> lk(m->mu);
  m->bc_y = bc_y;
  m->bc_z = bc_z;
  return MC_OK;
}

int mc_map_set_tuning(mc_map *m, int32_t batch, int32_t rows_per_group, int32_t lds_bytes) {
  MC_CHECK(m, MC_ERR_ARG, "null map");
  std::lock_guard<std::mutex> lk(m->mu);
  if (batch) {
    MC_CHECK(batch == 8 || batch == 16 || batch == 32 || batch == 64 || batch == 128,
             MC_ERR_ARG, "batch %d not in {8,16,32,64,128}", batch);
    m->batch = batch;
  }
  if (rows_per_group) {
    MC_CHECK(rows_per_group == 16 || rows_per_group == 32, MC_ERR_ARG,
             "rows per band %d not in {16, 32}", rows_per_group);
    m->rows_per_group = rows_per_group;
  }
  if (lds_bytes) {
    MC_CHECK(lds_bytes >= 8192 && lds_bytes <= 160 * 1024, MC_ERR_ARG, "lds_bytes %d",
             lds_bytes);
    m->lds_bytes = lds_bytes;
  }
  return MC_OK;
}

int mc_map_set_tile_shape(mc_map *m, int32_t bins, int32_t entries) {
  MC_CHECK(m, MC_ERR_ARG, "null map");
  std::lock_guard<std::mutex> lk(m->mu);
  if (bins) {
    MC_CHECK(bins >= 32 && bins <= 1024 && bins % 16 == 0, MC_ERR_ARG,
             "bins per tile %d not a multiple of 16 in 32..1024", bins);
    m->tile_bins = bins;
  }
  if (entries) {
    MC_CHECK(entries >= 256 && entries <= 16384 && entries % 64 == 0, MC_ERR_ARG,
             "entries per tile %d not a multiple of 64 in 256..16384", entries);
    m->tile_ents = entries;
  }
  return MC_OK;
}

int mc_map_get_tile_shape(mc_map *m, int32_t *bins, int32_t *entries) {
  MC_CHECK(m, MC_ERR_ARG, "null map");
  if (bins) *bins = m->tile_bins;
  if (entries) *entries = m->tile_ents;
  return MC_OK;
}

int mc_map_set_plain_lanes(mc_map *m, int32_t lanes) {
  MC_CHECK(m, MC_ERR_ARG, "null map");
  MC_CHECK(lanes == 4 || lanes == 8 || lanes == 16, MC_ERR_ARG,
           "lanes per bin %d not in {4, 8, 16}", lanes);
  std::lock_guard<std::mutex> lk(m->mu);
  m->plain_lanes = lanes;
  return MC_OK;
}

int mc_map_get_plain_lanes(mc_map *m, int32_t *lanes) {
  MC_CHECK(m && lanes, MC_ERR_ARG, "null argument");
  *lanes = m->plain_lanes;
  return MC_OK;
}

int mc_map_get_tuning(mc_map *m, int32_t *batch, int32_t *rows_per_group, int32_t *lds_bytes) {
  MC_CHECK(m, MC_ERR_ARG, "null map");
  if (batch) *batch = m->batch;
  if (rows_per_group) *rows_per_group = m->rows_per_group;
  if (lds_bytes) *lds_bytes = m->lds_bytes;
  return MC_OK;
}

int mc_map_layout_info(mc_map *m, int32_t mode, int32_t dtype, int64_t *out) {
  MC_CHECK(m && out, MC_ERR_ARG, "null argument");
  MC_CHECK(mode >= 0 && mode <= 2, MC_ERR_ARG, "mode %d", mode);
  const size_t es = dtype_size(dtype);
  MC_CHECK(es > 0, MC_ERR_ARG, "unsupported dtype %d", dtype);
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  std::lock_guard<std::mutex> lk2(m->mu);
  MC_HIP(hipSetDevice(m->ctx->device));
  Tiles *t = nullptr;
  int rc = ensure_tiles(m, mode, batch_elem(dtype), &t);
  if (rc) return rc;
  const int64_t v[9] = {t->n_tiles, t->n_chunks,  t->n_direct, t->staged_px, t->n_ents,
                        (int64_t)t->lds_needed, t->n_bands, t->in_size, t->n_blocks64};
  for (int i = 0; i < 9; i++) out[i] = v[i];
  return MC_OK;
}

// Hybrid frames: F messages back to back (u32 count, u16 base[npx], i32
// index[count], i64 value[count]); the counts are read on the host (copied
// from the device for device messages), the frames decoded on the GPU to
// int32 and integrated as such.
static int integrate_hybrid(mc_map *m, const void *msgs, int32_t n_frames, int32_t on_device,
                            int32_t mode, uint32_t flags, float *cd, float *ld) {
  hipStream_t s = m->ctx->stream;
  const int64_t npx = m->npx();
  std::vector<int64_t> off(n_frames + 1, 0), cnt(n_frames, 0);
  for (int f = 0; f < n_frames; f++) {
    uint32_t c = 0;
    const char *p = (const char *)msgs + off[f];
    if (on_device) MC_HIP(hipMemcpy(&c, p, 4, hipMemcpyDeviceToHost));
    else memcpy(&c, p, 4);
    MC_CHECK((int64_t)c <= npx, MC_ERR_FORMAT, "hybrid frame %d: %u overflow patches for %lld "
             "pixels", f, c, (long long)npx);
    cnt[f] = c;
    off[f + 1] = off[f] + 4 + 2 * npx + 12 * (int64_t)c;
  }
  DBuf dmsg, dec;
  int rc;
  const char *dm = (const char *)msgs;
  if (!on_device) {
    if ((rc = dmsg.alloc(off[n_frames]))) return rc;
    MC_HIP(hipMemcpyAsync(dmsg.p, msgs, off[n_frames], hipMemcpyHostToDevice, s));
    dm = dmsg.as<char>();
  }
  if ((rc = dec.alloc((size_t)n_frames * npx * 4))) return rc;
  for (int f = 0; f < n_frames; f++)
    if ((rc = launch_hybrid_decode(s, dm + off[f], npx, cnt[f], dec.as<int32_t>() + f * npx)))
      return rc;
  return run_integrate(m, s, dec.p, MC_DTYPE_I32, n_frames, mode, flags, cd, ld, false, false);
}

int mc_integrate(mc_map *m, const void *frames, int32_t dtype, int32_t n_frames,
                 int32_t frames_on_device, int32_t mode, uint32_t flags, float *cakes,
                 int32_t cakes_on_device, float *lineouts, int32_t lineouts_on_device) {
  MC_CHECK(m && frames && cakes && n_frames >= 1, MC_ERR_ARG, "bad arguments");
  MC_CHECK(mode >= 0 && mode <= 2, MC_ERR_ARG, "mode %d", mode);
  const bool hybrid = dtype == MC_DTYPE_HYBRID;
  const size_t es = dtype_size(dtype);
  MC_CHECK(es > 0 || hybrid, MC_ERR_ARG, "unsupported frame dtype %d", dtype);
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  std::lock_guard<std::mutex> lk2(m->mu);
  MC_HIP(hipSetDevice(m->ctx->device));
  hipStream_t s = m->ctx->stream;
  const size_t fbytes = (size_t)n_frames * m->npx() * es;
  const size_t cbytes = (size_t)n_frames * m->nb() * 4;
  const size_t lbytes = (size_t)n_frames * 3 * m->n_r * 4;
  DBuf df, dc, dl;
  int rc;
  const void *fd = frames;
  if (!frames_on_device && !hybrid) {
    if ((rc = df.alloc(fbytes))) return rc;
    MC_HIP(hipMemcpyAsync(df.p, frames, fbytes, hipMemcpyHostToDevice, s));
    fd = df.p;
  }
  float *cd = cakes;
  if (!cakes_on_device) {
    if ((rc = dc.alloc(cbytes))) return rc;
    cd = dc.as<float>();
  }
  float *ld = lineouts;
  if (lineouts && !lineouts_on_device) {
    if ((rc = dl.alloc(lbytes))) return rc;
    ld = dl.as<float>();
  }
  rc = hybrid ? integrate_hybrid(m, frames, n_frames, frames_on_device, mode, flags, cd, ld)
              : run_integrate(m, s, fd, dtype, n_frames, mode, flags, cd, ld, false, false);
  if (rc) return rc;
  if (!cakes_on_device) MC_HIP(hipMemcpyAsync(cakes, cd, cbytes, hipMemcpyDeviceToHost, s));
  if (lineouts && !lineouts_on_device)
    MC_HIP(hipMemcpyAsync(lineouts, ld, lbytes, hipMemcpyDeviceToHost, s));
  MC_HIP(hipStreamSynchronize(s));
  return MC_OK;
}

int mc_integrate_async(mc_map *m, const void *frames_dev, int32_t dtype, int32_t n_frames,
                       int32_t mode, uint32_t flags, float *cakes_dev, float *lineouts_dev) {
  MC_CHECK(m && frames_dev && cakes_dev && n_frames >= 1, MC_ERR_ARG, "bad arguments");
  MC_CHECK(mode >= 0 && mode <= 2, MC_ERR_ARG, "mode %d", mode);
  MC_CHECK(dtype_size(dtype) > 0, MC_ERR_ARG,
           "unsupported frame dtype %d (hybrid frames: mc_integrate)", dtype);
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  std::lock_guard<std::mutex> lk2(m->mu);
  MC_HIP(hipSetDevice(m->ctx->device));
  return run_integrate(m, m->ctx->stream, frames_dev, dtype, n_frames, mode, flags, cakes_dev,
                       lineouts_dev, false, false);
}

// Frames from which mc_integrate_variance sums Poisson variances in
// batched passes (fewer: the one-frame kernel per frame).
constexpr int kVarBatchMin = 3;

int mc_integrate_variance(mc_map *m, const void *frames, int32_t dtype, int32_t n_frames,
                          int32_t frames_on_device, const float *variance,
                          int32_t variance_on_device, int32_t mode, uint32_t flags,
                          float *cakes, float *variances, int32_t out_on_device) {
  MC_CHECK(m && frames && cakes && variances && n_frames >= 1, MC_ERR_ARG, "bad arguments");
  MC_CHECK(mode >= 0 && mode <= 2, MC_ERR_ARG, "mode %d", mode);
  const size_t es = dtype_size(dtype);
  MC_CHECK(es > 0 && dtype != MC_DTYPE_HYBRID, MC_ERR_ARG, "unsupported frame dtype %d", dtype);
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  std::lock_guard<std::mutex> lk2(m->mu);
  MC_HIP(hipSetDevice(m->ctx->device));
  hipStream_t s = m->ctx->stream;
  const float2 *binfo = nullptr;
  int rc = ensure_binfo(m, mode, flags, &binfo);
  if (rc) return rc;
  Csr &c = m->csr[mode];
  const float *d2 = nullptr;
  if (!variance && m->has_dark) {
    if (!c.d2_valid) {
      if (!c.d2term.p && (rc = c.d2term.alloc(m->nb() * 4))) return rc;
      if ((rc = launch_dark_term2(s, c.row.as<int32_t>(), c.pix.as<int32_t>(), c.w.as<float>(),
                                  m->d_dark.as<double>(), m->d_dint.as<uint16_t>(), m->nb(),
                                  c.d2term.as<float>())))
        return rc;
      c.d2_valid = true;
    }
    d2 = c.d2term.as<float>();
  }
  const size_t fbytes = (size_t)n_frames * m->npx() * es;
  const size_t vbytes = (size_t)n_frames * m->npx() * 4;
  const size_t cbytes = (size_t)n_frames * m->nb() * 4;
  DBuf df, dv, dc, dvar;
  const void *fd = frames;
  if (!frames_on_device) {
    if ((rc = df.alloc(fbytes))) return rc;
    MC_HIP(hipMemcpyAsync(df.p, frames, fbytes, hipMemcpyHostToDevice, s));
    fd = df.p;
  }
  const float *vd = variance;
  if (variance && !variance_on_device) {
    if ((rc = dv.alloc(vbytes))) return rc;
    MC_HIP(hipMemcpyAsync(dv.p, variance, vbytes, hipMemcpyHostToDevice, s));
    vd = dv.as<float>();
  }
  float *cd = cakes, *vard = variances;
  if (!out_on_device) {
    if ((rc = dc.alloc(cbytes)) || (rc = dvar.alloc(cbytes))) return rc;
    cd = dc.as<float>();
    vard = dvar.as<float>();
  }
  if (!variance && n_frames >= kVarBatchMin && m->rows_per_group == 16) {
    // Poisson statistics over 3 frames or more: the batched kernel sums
    // w x and w^2 x from the same staged pixel rows, 32 frames a pass.
    Tiles *t = nullptr;
    if ((rc = ensure_tiles(m, mode, batch_elem(dtype), &t, 32)) ||
        (rc = ensure_recs(m, s, t, mode, flags, binfo)) ||
        (rc = ensure_xt(m->work.xt, *m, dtype)) || (rc = ensure_part(m->work.part, *t)))
      return rc;
    const int mask_nan = (flags & MC_F_MASK_NAN) ? 1 : 0;
    for (int f0 = 0; f0 < n_frames; f0 += 32) {
      const int nf = std::min(32, n_frames - f0);
      if ((rc = launch_batched(s, *m, *t, binfo, mask_nan,
                               (const char *)fd + (size_t)f0 * m->npx() * es, dtype, nf,
                               cd + (size_t)f0 * m->nb(), nullptr, m->work.xt.p,
                               m->work.part.as<float>(), vard + (size_t)f0 * m->nb(), d2)))
        return rc;
    }
  } else {
    int64_t sub_frames = 0;
    if ((rc = ensure_sub(m->work.xt, *m, n_frames, &sub_frames))) return rc;
    if ((rc = launch_plain_var(s, *m, c, binfo, d2, fd, dtype, vd, cd, vard, n_frames,
                               m->work.xt.as<float>(), sub_frames)))
      return rc;
  }
  if (!out_on_device) {
    MC_HIP(hipMemcpyAsync(cakes, cd, cbytes, hipMemcpyDeviceToHost, s));
    MC_HIP(hipMemcpyAsync(variances, vard, cbytes, hipMemcpyDeviceToHost, s));
  }
  MC_HIP(hipStreamSynchronize(s));
  return MC_OK;
}

int mc_profile(mc_map *m, const float *cakes, int32_t n_frames, int32_t cakes_on_device,
               float *lineouts, int32_t lineouts_on_device) {
  MC_CHECK(m && cakes && lineouts && n_frames >= 1, MC_ERR_ARG, "bad arguments");
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  MC_HIP(hipSetDevice(m->ctx->device));
  hipStream_t s = m->ctx->stream;
  const size_t cbytes = (size_t)n_frames * m->nb() * 4, lbytes = (size_t)n_frames * 3 * m->n_r * 4;
  DBuf dc, dl;
  int rc;
  const float *cd = cakes;
  if (!cakes_on_device) {
    if ((rc = dc.alloc(cbytes))) return rc;
    MC_HIP(hipMemcpyAsync(dc.p, cakes, cbytes, hipMemcpyHostToDevice, s));
    cd = dc.as<float>();
  }
  float *ld = lineouts;
  if (!lineouts_on_device) {
    if ((rc = dl.alloc(lbytes))) return rc;
    ld = dl.as<float>();
  }
  // NaN bins are missing values: the profile uses the unflagged row sums
  // when any NaN can be present, which is what the fused path does.
  if ((rc = launch_profile(s, *m, cd, n_frames, 1, ld))) return rc;
  if (!lineouts_on_device) MC_HIP(hipMemcpyAsync(lineouts, ld, lbytes, hipMemcpyDeviceToHost, s));
  MC_HIP(hipStreamSynchronize(s));
  return MC_OK;
}

int mc_profile_variance(mc_map *m, const float *cakes, const float *variances,
                        int32_t n_frames, int32_t in_on_device, int32_t mode, float *profile,
                        float *profile_var, int32_t out_on_device) {
  MC_CHECK(m && cakes && variances && profile && profile_var && n_frames >= 1, MC_ERR_ARG,
           "bad arguments");
  MC_CHECK(mode == MC_LINEOUT_AREA || mode == MC_LINEOUT_MEAN, MC_ERR_ARG,
           "profile mode %d: area-weighted (%d) or simple mean (%d)", mode,
           MC_LINEOUT_AREA, MC_LINEOUT_MEAN);
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  MC_HIP(hipSetDevice(m->ctx->device));
  hipStream_t s = m->ctx->stream;
  const size_t cbytes = (size_t)n_frames * m->nb() * 4, pbytes = (size_t)n_frames * m->n_r * 4;
  DBuf dc, dv, dp, dq;
  int rc;
  const float *cd = cakes, *vd = variances;
  if (!in_on_device) {
    if ((rc = dc.alloc(cbytes)) || (rc = dv.alloc(cbytes))) return rc;
    MC_HIP(hipMemcpyAsync(dc.p, cakes, cbytes, hipMemcpyHostToDevice, s));
    MC_HIP(hipMemcpyAsync(dv.p, variances, cbytes, hipMemcpyHostToDevice, s));
    cd = dc.as<float>();
    vd = dv.as<float>();
  }
  float *pd = profile, *qd = profile_var;
  if (!out_on_device) {
    if ((rc = dp.alloc(pbytes)) || (rc = dq.alloc(pbytes))) return rc;
    pd = dp.as<float>();
    qd = dq.as<float>();
  }
  if ((rc = launch_profile_var(s, *m, cd, vd, n_frames, mode == MC_LINEOUT_MEAN ? 1 : 0,
                               pd, qd)))
    return rc;
  if (!out_on_device) {
    MC_HIP(hipMemcpyAsync(profile, pd, pbytes, hipMemcpyDeviceToHost, s));
    MC_HIP(hipMemcpyAsync(profile_var, qd, pbytes, hipMemcpyDeviceToHost, s));
  }
  MC_HIP(hipStreamSynchronize(s));
  return MC_OK;
}

// ---- the batched passes of one call, captured in a HIP graph ----
struct mc_graph {
  mc_map *m = nullptr;
  Ctx *ctx = nullptr;         // outlives the map (destroy needs only the stream)
  int mode = 0, dtype = 0;
  uint32_t flags = 0;
  uint64_t tiles_id = 0;      // the tile layout the graph reads
  Tiles *t = nullptr;
  Work work;                  // the graph's own scratch (its passes write it)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int mc_graph_create(mc_map *m, const void *frames_dev, int32_t dtype, int32_t n_frames,
                    int32_t mode, uint32_t flags, float *cakes_dev, float *lineouts_dev,
                    mc_graph **out) {
  MC_CHECK(m && frames_dev && cakes_dev && out && n_frames >= 1, MC_ERR_ARG, "bad arguments");
  MC_CHECK(mode >= 0 && mode <= 2, MC_ERR_ARG, "mode %d", mode);
  MC_CHECK(dtype_size(dtype) > 0 && dtype != MC_DTYPE_HYBRID, MC_ERR_ARG,
           "unsupported frame dtype %d", dtype);
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  std::lock_guard<std::mutex> lk2(m->mu);
  MC_HIP(hipSetDevice(m->ctx->device));
  hipStream_t s = m->ctx->stream;
  int rc;
  auto *g = new mc_graph();
  std::unique_ptr<mc_graph> guard(g);
  // One eager call builds every table and buffer the passes read (nothing
  // is allocated or copied while capturing) and sets the kernels' LDS
  // attributes.
  if ((rc = run_integrate(m, s, frames_dev, dtype, n_frames, mode, flags, cakes_dev,
                          lineouts_dev, false, true, &g->work)))
    return rc;
  MC_HIP(hipStreamSynchronize(s));
  Tiles *t = nullptr;
  if ((rc = ensure_tiles(m, mode, batch_elem(dtype), &t))) return rc;
  g->m = m;
  g->ctx = m->ctx;
  g->mode = mode;
  g->dtype = dtype;
  g->flags = flags;
  g->t = t;
  g->tiles_id = t->id;
  MC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  rc = run_integrate(m, s, frames_dev, dtype, n_frames, mode, flags, cakes_dev, lineouts_dev,
                     false, true, &g->work);
  const hipError_t e = hipStreamEndCapture(s, &g->graph);
  if (rc) return rc;
  MC_CHECK(e == hipSuccess && g->graph, MC_ERR_HIP, "graph capture failed: %s",
           hipGetErrorString(e));
  MC_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  *out = guard.release();
  return MC_OK;
}

int mc_graph_launch(mc_graph *g) {
  MC_CHECK(g && g->exec, MC_ERR_ARG, "null graph");
  mc_map *m = g->m;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  std::lock_guard<std::mutex> lk2(m->mu);
  MC_HIP(hipSetDevice(m->ctx->device));
  // The graph reads the map's buffers in place: the tile layout must still
  // be the one captured (a new beam centre in gradient mode rebuilds it); a
  // new dark is folded into the same record and entry buffers before the
  // launch.  The scratch buffers are the graph's own.
  const float2 *binfo = nullptr;
  int rc;
  if ((rc = ensure_binfo(m, g->mode, g->flags, &binfo))) return rc;
  bool alive = false;
  for (auto &tp : m->tiles[g->mode]) alive |= tp.get() == g->t && tp->id == g->tiles_id;
  MC_CHECK(alive, MC_ERR_STATE, "the map's batched layout changed since the graph was captured");
  if ((rc = ensure_recs(m, m->ctx->stream, g->t, g->mode, g->flags, binfo))) return rc;
  MC_HIP(hipGraphLaunch(g->exec, m->ctx->stream));
  return MC_OK;
}

int mc_graph_destroy(mc_graph *g) {
  if (!g) return MC_OK;
  {
    std::lock_guard<std::recursive_mutex> lk(g->ctx->mu);
    (void)hipStreamSynchronize(g->ctx->stream);
  }
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return MC_OK;
}

int mc_time_integrate(mc_map *m, const void *frames_dev, int32_t dtype, int32_t n_frames,
                      int32_t mode, uint32_t flags, int32_t kind, int32_t iters, int32_t warmup,
                      float *cakes_dev, float *lineouts_dev, double *ms_per_launch,
                      double *ms_per_iter) {
  MC_CHECK(m && frames_dev && cakes_dev && n_frames >= 1 && iters >= 1, MC_ERR_ARG,
           "bad arguments");
  MC_CHECK(kind >= 0 && kind <= 2, MC_ERR_ARG, "timing kind %d", kind);
  if (kind == 2) {  // the batched passes replayed as one graph per iteration
    mc_graph *g = nullptr;
    int rc = mc_graph_create(m, frames_dev, dtype, n_frames, mode, flags, cakes_dev,
                             lineouts_dev, &g);
    if (rc) return rc;
    std::unique_ptr<mc_graph, int (*)(mc_graph *)> keep(g, mc_graph_destroy);
    for (int w = 0; w < warmup; w++)
      if ((rc = mc_graph_launch(g))) return rc;
    hipStream_t s = m->ctx->stream;
    MC_HIP(hipStreamSynchronize(s));
    hipEvent_t e0, e1;
    MC_HIP(hipEventCreate(&e0));
    MC_HIP(hipEventCreate(&e1));
    MC_HIP(hipEventRecord(e0, s));
    for (int it = 0; it < iters && !rc; it++) rc = mc_graph_launch(g);
    MC_HIP(hipEventRecord(e1, s));
    MC_HIP(hipEventSynchronize(e1));
    float tot = 0;
    MC_HIP(hipEventElapsedTime(&tot, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    const int step = pass_frames(*g->t);
    const int launches = (n_frames + step - 1) / step;
    if (ms_per_launch) *ms_per_launch = tot / iters / launches;
    if (ms_per_iter) *ms_per_iter = tot / iters;
    return MC_OK;
  }
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  std::lock_guard<std::mutex> lk2(m->mu);
  MC_HIP(hipSetDevice(m->ctx->device));
  hipStream_t s = m->ctx->stream;
  int rc;
  const float2 *binfo = nullptr;
  if ((rc = ensure_binfo(m, mode, flags, &binfo))) return rc;
  const int mask_nan = (flags & MC_F_MASK_NAN) ? 1 : 0;
  Tiles *t = nullptr;
  if (kind == 1 && ((rc = ensure_tiles(m, mode, batch_elem(dtype), &t)) ||
                    (rc = ensure_recs(m, s, t, mode, flags, binfo)) ||
                    (rc = ensure_xt(m->work.xt, *m, dtype)) ||
                    (rc = ensure_part(m->work.part, *t))))
    return rc;
  int64_t sub_frames = 0;
  if (kind == 0 && (rc = ensure_sub(m->work.xt, *m, 1, &sub_frames))) return rc;
  const size_t es = dtype_size(dtype);
  auto one_iter = [&](std::vector<hipEvent_t> *ev) -> int {
    const int step = kind == 1 ? pass_frames(*t) : 1;
    for (int f0 = 0; f0 < n_frames; f0 += step) {
      const int nf = std::min(step, n_frames - f0);
      const char *fb = (const char *)frames_dev + (size_t)f0 * m->npx() * es;
      float *cb = cakes_dev + (size_t)f0 * m->nb();
      float *lb = lineouts_dev ? lineouts_dev + (size_t)f0 * 3 * m->n_r : nullptr;
      if (ev) MC_HIP(hipEventRecord((*ev)[2 * (f0 / step)], s));
      int r2 = kind == 1 ? launch_batched(s, *m, *t, binfo, mask_nan, fb, dtype, nf, cb, lb,
                                          m->work.xt.p, m->work.part.as<float>())
                         : launch_plain(s, *m, m->csr[mode], binfo, mask_nan, fb, dtype, cb, lb,
                                        1, m->npx(), m->work.xt.as<float>(), sub_frames);
      if (r2) return r2;
      if (ev) MC_HIP(hipEventRecord((*ev)[2 * (f0 / step) + 1], s));
    }
    return MC_OK;
  };
  for (int w = 0; w < warmup; w++)
    if ((rc = one_iter(nullptr))) return rc;
  MC_HIP(hipStreamSynchronize(s));
  const int step = kind == 1 ? pass_frames(*t) : 1;
  const int launches = (n_frames + step - 1) / step;
  std::vector<hipEvent_t> ev(2 * (size_t)launches * iters);
  for (auto &e : ev) MC_HIP(hipEventCreate(&e));
  hipEvent_t e0, e1;
  MC_HIP(hipEventCreate(&e0));
  MC_HIP(hipEventCreate(&e1));
  MC_HIP(hipEventRecord(e0, s));
  for (int it = 0; it < iters; it++) {
    std::vector<hipEvent_t> sub(ev.begin() + 2 * (size_t)launches * it,
                                ev.begin() + 2 * (size_t)launches * (it + 1));
    if ((rc = one_iter(&sub))) return rc;
  }
  MC_HIP(hipEventRecord(e1, s));
  MC_HIP(hipEventSynchronize(e1));
  double kern_ms = 0;
  for (size_t l = 0; l < ev.size() / 2; l++) {
    float ms = 0;
    MC_HIP(hipEventElapsedTime(&ms, ev[2 * l], ev[2 * l + 1]));
    kern_ms += ms;
  }
  float tot = 0;
  MC_HIP(hipEventElapsedTime(&tot, e0, e1));
  for (auto &e : ev) (void)hipEventDestroy(e);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms_per_launch) *ms_per_launch = kern_ms / ((double)iters * launches);
  if (ms_per_iter) *ms_per_iter = tot / iters;
  return MC_OK;
}

// Times the batched kernel for each candidate (batch, rows per workgroup,
// LDS budget) on device frames of the map's shape (constant content: the
// kernel's work does not depend on pixel values) and keeps the fastest.
}  // extern "C"

namespace {

struct Plan {
  int batch, rows, lds, bins, ents, lanes;
  double fps, plain_fps;
};

// Times the candidate plans on synthetic frames of the map's shape (128
// frames per batched timing, 16 per one-frame timing) and applies the
// fastest; the table of every candidate goes to *table.
int autotune_run(mc_map *m, int32_t dtype, Plan *best_out, std::string *table) {
  const size_t es = dtype_size(dtype);
  MC_CHECK(es > 0, MC_ERR_ARG, "unsupported dtype %d", dtype);
  DBuf frames, cakes;
  int rc;
  constexpr int kFrames = 128, kPlainFrames = 16;
  if ((rc = frames.alloc((size_t)kFrames * m->npx() * es)) ||
      (rc = cakes.alloc((size_t)kFrames * m->nb() * 4)))
    return rc;
  MC_HIP(hipMemset(frames.p, 0x11, frames.bytes));
  MC_HIP(hipDeviceSynchronize());
  const int Bs[] = {16, 32, 64}, Hs[] = {16, 32}, Ls[] = {81920, 163840};
  const int Shapes[][2] = {{256, 2048}, {128, 1024}, {512, 4096}};
  Plan best{m->batch, m->rows_per_group, m->lds_bytes, m->tile_bins, m->tile_ents,
            m->plain_lanes, -1.0, -1.0};
  char line[160];
  for (int B : Bs)
    for (int H : Hs)
      for (int L : Ls)
        for (const auto &sh : Shapes) {
          if (mc_map_set_tuning(m, B, H, L) || mc_map_set_tile_shape(m, sh[0], sh[1])) continue;
          double per_launch = 0, per_iter = 0;
          if (mc_time_integrate(m, frames.p, dtype, kFrames, MC_MODE_BILINEAR,
                                MC_F_NORMALIZE | MC_F_MASK_NAN, 1, 3, 1, cakes.as<float>(),
                                nullptr, &per_launch, &per_iter))
            continue;  // e.g. an LDS budget the device does not grant
          const double fps = kFrames / (per_iter / 1e3);
          snprintf(line, sizeof line, "# batched B %d rows %d lds %d tile %d/%d: %.0f frames/s\n",
                   B, H, L, sh[0], sh[1], fps);
          *table += line;
          if (fps > best.fps) best = Plan{B, H, L, sh[0], sh[1], m->plain_lanes, fps, -1.0};
        }
  MC_CHECK(best.fps > 0, MC_ERR_STATE, "no candidate tuning ran: %s", g_err.c_str());
  mc_map_set_tuning(m, best.batch, best.rows, best.lds);
  mc_map_set_tile_shape(m, best.bins, best.ents);
  for (int lanes : {4, 8, 16}) {
    mc_map_set_plain_lanes(m, lanes);
    double per_launch = 0, per_iter = 0;
    if (mc_time_integrate(m, frames.p, dtype, kPlainFrames, MC_MODE_BILINEAR,
                          MC_F_NORMALIZE | MC_F_MASK_NAN, 0, 3, 1, cakes.as<float>(), nullptr,
                          &per_launch, &per_iter))
      continue;
    const double fps = kPlainFrames / (per_iter / 1e3);
    snprintf(line, sizeof line, "# one-frame lanes %d: %.0f frames/s\n", lanes, fps);
    *table += line;
    if (fps > best.plain_fps) {
      best.plain_fps = fps;
      best.lanes = lanes;
    }
  }
  mc_map_set_plain_lanes(m, best.lanes);
  {
    // Drop the layouts of the losers.
    std::lock_guard<std::mutex> lk(m->mu);
    for (auto &v : m->tiles) {
      std::vector<std::unique_ptr<Tiles>> keep;
      for (auto &t : v)
        if (t->width == width_of(best.batch) && t->rows == best.rows && t->lds_bytes == best.lds &&
            t->bin_cap == best.bins && t->ent_cap == best.ents)
          keep.push_back(std::move(t));
      v.swap(keep);
    }
  }
  *best_out = best;
  return MC_OK;
}

// FNV-1a, 64 bits.
struct Fnv {
  uint64_t h = 1469598103934665603ull;
  void add(const void *p, size_t n) {
    const unsigned char *c = static_cast<const unsigned char *>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
  }
  template <typename T> void add(const T &v) { add(&v, sizeof v); }
};

int plan_key(mc_map *m, int32_t dtype, const char *key_extra, std::string *key) {
  hipDeviceProp_t prop;
  MC_HIP(hipGetDeviceProperties(&prop, m->ctx->device));
  Fnv f;
  const char tag[] = "midas_amd batched plan 1";
  f.add(tag, sizeof tag);
  f.add(m->ny);
  f.add(m->nz);
  f.add(m->n_r);
  f.add(m->n_eta);
  f.add(m->nnz);
  f.add(m->q_mode);
  f.add(m->fused);
  f.add(m->h_pairs.data(), m->h_pairs.size() * sizeof(int32_t));
  f.add(m->h_flags.data(), m->h_flags.size() * sizeof(int32_t));
  f.add(dtype);
  f.add(prop.gcnArchName, strnlen(prop.gcnArchName, sizeof prop.gcnArchName));
  f.add(prop.multiProcessorCount);
  if (key_extra) f.add(key_extra, strlen(key_extra));
  char hex[17];
  snprintf(hex, sizeof hex, "%016llx", (unsigned long long)f.h);
  *key = hex;
  return MC_OK;
}

bool read_plan(const std::string &path, Plan *p) {
  FILE *fp = fopen(path.c_str(), "r");
  if (!fp) return false;
  Plan q{0, 0, 0, 0, 0, 0, 0.0, 0.0};
  int seen = 0;
  char line[256];
  while (fgets(line, sizeof line, fp)) {
    char k[64];
    double v;
    if (line[0] == '#' || sscanf(line, "%63s %lf", k, &v) != 2) continue;
    const std::string key(k);
    if (key == "batch") q.batch = (int)v, seen |= 1;
    else if (key == "rows_per_group") q.rows = (int)v, seen |= 2;
    else if (key == "lds_bytes") q.lds = (int)v, seen |= 4;
    else if (key == "tile_bins") q.bins = (int)v, seen |= 8;
    else if (key == "tile_entries") q.ents = (int)v, seen |= 16;
    else if (key == "plain_lanes") q.lanes = (int)v, seen |= 32;
    else if (key == "frames_per_s") q.fps = v;
    else if (key == "plain_frames_per_s") q.plain_fps = v;
  }
  fclose(fp);
  if (seen != 63) return false;
  *p = q;
  return true;
}

}  // namespace

extern "C" {

int mc_map_autotune(mc_map *m, int32_t dtype, double *best_fps) {
  MC_CHECK(m, MC_ERR_ARG, "null map");
  Plan best;
  std::string table;
  int rc = autotune_run(m, dtype, &best, &table);
  if (rc) return rc;
  if (best_fps) *best_fps = best.fps;
  return MC_OK;
}

int mc_map_plan_key(mc_map *m, int32_t dtype, const char *key_extra, char *out, int32_t cap) {
  MC_CHECK(m && out && cap >= 17, MC_ERR_ARG, "bad arguments");
  MC_CHECK(dtype_size(dtype) > 0, MC_ERR_ARG, "unsupported dtype %d", dtype);
  std::string key;
  int rc = plan_key(m, dtype, key_extra, &key);
  if (rc) return rc;
  snprintf(out, (size_t)cap, "%s", key.c_str());
  return MC_OK;
}

int mc_map_autotune_plan(mc_map *m, int32_t dtype, const char *plan_dir, const char *key_extra,
                         double *best_fps, int32_t *loaded) {
  MC_CHECK(m && plan_dir && *plan_dir, MC_ERR_ARG, "bad arguments");
  MC_CHECK(dtype_size(dtype) > 0, MC_ERR_ARG, "unsupported dtype %d", dtype);
  std::string key;
  int rc = plan_key(m, dtype, key_extra, &key);
  if (rc) return rc;
  const std::string path = std::string(plan_dir) + "/" + key + ".plan";
  Plan p;
  if (read_plan(path, &p)) {
    if (!(rc = mc_map_set_tuning(m, p.batch, p.rows, p.lds)) &&
        !(rc = mc_map_set_tile_shape(m, p.bins, p.ents)) &&
        !(rc = mc_map_set_plain_lanes(m, p.lanes))) {
      if (best_fps) *best_fps = p.fps;
      if (loaded) *loaded = 1;
      return MC_OK;
    }
    return rc;  // a plan file with values this library refuses
  }
  std::string table;
  if ((rc = autotune_run(m, dtype, &p, &table))) return rc;
  mkdir(plan_dir, 0755);  // may exist already
  const std::string tmp = path + ".tmp";
  FILE *fp = fopen(tmp.c_str(), "w");
  MC_CHECK(fp, MC_ERR_IO, "cannot write %s", tmp.c_str());
  fprintf(fp,
          "# tuned plan of the batched and one-frame kernels, key %s\n"
          "# (map size, bins, records per bin, mask flags, pixel type, device)\n"
          "batch %d\nrows_per_group %d\nlds_bytes %d\ntile_bins %d\ntile_entries %d\n"
          "plain_lanes %d\nframes_per_s %.1f\nplain_frames_per_s %.1f\n%s",
          key.c_str(), p.batch, p.rows, p.lds, p.bins, p.ents, p.lanes, p.fps, p.plain_fps,
          table.c_str());
  const bool ok = fclose(fp) == 0 && rename(tmp.c_str(), path.c_str()) == 0;
  MC_CHECK(ok, MC_ERR_IO, "cannot write %s", path.c_str());
  if (best_fps) *best_fps = p.fps;
  if (loaded) *loaded = 0;
  return MC_OK;
}

}  // extern "C"
