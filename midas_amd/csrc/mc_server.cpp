// mc_server.cpp -- continuous-batching frame server (the MI355X analogue of
// the frame server in IntegratorFitPeaksGPUStream.cu:567-734, 2567-2883).
//
//  * Geometries (map + frame dtype + mode + flags) stay resident on the
//    device; each is admitted only if the maps, the in-flight batch buffers
//    and the new geometry's tables fit the HBM budget.
//  * submit() copies a frame into a pinned host buffer (frames of 4 MiB
//    and more on the caller and 3 persistent helper threads) and queues it; a full queue blocks the caller
//    (backpressure) up to its timeout.  acquire_buffer()/submit_buffer()
//    let a producer write straight into the pinned buffer (zero copy, e.g.
//    a socket receive).
//  * A scheduler thread forms batches while earlier batches are on the GPU:
//    it takes the oldest queued request and up to max_batch queued frames
//    of the same geometry (frames are grouped by map), copies them to the
//    device on one of several in-flight slots, each with its own stream,
//    and launches the one-frame kernel (up to 6 frames) or the batched
//    kernel.
//  * A completion thread waits for each slot's event in launch order, hands
//    the cakes and lineouts to the caller's callback and records the
//    latency from submit to result (p50/p99).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "mc_internal.h"

namespace mc {
namespace {

using Clock = std::chrono::steady_clock;

// Batches of at most this many frames run the one-frame kernel per frame;
// larger ones run batched passes of the narrowest width (16, 32 or 64
// frames) that holds them.  Measured on 2048^2 uint16 with a dark: 113 us
// per frame one at a time; one pass of 16 frames 296 us whatever its frame
// count, of 32 frames 360-410 us, of 64 frames 500-700 us.
constexpr int kPlainMaxFrames = 2;

// In-flight batches (slots, each with its own stream and buffers).  Serving
// 5000 frames/s of the three-geometry mix (host frames over PCIe), p99
// latency fell from about 11 ms with 3 slots to 7.5 ms with 6: one slot's
// device-to-host copy of cakes overlaps another's frames going up.
constexpr int kSlots = 6;

double ms_since(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Geo {
  mc_map *m = nullptr;
  int dtype = 1, mode = 1;
  uint32_t flags = 0;
  size_t frame_bytes = 0, cake_floats = 0, lo_floats = 0;
  size_t part_bytes = 0;  // lineout partials of one batched pass
};

struct Req {
  int geo = 0;
  int64_t id = 0;
  void *host = nullptr;  // pinned copy of the frame
  Clock::time_point t_submit;
};

struct Slot {
  hipStream_t s = nullptr;
  hipEvent_t done = nullptr;
  DBuf frames, cakes, lo;
  Work work;  // scratch of the slot's batched passes (pixel-major batch, lineout partials)
  float *h_cakes = nullptr, *h_lo = nullptr;  // pinned
  size_t cap_frames = 0, cap_cakes = 0, cap_lo = 0;
  std::vector<Req> reqs;
  int geo = -1;
  bool busy = false;
  int rc = MC_OK;
};

// memcpy of large frames split over the caller and 3 persistent helper
// threads (one thread's memcpy into pinned memory is slower than PCIe, and
// starting threads per frame cost tens of microseconds each).  Helpers spin
// briefly on the job generation before sleeping, so back-to-back frames do
// not pay a wake-up each.  Copies of concurrent callers take turns.
class CopyPool {
 public:
  static constexpr size_t kMin = 4 << 20;  // smaller copies: the caller alone
  static constexpr int kHelpers = 3;

  CopyPool() {
    for (int k = 0; k < kHelpers; k++) th_[k] = std::thread([this, k] { helper(k); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
      gen_.fetch_add(1);
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  void copy(void *dst, const void *src, size_t n) {
    if (n < kMin) {
      memcpy(dst, src, n);
      return;
    }
    std::lock_guard<std::mutex> turn(run_);
    const size_t part = ((n / (kHelpers + 1)) + 4095) & ~(size_t)4095;
    {
      std::lock_guard<std::mutex> lk(m_);
      for (int k = 0; k < kHelpers; k++) {
        const size_t a = std::min(n, (size_t)(k + 1) * part);
        const size_t b = std::min(n, (size_t)(k + 2) * part);
        job_[k] = Job{(char *)dst + a, (const char *)src + a, b - a};
      }
      pending_.store(kHelpers);
      gen_.fetch_add(1);
    }
    cv_.notify_all();
    memcpy(dst, src, std::min(n, part));
    // The helpers' parts take about as long as the caller's.
    while (pending_.load() != 0) std::this_thread::yield();
  }

 private:
  struct Job {
    char *dst = nullptr;
    const char *src = nullptr;
    size_t n = 0;
  };
  void helper(int k) {
    uint64_t seen = 0;
    for (;;) {
      // Spin for about 200 us on the generation, then sleep on the
      // condition variable.
      const auto t_spin = Clock::now() + std::chrono::microseconds(200);
      while (gen_.load() == seen && Clock::now() < t_spin) std::this_thread::yield();
      Job j;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return gen_.load() != seen; });
        seen = gen_.load();
        if (stop_) return;
        j = job_[k];
      }
      if (j.n) memcpy(j.dst, j.src, j.n);
      pending_.fetch_sub(1);
    }
  }
  std::thread th_[kHelpers];
  std::mutex m_, run_;
  std::condition_variable cv_;
  Job job_[kHelpers];
  std::atomic<uint64_t> gen_{0};
  std::atomic<int> pending_{0};
  bool stop_ = false;
};

}  // namespace
}  // namespace mc

using namespace mc;

struct mc_server {
  Ctx *ctx = nullptr;
  int64_t budget = 0;
  int max_batch = 64, qdepth = 256;
  // Batching window: a free slot waits up to this long after the oldest
  // queued request's arrival for max_batch frames of its geometry (0: launch
  // whatever is queued as soon as a slot is free).
  int64_t window_us = 0;
  int draining = 0;  // callers in mc_server_drain: close every window
  mc_result_cb cb = nullptr;
  void *user = nullptr;
  std::vector<Geo> geos;
  std::mutex mu;
  std::condition_variable cv_work, cv_space, cv_idle, cv_done;
  std::deque<Req> queue;
  std::vector<Slot> slots;
  std::deque<int> launched;  // slot indices in launch order
  std::vector<std::vector<void *>> free_host;  // pinned frame buffers per geometry
  std::thread sched, compl_;
  std::unique_ptr<CopyPool> copier;  // frame copies of mc_server_submit
  bool stop = false;
  int in_flight = 0;
  // stats
  std::vector<double> lat;
  int64_t frames_done = 0, batches = 0, rejected = 0, peak_bytes = 0;
  Clock::time_point t_first, t_last;
  bool any = false;

  int64_t resident_bytes() {
    int64_t b = 0;
    std::vector<mc_map *> seen;
    for (auto &g : geos)
      if (std::find(seen.begin(), seen.end(), g.m) == seen.end()) {
        seen.push_back(g.m);
        mc_map_info info;
        mc_map_info_get(g.m, &info);
        b += info.hbm_bytes;
      }
    for (auto &s : slots)
      b += s.frames.bytes + s.cakes.bytes + s.lo.bytes + s.work.xt.bytes + s.work.part.bytes;
    return b;
  }
};

namespace {

int grow_slots(mc_server *sv, size_t fb, size_t cf, size_t lf, size_t xb, size_t pb) {
  for (auto &s : sv->slots) {
    int rc;
    if (s.work.xt.bytes < xb && (rc = s.work.xt.alloc(xb))) return rc;
    if (s.work.part.bytes < pb && (rc = s.work.part.alloc(pb))) return rc;
    if (s.cap_frames < fb) {
      if ((rc = s.frames.alloc(fb))) return rc;
      s.cap_frames = fb;
    }
    if (s.cap_cakes < cf) {
      if ((rc = s.cakes.alloc(cf * 4))) return rc;
      if (s.h_cakes) (void)hipHostFree(s.h_cakes);
      MC_HIP(hipHostMalloc((void **)&s.h_cakes, cf * 4, hipHostMallocDefault));
      s.cap_cakes = cf;
    }
    if (s.cap_lo < lf) {
      if ((rc = s.lo.alloc(lf * 4))) return rc;
      if (s.h_lo) (void)hipHostFree(s.h_lo);
      MC_HIP(hipHostMalloc((void **)&s.h_lo, lf * 4, hipHostMallocDefault));
      s.cap_lo = lf;
    }
  }
  return MC_OK;
}

void scheduler(mc_server *sv) {
  (void)hipSetDevice(sv->ctx->device);
  std::unique_lock<std::mutex> lk(sv->mu);
  for (;;) {
    int slot = -1;
    sv->cv_work.wait(lk, [&] {
      if (sv->stop && sv->queue.empty()) return true;
      if (sv->queue.empty()) return false;
      for (size_t i = 0; i < sv->slots.size(); i++)
        if (!sv->slots[i].busy) {
          slot = (int)i;
          return true;
        }
      return false;
    });
    if (sv->queue.empty()) break;
    if (sv->window_us > 0 && !sv->stop) {
      // Wait for a full batch of the oldest request's geometry, at most
      // until its window closes (a drain or stop ends the wait at once).
      auto full = [&] {
        if (sv->stop || sv->draining) return true;
        const int g0 = sv->queue.front().geo;
        int n = 0;
        for (const Req &r : sv->queue) n += r.geo == g0;
        return n >= sv->max_batch;
      };
      const auto until = sv->queue.front().t_submit + std::chrono::microseconds(sv->window_us);
      sv->cv_work.wait_until(lk, until, full);
      if (sv->queue.empty()) break;
    }
    Slot &S = sv->slots[slot];
    // The oldest request picks the geometry; take up to max_batch queued
    // frames of that geometry in arrival order.
    const int g = sv->queue.front().geo;
    S.reqs.clear();
    for (auto it = sv->queue.begin(); it != sv->queue.end() && (int)S.reqs.size() < sv->max_batch;) {
      if (it->geo == g) {
        S.reqs.push_back(*it);
        it = sv->queue.erase(it);
      } else {
        ++it;
      }
    }
    S.busy = true;
    S.geo = g;
    sv->in_flight++;
    sv->cv_space.notify_all();
    const Geo &G = sv->geos[g];
    const int nf = (int)S.reqs.size();
    lk.unlock();
    int rc = MC_OK;
    for (int i = 0; i < nf && !rc; i++)
      if (hipMemcpyAsync((char *)S.frames.p + (size_t)i * G.frame_bytes, S.reqs[i].host,
                         G.frame_bytes, hipMemcpyHostToDevice, S.s) != hipSuccess)
        rc = MC_ERR_HIP;
    if (!rc) {
      std::lock_guard<std::mutex> mlk(G.m->mu);
      // Small batches take the one-frame kernel per frame; the others one
      // pass of the narrowest width that holds them.
      rc = run_integrate(G.m, S.s, S.frames.p, G.dtype, nf, G.mode, G.flags, S.cakes.as<float>(),
                         S.lo.as<float>(), nf <= kPlainMaxFrames, false, &S.work,
                         width_of(nf));
    }
    if (!rc && hipMemcpyAsync(S.h_cakes, S.cakes.p, (size_t)nf * G.cake_floats * 4,
                              hipMemcpyDeviceToHost, S.s) != hipSuccess)
      rc = MC_ERR_HIP;
    if (!rc && hipMemcpyAsync(S.h_lo, S.lo.p, (size_t)nf * G.lo_floats * 4,
                              hipMemcpyDeviceToHost, S.s) != hipSuccess)
      rc = MC_ERR_HIP;
    if (!rc && hipEventRecord(S.done, S.s) != hipSuccess) rc = MC_ERR_HIP;
    S.rc = rc;
    lk.lock();
    sv->launched.push_back(slot);
    sv->batches++;
    sv->cv_done.notify_all();
  }
  sv->cv_done.notify_all();
}

void completer(mc_server *sv) {
  (void)hipSetDevice(sv->ctx->device);
  std::unique_lock<std::mutex> lk(sv->mu);
  for (;;) {
    sv->cv_done.wait(lk, [&] { return !sv->launched.empty() || (sv->stop && sv->in_flight == 0); });
    if (sv->launched.empty()) break;
    const int slot = sv->launched.front();
    sv->launched.pop_front();
    Slot &S = sv->slots[slot];
    lk.unlock();
    int rc = S.rc;
    if (!rc && hipEventSynchronize(S.done) != hipSuccess) rc = MC_ERR_HIP;
    const auto t = Clock::now();
    const Geo &G = sv->geos[S.geo];
    for (size_t i = 0; i < S.reqs.size(); i++)
      if (sv->cb)
        sv->cb(sv->user, S.reqs[i].id, S.geo, rc ? nullptr : S.h_cakes + i * G.cake_floats,
               rc ? nullptr : S.h_lo + i * G.lo_floats, rc);
    lk.lock();
    for (auto &r : S.reqs) {
      sv->lat.push_back(ms_since(r.t_submit, t));
      sv->free_host[r.geo].push_back(r.host);
      if (!sv->any) {
        sv->t_first = r.t_submit;
        sv->any = true;
      }
      if (r.t_submit < sv->t_first) sv->t_first = r.t_submit;
    }
    sv->t_last = t;
    sv->frames_done += (int64_t)S.reqs.size();
    S.reqs.clear();
    S.busy = false;
    sv->in_flight--;
    sv->cv_work.notify_all();
    sv->cv_space.notify_all();
    sv->cv_idle.notify_all();
  }
}

}  // namespace

extern "C" {

int mc_server_create(mc_ctx *ctx, int64_t hbm_budget_bytes, int32_t max_batch,
                     int32_t queue_depth, mc_result_cb cb, void *user, mc_server **out) {
  MC_CHECK(ctx && out && hbm_budget_bytes > 0, MC_ERR_ARG, "bad arguments");
  MC_CHECK(max_batch >= 1 && max_batch <= 128, MC_ERR_ARG, "max_batch %d not in 1..128",
           max_batch);
  MC_CHECK(queue_depth >= 1, MC_ERR_ARG, "queue_depth %d", queue_depth);
  MC_HIP(hipSetDevice(ctx->device));
  auto *sv = new mc_server();
  sv->ctx = ctx;
  sv->budget = hbm_budget_bytes;
  sv->max_batch = max_batch;
  sv->qdepth = queue_depth;
  sv->cb = cb;
  sv->user = user;
  sv->slots = std::vector<Slot>(kSlots);
  for (auto &s : sv->slots) {
    if (hipStreamCreateWithFlags(&s.s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
      set_error("server: stream/event creation failed");
      delete sv;
      return MC_ERR_HIP;
    }
  }
  sv->copier = std::make_unique<CopyPool>();
  sv->sched = std::thread(scheduler, sv);
  sv->compl_ = std::thread(completer, sv);
  *out = sv;
  return MC_OK;
}

int mc_server_add_geometry(mc_server *sv, mc_map *m, int32_t dtype, int32_t mode, uint32_t flags,
                           int32_t *geometry) {
  MC_CHECK(sv && m && geometry, MC_ERR_ARG, "null argument");
  MC_CHECK(dtype_size(dtype) > 0, MC_ERR_ARG, "unsupported dtype %d", dtype);
  MC_CHECK(mode >= 0 && mode <= 2 && flags < 4, MC_ERR_ARG, "mode %d flags %u", mode, flags);
  MC_CHECK(m->ctx == sv->ctx, MC_ERR_ARG, "map belongs to another context");
  std::unique_lock<std::mutex> lk(sv->mu);
  // Wait for the device to be idle: slot buffers may have to grow.
  sv->cv_idle.wait(lk, [&] { return sv->in_flight == 0 && sv->queue.empty(); });
  int rc;
  size_t pb = 0;  // lineout partials of this geometry's batched passes
  {
    // The layouts of every pass width a batch of up to max_batch frames
    // may use are built now, not on a request's path.
    std::lock_guard<std::recursive_mutex> clk(sv->ctx->mu);
    std::lock_guard<std::mutex> mlk(m->mu);
    for (int w : {16, 32, 64}) {
      if (w > width_of(sv->max_batch)) break;
      if ((rc = prepare_batched(m, mode, flags, dtype, w))) return rc;
      Tiles *t = nullptr;
      if ((rc = ensure_tiles(m, mode, dtype_size(dtype) <= 2 ? 2 : 4, &t, w))) return rc;
      pb = std::max(pb, t->part_bytes());
    }
  }
  Geo G;
  G.m = m;
  G.dtype = dtype;
  G.mode = mode;
  G.flags = flags;
  G.frame_bytes = (size_t)m->npx() * dtype_size(dtype);
  G.cake_floats = (size_t)m->nb();
  G.lo_floats = (size_t)3 * m->n_r;
  G.part_bytes = pb;
  size_t fb = 0, cf = 0, lf = 0, xb = batch_xt_bytes(*m, dtype);
  for (auto &g : sv->geos) {
    fb = std::max(fb, g.frame_bytes);
    cf = std::max(cf, g.cake_floats);
    lf = std::max(lf, g.lo_floats);
    xb = std::max(xb, batch_xt_bytes(*g.m, g.dtype));
    pb = std::max(pb, g.part_bytes);
  }
  fb = std::max(fb, G.frame_bytes) * sv->max_batch;
  cf = std::max(cf, G.cake_floats) * sv->max_batch;
  lf = std::max(lf, G.lo_floats) * sv->max_batch;
  // Admission: resident maps + in-flight buffers after growth.
  int64_t need = sv->resident_bytes();
  {
    mc_map_info info;
    mc_map_info_get(m, &info);
    bool known = false;
    for (auto &g : sv->geos) known |= g.m == m;
    if (!known) need += info.hbm_bytes;
  }
  for (auto &s : sv->slots)
    need += (int64_t)(std::max(fb, s.cap_frames) - s.frames.bytes) +
            (int64_t)(std::max(cf, s.cap_cakes) * 4 - s.cakes.bytes) +
            (int64_t)(std::max(lf, s.cap_lo) * 4 - s.lo.bytes) +
            (int64_t)(std::max(xb, s.work.xt.bytes) - s.work.xt.bytes) +
            (int64_t)(std::max(pb, s.work.part.bytes) - s.work.part.bytes);
  if (need > sv->budget) {
    sv->rejected++;
    set_error("geometry needs %lld bytes resident, budget is %lld", (long long)need,
              (long long)sv->budget);
    return MC_ERR_BUDGET;
  }
  MC_HIP(hipSetDevice(sv->ctx->device));
  if ((rc = grow_slots(sv, fb, cf, lf, xb, pb))) return rc;
  sv->geos.push_back(G);
  sv->free_host.emplace_back();
  // Pinned staging for the first frames of this geometry, so that early
  // requests do not pay for page-locked allocations.
  for (int i = 0; i < std::min(sv->max_batch, 16); i++) {
    void *buf = nullptr;
    if (hipHostMalloc(&buf, G.frame_bytes, hipHostMallocDefault) != hipSuccess) break;
    sv->free_host.back().push_back(buf);
  }
  sv->peak_bytes = std::max(sv->peak_bytes, sv->resident_bytes());
  *geometry = (int32_t)sv->geos.size() - 1;
  return MC_OK;
}

int mc_server_submit(mc_server *sv, int32_t geometry, const void *frame, int64_t request_id,
                     int32_t timeout_ms) {
  MC_CHECK(sv && frame, MC_ERR_ARG, "null argument");
  std::unique_lock<std::mutex> lk(sv->mu);
  MC_CHECK(geometry >= 0 && geometry < (int)sv->geos.size(), MC_ERR_ARG, "geometry %d",
           geometry);
  MC_CHECK(!sv->stop, MC_ERR_STATE, "server is shutting down");
  const auto t0 = Clock::now();
  const auto deadline = t0 + std::chrono::milliseconds(timeout_ms < 0 ? 0 : timeout_ms);
  if (!sv->cv_space.wait_until(lk, deadline, [&] { return (int)sv->queue.size() < sv->qdepth; })) {
    set_error("queue full for %d ms", timeout_ms);
    return MC_ERR_TIMEOUT;
  }
  const Geo &G = sv->geos[geometry];
  void *buf = nullptr;
  auto &pool = sv->free_host[geometry];
  if (!pool.empty()) {
    buf = pool.back();
    pool.pop_back();
  } else {
    lk.unlock();
    MC_HIP(hipHostMalloc(&buf, G.frame_bytes, hipHostMallocDefault));
    lk.lock();
  }
  lk.unlock();
  sv->copier->copy(buf, frame, G.frame_bytes);
  lk.lock();
  Req r;
  r.geo = geometry;
  r.id = request_id;
  r.host = buf;
  r.t_submit = t0;
  sv->queue.push_back(r);
  sv->cv_work.notify_one();
  return MC_OK;
}

int mc_server_acquire_buffer(mc_server *sv, int32_t geometry, void **buf) {
  MC_CHECK(sv && buf, MC_ERR_ARG, "null argument");
  std::unique_lock<std::mutex> lk(sv->mu);
  MC_CHECK(geometry >= 0 && geometry < (int)sv->geos.size(), MC_ERR_ARG, "geometry %d",
           geometry);
  auto &pool = sv->free_host[geometry];
  if (!pool.empty()) {
    *buf = pool.back();
    pool.pop_back();
    return MC_OK;
  }
  const size_t fb = sv->geos[geometry].frame_bytes;
  lk.unlock();
  MC_HIP(hipHostMalloc(buf, fb, hipHostMallocDefault));
  return MC_OK;
}

int mc_server_release_buffer(mc_server *sv, int32_t geometry, void *buf) {
  MC_CHECK(sv && buf, MC_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  MC_CHECK(geometry >= 0 && geometry < (int)sv->geos.size(), MC_ERR_ARG, "geometry %d",
           geometry);
  sv->free_host[geometry].push_back(buf);
  return MC_OK;
}

int mc_server_submit_buffer(mc_server *sv, int32_t geometry, void *buf, int64_t request_id,
                            int32_t timeout_ms) {
  MC_CHECK(sv && buf, MC_ERR_ARG, "null argument");
  std::unique_lock<std::mutex> lk(sv->mu);
  MC_CHECK(geometry >= 0 && geometry < (int)sv->geos.size(), MC_ERR_ARG, "geometry %d",
           geometry);
  MC_CHECK(!sv->stop, MC_ERR_STATE, "server is shutting down");
  const auto t0 = Clock::now();
  const auto deadline = t0 + std::chrono::milliseconds(timeout_ms < 0 ? 0 : timeout_ms);
  if (!sv->cv_space.wait_until(lk, deadline, [&] { return (int)sv->queue.size() < sv->qdepth; })) {
    set_error("queue full for %d ms", timeout_ms);
    return MC_ERR_TIMEOUT;
  }
  Req r;
  r.geo = geometry;
  r.id = request_id;
  r.host = buf;
  r.t_submit = t0;
  sv->queue.push_back(r);
  sv->cv_work.notify_one();
  return MC_OK;
}

int mc_server_drain(mc_server *sv, int32_t timeout_ms) {
  MC_CHECK(sv, MC_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(sv->mu);
  const auto deadline = Clock::now() + std::chrono::milliseconds(timeout_ms);
  sv->draining++;
  sv->cv_work.notify_all();
  const bool idle = sv->cv_idle.wait_until(
      lk, deadline, [&] { return sv->queue.empty() && sv->in_flight == 0; });
  sv->draining--;
  if (!idle) {
    set_error("drain timed out with %zu queued, %d in flight", sv->queue.size(), sv->in_flight);
    return MC_ERR_TIMEOUT;
  }
  return MC_OK;
}

int mc_server_set_batch_window(mc_server *sv, int64_t window_us) {
  MC_CHECK(sv, MC_ERR_ARG, "null server");
  MC_CHECK(window_us >= 0 && window_us <= 10000000, MC_ERR_ARG,
           "batch window %lld us outside 0..10 s", (long long)window_us);
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->window_us = window_us;
  sv->cv_work.notify_all();
  return MC_OK;
}

int mc_server_stats_get(mc_server *sv, mc_server_stats *st) {
  MC_CHECK(sv && st, MC_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(sv->mu);
  memset(st, 0, sizeof(*st));
  st->frames_done = sv->frames_done;
  st->batches = sv->batches;
  st->rejected = sv->rejected;
  st->admitted_bytes = sv->resident_bytes();
  st->peak_bytes = sv->peak_bytes;
  if (!sv->lat.empty()) {
    std::vector<double> v = sv->lat;
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) {
      double idx = p * (v.size() - 1);
      size_t a = (size_t)idx;
      size_t b = std::min(a + 1, v.size() - 1);
      return v[a] + (v[b] - v[a]) * (idx - a);
    };
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    st->max_ms = v.back();
    double s = 0;
    for (double x : v) s += x;
    st->mean_ms = s / v.size();
    const double span = ms_since(sv->t_first, sv->t_last);
    st->frames_per_s = span > 0 ? sv->frames_done / (span / 1e3) : 0;
  }
  return MC_OK;
}

int mc_server_destroy(mc_server *sv) {
  if (!sv) return MC_OK;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv_work.notify_all();
  sv->cv_done.notify_all();
  if (sv->sched.joinable()) sv->sched.join();
  {
    std::lock_guard<std::mutex> lk(sv->mu);
  }
  sv->cv_done.notify_all();
  if (sv->compl_.joinable()) sv->compl_.join();
  for (auto &s : sv->slots) {
    if (s.s) (void)hipStreamSynchronize(s.s);
    if (s.h_cakes) (void)hipHostFree(s.h_cakes);
    if (s.h_lo) (void)hipHostFree(s.h_lo);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.s) (void)hipStreamDestroy(s.s);
  }
  for (auto &pool : sv->free_host)
    for (void *p : pool) (void)hipHostFree(p);
  for (auto &r : sv->queue) (void)hipHostFree(r.host);
  delete sv;
  return MC_OK;
}

}  // extern "C"
