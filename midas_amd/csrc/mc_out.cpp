// Result files of an integration run: the caked HDF5 file, the lineout and
// Int2D binaries, and the peak-fit outputs, written by a writer thread off
// the integration path (the reference's writer thread is
// IntegratorFitPeaksGPUStream.cu:506-565).  Formats follow
//   <name>.caked.hdf          IntegratorZarrOMP.c:1671-1687, 2084-2168
//   <prefix>_lineout.bin/.xy  IntegratorZarrOMP.c:1514-1527, 1829-1863
//   <prefix>_fit.bin          IntegratorZarrOMP.c:1530-1539, 1879-1883
//   <prefix>_fit_per_eta.csv  IntegratorZarrOMP.c:1540-1549, 2038-2058
//   <prefix>_caked_peaks.h5   PeakFitIO.c:52-267, IntegratorZarrOMP.c:2176-2195
//   lineout.bin, lineout_simple_mean.bin, Int2D.bin
//                             IntegratorFitPeaksGPUStream.cu:506-565, 2465-2467
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mc_h5.h"
#include "mc_zarr.h"
#include "mc_internal.h"

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr size_t kMaxQueued = 64;

double atand(double x) { return std::atan(x) * 180.0 / kPi; }

struct PeakRow {  // PeakH5Row, PeakFitIO.h:27-40
  int32_t frame, eta_idx, peak_nr;
  double eta_deg, center_2theta, area, sig, gam, fwhm, eta_mix, d_spacing, chi_sq;
};

// The caked results go to the HDF5 file and, when asked for, to its Zarr
// copy (<name>.caked.hdf.zarr.zip): the same groups, datasets and
// attributes through both writers.
struct CakedSink {
  mc::h5::Writer h5;
  mc::zarr::Writer zr;
  bool zarr = false;
  std::string err;
  int check(int a, const std::string &e) {
    if (a && err.empty()) err = e;
    return a;
  }
  int open(const std::string &h5_path, bool with_zarr) {
    zarr = with_zarr;
    if (check(h5.open(h5_path), h5.error())) return -1;
    return zarr ? check(zr.open(h5_path + ".zarr.zip"), zr.error()) : 0;
  }
  int group(const std::string &p) {
    return check(h5.group(p), h5.error()) || (zarr && check(zr.group(p), zr.error()));
  }
  int dataset(const std::string &p, mc::h5::Kind k, const std::vector<uint64_t> &dims,
              const void *data) {
    return check(h5.dataset(p, k, dims, data), h5.error()) ||
           (zarr && check(zr.dataset(p, k, dims, data), zr.error()));
  }
  int attr_f64(const std::string &p, const std::string &n, const double *v, int c) {
    return check(h5.attr_f64(p, n, v, c), h5.error()) ||
           (zarr && check(zr.attr_f64(p, n, v, c), zr.error()));
  }
  int attr_i32(const std::string &p, const std::string &n, const int32_t *v, int c) {
    return check(h5.attr_i32(p, n, v, c), h5.error()) ||
           (zarr && check(zr.attr_i32(p, n, v, c), zr.error()));
  }
  int attr_str(const std::string &p, const std::string &n, const std::string &v) {
    return check(h5.attr_str(p, n, v), h5.error()) ||
           (zarr && check(zr.attr_str(p, n, v), zr.error()));
  }
  int close() {
    const int a = check(h5.close(), h5.error());
    const int b = zarr ? check(zr.close(), zr.error()) : 0;
    return a || b;
  }
  bool is_open() const { return h5.is_open(); }
  const std::string &error() const { return err; }
};

}  // namespace

struct mc_out {
  mc_out_config cfg{};
  std::vector<double> reta, rc, eta_lo, eta_hi, instrument;
  std::string source, dir, name, prefix;
  CakedSink caked;
  FILE *f_lineout = nullptr, *f_xy = nullptr, *f_stream = nullptr, *f_stream_mean = nullptr,
       *f_int2d = nullptr, *f_fit = nullptr, *f_csv = nullptr;
  // OmegaSumFrame accumulation (IntegratorZarrOMP.c:1633-1642, 2093-2132)
  std::vector<double> chunk;
  int chunk_frames = 0, chunk_last = 0;
  double chunk_first_ome = 0, chunk_last_ome = 0, chunk_meta[4] = {0, 0, 0, 0};
  std::vector<double> sum;
  std::vector<double> omegas;
  int n_frames = 0;
  std::vector<PeakRow> peaks;
  // writer thread
  std::mutex mu;
  std::condition_variable cv_push, cv_pop;
  std::deque<std::function<int()>> q;
  bool stop = false;
  int err = MC_OK;
  std::string err_msg;
  std::thread th;

  // The caked file is written when any of its parts is asked for; the
  // OmegaSumFrame sums are part of it.
  bool hdf() const {
    return cfg.files & (MC_OUT_CAKED_HDF | MC_OUT_PER_FRAME | MC_OUT_SUM_FRAMES | MC_OUT_OMEGAS);
  }
  bool omega_sums() const { return hdf() && cfg.omega_sum_frames > 0; }
  // Records the first error (err and err_msg are guarded by mu; jobs run
  // without holding it).
  int fail(int code, const std::string &m) {
    std::lock_guard<std::mutex> lk(mu);
    if (err == MC_OK) {
      err = code;
      err_msg = m;
    }
    return code;
  }
  int status() {
    std::lock_guard<std::mutex> lk(mu);
    return err;
  }
  int write_chunk();
  int frame(int32_t frame_nr, double omega, const std::vector<float> &cake,
            const std::vector<float> &lo, const double *meta);
  int finish();
  int write_peaks_h5();
};

int mc_out::write_chunk() {
  // /OmegaSumFrame/LastFrameNumber_<i> with the attributes of
  // IntegratorZarrOMP.c:2101-2130; the meta values are chunk means.
  const int n_r = cfg.n_r, n_eta = cfg.n_eta;
  const std::string p = "/OmegaSumFrame/LastFrameNumber_" + std::to_string(chunk_last);
  const int32_t last = chunk_last, ns = chunk_frames;
  const double first = chunk_first_ome, lo = chunk_last_ome;
  double m[4];
  for (int k = 0; k < 4; k++) m[k] = chunk_meta[k] / ns;
  if (caked.dataset(p, mc::h5::Kind::F64, {(uint64_t)n_r, (uint64_t)n_eta}, chunk.data()) ||
      caked.attr_i32(p, "LastFrameNumber", &last, 1) ||
      caked.attr_i32(p, "Number Of Frames Summed", &ns, 1) ||
      caked.attr_f64(p, "FirstOme", &first, 1) || caked.attr_f64(p, "LastOme", &lo, 1) ||
      caked.attr_f64(p, "Temperature", &m[0], 1) || caked.attr_f64(p, "Pressure", &m[1], 1) ||
      caked.attr_f64(p, "I", &m[2], 1) || caked.attr_f64(p, "I0", &m[3], 1))
    return fail(MC_ERR_IO, caked.error());
  chunk_frames = 0;
  return MC_OK;
}

int mc_out::frame(int32_t frame_nr, double omega, const std::vector<float> &cake,
                  const std::vector<float> &lo, const double *meta) {
  const int n_r = cfg.n_r, n_eta = cfg.n_eta;
  const size_t nb = (size_t)n_r * n_eta;
  std::vector<double> cd(cake.begin(), cake.end());
  if (cfg.files & MC_OUT_PER_FRAME) {
    const std::string p = "/IntegrationResult/FrameNr_" + std::to_string(frame_nr);
    if (caked.dataset(p, mc::h5::Kind::F64, {(uint64_t)n_r, (uint64_t)n_eta}, cd.data()) ||
        caked.attr_f64(p, "omega", &omega, 1) || caked.attr_str(p, "Header", "Radius,Eta") ||
        caked.attr_str(p, "Units", "Pixels,Degrees"))
      return fail(MC_ERR_IO, caked.error());
  }
  if (omega_sums()) {
    if (n_frames % cfg.omega_sum_frames == 0) {
      std::fill(chunk.begin(), chunk.end(), 0.0);
      chunk_frames = 0;
      chunk_first_ome = omega;
      for (double &v : chunk_meta) v = 0;
    }
    for (size_t i = 0; i < nb; i++) chunk[i] += cd[i];
    if (meta)
      for (int k = 0; k < 4; k++) chunk_meta[k] += meta[k];
    chunk_frames++;
    chunk_last = frame_nr;
    chunk_last_ome = omega;
    if ((n_frames + 1) % cfg.omega_sum_frames == 0 && write_chunk()) return status();
  }
  if (cfg.files & MC_OUT_SUM_FRAMES)
    for (size_t i = 0; i < nb; i++) sum[i] += cd[i];
  if (cfg.files & MC_OUT_OMEGAS) omegas.push_back(omega);
  if (cfg.files & MC_OUT_LINEOUT) {
    // Mean over eta of the values that are nonzero and not NaN, in FP64
    // from the written cake (IntegratorZarrOMP.c:1829-1843).
    std::vector<double> l1(n_r, 0.0);
    for (int r = 0; r < n_r; r++) {
      double s = 0;
      int c = 0;
      for (int e = 0; e < n_eta; e++) {
        const double v = cd[(size_t)r * n_eta + e];
        if (v != 0 && !std::isnan(v)) {
          s += v;
          c++;
        }
      }
      l1[r] = c > 0 ? s / c : 0;
    }
    if (fwrite(l1.data(), sizeof(double), n_r, f_lineout) != (size_t)n_r)
      return fail(MC_ERR_IO, "write " + prefix + "_lineout.bin");
    fflush(f_lineout);
    for (int r = 0; r < n_r; r++) {
      if (std::isnan(l1[r])) continue;
      const double x = cfg.q_mode ? cfg.q_min + cfg.q_bin_size * (r + 0.5)
                                  : atand(rc[r] * cfg.px / cfg.lsd);
      fprintf(f_xy, "%.6f  %.6f\n", x, l1[r]);
    }
    fflush(f_xy);
  }
  if (cfg.files & MC_OUT_STREAM) {
    // (R, I) double pairs per R bin (IntegratorFitPeaksGPUStream.cu:534-551)
    std::vector<double> pr(2 * (size_t)n_r), pm(2 * (size_t)n_r);
    for (int r = 0; r < n_r; r++) {
      pr[2 * r] = pm[2 * r] = rc[r];
      pr[2 * r + 1] = lo[(size_t)MC_LINEOUT_AREA * n_r + r];
      pm[2 * r + 1] = lo[(size_t)MC_LINEOUT_MEAN * n_r + r];
    }
    if (fwrite(pr.data(), sizeof(double), pr.size(), f_stream) != pr.size() ||
        fwrite(pm.data(), sizeof(double), pm.size(), f_stream_mean) != pm.size())
      return fail(MC_ERR_IO, "write lineout.bin");
    fflush(f_stream);
    fflush(f_stream_mean);
  }
  if (cfg.files & MC_OUT_INT2D) {
    if (fwrite(cd.data(), sizeof(double), nb, f_int2d) != nb)
      return fail(MC_ERR_IO, "write Int2D.bin");
    fflush(f_int2d);
  }
  n_frames++;
  return MC_OK;
}

int mc_out::write_peaks_h5() {
  // PeakFitIO.c:124-267 (pfio_write_peaks_h5).  frame_keys holds one key
  // per frame written.
  mc::h5::Writer w;
  const std::string fn = dir + "/" + prefix + "_caked_peaks.h5";
  if (w.open(fn)) return fail(MC_ERR_IO, w.error());
  const int n_r = cfg.n_r, n_eta = cfg.n_eta;
  std::vector<double> tth(n_r), eta(n_eta);
  for (int r = 0; r < n_r; r++) tth[r] = std::atan(rc[r] * cfg.px / cfg.lsd) * 180.0 / kPi;
  for (int e = 0; e < n_eta; e++) eta[e] = (eta_lo[e] + eta_hi[e]) / 2.0;
  std::vector<std::string> keys;
  for (int i = 0; i < n_frames; i++) keys.push_back(std::to_string(i));
  int bad = w.group("/metadata") || w.attr_str("/metadata", "zarr_source", source) ||
            w.dataset("/metadata/tth_axis", mc::h5::Kind::F64, {(uint64_t)n_r}, tth.data()) ||
            w.dataset("/metadata/eta_axis", mc::h5::Kind::F64, {(uint64_t)n_eta}, eta.data()) ||
            w.vstr_dataset("/metadata/frame_keys", keys);
  const size_t n = peaks.size();
  if (!bad && n > 0) {
    std::vector<int32_t> fi(n), ei(n), pn(n);
    std::vector<double> col[9];
    std::vector<std::string> fk(n);
    for (auto &c : col) c.resize(n);
    for (size_t i = 0; i < n; i++) {
      const PeakRow &p = peaks[i];
      fi[i] = p.frame;
      ei[i] = p.eta_idx;
      pn[i] = p.peak_nr;
      const double v[9] = {p.eta_deg, p.center_2theta, p.area, p.sig, p.gam,
                           p.fwhm, p.eta_mix, p.d_spacing, p.chi_sq};
      for (int k = 0; k < 9; k++) col[k][i] = v[k];
      fk[i] = std::to_string(p.frame);
    }
    static const char *names[9] = {"eta_deg", "center_2theta", "area", "sig", "gam",
                                   "FWHM_deg", "eta_mix", "d_spacing_A", "chi_sq"};
    bad = w.group("/peaks") ||
          w.dataset("/peaks/frame_idx", mc::h5::Kind::I32, {n}, fi.data()) ||
          w.dataset("/peaks/eta_idx", mc::h5::Kind::I32, {n}, ei.data()) ||
          w.dataset("/peaks/peak_nr", mc::h5::Kind::I32, {n}, pn.data());
    for (int k = 0; k < 9 && !bad; k++)
      bad = w.dataset(std::string("/peaks/") + names[k], mc::h5::Kind::F64, {n}, col[k].data());
    if (!bad) bad = w.vstr_dataset("/peaks/frame_key", fk);
  }
  if (bad) {
    const int e = fail(MC_ERR_IO, w.error());
    w.close();
    return e;
  }
  if (w.close()) return fail(MC_ERR_IO, w.error());
  return MC_OK;
}

int mc_out::finish() {
  const int n_r = cfg.n_r, n_eta = cfg.n_eta;
  if (hdf() && caked.is_open()) {
    // The last, partial OmegaSumFrame chunk (IntegratorZarrOMP.c:2097).
    if (omega_sums() && chunk_frames > 0 && status() == MC_OK) write_chunk();
    if (status() == MC_OK && (cfg.files & MC_OUT_OMEGAS) &&
        (caked.dataset("/Omegas", mc::h5::Kind::F64, {omegas.size()}, omegas.data()) ||
         caked.attr_str("/Omegas", "Units", "Degrees")))
      fail(MC_ERR_IO, caked.error());
    if (status() == MC_OK && (cfg.files & MC_OUT_SUM_FRAMES)) {
      const int32_t nf = n_frames;
      if (caked.dataset("/SumFrames", mc::h5::Kind::F64, {(uint64_t)n_r, (uint64_t)n_eta},
                        sum.data()) ||
          caked.attr_str("/SumFrames", "Header", "Radius,Eta") ||
          caked.attr_str("/SumFrames", "Units", "Pixels,Degrees") ||
          caked.attr_i32("/SumFrames", "nFrames", &nf, 1))
        fail(MC_ERR_IO, caked.error());
    }
    if (status() == MC_OK && !reta.empty()) {
      const int32_t ne = n_eta, nr = n_r;
      if (caked.dataset("/REtaMap", mc::h5::Kind::F64, {5, (uint64_t)n_r, (uint64_t)n_eta},
                        reta.data()) ||
          caked.attr_i32("/REtaMap", "nEtaBins", &ne, 1) ||
          caked.attr_i32("/REtaMap", "nRBins", &nr, 1) ||
          caked.attr_str("/REtaMap", "Header", "Radius,2Theta,Eta,BinArea,Q") ||
          caked.attr_str("/REtaMap", "Units", "Pixels,Degrees,Degrees,Pixels,InvAngstrom"))
        fail(MC_ERR_IO, caked.error());
    }
    if (status() == MC_OK && !instrument.empty()) {
      static const char *names[10] = {"Polariz", "Lam", "SH_L", "U", "V",
                                      "W", "X", "Y", "Z", "Distance"};
      for (int k = 0; k < 10 && status() == MC_OK; k++)
        if (caked.dataset(std::string("/InstrumentParameters/") + names[k], mc::h5::Kind::F64,
                          {1}, &instrument[k]))
          fail(MC_ERR_IO, caked.error());
    }
    if (caked.close() && status() == MC_OK) fail(MC_ERR_IO, caked.error());
  }
  if ((cfg.files & MC_OUT_PEAKS) && status() == MC_OK) write_peaks_h5();
  for (FILE **f : {&f_lineout, &f_xy, &f_stream, &f_stream_mean, &f_int2d, &f_fit, &f_csv})
    if (*f) {
      if (fclose(*f) != 0) fail(MC_ERR_IO, "closing an output file failed");
      *f = nullptr;
    }
  return status();
}

namespace {

// Pushes a job; blocks while the queue is full.  Returns the output set's
// first error.
int push(mc_out *o, std::function<int()> job) {
  std::unique_lock<std::mutex> lk(o->mu);
  o->cv_push.wait(lk, [&] { return o->q.size() < kMaxQueued || o->err != MC_OK; });
  if (o->err != MC_OK) {
    mc::set_error("%s", o->err_msg.c_str());
    return o->err;
  }
  o->q.push_back(std::move(job));
  o->cv_pop.notify_one();
  return MC_OK;
}

void writer_loop(mc_out *o) {
  for (;;) {
    std::function<int()> job;
    bool failed;
    {
      std::unique_lock<std::mutex> lk(o->mu);
      o->cv_pop.wait(lk, [&] { return !o->q.empty() || o->stop; });
      if (o->q.empty()) return;
      job = std::move(o->q.front());
      o->q.pop_front();
      failed = o->err != MC_OK;  // after an error the remaining jobs are dropped
    }
    if (!failed) job();
    o->cv_push.notify_all();
  }
}

FILE *open_file(mc_out *o, const std::string &fn, const char *mode) {
  FILE *f = fopen(fn.c_str(), mode);
  if (!f) o->fail(MC_ERR_IO, "cannot create " + fn);
  return f;
}

}  // namespace

extern "C" {

int mc_out_open(const char *dir, const char *name, const mc_out_config *cfg, mc_out **out) {
  MC_CHECK(dir && name && cfg && out, MC_ERR_ARG, "null argument");
  MC_CHECK(cfg->n_r > 0 && cfg->n_eta > 0, MC_ERR_ARG, "bin counts %d x %d", cfg->n_r,
           cfg->n_eta);
  MC_CHECK(cfg->omega_sum_frames >= 0, MC_ERR_ARG, "omega_sum_frames %d", cfg->omega_sum_frames);
  MC_CHECK(cfg->r_centers && cfg->eta_lo && cfg->eta_hi, MC_ERR_ARG,
           "r_centers, eta_lo and eta_hi are required");
  MC_CHECK(!(cfg->files & (MC_OUT_LINEOUT | MC_OUT_PEAKS)) || cfg->q_mode ||
               (cfg->px > 0 && cfg->lsd > 0),
           MC_ERR_ARG, "px and lsd must be positive for 2-theta axes");
  auto o = std::make_unique<mc_out>();
  o->cfg = *cfg;
  const size_t n_r = cfg->n_r, n_eta = cfg->n_eta, nb = n_r * n_eta;
  if (cfg->reta) o->reta.assign(cfg->reta, cfg->reta + 5 * nb);
  o->rc.assign(cfg->r_centers, cfg->r_centers + n_r);
  o->eta_lo.assign(cfg->eta_lo, cfg->eta_lo + n_eta);
  o->eta_hi.assign(cfg->eta_hi, cfg->eta_hi + n_eta);
  if (cfg->instrument) o->instrument.assign(cfg->instrument, cfg->instrument + 10);
  o->source = cfg->source ? cfg->source : "";
  o->cfg.reta = o->cfg.r_centers = o->cfg.eta_lo = o->cfg.eta_hi = o->cfg.instrument = nullptr;
  o->cfg.source = nullptr;
  o->dir = dir;
  o->name = name;
  {
    std::string base = o->name;
    const size_t slash = base.find_last_of('/');
    if (slash != std::string::npos) base = base.substr(slash + 1);
    o->name = base;
    const size_t dot = base.find_last_of('.');
    o->prefix = dot == std::string::npos ? base : base.substr(0, dot);
  }
  if (o->omega_sums()) o->chunk.assign(nb, 0.0);
  if (cfg->files & MC_OUT_SUM_FRAMES) o->sum.assign(nb, 0.0);
  const std::string d = o->dir + "/";
  if (o->hdf()) {
    if (o->caked.open(d + o->name + ".caked.hdf", (cfg->files & MC_OUT_CAKED_ZARR) != 0)) {
      mc::set_error("%s", o->caked.error().c_str());
      return MC_ERR_IO;
    }
    if ((cfg->files & MC_OUT_PER_FRAME) && o->caked.group("/IntegrationResult"))
      o->fail(MC_ERR_IO, o->caked.error());
    if (o->omega_sums() && o->caked.group("/OmegaSumFrame"))
      o->fail(MC_ERR_IO, o->caked.error());
  }
  if (cfg->files & MC_OUT_LINEOUT) {
    o->f_lineout = open_file(o.get(), d + o->prefix + "_lineout.bin", "wb");
    o->f_xy = open_file(o.get(), d + o->prefix + "_lineout.xy", "w");
  }
  if (cfg->files & MC_OUT_STREAM) {
    o->f_stream = open_file(o.get(), d + "lineout.bin", "wb");
    o->f_stream_mean = open_file(o.get(), d + "lineout_simple_mean.bin", "wb");
  }
  if (cfg->files & MC_OUT_INT2D) o->f_int2d = open_file(o.get(), d + "Int2D.bin", "wb");
  if (cfg->files & MC_OUT_PEAKS) {
    o->f_fit = open_file(o.get(), d + o->prefix + "_fit.bin", "wb");
    o->f_csv = open_file(o.get(), d + o->prefix + "_fit_per_eta.csv", "w");
    if (o->f_csv)
      fprintf(o->f_csv,
              "Frame,EtaBin,EtaCen,PeakIdx,R_px,R_um,TwoTheta_deg,"
              "Area,Sig_centideg2,Gam_centideg,FWHM_deg,Eta,ChiSq,Area2\n");
  }
  if (o->err != MC_OK) {
    const int e = o->err;
    mc::set_error("%s", o->err_msg.c_str());
    o->finish();
    return e;
  }
  mc_out *p = o.release();
  p->th = std::thread(writer_loop, p);
  *out = p;
  return MC_OK;
}

int mc_out_frame(mc_out *o, int32_t frame, double omega, const float *cake,
                 const float *lineouts, const double *meta) {
  MC_CHECK(o && cake, MC_ERR_ARG, "null argument");
  MC_CHECK(lineouts || !(o->cfg.files & MC_OUT_STREAM), MC_ERR_ARG,
           "lineout.bin needs the lineouts");
  const size_t nb = (size_t)o->cfg.n_r * o->cfg.n_eta;
  std::vector<float> c(cake, cake + nb), l;
  if (lineouts) l.assign(lineouts, lineouts + 3 * (size_t)o->cfg.n_r);
  std::vector<double> m;
  if (meta) m.assign(meta, meta + 4);
  return push(o, [o, frame, omega, c = std::move(c), l = std::move(l), m = std::move(m)]() {
    return o->frame(frame, omega, c, l, m.empty() ? nullptr : m.data());
  });
}

int mc_out_lineout_fits(mc_out *o, int32_t frame, int32_t n_peaks, const double *params) {
  MC_CHECK(o && (params || n_peaks == 0) && n_peaks >= 0, MC_ERR_ARG, "bad arguments");
  MC_CHECK(o->cfg.files & MC_OUT_PEAKS, MC_ERR_STATE, "output set has no peak files");
  std::vector<double> v(params, params + 7 * (size_t)n_peaks);
  (void)frame;
  return push(o, [o, v = std::move(v)]() {
    if (!v.empty() && fwrite(v.data(), sizeof(double), v.size(), o->f_fit) != v.size())
      return o->fail(MC_ERR_IO, "write " + o->prefix + "_fit.bin");
    fflush(o->f_fit);
    return MC_OK;
  });
}

int mc_out_peaks(mc_out *o, int32_t n_rows, const mc_peak_row *rows) {
  MC_CHECK(o && (rows || n_rows == 0) && n_rows >= 0, MC_ERR_ARG, "bad arguments");
  MC_CHECK(o->cfg.files & MC_OUT_PEAKS, MC_ERR_STATE, "output set has no peak files");
  for (int32_t i = 0; i < n_rows; i++)
    MC_CHECK(rows[i].eta_bin >= 0 && rows[i].eta_bin < o->cfg.n_eta, MC_ERR_ARG,
             "eta bin %d out of range", rows[i].eta_bin);
  std::vector<mc_peak_row> v(rows, rows + n_rows);
  return push(o, [o, v = std::move(v)]() {
    const mc_out_config &c = o->cfg;
    for (const mc_peak_row &r : v) {
      const double center_px = r.p[1];
      if (center_px == 0) continue;  // peak not fitted
      const double eta_cen = (o->eta_lo[r.eta_bin] + o->eta_hi[r.eta_bin]) / 2.0;
      const double r_um = center_px * c.px;
      const double tth = std::atan(r_um / c.lsd) * 180.0 / kPi;
      fprintf(o->f_csv, "%d,%d,%.4f,%d,%.6f,%.4f,%.6f,%.4f,%.6f,%.6f,%.4f,%.4f,%.4f,%.4f\n",
              r.frame, r.eta_bin, eta_cen, r.peak, center_px, r_um, tth, r.p[0], r.p[2], r.p[3],
              r.p[4], r.p[5], r.p[6], r.p[0]);
      // pfio_make_row, PeakFitIO.c:52-86
      PeakRow h;
      h.frame = r.frame;
      h.eta_idx = r.eta_bin;
      h.peak_nr = r.peak;
      h.eta_deg = eta_cen;
      h.area = r.p[0];
      h.sig = r.p[2];
      h.gam = r.p[3];
      h.fwhm = r.p[4];
      h.eta_mix = r.p[5];
      h.chi_sq = r.p[6];
      h.center_2theta = tth;
      const double theta = h.center_2theta * kPi / 360.0;
      h.d_spacing = theta > 1e-10 ? c.wavelength / (2.0 * std::sin(theta)) : 0.0;
      o->peaks.push_back(h);
    }
    fflush(o->f_csv);
    return MC_OK;
  });
}

int mc_out_close(mc_out *o) {
  MC_CHECK(o, MC_ERR_ARG, "null argument");
  {
    std::lock_guard<std::mutex> lk(o->mu);
    o->stop = true;
  }
  o->cv_pop.notify_all();
  if (o->th.joinable()) o->th.join();
  const int rc = o->finish();
  if (rc != MC_OK) mc::set_error("%s", o->err_msg.c_str());
  delete o;
  return rc;
}

}  // extern "C"
