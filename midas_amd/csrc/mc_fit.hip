// mc_fit.hip -- batched Levenberg-Marquardt peak fits, one thread per
// window (gfx950).
//
// Models (PeakFit.c):
//   pV  (MC_FIT_PV, :882-920): f = BG + Imax (mu L + (1 - mu) G),
//        L = 1 / (1 + 4 dr^2 / Gamma^2), G = exp(-4 ln2 dr^2 / Gamma^2);
//        params [Rcen, mu, Gamma, Imax, BG], bounds and start of
//        pf_fit_singlet_pv (:922-943).
//   TCH (MC_FIT_TCH, :17-90): f = bg0 + bg1 xn + A (eta L + (1 - eta) G),
//        area-normalised Gaussian/Lorentzian of total width F(sig, gam)
//        and mixing eta(sig, gam) by Thompson-Cox-Hastings; params
//        [A, center, sig, gam, bg0, bg1], start and bounds of fitPeaks for
//        one peak whose ROI is the window (:356-469).
//   pV doublet (MC_FIT_PV2, :984-1101): two pV peaks sharing mu and BG,
//        params [R1, R2, mu, Gamma1, Imax1, Gamma2, Imax2, BG], start and
//        bounds of pf_fit_doublet_pv (:1037-1078): R1 in [x0, Rmid], R2 in
//        [Rmid, x_last], Rmid the guesses' midpoint.
//   TCH with N = 2..8 peaks (MC_FIT_TCH_PEAKS(N)): a merged fitPeaks job
//        (:321-351, 356-544): N TCH peaks [A, center, sig, gam] and one
//        Chebyshev background [bg0, bg1] (4N + 2 params), each peak started
//        and bounded as fitPeaks does (:380-445) at the window sample
//        nearest its guess, the background from the first peak's estimate;
//        the TCH doublet of the per-eta fits (pf_fit_doublet_tch,
//        :1138-1170) is N = 2 over the whole window.
// The objective is the reference's: the sum of squared residuals.  The
// reference minimises it with NLopt (L-BFGS, Nelder-Mead fallback); this
// kernel uses Levenberg-Marquardt with Marquardt damping and box bounds by
// projection; a parameter on a bound whose descent direction leaves the box
// is held there for the step (active set).  The TCH peaks are fitted in
// u = sqrt(sig) in place of sig (the Gaussian FWHM is linear in u, while
// dF/dsig grows without bound as sig -> 0 and stalls the damped steps near
// a Lorentzian shape); bounds are the square roots of the sig bounds and
// results are reported in sig.  Residuals and Jacobians are FP32, evaluated
// on coordinates relative to the window's first sample; the normal
// equations are accumulated and solved (Cholesky) in FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>

#include "mc_internal.h"

namespace mc {
namespace {

// Peaks of a TCH-family model (1 for TCH, N for TCH_PEAKS(N); 0 otherwise).
template <int MODEL> constexpr int tch_peaks() {
  return MODEL == MC_FIT_TCH ? 1
         : (MODEL >= MC_FIT_TCH2 && MODEL <= MC_FIT_TCH_PEAKS(MC_FIT_MAX_PEAKS)) ? MODEL - 1
                                                                                  : 0;
}
template <int MODEL> constexpr int n_params() {
  return MODEL == MC_FIT_PV ? 5 : MODEL == MC_FIT_PV2 ? 8 : 4 * tch_peaks<MODEL>() + 2;
}
// Guesses a window takes and result rows it writes (MC_FIT_OUT each).
template <int MODEL> constexpr int n_guesses() {
  return MODEL == MC_FIT_PV ? 1 : MODEL == MC_FIT_PV2 ? 2 : tch_peaks<MODEL>();
}
template <int MODEL> constexpr int n_rows() {
  return MODEL == MC_FIT_PV || MODEL == MC_FIT_PV2 ? 1 : tch_peaks<MODEL>();
}
constexpr float kLn2 = 0.69314718055994530942f;

// The normal equations are symmetric: their lower triangle is kept packed,
// row after row (element (a, b), b <= a, at a (a + 1) / 2 + b).
template <int NP> constexpr int n_tri() { return NP * (NP + 1) / 2; }
__device__ __forceinline__ constexpr int tri(int a, int b) { return a * (a + 1) / 2 + b; }
// Loops over parameters are unrolled (compile-time indices keep the arrays
// in registers) up to this many parameters.
template <int NP> constexpr int unroll_np() { return NP <= 10 ? NP : 1; }

struct Win {
  const float *xr;  // sample abscissae relative to x0 (computed in FP64)
  const float *y;   // samples
  int n;
  double x0;        // absolute abscissa of the first sample
  __device__ double xa(int i) const { return x0 + (double)xr[i]; }
};

// pV value and Jacobian at window-relative coordinate xr.
__device__ __forceinline__ float pv_eval(const double *p, float xr, float cr, float *J) {
  const float mu = (float)p[1], gam = (float)p[2], imax = (float)p[3], bg = (float)p[4];
  const float C0 = 4.0f * kLn2;
  const float dr = xr - cr, dr2 = dr * dr;
  const float ig2 = 1.0f / (gam * gam);
  const float den = 1.0f + 4.0f * dr2 * ig2;
  const float L = 1.0f / den;
  const float G = __expf(-C0 * dr2 * ig2);
  if (J) {
    const float dLdc = 8.0f * dr * ig2 / (den * den);
    const float dGdc = G * 2.0f * C0 * dr * ig2;
    const float dLdg = 8.0f * dr2 * ig2 / (gam * den * den);
    const float dGdg = G * 2.0f * C0 * dr2 * ig2 / gam;
    J[0] = imax * (mu * dLdc + (1.0f - mu) * dGdc);
    J[1] = imax * (L - G);
    J[2] = imax * (mu * dLdg + (1.0f - mu) * dGdg);
    J[3] = mu * L + (1.0f - mu) * G;
    J[4] = 1.0f;
  }
  return bg + imax * (mu * L + (1.0f - mu) * G);
}

// pV doublet value and Jacobian (pf_pv_doublet_obj, PeakFit.c:984-1035).
__device__ __forceinline__ float pv2_eval(const double *p, float xr, float c1, float c2,
                                          float *J) {
  const float mu = (float)p[2], g1 = (float)p[3], i1 = (float)p[4], g2 = (float)p[5],
              i2 = (float)p[6], bg = (float)p[7];
  const float C0 = 4.0f * kLn2;
  const float d1 = xr - c1, d2 = xr - c2, d1s = d1 * d1, d2s = d2 * d2;
  const float ig1 = 1.0f / (g1 * g1), ig2 = 1.0f / (g2 * g2);
  const float den1 = 1.0f + 4.0f * d1s * ig1, den2 = 1.0f + 4.0f * d2s * ig2;
  const float L1 = 1.0f / den1, L2 = 1.0f / den2;
  const float G1 = __expf(-C0 * d1s * ig1), G2 = __expf(-C0 * d2s * ig2);
  if (J) {
    J[0] = i1 * (mu * 8.0f * d1 * ig1 / (den1 * den1) + (1.0f - mu) * G1 * 2.0f * C0 * d1 * ig1);
    J[1] = i2 * (mu * 8.0f * d2 * ig2 / (den2 * den2) + (1.0f - mu) * G2 * 2.0f * C0 * d2 * ig2);
    J[2] = i1 * (L1 - G1) + i2 * (L2 - G2);
    J[3] = i1 * (mu * 8.0f * d1s * ig1 / (g1 * den1 * den1) +
                 (1.0f - mu) * G1 * 2.0f * C0 * d1s * ig1 / g1);
    J[4] = mu * L1 + (1.0f - mu) * G1;
    J[5] = i2 * (mu * 8.0f * d2s * ig2 / (g2 * den2 * den2) +
                 (1.0f - mu) * G2 * 2.0f * C0 * d2s * ig2 / g2);
    J[6] = mu * L2 + (1.0f - mu) * G2;
    J[7] = 1.0f;
  }
  return bg + i1 * (mu * L1 + (1.0f - mu) * G1) + i2 * (mu * L2 + (1.0f - mu) * G2);
}

// TCH width/mixing with derivatives (PeakFit.c:17-42), in u = sqrt(sig).
struct Tch {
  float F, eta, dF_du, dF_dgam, deta_du, deta_dgam;
};

__device__ Tch tch(double u_d, double gam_d) {
  const float u = (float)fmax(u_d, 1e-6), gam = (float)fmax(gam_d, 1e-6);
  const float kfg = 0.0235482004503094933f;  // sqrt(8 ln 2) / 100
  const float fg = kfg * u;
  const float fl = gam / 100.0f;
  const float g2 = fg * fg, g3 = g2 * fg, g4 = g3 * fg, g5 = g4 * fg;
  const float l2 = fl * fl, l3 = l2 * fl, l4 = l3 * fl, l5 = l4 * fl;
  const float P5 = g5 + 2.69269f * g4 * fl + 2.42843f * g3 * l2 + 4.47163f * g2 * l3 +
                   0.07842f * fg * l4 + l5;
  Tch t;
  t.F = fmaxf(powf(P5, 0.2f), 1e-15f);
  const float dP_dfg = 5.0f * g4 + 4.0f * 2.69269f * g3 * fl + 3.0f * 2.42843f * g2 * l2 +
                       2.0f * 4.47163f * fg * l3 + 0.07842f * l4;
  const float dP_dfl = 2.69269f * g4 + 2.0f * 2.42843f * g3 * fl + 3.0f * 4.47163f * g2 * l2 +
                       4.0f * 0.07842f * fg * l3 + 5.0f * l4;
  const float dF_dP = t.F / (5.0f * fmaxf(P5, 1e-30f));
  const float dfl_dgam = 0.01f;
  const float dF_dfg = dF_dP * dP_dfg, dF_dfl = dF_dP * dP_dfl;
  const float q = fl / t.F;
  float eta = 1.36603f * q - 0.47719f * q * q + 0.11116f * q * q * q;
  float deta_dq = 1.36603f - 2.0f * 0.47719f * q + 3.0f * 0.11116f * q * q;
  if (eta < 0.0f) { eta = 0.0f; deta_dq = 0.0f; }
  if (eta > 1.0f) { eta = 1.0f; deta_dq = 0.0f; }
  t.eta = eta;
  const float dq_dfg = -fl / (t.F * t.F) * dF_dfg;
  const float dq_dfl = 1.0f / t.F - fl / (t.F * t.F) * dF_dfl;
  t.dF_du = dF_dfg * kfg;
  t.dF_dgam = dF_dfl * dfl_dgam;
  t.deta_du = deta_dq * dq_dfg * kfg;
  t.deta_dgam = deta_dq * dq_dfl * dfl_dgam;
  return t;
}

// One TCH peak (params p[0..3] = A, center, u = sqrt(sig), gam): A * shape
// and, with J, its 4 derivatives.
__device__ __forceinline__ float tch_peak(const double *p, const Tch &t, float xr, float cr,
                                          float *J) {
  const float A = (float)p[0];
  const float F = t.F;
  const float k = 1.0f / (2.0f * sqrtf(2.0f * kLn2));
  const float sg = F * k;
  const float dx = xr - cr, dx2 = dx * dx;
  const float G = __expf(-0.5f * dx2 / (sg * sg)) / (sg * 2.50662827463100050f);
  const float h = 0.5f * F;
  const float D = dx2 + h * h;
  const float L = (h * 0.31830988618379067f) / D;
  const float shape = t.eta * L + (1.0f - t.eta) * G;
  if (J) {
    const float dG_dF = G * (-1.0f / F + dx2 / (sg * sg * F));
    const float dL_dF = L / F - L * h / D;
    const float ds_dF = t.eta * dL_dF + (1.0f - t.eta) * dG_dF;
    const float ds_deta = L - G;
    const float dG_dc = G * dx / (sg * sg);
    const float dL_dc = L * 2.0f * dx / D;
    J[0] = shape;
    J[1] = A * (t.eta * dL_dc + (1.0f - t.eta) * dG_dc);
    J[2] = A * (ds_dF * t.dF_du + ds_deta * t.deta_du);
    J[3] = A * (ds_dF * t.dF_dgam + ds_deta * t.deta_dgam);
  }
  return A * shape;
}

// Sum of squared residuals and (optionally) the FP64 normal equations
// (JtJ packed lower triangle).
template <int MODEL>
__device__ double chi2_normal(const Win &w, const double *p, double *JtJ, double *Jtr) {
  constexpr int NP = n_params<MODEL>();
  constexpr int UF = unroll_np<NP>();
  constexpr int NPK = tch_peaks<MODEL>();
  const float cr = (float)(p[0] - w.x0);
  const float cr2 = MODEL == MC_FIT_PV2 ? (float)(p[1] - w.x0) : 0.0f;
  const float xrange = fmaxf(w.xr[w.n - 1] - w.xr[0], 1e-12f);
  Tch t[NPK > 0 ? NPK : 1];
  float tc[NPK > 0 ? NPK : 1];
#pragma unroll
  for (int k = 0; k < NPK; k++) {
    t[k] = tch(p[4 * k + 2], p[4 * k + 3]);
    tc[k] = (float)(p[4 * k + 1] - w.x0);
  }
  if (JtJ) {
#pragma unroll UF
    for (int a = 0; a < n_tri<NP>(); a++) JtJ[a] = 0;
#pragma unroll UF
    for (int a = 0; a < NP; a++) Jtr[a] = 0;
  }
  double chi = 0;
  for (int i = 0; i < w.n; i++) {
    const float yi = w.y[i];
    if (isnan(yi)) continue;  // masked cake bins carry no information
    const float xr = w.xr[i];
    float J[NP];
    float f;
    if (MODEL == MC_FIT_PV) {
      f = pv_eval(p, xr, cr, JtJ ? J : nullptr);
    } else if (MODEL == MC_FIT_PV2) {
      f = pv2_eval(p, xr, cr, cr2, JtJ ? J : nullptr);
    } else {  // TCH family: Chebyshev T0 + T1 background plus the peaks
      const float xn = 2.0f * (xr - w.xr[0]) / xrange - 1.0f;
      f = (float)p[4 * NPK] + (float)p[4 * NPK + 1] * xn;
#pragma unroll
      for (int k = 0; k < NPK; k++) f += tch_peak(p + 4 * k, t[k], xr, tc[k], JtJ ? J + 4 * k : nullptr);
      if (JtJ) {
        J[4 * NPK] = 1.0f;
        J[4 * NPK + 1] = xn;
      }
    }
    const float r = yi - f;
    chi += (double)r * r;
    if (JtJ) {
#pragma unroll
      for (int a = 0; a < NP; a++) {
        Jtr[a] += (double)J[a] * r;
#pragma unroll
        for (int b = 0; b <= a; b++) JtJ[tri(a, b)] += (double)J[a] * J[b];
      }
    }
  }
  return chi;
}

// Cholesky factorisation and solve of the packed matrix A (overwritten by
// its factor) with right-hand side b (overwritten by the solution).
template <int NP> __device__ bool cholesky_solve(double *A, double *b) {
  constexpr int UF = unroll_np<NP>();
#pragma unroll UF
  for (int j = 0; j < NP; j++) {
    double s = A[tri(j, j)];
#pragma unroll UF
    for (int k = 0; k < j; k++) s -= A[tri(j, k)] * A[tri(j, k)];
    if (!(s > 0)) return false;
    const double dd = sqrt(s);
    A[tri(j, j)] = dd;
#pragma unroll UF
    for (int i = j + 1; i < NP; i++) {
      double t = A[tri(i, j)];
#pragma unroll UF
      for (int k = 0; k < j; k++) t -= A[tri(i, k)] * A[tri(j, k)];
      A[tri(i, j)] = t / dd;
    }
  }
#pragma unroll UF
  for (int i = 0; i < NP; i++) {
    double t = b[i];
#pragma unroll UF
    for (int k = 0; k < i; k++) t -= A[tri(i, k)] * b[k];
    b[i] = t / A[tri(i, i)];
  }
#pragma unroll UF
  for (int i = NP - 1; i >= 0; i--) {
    double t = b[i];
#pragma unroll UF
    for (int k = i + 1; k < NP; k++) t -= A[tri(k, i)] * b[k];
    b[i] = t / A[tri(i, i)];
  }
  return true;
}

// Start values and bounds (see the file comment).  guess: the window's
// n_guesses<MODEL>() expected centres.
template <int MODEL>
__device__ void init_params(const Win &w, double rstep, const double *guess, double *p,
                            double *lo, double *hi) {
  const int n = w.n;
  const double rmean = guess[0];
  const double rmean2 = MODEL == MC_FIT_PV2 ? (double)guess[1] : 0.0;
  if (MODEL == MC_FIT_PV2) {  // pf_fit_doublet_pv, PeakFit.c:1046-1066
    double bg = ((double)w.y[0] + w.y[n - 1]) / 2;
    if (bg < 0) bg = 0;
    double mx = -1e20;
    for (int i = 0; i < n; i++) mx = fmax(mx, (double)w.y[i]);
    double ig = mx - bg;
    if (ig < 1e-6) ig = 1e-6;
    const double mid = (rmean + rmean2) / 2;
    p[0] = rmean; lo[0] = w.xa(0); hi[0] = mid;
    p[1] = rmean2; lo[1] = mid; hi[1] = w.xa(n - 1);
    p[2] = 0.5; lo[2] = 0; hi[2] = 1;
    p[3] = rstep * 3; lo[3] = rstep / 2; hi[3] = rstep * n / 2;
    p[4] = ig / 2; lo[4] = ig / 100; hi[4] = mx * 1.5;
    p[5] = rstep * 3; lo[5] = rstep / 2; hi[5] = rstep * n / 2;
    p[6] = ig / 2; lo[6] = ig / 100; hi[6] = mx * 1.5;
    p[7] = bg; lo[7] = 0; hi[7] = (bg > 0) ? bg * 1.5 : mx * 0.5;
    return;
  }
  if (MODEL == MC_FIT_PV) {
    double bg = ((double)w.y[0] + w.y[n - 1]) / 2;
    if (bg < 0) bg = 0;
    double mx = -1e20;
    for (int i = 0; i < n; i++) mx = fmax(mx, (double)w.y[i]);
    p[0] = rmean; lo[0] = w.xa(0); hi[0] = w.xa(n - 1);
    p[1] = 0.5; lo[1] = 0; hi[1] = 1;
    p[2] = rstep * 3; lo[2] = rstep / 2; hi[2] = rstep * n / 2;
    p[3] = mx - bg; lo[3] = (mx - bg) / 100; hi[3] = mx * 1.5;
    p[4] = bg; lo[4] = 0; hi[4] = (bg > 0) ? bg * 1.5 : mx * 0.5;
    return;
  }
  // TCH family: fitPeaks' per-peak start and bounds (PeakFit.c:380-437)
  constexpr int NPK = tch_peaks<MODEL>();
  double mx = -1e30, mn = 1e30;
  for (int i = 0; i < n; i++) { mx = fmax(mx, (double)w.y[i]); mn = fmin(mn, (double)w.y[i]); }
  const double range = fmax(mx - mn, 1.0);
  int bw = (int)fmin(5.0, (double)(n / 4));
  if (bw < 1) bw = 1;
  double bs = 0;
  for (int i = 0; i < bw; i++) bs += (double)w.y[i] + (double)w.y[n - 1 - i];
  const double bgg = bs / (2.0 * bw);  // pf_estimate_initial_params' background (:215-243)
  for (int k = 0; k < NPK; k++) {
    int pk = 0;
    double best = 1e10;
    for (int i = 0; i < n; i++)
      if (fabs(w.xa(i) - (double)guess[k]) < best) {
        best = fabs(w.xa(i) - (double)guess[k]);
        pk = i;
      }
    double amp = w.y[pk] - bgg;
    if (amp <= 0) amp = w.y[pk];
    const double half = bgg + amp / 2.0;
    int l = pk, r = pk;
    while (l > 0 && w.y[l] > half) l--;
    while (r < n - 1 && w.y[r] > half) r++;
    double fwb = (double)(r - l);
    if (!(fwb > 1.0)) fwb = 2.0;
    const double fw = fwb * rstep, fgi = fw * 0.5;
    double sig0 = (fgi * 100.0) * (fgi * 100.0) / (8.0 * log(2.0));
    if (sig0 < 1e-6) sig0 = 1.0;
    double gam0 = fw * 0.5 * 100.0;
    if (gam0 < 0.01) gam0 = 1.0;
    double a0 = amp * fw * 1.5;
    if (a0 <= 0) a0 = range * rstep;
    const double crange = fmax(fw, (double)n * rstep * 0.5);
    double ahi = range * (double)n * rstep * 2.0;
    if (ahi < 1.0) ahi = 1.0;
    const double shi = (rstep * 200.0 * 100.0) * (rstep * 200.0 * 100.0) / (8.0 * log(2.0));
    const double ghi = rstep * 200.0 * 100.0;
    double *q = p + 4 * k, *ql = lo + 4 * k, *qh = hi + 4 * k;
    q[0] = fmax(fmin(a0, ahi * 0.9), 0.01); ql[0] = 0; qh[0] = ahi;
    q[1] = w.xa(pk); ql[1] = w.xa(pk) - crange; qh[1] = w.xa(pk) + crange;
    q[2] = fmax(fmin(sig0, shi * 0.5), 2e-6); ql[2] = 1e-6; qh[2] = shi;
    q[3] = fmax(fmin(gam0, ghi * 0.5), 0.02); ql[3] = 0.01; qh[3] = ghi;
    for (int j = 0; j < 4; j++) {
      if (q[j] <= ql[j]) q[j] = ql[j] + (qh[j] - ql[j]) * 0.01;
      if (q[j] >= qh[j]) q[j] = qh[j] - (qh[j] - ql[j]) * 0.01;
    }
  }
  double *b = p + 4 * NPK, *bl = lo + 4 * NPK, *bh = hi + 4 * NPK;
  b[0] = bgg; bl[0] = fmin(mn - range * 0.5, -1.0); bh[0] = mx + range;
  if (b[0] <= bl[0]) b[0] = bl[0] + 0.01;
  if (b[0] >= bh[0]) b[0] = bh[0] - 0.01;
  b[1] = 0.0; bl[1] = -range; bh[1] = range;
}

template <int MODEL>
__device__ void fit_one(const Win &w, double rstep, const double *guess, int max_iter,
                        double *out) {
  constexpr int NP = n_params<MODEL>();
  constexpr int NR = n_rows<MODEL>();
  double p[NP], lo[NP], hi[NP];
  for (int k = 0; k < MC_FIT_OUT * NR; k++) out[k] = 0;
  int valid = 0;
  for (int i = 0; i < w.n; i++) valid += !isnan(w.y[i]);
  if (w.n < NP || valid < NP) {
    for (int k = 0; k < NR; k++) out[MC_FIT_OUT * k + 7] = -1;
    return;
  }
  init_params<MODEL>(w, rstep, guess, p, lo, hi);
  constexpr int NPK = tch_peaks<MODEL>();
#pragma unroll
  for (int k = 0; k < NPK; k++) {  // sig -> u = sqrt(sig)
    p[4 * k + 2] = sqrt(p[4 * k + 2]);
    lo[4 * k + 2] = sqrt(lo[4 * k + 2]);
    hi[4 * k + 2] = sqrt(hi[4 * k + 2]);
  }
  constexpr int UF = unroll_np<NP>();
  double JtJ[n_tri<NP>()], Jtr[NP], M[n_tri<NP>()], d[NP], q[NP];
  double chi = chi2_normal<MODEL>(w, p, JtJ, Jtr);
  double lam = 1e-3;
  int it = 0, stall = 0;
  for (; it < max_iter; it++) {
    bool accepted = false;
    for (int tries = 0; tries < 12 && !accepted; tries++) {
#pragma unroll UF
      for (int a = 0; a < n_tri<NP>(); a++) M[a] = JtJ[a];
#pragma unroll UF
      for (int a = 0; a < NP; a++) {
        M[tri(a, a)] += lam * fmax(JtJ[tri(a, a)], 1e-30);
        d[a] = Jtr[a];
      }
#pragma unroll UF
      for (int a = 0; a < NP; a++) {  // active set: hold a bound that pushes outwards
        if ((p[a] <= lo[a] && Jtr[a] < 0) || (p[a] >= hi[a] && Jtr[a] > 0)) {
#pragma unroll UF
          for (int b = 0; b < a; b++) M[tri(a, b)] = 0;
#pragma unroll UF
          for (int b = a + 1; b < NP; b++) M[tri(b, a)] = 0;
          M[tri(a, a)] = 1;
          d[a] = 0;
        }
      }
      if (!cholesky_solve<NP>(M, d)) {
        lam *= 10;
        continue;
      }
      double step = 0, mag = 0;
#pragma unroll UF
      for (int a = 0; a < NP; a++) {
        q[a] = fmin(fmax(p[a] + d[a], lo[a]), hi[a]);
        step += (q[a] - p[a]) * (q[a] - p[a]);
        mag += p[a] * p[a];
      }
      const double c2 = chi2_normal<MODEL>(w, q, nullptr, nullptr);
      if (c2 <= chi) {
        const double dchi = chi - c2;
#pragma unroll UF
        for (int a = 0; a < NP; a++) p[a] = q[a];
        chi = chi2_normal<MODEL>(w, p, JtJ, Jtr);
        lam = fmax(lam / 10, 1e-12);
        accepted = true;
        // FP32 residuals: converged when a step no longer moves the
        // parameters or the objective beyond single-precision noise.
        stall = (step <= 1e-20 * (mag + 1e-30) || dchi <= 1e-9 * chi) ? stall + 1 : 0;
      } else {
        lam *= 10;
      }
    }
    if (!accepted || stall >= 2) break;
  }
  if (!isfinite(chi)) {
    for (int k = 0; k < NR; k++) out[MC_FIT_OUT * k + 7] = -2;
    return;
  }
  const int n_used = valid;
  if (MODEL == MC_FIT_PV2) {  // pf_fit_doublet_pv outputs, PeakFit.c:1093-1100
    const double rms = sqrt(chi / n_used);
    out[0] = p[0];
    out[1] = p[1];
    out[2] = rms > 0 ? p[4] / rms : 1.0;
    out[3] = rms > 0 ? p[6] / rms : 1.0;
    out[4] = p[3];
    out[5] = p[5];
    out[6] = p[2];
  } else if (MODEL == MC_FIT_PV) {
    const double G = p[2], mu = p[1], imax = p[3];
    out[0] = imax * G * (mu * M_PI / 2 + (1 - mu) * sqrt(M_PI / (4 * log(2.0))));
    out[1] = p[0];
    out[4] = G;
    out[5] = mu;
    const double rms = sqrt(chi / n_used);
    out[6] = rms > 0 ? imax / rms : 1.0;  // SNR, pf_fit_singlet_pv :973-974
  } else {  // fitPeaks' rows, one per peak (PeakFit.c:505-527)
    const double dof_chi = chi / fmax((double)(n_used - NP), 1.0);  // chi^2 / dof
    for (int k = 0; k < NR; k++) {
      const Tch t = tch(p[4 * k + 2], p[4 * k + 3]);
      double *o = out + MC_FIT_OUT * k;
      o[0] = p[4 * k];
      o[1] = p[4 * k + 1];
      o[2] = p[4 * k + 2] * p[4 * k + 2];
      o[3] = p[4 * k + 3];
      o[4] = t.F;
      o[5] = t.eta;
      o[6] = dof_chi;
      o[7] = it + 1;
    }
    return;
  }
  out[7] = it + 1;
}

template <int MODEL>
__global__ __launch_bounds__(64) void k_fit_windows(int nw, int npts, const float *__restrict__ x,
                                                    const float *__restrict__ y,
                                                    const float *__restrict__ guess, double rstep,
                                                    int max_iter, float *__restrict__ xr,
                                                    double *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nw) return;
  const float *xi = x + (int64_t)i * npts;
  float *ri = xr + (int64_t)i * npts;
  const double x0 = xi[0];
  for (int q = 0; q < npts; q++) ri[q] = (float)((double)xi[q] - x0);
  Win w;
  w.xr = ri;
  w.y = y + (int64_t)i * npts;
  w.n = npts;
  w.x0 = x0;
  double g[n_guesses<MODEL>()];
  for (int k = 0; k < n_guesses<MODEL>(); k++) g[k] = guess[(int64_t)n_guesses<MODEL>() * i + k];
  fit_one<MODEL>(w, rstep, g, max_iter, out + (int64_t)i * MC_FIT_OUT * n_rows<MODEL>());
}

// Windows cut from cakes: one thread per (frame, eta bin, ring).  The
// ROI of ring k is R bins [lo_k, hi_k] (host-computed like
// IntegratorZarrOMP.c:1966-1988); eta profiles with fewer than 5 nonzero
// values are skipped (:1943-1951).
template <int MODEL>
__global__ __launch_bounds__(64) void k_fit_rings(
    const float *__restrict__ cakes, int n_frames, int n_r, int n_eta,
    const float *__restrict__ xrel, const double *__restrict__ x0s,
    const int32_t *__restrict__ roi, const double *__restrict__ ring,
    int n_rings, int max_pts, const int32_t *__restrict__ nonzero, double rstep, int max_iter,
    float *__restrict__ scratch, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)n_frames * n_eta * n_rings;
  if (i >= total) return;
  const int k = (int)(i % n_rings);
  const int e = (int)((i / n_rings) % n_eta);
  const int f = (int)(i / ((int64_t)n_rings * n_eta));
  double *o = out + i * MC_FIT_OUT;
  const int lo = roi[2 * k], hi = roi[2 * k + 1];
  if (lo < 0 || hi - lo < 5 || nonzero[(int64_t)f * n_eta + e] < 5) {
    for (int q = 0; q < MC_FIT_OUT; q++) o[q] = 0;
    o[7] = -3;
    return;
  }
  const int n = hi - lo + 1;
  float *ys = scratch + i * max_pts;
  const float *c = cakes + (int64_t)f * n_r * n_eta;
  for (int q = 0; q < n; q++) ys[q] = c[(int64_t)(lo + q) * n_eta + e];
  Win w;
  w.xr = xrel + (int64_t)k * max_pts;
  w.y = ys;
  w.n = n;
  w.x0 = x0s[k];
  fit_one<MODEL>(w, rstep, ring + k, max_iter, o);
}

__global__ void k_nonzero(const float *__restrict__ cakes, int n_frames, int n_r, int n_eta,
                          int32_t *__restrict__ nz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n_frames * n_eta) return;
  const int e = (int)(i % n_eta), f = (int)(i / n_eta);
  const float *c = cakes + (int64_t)f * n_r * n_eta + e;
  int cnt = 0;
  for (int r = 0; r < n_r; r++) cnt += c[(int64_t)r * n_eta] != 0.f;
  nz[i] = cnt;
}

}  // namespace
}  // namespace mc

using namespace mc;

// ---- the allocation-free window fits ----
// Scratch of mc_fit_windows_noalloc: the window-relative x values of every
// window, [n_windows][n_points] float, kMaxScratchFloats at most (32 MiB).
// Each GPU holds its own copy for the life of the process; host code uses
// only the per-device address from hipGetSymbolAddress, never the name.
namespace {
constexpr int64_t kMaxScratchFloats = int64_t(8) << 20;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch_dev[kMaxDevices] = {};
}  // namespace
static __device__ __align__(256) float scratch[kMaxScratchFloats];

// The device address of this device's scratch array (looked up once per
// device, with the device current).
static int scratch_ptr(int device, float **out) {
  MC_CHECK(device >= 0 && device < kMaxDevices, MC_ERR_ARG, "device %d", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void *p = nullptr;
    MC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MC_CHECK(p, MC_ERR_HIP, "no device address for the fit scratch array");
    g_scratch_dev[device] = static_cast<float *>(p);
  }
  *out = g_scratch_dev[device];
  return MC_OK;
}

// Guesses a window of a model takes, result rows it writes; 0 for an
// unknown model.
static int fit_guesses(int model) {
  if (model == MC_FIT_PV || model == MC_FIT_TCH) return 1;
  if (model == MC_FIT_PV2) return 2;
  if (model >= MC_FIT_TCH2 && model <= MC_FIT_TCH_PEAKS(MC_FIT_MAX_PEAKS)) return model - 1;
  return 0;
}
static int fit_rows(int model) { return model == MC_FIT_PV2 ? 1 : fit_guesses(model); }

static int launch_fit_windows(hipStream_t s, int model, int nw, int np, const float *x,
                              const float *y, const float *g, double rstep, int max_iter,
                              float *xr, double *out) {
  const unsigned blocks = (unsigned)((nw + 63) / 64);
#define MC_FIT_LAUNCH(M)                                                                  \
  hipLaunchKernelGGL(k_fit_windows<M>, dim3(blocks), dim3(64), 0, s, nw, np, x, y, g,  \
                     rstep, max_iter, xr, out)
  switch (model) {
    case MC_FIT_PV: MC_FIT_LAUNCH(MC_FIT_PV); break;
    case MC_FIT_PV2: MC_FIT_LAUNCH(MC_FIT_PV2); break;
    case MC_FIT_TCH: MC_FIT_LAUNCH(MC_FIT_TCH); break;
    case MC_FIT_TCH_PEAKS(2): MC_FIT_LAUNCH(MC_FIT_TCH_PEAKS(2)); break;
    case MC_FIT_TCH_PEAKS(3): MC_FIT_LAUNCH(MC_FIT_TCH_PEAKS(3)); break;
    case MC_FIT_TCH_PEAKS(4): MC_FIT_LAUNCH(MC_FIT_TCH_PEAKS(4)); break;
    case MC_FIT_TCH_PEAKS(5): MC_FIT_LAUNCH(MC_FIT_TCH_PEAKS(5)); break;
    case MC_FIT_TCH_PEAKS(6): MC_FIT_LAUNCH(MC_FIT_TCH_PEAKS(6)); break;
    case MC_FIT_TCH_PEAKS(7): MC_FIT_LAUNCH(MC_FIT_TCH_PEAKS(7)); break;
    case MC_FIT_TCH_PEAKS(8): MC_FIT_LAUNCH(MC_FIT_TCH_PEAKS(8)); break;
    default: set_error("model %d", model); return MC_ERR_ARG;
  }
#undef MC_FIT_LAUNCH
  MC_HIP(hipGetLastError());
  return MC_OK;
}

extern "C" int mc_fit_windows_noalloc(mc_ctx *ctx, int32_t model, int32_t n_windows,
                                      int32_t n_points, const float *x_dev, const float *y_dev,
                                      const float *guess_dev, double rstep, int32_t max_iter,
                                      double *out_dev) {
  MC_CHECK(ctx && x_dev && y_dev && guess_dev && out_dev, MC_ERR_ARG, "null argument");
  MC_CHECK(fit_guesses(model) > 0, MC_ERR_ARG, "model %d", model);
  MC_CHECK(n_windows >= 0 && n_points >= 2 && n_points <= 4096 && max_iter >= 1, MC_ERR_ARG,
           "bad sizes n_windows=%d n_points=%d", n_windows, n_points);
  MC_CHECK((int64_t)n_windows * n_points <= kMaxScratchFloats, MC_ERR_ARG,
           "%d windows of %d points need %lld scratch floats, the allocation-free fit holds "
           "%lld (use mc_fit_windows)",
           n_windows, n_points, (long long)n_windows * n_points, (long long)kMaxScratchFloats);
  if (n_windows == 0) return MC_OK;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  MC_HIP(hipSetDevice(ctx->device));
  float *xr = nullptr;
  int rc = scratch_ptr(ctx->device, &xr);
  if (rc) return rc;
  return launch_fit_windows(ctx->stream, model, n_windows, n_points, x_dev, y_dev, guess_dev,
                            rstep, max_iter, xr, out_dev);
}

extern "C" int mc_fit_windows(mc_ctx *ctx, int32_t model, int32_t n_windows, int32_t n_points,
                              const float *x, const float *y, const float *peak_guess,
                              double rstep, int32_t max_iter, double *out, int32_t on_device) {
  MC_CHECK(ctx && x && y && peak_guess && out, MC_ERR_ARG, "null argument");
  MC_CHECK(fit_guesses(model) > 0, MC_ERR_ARG, "model %d", model);
  MC_CHECK(n_windows >= 0 && n_points >= 2 && n_points <= 4096 && max_iter >= 1, MC_ERR_ARG,
           "bad sizes n_windows=%d n_points=%d", n_windows, n_points);
  if (n_windows == 0) return MC_OK;
  std::lock_guard<std::recursive_mutex> lk(ctx->mu);
  MC_HIP(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const size_t wb = (size_t)n_windows * n_points * 4;
  const size_t ob = (size_t)n_windows * fit_rows(model) * MC_FIT_OUT * 8;
  const size_t gb = (size_t)n_windows * fit_guesses(model) * 4;
  DBuf dx, dy, dg, dout, dxr;
  const float *px = x, *py = y, *pg = peak_guess;
  double *po = out;
  int rc;
  if (!on_device) {
    if ((rc = dx.alloc(wb)) || (rc = dy.alloc(wb)) || (rc = dg.alloc(gb)) ||
        (rc = dout.alloc(ob)))
      return rc;
    MC_HIP(hipMemcpyAsync(dx.p, x, wb, hipMemcpyHostToDevice, s));
    MC_HIP(hipMemcpyAsync(dy.p, y, wb, hipMemcpyHostToDevice, s));
    MC_HIP(hipMemcpyAsync(dg.p, peak_guess, gb, hipMemcpyHostToDevice, s));
    px = dx.as<float>();
    py = dy.as<float>();
    pg = dg.as<float>();
    po = dout.as<double>();
  }
  if ((rc = dxr.alloc(wb))) return rc;
  if ((rc = launch_fit_windows(s, model, n_windows, n_points, px, py, pg, rstep, max_iter,
                               dxr.as<float>(), po)))
    return rc;
  if (!on_device) MC_HIP(hipMemcpyAsync(out, po, ob, hipMemcpyDeviceToHost, s));
  MC_HIP(hipStreamSynchronize(s));
  return MC_OK;
}

// The x values of the cake rows are the R bin centres; the ROI padding is
// pad times the first bin width (RBinSize in R-mode).
extern "C" int mc_fit_cake_rings(mc_map *m, const float *cakes_dev, int32_t n_frames,
                                 const double *r_centers, const double *ring_r_bins,
                                 int32_t n_rings, int32_t pad, int32_t model, int32_t max_iter,
                                 double *out, int32_t out_on_device) {
  MC_CHECK(m && cakes_dev && r_centers && ring_r_bins && out && n_frames >= 1 &&
               n_rings >= 1 && pad >= 1,
           MC_ERR_ARG, "bad arguments");
  MC_CHECK(model == MC_FIT_PV || model == MC_FIT_TCH, MC_ERR_ARG, "model %d", model);
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  MC_HIP(hipSetDevice(m->ctx->device));
  hipStream_t s = m->ctx->stream;
  const int n_r = m->n_r, n_eta = m->n_eta;
  const std::vector<double> rcen(r_centers, r_centers + n_r);
  const double rstep = n_r > 1 ? rcen[1] - rcen[0] : 1.0;
  // ROI per ring, IntegratorZarrOMP.c:1966-1988.
  std::vector<int32_t> roi(2 * n_rings);
  std::vector<double> ring(n_rings), x0s(n_rings, 0.0);
  int max_pts = 1;
  for (int k = 0; k < n_rings; k++) {
    const double loR = ring_r_bins[k] - pad * rstep, hiR = ring_r_bins[k] + pad * rstep;
    int lo = -1, hi = -1;
    for (int r = 0; r < n_r; r++) {
      if (rcen[r] >= loR && lo < 0) lo = r;
      if (rcen[r] <= hiR) hi = r;
    }
    roi[2 * k] = lo;
    roi[2 * k + 1] = hi;
    ring[k] = ring_r_bins[k];
    if (lo >= 0 && hi >= lo) max_pts = std::max(max_pts, hi - lo + 1);
  }
  std::vector<float> xrel((size_t)n_rings * max_pts, 0.f);
  for (int k = 0; k < n_rings; k++) {
    const int lo = roi[2 * k], hi = roi[2 * k + 1];
    if (lo < 0 || hi < lo) continue;
    x0s[k] = rcen[lo];
    for (int q = 0; q <= hi - lo; q++) xrel[(size_t)k * max_pts + q] = (float)(rcen[lo + q] - rcen[lo]);
  }
  const int64_t total = (int64_t)n_frames * n_eta * n_rings;
  DBuf droi, dring, dxrel, dx0, dnz, dscr, dout;
  int rc;
  if ((rc = droi.alloc(roi.size() * 4)) || (rc = dring.alloc(n_rings * 8)) ||
      (rc = dxrel.alloc(xrel.size() * 4)) || (rc = dx0.alloc(n_rings * 8)) ||
      (rc = dnz.alloc((size_t)n_frames * n_eta * 4)) ||
      (rc = dscr.alloc((size_t)total * max_pts * 4)))
    return rc;
  MC_HIP(hipMemcpyAsync(droi.p, roi.data(), roi.size() * 4, hipMemcpyHostToDevice, s));
  MC_HIP(hipMemcpyAsync(dring.p, ring.data(), n_rings * 8, hipMemcpyHostToDevice, s));
  MC_HIP(hipMemcpyAsync(dxrel.p, xrel.data(), xrel.size() * 4, hipMemcpyHostToDevice, s));
  MC_HIP(hipMemcpyAsync(dx0.p, x0s.data(), n_rings * 8, hipMemcpyHostToDevice, s));
  double *po = out;
  if (!out_on_device) {
    if ((rc = dout.alloc((size_t)total * MC_FIT_OUT * 8))) return rc;
    po = dout.as<double>();
  }
  hipLaunchKernelGGL(k_nonzero, dim3((unsigned)(((int64_t)n_frames * n_eta + 255) / 256)),
                     dim3(256), 0, s, cakes_dev, n_frames, n_r, n_eta, dnz.as<int32_t>());
  const unsigned blocks = (unsigned)((total + 63) / 64);
  if (model == MC_FIT_PV)
    hipLaunchKernelGGL(k_fit_rings<MC_FIT_PV>, dim3(blocks), dim3(64), 0, s, cakes_dev,
                       n_frames, n_r, n_eta, dxrel.as<float>(), dx0.as<double>(),
                       droi.as<int32_t>(), dring.as<double>(), n_rings, max_pts,
                       dnz.as<int32_t>(), rstep, max_iter, dscr.as<float>(), po);
  else
    hipLaunchKernelGGL(k_fit_rings<MC_FIT_TCH>, dim3(blocks), dim3(64), 0, s, cakes_dev,
                       n_frames, n_r, n_eta, dxrel.as<float>(), dx0.as<double>(),
                       droi.as<int32_t>(), dring.as<double>(), n_rings, max_pts,
                       dnz.as<int32_t>(), rstep, max_iter, dscr.as<float>(), po);
  MC_HIP(hipGetLastError());
  if (!out_on_device)
    MC_HIP(hipMemcpyAsync(out, po, (size_t)total * MC_FIT_OUT * 8, hipMemcpyDeviceToHost, s));
  MC_HIP(hipStreamSynchronize(s));
  return MC_OK;
}
