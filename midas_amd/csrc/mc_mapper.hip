// mc_mapper.hip -- GPU FP64 DetectorMapper (gfx950).
//
// One thread per detector pixel computes every (pixel, sub-pixel, bin)
// overlap of the Green's-theorem mapping of MapperCore.c:45-424: pixel
// corners through the forward geometry (DetectorGeometry.c:95-149), the
// candidate bins that the pixel's (R, eta) box overlaps (MapperCore.c
// :180-209), the exact polygon clip of the pixel quad with each annular
// sector (:226-333, DetectorGeometry.c:432-609) and the solid-angle and
// polarisation corrections (:334-348).
//
// Bin membership must agree with the CPU mapper bit-for-bit, so this file
// is compiled with FP contraction off and keeps the CPU operation order;
// the candidate search replaces the reference's linear scan by a binary
// search for the first/last candidate followed by the reference's own
// predicate on each bin in that range, which selects the same set.
//
// Two passes: count entries per pixel, then (after a host scan) write them
// pixel-major.  The host turns the pixel-major list into the bin-major
// Map.bin order with a stable counting sort, which reproduces the
// reference's single-thread within-bin order (pixel loop order).
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include <array>
#include <thread>

#include "mc_internal.h"

namespace mc {
namespace {

constexpr double kEps = 1E-6;
constexpr double kD2R = M_PI / 180.0;
constexpr double kR2D = 180.0 / M_PI;
constexpr int kQ[4] = {0, 1, 3, 2};
constexpr int kMaxV = 40;

struct GeoDev {
  mc_geometry g;
  double T[9];
  const double *RLo, *RHi, *ELo, *EHi;
  int nR, nE;
  const double *mask, *dY, *dZ, *resid;
  const mc_panel *panels;
  int n_panels;
};

__host__ __device__ __forceinline__ double eta_of(double y, double z) {
  return kR2D * atan2(-y, z);
}
__host__ __device__ __forceinline__ double sgn(double x) {
  return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0);
}
__host__ __device__ __forceinline__ bool between(double v, double lo, double hi) {
  return v - kEps <= hi && v + kEps >= lo;
}

__host__ __device__ double residual_at(const GeoDev &G, double Y, double Z) {
  if (!G.resid) return 0.0;
  const int NY = G.g.NrPixelsY, NZ = G.g.NrPixelsZ;
  double y = Y, z = Z;
  if (y < 0.0) y = 0.0;
  if (z < 0.0) z = 0.0;
  if (y >= NY - 1.0) y = NY - 1.001;
  if (z >= NZ - 1.0) z = NZ - 1.001;
  int y0 = (int)y, z0 = (int)z;
  double fy = y - y0, fz = z - z0;
  const double *m = G.resid;
  return m[z0 * NY + y0] * (1 - fy) * (1 - fz) + m[z0 * NY + y0 + 1] * fy * (1 - fz) +
         m[(z0 + 1) * NY + y0] * (1 - fy) * fz + m[(z0 + 1) * NY + y0 + 1] * fy * fz;
}

// Forward geometry; operation order as DetectorGeometry.c:95-149.
__host__ __device__ void to_reta(const GeoDev &G, double Y, double Z, double dLsd, double dP2,
                        double &Rout, double &Eout) {
  const mc_geometry &g = G.g;
  const double px = g.pxY;
  const double *p = g.p;
  const double *T = G.T;
  double L = g.Lsd + dLsd;
  double p2 = p[2] + dP2;
  double Yc = (-Y + g.yCen) * px;
  double Zc = (Z - g.zCen) * px;
  double v1 = T[0] * 0 + T[1] * Yc + T[2] * Zc;
  double v2 = T[3] * 0 + T[4] * Yc + T[5] * Zc;
  double v3 = T[6] * 0 + T[7] * Yc + T[8] * Zc;
  double X = L + v1;
  double Rad = (L / X) * sqrt(v2 * v2 + v3 * v3);
  double et = eta_of(v2, v3);
  double rn = Rad / g.RhoD;
  double e90 = 90 - et;
  double D = (p[0] * pow(rn, 2.0) * cos(kD2R * (2 * e90 + p[6]))) +
             (p[1] * pow(rn, 4.0) * cos(kD2R * (4 * e90 + p[3]))) + (p2 * pow(rn, 2.0));
  D += p[4] * pow(rn, 6.0);
  D += p[5] * pow(rn, 4.0);
  D += p[7] * pow(rn, 4.0) * cos(kD2R * (e90 + p[8]));
  D += p[9] * pow(rn, 3.0) * cos(kD2R * (3 * e90 + p[10]));
  D += p[11] * pow(rn, 5.0) * cos(kD2R * (5 * e90 + p[12]));
  D += p[13] * pow(rn, 6.0) * cos(kD2R * (6 * e90 + p[14]));
  D += 1;
  double Rt = Rad * D / px;
  Rt = Rt * (g.Lsd / L);
  if (g.Parallax != 0.0) {
    double tth = atan(Rad / L);
    Rt += g.Parallax * sin(tth) / px;
  }
  Rt += residual_at(G, Y, Z);
  Rout = Rt;
  Eout = et;
}

__host__ __device__ __forceinline__ void to_yz(double R, double eta, double &y, double &z) {
  y = -R * sin(eta * kD2R);
  z = R * cos(eta * kD2R);
}

__host__ __device__ int circle_hits(double y1, double z1, double y2, double z2, double R,
                           double h[2][2]) {
  double dy = y2 - y1, dz = z2 - z1;
  double a = dy * dy + dz * dz;
  if (a < 1e-30) return 0;
  double b = 2.0 * (y1 * dy + z1 * dz);
  double c = y1 * y1 + z1 * z1 - R * R;
  double disc = b * b - 4.0 * a * c;
  if (disc < 0) return 0;
  double sd = sqrt(disc), i2a = 0.5 / a;
  int n = 0;
  double t1 = (-b - sd) * i2a;
  if (t1 >= -kEps && t1 <= 1.0 + kEps) {
    double t = fmax(0.0, fmin(1.0, t1));
    h[n][0] = y1 + t * dy;
    h[n][1] = z1 + t * dz;
    n++;
  }
  double t2 = (-b + sd) * i2a;
  if (t2 >= -kEps && t2 <= 1.0 + kEps && fabs(t2 - t1) > 1e-12) {
    double t = fmax(0.0, fmin(1.0, t2));
    h[n][0] = y1 + t * dy;
    h[n][1] = z1 + t * dz;
    n++;
  }
  return n;
}

__host__ __device__ bool ray_hit(double y1, double z1, double y2, double z2, double etaDeg,
                        double &hy, double &hz) {
  double er = etaDeg * kD2R;
  double ce = cos(er), se = sin(er);
  double dy = y2 - y1, dz = z2 - z1;
  double den = dy * ce + dz * se;
  if (fabs(den) < 1e-30) return false;
  double t = -(y1 * ce + z1 * se) / den;
  if (t < -kEps || t > 1.0 + kEps) return false;
  t = fmax(0.0, fmin(1.0, t));
  hy = y1 + t * dy;
  hz = z1 + t * dz;
  if ((-se) * hy + ce * hz < 0) return false;
  return true;
}

__host__ __device__ bool in_quad(double py, double pz, const double q[4][2]) {
  int pos = 0, neg = 0;
  for (int e = 0; e < 4; e++) {
    int a = kQ[e], b = kQ[(e + 1) % 4];
    double ey = q[b][0] - q[a][0], ez = q[b][1] - q[a][1];
    double cr = ey * (pz - q[a][1]) - ez * (py - q[a][0]);
    if (cr > 0) pos++;
    else if (cr < 0) neg++;
  }
  return pos == 0 || neg == 0;
}

// Angular comparator of the reference's polygon sort (positive: a after b).
__host__ __device__ __forceinline__ int cmp_angle(double ax_, double ay_, double bx_, double by_,
                                         double cx, double cy) {
  double ax = ax_ - cx, ay = ay_ - cy, bx = bx_ - cx, by = by_ - cy;
  if (ax >= 0 && bx < 0) return 1;
  if (ax < 0 && bx >= 0) return -1;
  if (ax == 0 && bx == 0) {
    if (ay >= 0 || by >= 0) return ay_ > by_ ? 1 : -1;
    return by_ > ay_ ? 1 : -1;
  }
  double det = ax * by - bx * ay;
  if (det < 0) return 1;
  if (det > 0) return -1;
  double d1 = ax * ax + ay * ay, d2 = bx * bx + by * by;
  return d1 > d2 ? 1 : -1;
}

// Exact area of the clipped polygon (arc edges integrated exactly).
__host__ __device__ double poly_area(double (*v)[2], int n, double Rlo, double Rhi) {
  double cx = 0, cy = 0;
  for (int i = 0; i < n; i++) {
    cx += v[i][0];
    cy += v[i][1];
  }
  cx /= n;
  cy /= n;
  // Stable insertion sort: same sequence as the reference's merge sort.
  for (int i = 1; i < n; i++) {
    double px = v[i][0], py = v[i][1];
    int j = i - 1;
    while (j >= 0 && cmp_angle(v[j][0], v[j][1], px, py, cx, cy) > 0) {
      v[j + 1][0] = v[j][0];
      v[j + 1][1] = v[j][1];
      j--;
    }
    v[j + 1][0] = px;
    v[j + 1][1] = py;
  }
  double r2lo = Rlo * Rlo, r2hi = Rhi * Rhi, tol = 1e-6, A = 0;
  for (int i = 0; i < n; i++) {
    int j = (i + 1) % n;
    double s1 = v[i][0] * v[i][0] + v[i][1] * v[i][1];
    double s2 = v[j][0] * v[j][0] + v[j][1] * v[j][1];
    bool onLo = fabs(s1 - r2lo) < tol * r2lo && fabs(s2 - r2lo) < tol * r2lo;
    bool onHi = fabs(s1 - r2hi) < tol * r2hi && fabs(s2 - r2hi) < tol * r2hi;
    if (onLo || onHi) {
      double R = onLo ? Rlo : Rhi;
      double a1 = atan2(v[i][1], v[i][0]);
      double a2 = atan2(v[j][1], v[j][0]);
      double d = a2 - a1;
      if (d > M_PI) d -= 2.0 * M_PI;
      if (d < -M_PI) d += 2.0 * M_PI;
      A += (R * R / 2.0) * d;
    } else {
      A += 0.5 * (v[i][0] * v[j][1] - v[j][0] * v[i][1]);
    }
  }
  return A;
}

// Overlap area of pixel quad cyz with bin [Rlo,Rhi]x[Elo,Ehi]; negative
// return = rejected (-1: <3 vertices, -2: <3 unique, -3: area < 1e-5).
__host__ __device__ double clip_area(const double cyz[4][2], double Rlo, double Rhi, double Elo,
                            double Ehi) {
  double box[4][2];
  to_yz(Rlo, Elo, box[0][0], box[0][1]);
  to_yz(Rlo, Ehi, box[1][0], box[1][1]);
  to_yz(Rhi, Elo, box[2][0], box[2][1]);
  to_yz(Rhi, Ehi, box[3][0], box[3][1]);
  double V[kMaxV][2];
  int nv = 0;
  for (int m = 0; m < 4; m++) {
    double rr = sqrt(cyz[m][0] * cyz[m][0] + cyz[m][1] * cyz[m][1]);
    double ee = eta_of(cyz[m][0], cyz[m][1]);
    if (Elo < -180 && sgn(ee) != sgn(Elo)) ee -= 360;
    if (Ehi > 180 && sgn(ee) != sgn(Ehi)) ee += 360;
    if (rr >= Rlo && rr <= Rhi && ee >= Elo && ee <= Ehi) {
      V[nv][0] = cyz[m][0];
      V[nv][1] = cyz[m][1];
      nv++;
    }
  }
  for (int m = 0; m < 4; m++)
    if (in_quad(box[m][0], box[m][1], cyz)) {
      V[nv][0] = box[m][0];
      V[nv][1] = box[m][1];
      nv++;
    }
  if (nv < 4) {
    for (int e = 0; e < 4; e++) {
      double y1 = cyz[kQ[e]][0], z1 = cyz[kQ[e]][1];
      double y2 = cyz[kQ[(e + 1) % 4]][0], z2 = cyz[kQ[(e + 1) % 4]][1];
      double hh[2][2];
      for (int side = 0; side < 2; side++) {
        int nh = circle_hits(y1, z1, y2, z2, side ? Rhi : Rlo, hh);
        for (int h = 0; h < nh; h++) {
          double eh = eta_of(hh[h][0], hh[h][1]);
          if (Elo < -180 && sgn(eh) != sgn(Elo)) eh -= 360;
          if (Ehi > 180 && sgn(eh) != sgn(Ehi)) eh += 360;
          if (eh >= Elo - kEps && eh <= Ehi + kEps && nv < kMaxV) {
            V[nv][0] = hh[h][0];
            V[nv][1] = hh[h][1];
            nv++;
          }
        }
      }
      for (int side = 0; side < 2; side++) {
        double hy, hz;
        if (ray_hit(y1, z1, y2, z2, side ? Ehi : Elo, hy, hz)) {
          double rh = sqrt(hy * hy + hz * hz);
          if (rh >= Rlo - kEps && rh <= Rhi + kEps && nv < kMaxV) {
            V[nv][0] = hy;
            V[nv][1] = hz;
            nv++;
          }
        }
      }
    }
  }
  if (nv < 3) return -1;
  // Drop duplicates and vertices outside the bin (DetectorGeometry.c:572-609).
  double U[kMaxV][2];
  int nu = 0;
  for (int i = 0; i < nv; i++) {
    bool dup = false;
    for (int j = i + 1; j < nv; j++) {
      double len = sqrt((V[i][0] - V[j][0]) * (V[i][0] - V[j][0]) +
                        (V[i][1] - V[j][1]) * (V[i][1] - V[j][1]));
      if (len == 0) dup = true;
    }
    double rt = sqrt(V[i][0] * V[i][0] + V[i][1] * V[i][1]);
    double et = eta_of(V[i][0], V[i][1]);
    if (!between(et, Elo, Ehi)) {
      if (between(et + 360, Elo, Ehi)) et += 360;
      else if (between(et - 360, Elo, Ehi)) et -= 360;
    }
    if (!between(rt, Rlo, Rhi)) dup = true;
    if (!between(et, Elo, Ehi)) dup = true;
    if (!dup) {
      U[nu][0] = V[i][0];
      U[nu][1] = V[i][1];
      nu++;
    }
  }
  if (nu < 3) return -2;
  double area = poly_area(U, nu, Rlo, Rhi);
  if (area < 1E-5) return -3;
  return area;
}

// First index in [0,n) with a[k] >= x (n if none).
__host__ __device__ int lower_ge(const double *a, int n, double x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] >= x) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}
// Last index with a[k] <= x (-1 if none).
__host__ __device__ int last_le(const double *a, int n, double x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (a[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  return lo - 1;
}

// One pixel of the map (the reference's per-pixel loop body,
// MapperCore.c:45-424): sub-pixel corners, candidate bins, exact clipped
// areas and the record of each overlap.  Out receives flag(bin) for the
// bins a masked pixel contaminates (count pass), skip(kind) for areas the
// clipper rejects, and rec(n, bin, record) for the records (write pass,
// while room(n)).  Returns the pixel's record count.  Compiled for the GPU
// (k_map) and for the host (the verification pass of map_build_gpu).
template <bool WRITE, typename Out>
__host__ __device__ int map_pixel(const GeoDev &G, int64_t p, Out &out) {
  const mc_geometry &g = G.g;
  const int NY = g.NrPixelsY, NZ = g.NrPixelsZ;
  // Reference loop order: Y outer (i), Z inner (j).
  const int i = (int)(p / NZ), j = (int)(p % NZ);
  const int64_t pos = (int64_t)j * NY + i;
  const bool masked = G.mask != nullptr && G.mask[pos] == 1.0;
  double pdY = 0, pdZ = 0, dLsd = 0, dP2 = 0;
  int pi = -1;
  for (int q = 0; q < G.n_panels; q++) {
    const mc_panel &pp = G.panels[q];
    if ((double)i >= pp.yMin && (double)i <= pp.yMax && (double)j >= pp.zMin &&
        (double)j <= pp.zMax) {
      pi = q;
      break;
    }
  }
  if (pi >= 0) {
    const mc_panel &pp = G.panels[pi];
    double yo, zo;
    if (pp.dTheta != 0.0) {
      double a = (M_PI / 180.0) * pp.dTheta;
      double ct = cos(a), st = sin(a);
      double ddy = (double)i - pp.centerY, ddz = (double)j - pp.centerZ;
      yo = pp.centerY + ddy * ct - ddz * st + pp.dY;
      zo = pp.centerZ + ddy * st + ddz * ct + pp.dZ;
    } else {
      yo = (double)i + pp.dY;
      zo = (double)j + pp.dZ;
    }
    pdY = yo - (double)i;
    pdZ = zo - (double)j;
    dLsd = pp.dLsd;
    dP2 = pp.dP2;
  }
  const double ypr = (double)i + (G.dY ? G.dY[pos] : 0.0) + pdY;
  const double zpr = (double)j + (G.dZ ? G.dZ[pos] : 0.0) + pdZ;
  int sp = 1;
  if (g.SubPixelLevel > 1) {
    double rc, ec;
    to_reta(G, ypr, zpr, dLsd, dP2, rc, ec);
    double ae = fabs(ec), W = g.SubPixelCardinalWidth;
    if (ae <= W || fabs(ae - 90.0) <= W || fabs(ae - 180.0) <= W) sp = g.SubPixelLevel;
  }
  // Raw (un-transformed) pixel coordinates stored in Map.bin.
  int ry = i, rz = j;
  for (int t = g.NrTransOpt - 1; t >= 0; t--) {
    int o = g.TransOpt[t];
    if (o == 1) ry = NY - 1 - ry;
    else if (o == 2) rz = NZ - 1 - rz;
    else if (o == 3) {
      int s = ry;
      ry = rz;
      rz = s;
    }
  }
  int n_out = 0;
  for (int si = 0; si < sp; si++) {
    for (int sj = 0; sj < sp; sj++) {
      double oy[2] = {(double)si / sp - 0.5, (double)(si + 1) / sp - 0.5};
      double oz[2] = {(double)sj / sp - 0.5, (double)(sj + 1) / sp - 0.5};
      double cy = ((double)(2 * si + 1) / (2.0 * sp)) - 0.5;
      double cz = ((double)(2 * sj + 1) / (2.0 * sp)) - 0.5;
      double eMin = 1800, eMax = -1800, rMin = 1E8, rMax = -1000;
      double cyz[4][2];
      double Rt, Et;
      for (int k = 0; k < 2; k++)
        for (int l = 0; l < 2; l++) {
          to_reta(G, ypr + oy[k], zpr + oz[l], dLsd, dP2, Rt, Et);
          if (Et < eMin) eMin = Et;
          if (Et > eMax) eMax = Et;
          if (Rt < rMin) rMin = Rt;
          if (Rt > rMax) rMax = Rt;
          to_yz(Rt, Et, cyz[k * 2 + l][0], cyz[k * 2 + l][1]);
        }
      to_reta(G, ypr + cy, zpr + cz, dLsd, dP2, Rt, Et);
      // Candidate R bins: [first RHi >= rMin, last RLo <= rMax].
      int ra = lower_ge(G.RHi, G.nR, rMin), rb = last_le(G.RLo, G.nR, rMax);
      double lo = eMin, hi = eMax;
      if (hi - lo > 180) {
        double s = hi;
        hi = 360 + lo;
        lo = s;
      }
      // Eta candidates: union of the three shifted windows.
      int ea = G.nE, eb = -1;
      for (int sh = -1; sh <= 1; sh++) {
        double l2 = lo + 360.0 * sh, h2 = hi + 360.0 * sh;
        if (sh == 0) {
          l2 = lo;
          h2 = hi;
        }
        int a = lower_ge(G.EHi, G.nE, l2), b = last_le(G.ELo, G.nE, h2);
        if (a <= b) {
          if (a < ea) ea = a;
          if (b > eb) eb = b;
        }
      }
      if (masked) {
        if (si == 0 && sj == 0 && !WRITE)
          for (int r = ra; r <= rb; r++) {
            if (!(G.RHi[r] >= rMin && G.RLo[r] <= rMax)) continue;
            for (int e = ea; e <= eb; e++) {
              bool sel = (G.EHi[e] >= lo && G.ELo[e] <= hi) ||
                         (G.EHi[e] >= lo + 360 && G.ELo[e] <= hi + 360) ||
                         (G.EHi[e] >= lo - 360 && G.ELo[e] <= hi - 360);
              if (sel) out.flag((int64_t)r * G.nE + e);
            }
          }
        continue;
      }
      for (int r = ra; r <= rb; r++) {
        if (!(G.RHi[r] >= rMin && G.RLo[r] <= rMax)) continue;
        const double Rlo = G.RLo[r], Rhi = G.RHi[r];
        for (int e = ea; e <= eb; e++) {
          bool sel = (G.EHi[e] >= lo && G.ELo[e] <= hi) ||
                     (G.EHi[e] >= lo + 360 && G.ELo[e] <= hi + 360) ||
                     (G.EHi[e] >= lo - 360 && G.ELo[e] <= hi - 360);
          if (!sel) continue;
          double area = clip_area(cyz, Rlo, Rhi, G.ELo[e], G.EHi[e]);
          if (area < 0) {
            out.skip(area == -1 ? 0 : area == -2 ? 1 : 2);
            continue;
          }
          if (WRITE && out.room(n_out)) {
            double corr = area;
            if (g.SolidAngleCorrection || g.PolarizationCorrection) {
              double tth = atan(Rt * g.pxY / g.Lsd);
              if (g.SolidAngleCorrection) {
                double c2 = cos(tth);
                corr /= (c2 * c2 * c2);
              }
              if (g.PolarizationCorrection) {
                double s2 = sin(tth);
                double ce = cos(Et * kD2R);
                double pf = 1.0 - g.PolarizationFraction * s2 * s2 * ce * ce;
                if (pf > 1e-6) corr /= pf;
              }
            }
            Rec rec;
            rec.y = (float)ry + (float)cy;
            rec.z = (float)rz + (float)cz;
            rec.frac = corr;
            rec.deltaR = (float)(Rt - (Rlo + Rhi) * 0.5);
            rec.areaWeight = (float)area;
            out.rec(n_out, (int32_t)(r * G.nE + e), rec);
          }
          n_out++;
        }
      }
    }
  }
  return n_out;
}

// Device side of map_pixel's output.
struct DevOut {
  int32_t *flags;
  int32_t *bin;
  Rec *rec_out;
  int cap;
  unsigned long long sk[3] = {0, 0, 0};
  __device__ void flag(int64_t b) { flags[b] = 1; }
  __device__ void skip(int k) { sk[k]++; }
  __device__ bool room(int n) const { return n < cap; }
  __device__ void rec(int n, int32_t b, const Rec &r) {
    bin[n] = b;
    rec_out[n] = r;
  }
};

template <bool WRITE>
__global__ __launch_bounds__(128) void k_map(GeoDev G, int32_t *__restrict__ counts,
                                             const int64_t *__restrict__ offsets,
                                             int32_t *__restrict__ out_bin,
                                             Rec *__restrict__ out_rec,
                                             int32_t *__restrict__ flags,
                                             unsigned long long *__restrict__ skipped) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= (int64_t)G.g.NrPixelsY * G.g.NrPixelsZ) return;
  // The write pass never exceeds the slots the count pass reserved.
  const int64_t w_at = WRITE ? offsets[p] : 0;
  DevOut out{flags, WRITE ? out_bin + w_at : nullptr, WRITE ? out_rec + w_at : nullptr,
             WRITE ? counts[p] : 0};
  const int n_out = map_pixel<WRITE>(G, p, out);
  if (!WRITE) {
    counts[p] = n_out;
    if (out.sk[0]) atomicAdd(&skipped[0], out.sk[0]);
    if (out.sk[1]) atomicAdd(&skipped[1], out.sk[1]);
    if (out.sk[2]) atomicAdd(&skipped[2], out.sk[2]);
  } else if (n_out != counts[p]) {
    atomicAdd(&skipped[3], 1ull);  // count/write passes disagree: reported
  }
}

// Host side of map_pixel's output (the verification pass).
struct HostOut {
  std::vector<int64_t> flagged;
  std::vector<int32_t> bins;
  std::vector<Rec> recs;
  void flag(int64_t b) { flagged.push_back(b); }
  void skip(int) {}
  bool room(int) const { return true; }
  void rec(int, int32_t b, const Rec &r) {
    bins.push_back(b);
    recs.push_back(r);
  }
};

void apply_transform(const mc_geometry *g, std::vector<double> &img) {
  const int NY = g->NrPixelsY, NZ = g->NrPixelsZ;
  for (int t = 0; t < g->NrTransOpt; t++) {
    int o = g->TransOpt[t];
    if (o == 1) {
      for (int r = 0; r < NZ; r++)
        for (int c = 0; c < NY / 2; c++) std::swap(img[(size_t)r * NY + c], img[(size_t)r * NY + NY - c - 1]);
    } else if (o == 2) {
      for (int r = 0; r < NZ / 2; r++)
        for (int c = 0; c < NY; c++) std::swap(img[(size_t)r * NY + c], img[(size_t)(NZ - r - 1) * NY + c]);
    } else if (o == 3 && NY == NZ) {
      for (int r = 0; r < NZ; r++)
        for (int c = r + 1; c < NY; c++) std::swap(img[(size_t)r * NY + c], img[(size_t)c * NY + r]);
    }
  }
}

}  // namespace

// Bin edges as DetectorMapper.c:712-755 (Q-mode rewrites RMin/RMax).
static void bin_edges(mc_geometry *g, int &nR, int &nE, std::vector<double> &RLo,
                      std::vector<double> &RHi, std::vector<double> &ELo,
                      std::vector<double> &EHi) {
  nE = (int)ceil((g->EtaMax - g->EtaMin) / g->EtaBinSize);
  const bool q = g->QBinSize > 0 && g->Wavelength > 0;
  if (q) {
    if (g->QMin > 0) g->RMin = (g->Lsd / g->pxY) * tan(2.0 * asin(g->QMin * g->Wavelength / (4.0 * M_PI)));
    if (g->QMax > 0) g->RMax = (g->Lsd / g->pxY) * tan(2.0 * asin(g->QMax * g->Wavelength / (4.0 * M_PI)));
    nR = (int)ceil((g->QMax - g->QMin) / g->QBinSize);
  } else {
    nR = (int)ceil((g->RMax - g->RMin) / g->RBinSize);
  }
  RLo.resize(nR);
  RHi.resize(nR);
  ELo.resize(nE);
  EHi.resize(nE);
  for (int e = 0; e < nE; e++) {
    ELo[e] = g->EtaBinSize * e + g->EtaMin;
    EHi[e] = g->EtaBinSize * (e + 1) + g->EtaMin;
  }
  for (int r = 0; r < nR; r++) {
    if (q) {
      double ql = g->QMin + g->QBinSize * r, qh = g->QMin + g->QBinSize * (r + 1);
      RLo[r] = (g->Lsd / g->pxY) * tan(2.0 * asin(ql * g->Wavelength / (4.0 * M_PI)));
      RHi[r] = (g->Lsd / g->pxY) * tan(2.0 * asin(qh * g->Wavelength / (4.0 * M_PI)));
    } else {
      RLo[r] = g->RBinSize * r + g->RMin;
      RHi[r] = g->RBinSize * (r + 1) + g->RMin;
    }
  }
}

int map_build_gpu(Ctx *ctx, const mc_geometry *gin, const double *mask,
                  const double *distY, const double *distZ, const double *residual,
                  const mc_panel *panels, int n_panels, Map *m) {
  MC_CHECK(gin && m, MC_ERR_ARG, "map_build: null argument");
  mc_geometry g = *gin;
  MC_CHECK(g.NrPixelsY > 1 && g.NrPixelsZ > 1, MC_ERR_ARG, "detector must be at least 2x2");
  MC_CHECK(g.SubPixelLevel >= 1 && g.SubPixelLevel <= 16, MC_ERR_ARG,
           "SubPixelLevel %d outside 1..16", g.SubPixelLevel);
  MC_CHECK(g.EtaBinSize > 0 && (g.RBinSize > 0 || g.QBinSize > 0), MC_ERR_ARG,
           "bin sizes must be positive");
  MC_CHECK(g.NrTransOpt >= 0 && g.NrTransOpt <= 10, MC_ERR_ARG, "NrTransOpt > 10");
  int nR, nE;
  std::vector<double> RLo, RHi, ELo, EHi;
  bin_edges(&g, nR, nE, RLo, RHi, ELo, EHi);
  MC_CHECK(nR > 0 && nE > 0, MC_ERR_ARG, "empty bin grid (%d x %d)", nR, nE);
  const int NY = g.NrPixelsY, NZ = g.NrPixelsZ;
  const int64_t npx = (int64_t)NY * NZ, nb = (int64_t)nR * nE;
  hipStream_t s = ctx->stream;

  GeoDev G{};
  G.g = g;
  {
    // Tilt matrix Rx(tx) * (Ry(ty) * Rz(tz)), DetectorGeometry.c:64-78.
    double a = kD2R * g.tx, b = kD2R * g.ty, c = kD2R * g.tz;
    double Rx[9] = {1, 0, 0, 0, cos(a), -sin(a), 0, sin(a), cos(a)};
    double Ry[9] = {cos(b), 0, sin(b), 0, 1, 0, -sin(b), 0, cos(b)};
    double Rz[9] = {cos(c), -sin(c), 0, sin(c), cos(c), 0, 0, 0, 1};
    double tmp[9];
    for (int r = 0; r < 3; r++)
      for (int q = 0; q < 3; q++)
        tmp[r * 3 + q] = Ry[r * 3 + 0] * Rz[0 * 3 + q] + Ry[r * 3 + 1] * Rz[1 * 3 + q] + Ry[r * 3 + 2] * Rz[2 * 3 + q];
    for (int r = 0; r < 3; r++)
      for (int q = 0; q < 3; q++)
        G.T[r * 3 + q] = Rx[r * 3 + 0] * tmp[0 * 3 + q] + Rx[r * 3 + 1] * tmp[1 * 3 + q] + Rx[r * 3 + 2] * tmp[2 * 3 + q];
  }
  DBuf dRLo, dRHi, dELo, dEHi, dMask, dDY, dDZ, dRes, dPan, dCounts, dSkip, dFlags;
  int rc;
  if ((rc = dRLo.alloc(nR * 8)) || (rc = dRHi.alloc(nR * 8)) || (rc = dELo.alloc(nE * 8)) ||
      (rc = dEHi.alloc(nE * 8)))
    return rc;
  MC_HIP(hipMemcpy(dRLo.p, RLo.data(), nR * 8, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(dRHi.p, RHi.data(), nR * 8, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(dELo.p, ELo.data(), nE * 8, hipMemcpyHostToDevice));
  MC_HIP(hipMemcpy(dEHi.p, EHi.data(), nE * 8, hipMemcpyHostToDevice));
  G.RLo = dRLo.as<double>();
  G.RHi = dRHi.as<double>();
  G.ELo = dELo.as<double>();
  G.EHi = dEHi.as<double>();
  G.nR = nR;
  G.nE = nE;
  // Host copies (transformed like the device ones) serve the verification.
  std::vector<double> hMask, hDY, hDZ, hRes;
  auto upload_img = [&](const double *src, bool transform, DBuf &dst, const double *&ptr,
                        std::vector<double> &host) -> int {
    ptr = nullptr;
    if (!src) return MC_OK;
    host.assign(src, src + npx);
    if (transform) apply_transform(&g, host);
    int r2 = dst.alloc(npx * 8);
    if (r2) return r2;
    MC_HIP(hipMemcpy(dst.p, host.data(), npx * 8, hipMemcpyHostToDevice));
    ptr = dst.as<double>();
    return MC_OK;
  };
  if ((rc = upload_img(mask, true, dMask, G.mask, hMask))) return rc;
  if ((rc = upload_img(distY, true, dDY, G.dY, hDY))) return rc;
  if ((rc = upload_img(distZ, true, dDZ, G.dZ, hDZ))) return rc;
  if ((rc = upload_img(residual, false, dRes, G.resid, hRes))) return rc;
  G.panels = nullptr;
  G.n_panels = 0;
  if (panels && n_panels > 0) {
    if ((rc = dPan.alloc(sizeof(mc_panel) * n_panels))) return rc;
    MC_HIP(hipMemcpy(dPan.p, panels, sizeof(mc_panel) * n_panels, hipMemcpyHostToDevice));
    G.panels = dPan.as<mc_panel>();
    G.n_panels = n_panels;
  }
  if ((rc = dCounts.alloc(npx * 4)) || (rc = dSkip.alloc(4 * 8)) || (rc = dFlags.alloc(nb * 4)))
    return rc;
  MC_HIP(hipMemsetAsync(dSkip.p, 0, 32, s));
  MC_HIP(hipMemsetAsync(dFlags.p, 0, nb * 4, s));
  const unsigned blocks = (unsigned)((npx + 127) / 128);
  hipLaunchKernelGGL(k_map<false>, dim3(blocks), dim3(128), 0, s, G, dCounts.as<int32_t>(),
                     nullptr, nullptr, nullptr, dFlags.as<int32_t>(),
                     dSkip.as<unsigned long long>());
  MC_HIP(hipGetLastError());
  std::vector<int32_t> counts(npx);
  MC_HIP(hipMemcpyAsync(counts.data(), dCounts.p, npx * 4, hipMemcpyDeviceToHost, s));
  unsigned long long sk[3];
  MC_HIP(hipMemcpyAsync(sk, dSkip.p, 24, hipMemcpyDeviceToHost, s));
  MC_HIP(hipStreamSynchronize(s));
  std::vector<int64_t> offs(npx);
  int64_t nnz = 0;
  for (int64_t q = 0; q < npx; q++) {
    offs[q] = nnz;
    nnz += counts[q];
  }
  MC_CHECK(nnz < (int64_t)INT32_MAX, MC_ERR_ARG, "map has %lld entries; nMap offsets are int32",
           (long long)nnz);
  DBuf dOffs, dBin, dRec;
  if ((rc = dOffs.alloc(npx * 8)) || (rc = dBin.alloc(nnz * 4)) || (rc = dRec.alloc(nnz * sizeof(Rec))))
    return rc;
  MC_HIP(hipMemcpyAsync(dOffs.p, offs.data(), npx * 8, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_map<true>, dim3(blocks), dim3(128), 0, s, G, dCounts.as<int32_t>(),
                     dOffs.as<int64_t>(), dBin.as<int32_t>(), dRec.as<Rec>(), nullptr,
                     dSkip.as<unsigned long long>());
  MC_HIP(hipGetLastError());
  std::vector<int32_t> bins(nnz);
  std::vector<Rec> recs(nnz);
  m->h_flags.assign(nb, 0);
  MC_HIP(hipMemcpyAsync(bins.data(), dBin.p, nnz * 4, hipMemcpyDeviceToHost, s));
  MC_HIP(hipMemcpyAsync(recs.data(), dRec.p, nnz * sizeof(Rec), hipMemcpyDeviceToHost, s));
  MC_HIP(hipMemcpyAsync(m->h_flags.data(), dFlags.p, nb * 4, hipMemcpyDeviceToHost, s));
  unsigned long long mismatch = 0;
  MC_HIP(hipMemcpyAsync(&mismatch, dSkip.as<unsigned long long>() + 3, 8, hipMemcpyDeviceToHost, s));
  MC_HIP(hipStreamSynchronize(s));
  MC_CHECK(mismatch == 0, MC_ERR_STATE, "mapper passes disagree on %llu pixels", mismatch);
  // Host verification (the same per-pixel code built for the host, with
  // the host's libm): a spread sample of pixels is recomputed and compared
  // with the GPU's pixel-major output.  Bin membership must agree; values
  // within 1e-6 px^2 (areas, corrected areas) and 1e-5 px (deltaR).  Where a
  // vertex lies exactly on a bin boundary, the two libms' last-bit
  // differences can flip a case: those pixels are counted, not refused.
  {
    const int64_t nver = std::min<int64_t>(npx, ctx->map_verify);
    int64_t checked = 0, compared = 0, member_bad = 0, value_bad = 0;
    GeoDev H = G;
    H.RLo = RLo.data();
    H.RHi = RHi.data();
    H.ELo = ELo.data();
    H.EHi = EHi.data();
    H.mask = hMask.empty() ? nullptr : hMask.data();
    H.dY = hDY.empty() ? nullptr : hDY.data();
    H.dZ = hDZ.empty() ? nullptr : hDZ.data();
    H.resid = hRes.empty() ? nullptr : hRes.data();
    H.panels = (panels && n_panels > 0) ? panels : nullptr;
    H.n_panels = (panels && n_panels > 0) ? n_panels : 0;
    // Spread over host threads; each keeps its own counts.
    const int nt = nver < 4096 ? 1
                               : (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::array<int64_t, 4>> part(nt, std::array<int64_t, 4>{0, 0, 0, 0});
    auto run = [&](int t) {
      int64_t &chk = part[t][0], &cmp = part[t][1], &mbad = part[t][2], &vbad = part[t][3];
      for (int64_t k = nver * t / nt; k < nver * (t + 1) / nt; k++) {
        const int64_t p =
            nver == npx ? k : (int64_t)(((uint64_t)k * 2654435761ull + 12345ull) % (uint64_t)npx);
        const int pi = (int)(p / NZ), pj = (int)(p % NZ);
        const bool masked = H.mask && H.mask[(int64_t)pj * NY + pi] == 1.0;
        chk++;
        if (masked) {
          HostOut o;
          map_pixel<false>(H, p, o);
          for (int64_t bq : o.flagged)
            if (m->h_flags[bq] != 1) {
              mbad++;
              break;
            }
          continue;
        }
        HostOut o;
        const int n = map_pixel<true>(H, p, o);
        if (n != counts[p]) {
          mbad++;
          continue;
        }
        bool same = true;
        for (int q = 0; q < n; q++) same = same && o.bins[q] == bins[offs[p] + q];
        if (!same) {
          mbad++;
          continue;
        }
        for (int q = 0; q < n; q++) {
          const Rec &x = o.recs[q], &y = recs[offs[p] + q];
          cmp++;
          if (x.y != y.y || x.z != y.z || fabs(x.frac - y.frac) > 1e-6 ||
              fabs((double)x.areaWeight - (double)y.areaWeight) > 1e-6 ||
              fabs((double)x.deltaR - (double)y.deltaR) > 1e-5)
            vbad++;
        }
      }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(run, t);
    run(0);
    for (auto &x : th) x.join();
    for (const auto &q : part) {
      checked += q[0];
      compared += q[1];
      member_bad += q[2];
      value_bad += q[3];
    }
    m->verify[0] = checked;
    m->verify[1] = compared;
    m->verify[2] = member_bad;
    m->verify[3] = value_bad;
  }
  // Stable counting sort pixel-major -> bin-major.
  m->h_pairs.assign(2 * nb, 0);
  for (int64_t q = 0; q < nnz; q++) m->h_pairs[2 * (int64_t)bins[q]]++;
  int64_t run = 0;
  std::vector<int64_t> cur(nb);
  for (int64_t b = 0; b < nb; b++) {
    m->h_pairs[2 * b + 1] = (int32_t)run;
    cur[b] = run;
    run += m->h_pairs[2 * b];
  }
  m->h_rec.resize(nnz);
  for (int64_t q = 0; q < nnz; q++) m->h_rec[cur[bins[q]]++] = recs[q];
  m->ny = NY;
  m->nz = NZ;
  m->n_r = nR;
  m->n_eta = nE;
  m->nnz = nnz;
  m->q_mode = (g.QBinSize > 0 && g.Wavelength > 0) ? 1 : 0;
  for (int q = 0; q < 3; q++) m->skipped[q] = (int64_t)sk[q];
  return MC_OK;
}

}  // namespace mc
