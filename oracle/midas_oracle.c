/*
 * midas_oracle.c -- FP64 CPU restatement of the MIDAS far-field caking path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  See midas_oracle.h for the list
 * of entry points and what pins each of them.  The arithmetic below keeps
 * the reference's operation order on purpose: bin membership must match the
 * reference bit-for-bit, and membership is decided by FP64 comparisons on
 * values whose last bits depend on that order.  Built with gcc -O3 and no
 * -march flag, like the reference (CMakeLists.txt:529), so no FMA
 * contraction happens on x86-64.
 */
#define _GNU_SOURCE
#include "midas_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_EPS 1E-6
#define ORC_D2R (M_PI / 180.0)
#define ORC_R2D (180.0 / M_PI)

/* ------------------------------------------------------------------ */
/* Parameters: DetectorMapper.c:84-114 defaults, :116-375 prefix parse */
/* ------------------------------------------------------------------ */
void orc_params_default(orc_params *P) {
  memset(P, 0, sizeof(*P));
  P->pxY = 200;
  P->pxZ = 200;
  P->yCen = 1024;
  P->zCen = 1024;
  P->Lsd = 1e6;
  P->RhoD = 2e5;
  P->EtaBinSize = 5;
  P->RBinSize = 0.25;
  P->RMax = 1524;
  P->RMin = 10;
  P->EtaMax = 180;
  P->EtaMin = -180;
  P->NrPixelsY = 2048;
  P->NrPixelsZ = 2048;
  P->SubPixelLevel = 1;
  P->SubPixelCardinalWidth = 10.0;
  P->PolarizationFraction = 0.99;
}

static int starts(const char *line, const char *key) {
  return strncmp(line, key, strlen(key)) == 0;
}

/* Reads "<key> <double>" like sscanf("%s %lf"): value left unchanged when
 * the second token does not parse. */
static void rd_d(const char *line, double *v) {
  char k[4096];
  sscanf(line, "%s %lf", k, v);
}
static void rd_i(const char *line, int *v) {
  char k[4096];
  sscanf(line, "%s %d", k, v);
}
static void rd_s(const char *line, char *v) {
  char k[4096];
  sscanf(line, "%s %s", k, v);
}

int orc_parse_mapper_params(const char *path, orc_params *P) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char line[4096];
  static const char *pk[15] = {"p0 ", "p1 ", "p2 ",  "p3 ",  "p4 ",
                               "p5 ", "p6 ", "p7 ",  "p8 ",  "p9 ",
                               "p10 ", "p11 ", "p12 ", "p13 ", "p14 "};
  while (fgets(line, sizeof(line), f)) {
    if (starts(line, "tx ")) rd_d(line, &P->tx);
    if (starts(line, "ty ")) rd_d(line, &P->ty);
    if (starts(line, "tz ")) rd_d(line, &P->tz);
    if (starts(line, "pxY ")) rd_d(line, &P->pxY);
    if (starts(line, "pxZ ")) rd_d(line, &P->pxZ);
    if (starts(line, "NPanelsY ")) { rd_i(line, &P->NPanelsY); continue; }
    if (starts(line, "NPanelsZ ")) { rd_i(line, &P->NPanelsZ); continue; }
    if (starts(line, "PanelSizeY ")) { rd_i(line, &P->PanelSizeY); continue; }
    if (starts(line, "PanelSizeZ ")) { rd_i(line, &P->PanelSizeZ); continue; }
    if (starts(line, "PanelShiftsFile ")) { rd_s(line, P->PanelShiftsFile); continue; }
    if (starts(line, "PanelGapsY ")) {
      char *q = line + strlen("PanelGapsY ");
      for (int k = 0; k < P->NPanelsY - 1 && k < 64; k++)
        P->PanelGapsY[k] = (int)strtol(q, &q, 10);
      continue;
    }
    if (starts(line, "PanelGapsZ ")) {
      char *q = line + strlen("PanelGapsZ ");
      for (int k = 0; k < P->NPanelsZ - 1 && k < 64; k++)
        P->PanelGapsZ[k] = (int)strtol(q, &q, 10);
      continue;
    }
    if (starts(line, "px ")) { rd_d(line, &P->pxY); P->pxZ = P->pxY; }
    if (starts(line, "BC ")) {
      char k[4096];
      sscanf(line, "%s %lf %lf", k, &P->yCen, &P->zCen);
    }
    if (starts(line, "Lsd ")) rd_d(line, &P->Lsd);
    if (starts(line, "RhoD ")) rd_d(line, &P->RhoD);
    if (starts(line, "MaxRingRad ")) rd_d(line, &P->RhoD);
    for (int q = 0; q < 15; q++)
      if (starts(line, pk[q])) rd_d(line, &P->p[q]);
    if (starts(line, "DistortionOrder ")) continue;
    if (starts(line, "EtaBinSize ")) rd_d(line, &P->EtaBinSize);
    if (starts(line, "RBinSize ")) rd_d(line, &P->RBinSize);
    if (starts(line, "RMax ")) rd_d(line, &P->RMax);
    if (starts(line, "RMin ")) rd_d(line, &P->RMin);
    if (starts(line, "EtaMax ")) rd_d(line, &P->EtaMax);
    if (starts(line, "EtaMin ")) rd_d(line, &P->EtaMin);
    if (starts(line, "QBinSize ")) rd_d(line, &P->QBinSize);
    if (starts(line, "SubPixelLevel ")) rd_i(line, &P->SubPixelLevel);
    if (starts(line, "SubPixelCardinalWidth ")) rd_d(line, &P->SubPixelCardinalWidth);
    if (starts(line, "QMin ")) rd_d(line, &P->QMin);
    if (starts(line, "QMax ")) rd_d(line, &P->QMax);
    if (starts(line, "Wavelength ")) rd_d(line, &P->Wavelength);
    if (starts(line, "NrPixelsY ")) rd_i(line, &P->NrPixelsY);
    if (starts(line, "NrPixelsZ ")) rd_i(line, &P->NrPixelsZ);
    if (starts(line, "NrPixels ")) { rd_i(line, &P->NrPixelsY); rd_i(line, &P->NrPixelsZ); }
    if (starts(line, "DistortionFile ")) { P->hasDistortion = 1; rd_s(line, P->DistortionFile); }
    if (starts(line, "ImTransOpt ")) {
      if (P->NrTransOpt < 10) rd_i(line, &P->TransOpt[P->NrTransOpt]);
      P->NrTransOpt++;
      continue;
    }
    if (starts(line, "MaskFile ")) { rd_s(line, P->MaskFile); P->hasMask = 1; }
    if (starts(line, "Parallax ")) rd_d(line, &P->Parallax);
    if (starts(line, "SolidAngleCorrection ")) rd_i(line, &P->SolidAngleCorrection);
    if (starts(line, "PolarizationCorrection ")) rd_i(line, &P->PolarizationCorrection);
    if (starts(line, "PolarizationFraction ")) rd_d(line, &P->PolarizationFraction);
    if (starts(line, "ResidualCorrectionMap ")) { rd_s(line, P->ResidualCorrectionMap); P->hasResidual = 1; }
  }
  fclose(f);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Bin edges: DetectorMapper.c:712-755 (R-mode and Q-mode)             */
/* ------------------------------------------------------------------ */
static int orc_qmode(const orc_params *P) {
  return P->QBinSize > 0 && P->Wavelength > 0;
}

int orc_bin_counts(const orc_params *P, int *nR, int *nEta) {
  *nEta = (int)ceil((P->EtaMax - P->EtaMin) / P->EtaBinSize);
  if (orc_qmode(P))
    *nR = (int)ceil((P->QMax - P->QMin) / P->QBinSize);
  else
    *nR = (int)ceil((P->RMax - P->RMin) / P->RBinSize);
  return 0;
}

/* Q-mode rewrites RMin/RMax in P (the header hash sees the rewritten
 * values, DetectorMapper.c:716-719 then :837). */
int orc_bin_edges(orc_params *P, double *RLo, double *RHi, double *ELo,
                  double *EHi) {
  int nR, nE;
  orc_bin_counts(P, &nR, &nE);
  for (int e = 0; e < nE; e++) {
    ELo[e] = P->EtaBinSize * e + P->EtaMin;
    EHi[e] = P->EtaBinSize * (e + 1) + P->EtaMin;
  }
  if (orc_qmode(P)) {
    const double lam = P->Wavelength;
    if (P->QMin > 0)
      P->RMin = (P->Lsd / P->pxY) * tan(2.0 * asin(P->QMin * lam / (4.0 * M_PI)));
    if (P->QMax > 0)
      P->RMax = (P->Lsd / P->pxY) * tan(2.0 * asin(P->QMax * lam / (4.0 * M_PI)));
    for (int r = 0; r < nR; r++) {
      double ql = P->QMin + P->QBinSize * r;
      double qh = P->QMin + P->QBinSize * (r + 1);
      RLo[r] = (P->Lsd / P->pxY) * tan(2.0 * asin(ql * lam / (4.0 * M_PI)));
      RHi[r] = (P->Lsd / P->pxY) * tan(2.0 * asin(qh * lam / (4.0 * M_PI)));
    }
  } else {
    for (int r = 0; r < nR; r++) {
      RLo[r] = P->RBinSize * r + P->RMin;
      RHi[r] = P->RBinSize * (r + 1) + P->RMin;
    }
  }
  return 0;
}

/* Canonical hash string of map_header_compute, MapHeader.h:216-234. */
void orc_header_string(const orc_params *P, int qMode, char *buf, int len) {
  int n = snprintf(buf, len,
                   "BC=%.6f,%.6f|EtaBinSize=%.6f|EtaMax=%.6f|EtaMin=%.6f|"
                   "Lsd=%.6f|NrPixelsY=%d|NrPixelsZ=%d|"
                   "RBinSize=%.6f|RMax=%.6f|RMin=%.6f|RhoD=%.6f|TransOpt=%d",
                   P->yCen, P->zCen, P->EtaBinSize, P->EtaMax, P->EtaMin,
                   P->Lsd, P->NrPixelsY, P->NrPixelsZ, P->RBinSize, P->RMax,
                   P->RMin, P->RhoD, P->NrTransOpt);
  for (int i = 0; i < P->NrTransOpt && i < 10; i++)
    n += snprintf(buf + n, len - n, ",%d", P->TransOpt[i]);
  n += snprintf(buf + n, len - n,
                "|p0=%.6f|p1=%.6f|p2=%.6f|p3=%.6f|p4=%.6f|p6=%.6f|"
                "pxY=%.6f|pxZ=%.6f|tx=%.6f|ty=%.6f|tz=%.6f",
                P->p[0], P->p[1], P->p[2], P->p[3], P->p[4], P->p[6], P->pxY,
                P->pxZ, P->tx, P->ty, P->tz);
  if (qMode) snprintf(buf + n, len - n, "|qMode=1|Wavelength=%.8f", P->Wavelength);
}

/* ------------------------------------------------------------------ */
/* Geometry: DetectorGeometry.c:26-154                                 */
/* ------------------------------------------------------------------ */
static double orc_eta(double y, double z) { return ORC_R2D * atan2(-y, z); }
static double orc_sgn(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : 0.0); }
static int orc_between(double v, double lo, double hi) {
  return (v - ORC_EPS <= hi && v + ORC_EPS >= lo) ? 1 : 0;
}

static void mm33(const double a[9], const double b[9], double c[9]) {
  for (int r = 0; r < 3; r++) {
    c[r * 3 + 0] = a[r * 3 + 0] * b[0] + a[r * 3 + 1] * b[3] + a[r * 3 + 2] * b[6];
    c[r * 3 + 1] = a[r * 3 + 0] * b[1] + a[r * 3 + 1] * b[4] + a[r * 3 + 2] * b[7];
    c[r * 3 + 2] = a[r * 3 + 0] * b[2] + a[r * 3 + 1] * b[5] + a[r * 3 + 2] * b[8];
  }
}

/* T = Rx(tx) * (Ry(ty) * Rz(tz)), DetectorGeometry.c:64-78. */
void orc_tilt_matrix(double tx, double ty, double tz, double T[9]) {
  double a = ORC_D2R * tx, b = ORC_D2R * ty, c = ORC_D2R * tz;
  double Rx[9] = {1, 0, 0, 0, cos(a), -sin(a), 0, sin(a), cos(a)};
  double Ry[9] = {cos(b), 0, sin(b), 0, 1, 0, -sin(b), 0, cos(b)};
  double Rz[9] = {cos(c), -sin(c), 0, sin(c), cos(c), 0, 0, 0, 1};
  double tmp[9];
  mm33(Ry, Rz, tmp);
  mm33(Rx, tmp, T);
}

/* Bilinear residual lookup, DetectorGeometry.h:38-56. */
static double orc_residual(const double *map, int NY, int NZ, double Y,
                           double Z) {
  if (!map) return 0.0;
  double y = Y, z = Z;
  if (y < 0.0) y = 0.0;
  if (z < 0.0) z = 0.0;
  if (y >= NY - 1.0) y = NY - 1.001;
  if (z >= NZ - 1.0) z = NZ - 1.001;
  int y0 = (int)y, z0 = (int)z;
  double fy = y - y0, fz = z - z0;
  double v00 = map[z0 * NY + y0], v10 = map[z0 * NY + y0 + 1];
  double v01 = map[(z0 + 1) * NY + y0], v11 = map[(z0 + 1) * NY + y0 + 1];
  return v00 * (1 - fy) * (1 - fz) + v10 * fy * (1 - fz) +
         v01 * (1 - fy) * fz + v11 * fy * fz;
}

/* Pixel (Y,Z) -> (R px, Eta deg), DetectorGeometry.c:95-149. */
void orc_pixel_to_reta(const orc_params *P, const double T[9], double Y,
                       double Z, double dLsd, double dP2,
                       const double *residual, double *R, double *Eta) {
  const double px = P->pxY;
  const double *p = P->p;
  double L = P->Lsd + dLsd;
  double p2 = p[2] + dP2;
  double Yc = (-Y + P->yCen) * px;
  double Zc = (Z - P->zCen) * px;
  double v1 = T[0] * 0 + T[1] * Yc + T[2] * Zc;
  double v2 = T[3] * 0 + T[4] * Yc + T[5] * Zc;
  double v3 = T[6] * 0 + T[7] * Yc + T[8] * Zc;
  double X = L + v1;
  double Rad = (L / X) * sqrt(v2 * v2 + v3 * v3);
  double et = orc_eta(v2, v3);
  double rn = Rad / P->RhoD;
  double e90 = 90 - et;
  double D = (p[0] * pow(rn, 2.0) * cos(ORC_D2R * (2 * e90 + p[6]))) +
             (p[1] * pow(rn, 4.0) * cos(ORC_D2R * (4 * e90 + p[3]))) +
             (p2 * pow(rn, 2.0));
  D += p[4] * pow(rn, 6.0);
  D += p[5] * pow(rn, 4.0);
  D += p[7] * pow(rn, 4.0) * cos(ORC_D2R * (e90 + p[8]));
  D += p[9] * pow(rn, 3.0) * cos(ORC_D2R * (3 * e90 + p[10]));
  D += p[11] * pow(rn, 5.0) * cos(ORC_D2R * (5 * e90 + p[12]));
  D += p[13] * pow(rn, 6.0) * cos(ORC_D2R * (6 * e90 + p[14]));
  D += 1;
  double Rt = Rad * D / px;
  Rt = Rt * (P->Lsd / L);
  if (P->Parallax != 0.0) {
    double tth = atan(Rad / L);
    Rt += P->Parallax * sin(tth) / px;
  }
  Rt += orc_residual(residual, P->NrPixelsY, P->NrPixelsZ, Y, Z);
  *R = Rt;
  *Eta = et;
}

static void orc_reta_to_yz(double R, double eta, double *y, double *z) {
  *y = -R * sin(eta * ORC_D2R);
  *z = R * cos(eta * ORC_D2R);
}

/* Panel.c:3-57 GeneratePanels. */
int orc_generate_panels(const orc_params *P, orc_panel *out, int maxPanels) {
  if (P->NPanelsY <= 0 || P->NPanelsZ <= 0) return 0;
  int n = 0, y0 = 0;
  for (int a = 0; a < P->NPanelsY; a++) {
    int y1 = y0 + P->PanelSizeY - 1;
    int z0 = 0;
    for (int b = 0; b < P->NPanelsZ; b++) {
      int z1 = z0 + P->PanelSizeZ - 1;
      if (n < maxPanels) {
        orc_panel *q = &out[n];
        memset(q, 0, sizeof(*q));
        q->id = n;
        q->yMin = y0; q->yMax = y1; q->zMin = z0; q->zMax = z1;
        q->centerY = (y0 + y1) / 2.0;
        q->centerZ = (z0 + z1) / 2.0;
      }
      n++;
      if (b < P->NPanelsZ - 1) z0 = z1 + 1 + P->PanelGapsZ[b];
    }
    if (a < P->NPanelsY - 1) y0 = y1 + 1 + P->PanelGapsY[a];
  }
  return n;
}

/* ImageUtils.c:20-73, applied in place. */
void orc_image_transform(int nOpt, const int *opt, double *img, int NY,
                         int NZ) {
  for (int t = 0; t < nOpt; t++) {
    if (opt[t] == 1) {
      for (int r = 0; r < NZ; r++)
        for (int c = 0; c < NY / 2; c++) {
          double s = img[r * NY + c];
          img[r * NY + c] = img[r * NY + NY - c - 1];
          img[r * NY + NY - c - 1] = s;
        }
    } else if (opt[t] == 2) {
      for (int r = 0; r < NZ / 2; r++)
        for (int c = 0; c < NY; c++) {
          double s = img[r * NY + c];
          img[r * NY + c] = img[(NZ - r - 1) * NY + c];
          img[(NZ - r - 1) * NY + c] = s;
        }
    } else if (opt[t] == 3 && NY == NZ) {
      for (int r = 0; r < NZ; r++)
        for (int c = r + 1; c < NY; c++) {
          double s = img[r * NY + c];
          img[r * NY + c] = img[c * NY + r];
          img[c * NY + r] = s;
        }
    }
  }
}

/* MapperCore.c:20-40: inverse of the transforms, applied in reverse. */
static void orc_inverse_transform(int y, int z, const orc_params *P, int *yo,
                                  int *zo) {
  for (int t = P->NrTransOpt - 1; t >= 0; t--) {
    int o = P->TransOpt[t];
    if (o == 1) y = P->NrPixelsY - 1 - y;
    else if (o == 2) z = P->NrPixelsZ - 1 - z;
    else if (o == 3) { int s = y; y = z; z = s; }
  }
  *yo = y;
  *zo = z;
}

/* ------------------------------------------------------------------ */
/* Clipping helpers: DetectorGeometry.c:432-609                        */
/* ------------------------------------------------------------------ */
static const int QORD[4] = {0, 1, 3, 2};

static int orc_circle_hits(double y1, double z1, double y2, double z2,
                           double R, double h[2][2]) {
  double dy = y2 - y1, dz = z2 - z1;
  double a = dy * dy + dz * dz;
  if (a < 1e-30) return 0;
  double b = 2.0 * (y1 * dy + z1 * dz);
  double c = y1 * y1 + z1 * z1 - R * R;
  double disc = b * b - 4.0 * a * c;
  if (disc < 0) return 0;
  double sd = sqrt(disc);
  double i2a = 0.5 / a;
  int n = 0;
  double t1 = (-b - sd) * i2a;
  if (t1 >= -ORC_EPS && t1 <= 1.0 + ORC_EPS) {
    double t = fmax(0.0, fmin(1.0, t1));
    h[n][0] = y1 + t * dy;
    h[n][1] = z1 + t * dz;
    n++;
  }
  double t2 = (-b + sd) * i2a;
  if (t2 >= -ORC_EPS && t2 <= 1.0 + ORC_EPS && fabs(t2 - t1) > 1e-12) {
    double t = fmax(0.0, fmin(1.0, t2));
    h[n][0] = y1 + t * dy;
    h[n][1] = z1 + t * dz;
    n++;
  }
  return n;
}

static int orc_ray_hit(double y1, double z1, double y2, double z2,
                       double etaDeg, double *hy, double *hz) {
  double er = etaDeg * ORC_D2R;
  double ce = cos(er), se = sin(er);
  double dy = y2 - y1, dz = z2 - z1;
  double den = dy * ce + dz * se;
  if (fabs(den) < 1e-30) return 0;
  double t = -(y1 * ce + z1 * se) / den;
  if (t < -ORC_EPS || t > 1.0 + ORC_EPS) return 0;
  t = fmax(0.0, fmin(1.0, t));
  *hy = y1 + t * dy;
  *hz = z1 + t * dz;
  if ((-se) * (*hy) + ce * (*hz) < 0) return 0;
  return 1;
}

static int orc_in_quad(double py, double pz, double q[4][2]) {
  int pos = 0, neg = 0;
  for (int e = 0; e < 4; e++) {
    int a = QORD[e], b = QORD[(e + 1) % 4];
    double ey = q[b][0] - q[a][0];
    double ez = q[b][1] - q[a][1];
    double cr = ey * (pz - q[a][1]) - ez * (py - q[a][0]);
    if (cr > 0) pos++;
    else if (cr < 0) neg++;
  }
  return (pos == 0 || neg == 0) ? 1 : 0;
}

static int orc_unique(double in[][2], double out[][2], int n, double Rlo,
                      double Rhi, double Elo, double Ehi) {
  int m = 0;
  for (int i = 0; i < n; i++) {
    int dup = 0;
    for (int j = i + 1; j < n; j++) {
      double len = sqrt((in[i][0] - in[j][0]) * (in[i][0] - in[j][0]) +
                        (in[i][1] - in[j][1]) * (in[i][1] - in[j][1]));
      if (len == 0) dup = 1;
    }
    double rt = sqrt(in[i][0] * in[i][0] + in[i][1] * in[i][1]);
    double et = orc_eta(in[i][0], in[i][1]);
    if (!orc_between(et, Elo, Ehi)) {
      if (orc_between(et + 360, Elo, Ehi)) et += 360;
      else if (orc_between(et - 360, Elo, Ehi)) et -= 360;
    }
    if (orc_between(rt, Rlo, Rhi) == 0) dup = 1;
    if (orc_between(et, Elo, Ehi) == 0) dup = 1;
    if (!dup) {
      out[m][0] = in[i][0];
      out[m][1] = in[i][1];
      m++;
    }
  }
  return m;
}

typedef struct { double x, y; } orc_pt;
typedef struct { double cx, cy; } orc_ctr;

/* Angular order around the centroid (DetectorGeometry.c:388-416). The
 * polygon is sorted with glibc qsort_r, as in the reference, so that the
 * summation order of the area is the same. */
static int orc_cmp_angle(const void *pa, const void *pb, void *ctx) {
  const orc_ctr *c = (const orc_ctr *)ctx;
  const orc_pt *a = (const orc_pt *)pa, *b = (const orc_pt *)pb;
  double ax = a->x - c->cx, ay = a->y - c->cy;
  double bx = b->x - c->cx, by = b->y - c->cy;
  if (ax >= 0 && bx < 0) return 1;
  if (ax < 0 && bx >= 0) return -1;
  if (ax == 0 && bx == 0) {
    if (ay >= 0 || by >= 0) return a->y > b->y ? 1 : -1;
    return b->y > a->y ? 1 : -1;
  }
  double det = ax * by - bx * ay;
  if (det < 0) return 1;
  if (det > 0) return -1;
  double d1 = ax * ax + ay * ay, d2 = bx * bx + by * by;
  return d1 > d2 ? 1 : -1;
}

/* Green's-theorem area with exact arc terms, DetectorGeometry.c:418-498. */
static double orc_poly_area(double v[][2], int n, double Rlo, double Rhi) {
  orc_pt pts[64];
  orc_ctr c = {0, 0};
  for (int i = 0; i < n; i++) {
    c.cx += v[i][0];
    c.cy += v[i][1];
    pts[i].x = v[i][0];
    pts[i].y = v[i][1];
  }
  c.cx /= n;
  c.cy /= n;
  qsort_r(pts, n, sizeof(orc_pt), orc_cmp_angle, &c);
  double r2lo = Rlo * Rlo, r2hi = Rhi * Rhi, tol = 1e-6, A = 0;
  for (int i = 0; i < n; i++) {
    int j = (i + 1) % n;
    double s1 = pts[i].x * pts[i].x + pts[i].y * pts[i].y;
    double s2 = pts[j].x * pts[j].x + pts[j].y * pts[j].y;
    int onLo = (fabs(s1 - r2lo) < tol * r2lo && fabs(s2 - r2lo) < tol * r2lo);
    int onHi = (fabs(s1 - r2hi) < tol * r2hi && fabs(s2 - r2hi) < tol * r2hi);
    if (onLo || onHi) {
      double R = onLo ? Rlo : Rhi;
      double a1 = atan2(pts[i].y, pts[i].x);
      double a2 = atan2(pts[j].y, pts[j].x);
      double d = a2 - a1;
      if (d > M_PI) d -= 2.0 * M_PI;
      if (d < -M_PI) d += 2.0 * M_PI;
      A += (R * R / 2.0) * d;
    } else {
      A += 0.5 * (pts[i].x * pts[j].y - pts[j].x * pts[i].y);
    }
  }
  return A;
}

/* ------------------------------------------------------------------ */
/* Mapper: MapperCore.c:45-424                                         */
/* ------------------------------------------------------------------ */
typedef struct {
  int bin;
  orc_entry e;
} orc_rec;

typedef struct {
  orc_rec *r;
  int64_t n, cap;
} orc_vec;

static void vec_push(orc_vec *v, int bin, const orc_entry *e) {
  if (v->n == v->cap) {
    v->cap = v->cap ? v->cap * 2 : 64;
    v->r = (orc_rec *)realloc(v->r, v->cap * sizeof(orc_rec));
  }
  v->r[v->n].bin = bin;
  v->r[v->n].e = *e;
  v->n++;
}

struct orc_map {
  int nR, nEta;
  int64_t nnz;
  orc_entry *entries;
  int32_t *pairs;
  int32_t *flags;
  int64_t skipped[3];
};

/* Emits one column (fixed Y index i) in the reference's loop order:
 * Z ascending, sub-pixels (si, sj), R candidates ascending, Eta candidates
 * ascending.  Single-threaded reference order = concatenation over i. */
static void orc_map_column(const orc_params *P, const double T[9], int i,
                           const double *RLo, const double *RHi,
                           const double *ELo, const double *EHi, int nR,
                           int nE, const double *mask, const double *dY,
                           const double *dZ, const double *resid,
                           const orc_panel *panels, int nPanels,
                           int *flags, orc_vec *out, int64_t sk[3],
                           int *rSel, int *eSel) {
  const int NY = P->NrPixelsY, NZ = P->NrPixelsZ;
  for (int j = 0; j < NZ; j++) {
    long long pos = (long long)j * NY + i;
    int masked = (mask != NULL && mask[pos] == 1.0);
    double pdY = 0, pdZ = 0, dLsd = 0, dP2 = 0;
    int pi = -1;
    for (int q = 0; q < nPanels; q++)
      if ((double)i >= panels[q].yMin && (double)i <= panels[q].yMax &&
          (double)j >= panels[q].zMin && (double)j <= panels[q].zMax) {
        pi = q;
        break;
      }
    if (pi >= 0) {
      const orc_panel *pp = &panels[pi];
      double yo, zo;
      if (pp->dTheta != 0.0) {
        double a = (M_PI / 180.0) * pp->dTheta;
        double ct = cos(a), st = sin(a);
        double ddy = (double)i - pp->centerY, ddz = (double)j - pp->centerZ;
        yo = pp->centerY + ddy * ct - ddz * st + pp->dY;
        zo = pp->centerZ + ddy * st + ddz * ct + pp->dZ;
      } else {
        yo = (double)i + pp->dY;
        zo = (double)j + pp->dZ;
      }
      pdY = yo - (double)i;
      pdZ = zo - (double)j;
      dLsd = pp->dLsd;
      dP2 = pp->dP2;
    }
    double ypr = (double)i + (dY ? dY[pos] : 0.0) + pdY;
    double zpr = (double)j + (dZ ? dZ[pos] : 0.0) + pdZ;
    int sp = 1;
    if (P->SubPixelLevel > 1) {
      double rc, ec;
      orc_pixel_to_reta(P, T, ypr, zpr, dLsd, dP2, resid, &rc, &ec);
      double ae = fabs(ec), W = P->SubPixelCardinalWidth;
      if (ae <= W || fabs(ae - 90.0) <= W || fabs(ae - 180.0) <= W)
        sp = P->SubPixelLevel;
    }
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
            orc_pixel_to_reta(P, T, ypr + oy[k], zpr + oz[l], dLsd, dP2,
                              resid, &Rt, &Et);
            if (Et < eMin) eMin = Et;
            if (Et > eMax) eMax = Et;
            if (Rt < rMin) rMin = Rt;
            if (Rt > rMax) rMax = Rt;
            orc_reta_to_yz(Rt, Et, &cyz[k * 2 + l][0], &cyz[k * 2 + l][1]);
          }
        orc_pixel_to_reta(P, T, ypr + cy, zpr + cz, dLsd, dP2, resid, &Rt,
                          &Et);
        int nr = 0, ne = 0;
        for (int k = 0; k < nR; k++)
          if (RHi[k] >= rMin && RLo[k] <= rMax) rSel[nr++] = k;
        double lo = eMin, hi = eMax;
        if (hi - lo > 180) {
          double s = hi;
          hi = 360 + lo;
          lo = s;
        }
        for (int k = 0; k < nE; k++) {
          if ((EHi[k] >= lo && ELo[k] <= hi) ||
              (EHi[k] >= lo + 360 && ELo[k] <= hi + 360) ||
              (EHi[k] >= lo - 360 && ELo[k] <= hi - 360))
            eSel[ne++] = k;
        }
        if (masked) {
          if (si == 0 && sj == 0)
            for (int a = 0; a < nr; a++)
              for (int b = 0; b < ne; b++) flags[(long long)rSel[a] * nE + eSel[b]] = 1;
          continue;
        }
        for (int a = 0; a < nr; a++) {
          double Rlo = RLo[rSel[a]], Rhi = RHi[rSel[a]];
          for (int b = 0; b < ne; b++) {
            double Elo = ELo[eSel[b]], Ehi = EHi[eSel[b]];
            double box[4][2];
            orc_reta_to_yz(Rlo, Elo, &box[0][0], &box[0][1]);
            orc_reta_to_yz(Rlo, Ehi, &box[1][0], &box[1][1]);
            orc_reta_to_yz(Rhi, Elo, &box[2][0], &box[2][1]);
            orc_reta_to_yz(Rhi, Ehi, &box[3][0], &box[3][1]);
            double V[64][2], U[64][2];
            int nv = 0;
            for (int m = 0; m < 4; m++) {
              double rr = sqrt(cyz[m][0] * cyz[m][0] + cyz[m][1] * cyz[m][1]);
              double ee = orc_eta(cyz[m][0], cyz[m][1]);
              if (Elo < -180 && orc_sgn(ee) != orc_sgn(Elo)) ee -= 360;
              if (Ehi > 180 && orc_sgn(ee) != orc_sgn(Ehi)) ee += 360;
              if (rr >= Rlo && rr <= Rhi && ee >= Elo && ee <= Ehi) {
                V[nv][0] = cyz[m][0];
                V[nv][1] = cyz[m][1];
                nv++;
              }
            }
            for (int m = 0; m < 4; m++)
              if (orc_in_quad(box[m][0], box[m][1], cyz)) {
                V[nv][0] = box[m][0];
                V[nv][1] = box[m][1];
                nv++;
              }
            if (nv < 4) {
              for (int e = 0; e < 4; e++) {
                double y1 = cyz[QORD[e]][0], z1 = cyz[QORD[e]][1];
                double y2 = cyz[QORD[(e + 1) % 4]][0], z2 = cyz[QORD[(e + 1) % 4]][1];
                double hh[2][2];
                for (int side = 0; side < 2; side++) {
                  int nh = orc_circle_hits(y1, z1, y2, z2, side ? Rhi : Rlo, hh);
                  for (int h = 0; h < nh; h++) {
                    double eh = orc_eta(hh[h][0], hh[h][1]);
                    if (Elo < -180 && orc_sgn(eh) != orc_sgn(Elo)) eh -= 360;
                    if (Ehi > 180 && orc_sgn(eh) != orc_sgn(Ehi)) eh += 360;
                    if (eh >= Elo - ORC_EPS && eh <= Ehi + ORC_EPS) {
                      V[nv][0] = hh[h][0];
                      V[nv][1] = hh[h][1];
                      nv++;
                    }
                  }
                }
                double hy, hz;
                for (int side = 0; side < 2; side++) {
                  if (orc_ray_hit(y1, z1, y2, z2, side ? Ehi : Elo, &hy, &hz)) {
                    double rh = sqrt(hy * hy + hz * hz);
                    if (rh >= Rlo - ORC_EPS && rh <= Rhi + ORC_EPS) {
                      V[nv][0] = hy;
                      V[nv][1] = hz;
                      nv++;
                    }
                  }
                }
              }
            }
            if (nv < 3) { sk[0]++; continue; }
            nv = orc_unique(V, U, nv, Rlo, Rhi, Elo, Ehi);
            if (nv < 3) { sk[1]++; continue; }
            double area = orc_poly_area(U, nv, Rlo, Rhi);
            if (area < 1E-5) { sk[2]++; continue; }
            double corr = area;
            if (P->SolidAngleCorrection || P->PolarizationCorrection) {
              double tth = atan(Rt * P->pxY / P->Lsd);
              if (P->SolidAngleCorrection) {
                double c2 = cos(tth);
                corr /= (c2 * c2 * c2);
              }
              if (P->PolarizationCorrection) {
                double s2 = sin(tth);
                double ce = cos(Et * ORC_D2R);
                double pf = 1.0 - P->PolarizationFraction * s2 * s2 * ce * ce;
                if (pf > 1e-6) corr /= pf;
              }
            }
            int ry, rz;
            orc_inverse_transform(i, j, P, &ry, &rz);
            orc_entry en;
            en.y = (float)ry + (float)cy;
            en.z = (float)rz + (float)cz;
            en.frac = corr;
            en.deltaR = (float)(Rt - (Rlo + Rhi) * 0.5);
            en.areaWeight = (float)area;
            vec_push(out, rSel[a] * nE + eSel[b], &en);
          }
        }
      }
    }
  }
}

orc_map *orc_map_build(const orc_params *Pin, const double *mask,
                       const double *distY, const double *distZ,
                       const double *residual, const orc_panel *panels,
                       int nPanels, int nThreads) {
  orc_params P = *Pin;
  int nR, nE;
  orc_bin_counts(&P, &nR, &nE);
  double *RLo = malloc(nR * sizeof(double)), *RHi = malloc(nR * sizeof(double));
  double *ELo = malloc(nE * sizeof(double)), *EHi = malloc(nE * sizeof(double));
  orc_bin_edges(&P, RLo, RHi, ELo, EHi);
  double T[9];
  orc_tilt_matrix(P.tx, P.ty, P.tz, T);
  const int NY = P.NrPixelsY;
  orc_map *m = calloc(1, sizeof(orc_map));
  m->nR = nR;
  m->nEta = nE;
  long long nb = (long long)nR * nE;
  m->flags = calloc(nb, sizeof(int32_t));
  orc_vec *cols = calloc(NY, sizeof(orc_vec));
  int64_t sk[3] = {0, 0, 0};
  if (nThreads > 0) omp_set_num_threads(nThreads);
#pragma omp parallel
  {
    int *rs = malloc(nR * sizeof(int)), *es = malloc(nE * sizeof(int));
    int64_t lsk[3] = {0, 0, 0};
#pragma omp for schedule(dynamic, 4)
    for (int i = 0; i < NY; i++)
      orc_map_column(&P, T, i, RLo, RHi, ELo, EHi, nR, nE, mask, distY, distZ,
                     residual, panels, nPanels, m->flags, &cols[i], lsk, rs, es);
#pragma omp critical
    {
      for (int q = 0; q < 3; q++) sk[q] += lsk[q];
    }
    free(rs);
    free(es);
  }
  for (int q = 0; q < 3; q++) m->skipped[q] = sk[q];
  /* Counting pass, then a stable scatter in column order: within a bin the
   * entries keep the single-threaded reference order. */
  int64_t *cnt = calloc(nb, sizeof(int64_t));
  int64_t nnz = 0;
  for (int i = 0; i < NY; i++)
    for (int64_t k = 0; k < cols[i].n; k++) cnt[cols[i].r[k].bin]++;
  m->pairs = malloc(2 * nb * sizeof(int32_t));
  int64_t *cur = malloc(nb * sizeof(int64_t));
  for (long long b = 0; b < nb; b++) {
    m->pairs[2 * b] = (int32_t)cnt[b];
    m->pairs[2 * b + 1] = (int32_t)nnz;
    cur[b] = nnz;
    nnz += cnt[b];
  }
  m->nnz = nnz;
  m->entries = malloc((nnz ? nnz : 1) * sizeof(orc_entry));
  for (int i = 0; i < NY; i++) {
    for (int64_t k = 0; k < cols[i].n; k++) {
      const orc_rec *r = &cols[i].r[k];
      m->entries[cur[r->bin]++] = r->e;
    }
    free(cols[i].r);
  }
  free(cols);
  free(cnt);
  free(cur);
  free(RLo);
  free(RHi);
  free(ELo);
  free(EHi);
  return m;
}

int64_t orc_map_nnz(const orc_map *m) { return m->nnz; }
int orc_map_nbins(const orc_map *m) { return m->nR * m->nEta; }
void orc_map_counters(const orc_map *m, int64_t out[3]) {
  for (int q = 0; q < 3; q++) out[q] = m->skipped[q];
}
void orc_map_copy(const orc_map *m, orc_entry *entries, int32_t *pairs,
                  int32_t *flags) {
  long long nb = (long long)m->nR * m->nEta;
  if (entries) memcpy(entries, m->entries, m->nnz * sizeof(orc_entry));
  if (pairs) memcpy(pairs, m->pairs, 2 * nb * sizeof(int32_t));
  if (flags) memcpy(flags, m->flags, nb * sizeof(int32_t));
}
void orc_map_free(orc_map *m) {
  if (!m) return;
  free(m->entries);
  free(m->pairs);
  free(m->flags);
  free(m);
}

/* ------------------------------------------------------------------ */
/* Integration: IntegratorZarrOMP.c:1710-1778 / .cu:975-1050           */
/* ------------------------------------------------------------------ */
void orc_integrate(const orc_entry *E, const int32_t *pairs,
                   const int32_t *flags, int nR, int nEta,
                   const double *img, int NY, int NZ, double BCy, double BCz,
                   int gradient, int normalize, int semantics, int nThreads,
                   double *cake, double *area) {
  if (nThreads > 0) omp_set_num_threads(nThreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int r = 0; r < nR; r++) {
    for (int e = 0; e < nEta; e++) {
      long long b = (long long)r * nEta + e;
      int n = pairs[2 * b], off = pairs[2 * b + 1];
      double I = 0, A = 0;
      for (int q = 0; q < n; q++) {
        const orc_entry *x = &E[off + q];
        double v;
        if (semantics == 1 && !gradient) {
          /* GPU stream: offset = (long long)z*NY + y in float arithmetic. */
          long long zz = (long long)x->z * NY;
          float o = (float)zz + x->y;
          v = img[(int)o];
        } else {
          double ry = x->y, rz = x->z;
          if (gradient && x->deltaR != 0.0f) {
            double dy = x->y - BCy, dz = x->z - BCz;
            double R = sqrt(dy * dy + dz * dz);
            if (R > 1.0) {
              ry -= x->deltaR * dy / R;
              rz -= x->deltaR * dz / R;
            }
          }
          int iy = (int)floorf((float)ry), iz = (int)floorf((float)rz);
          double fy = ry - iy, fz = rz - iz;
          if (iy < 0) { iy = 0; fy = 0; }
          if (iy >= NY - 1) { iy = NY - 2; fy = 1; }
          if (iz < 0) { iz = 0; fz = 0; }
          if (iz >= NZ - 1) { iz = NZ - 2; fz = 1; }
          v = img[(size_t)iz * NY + iy] * (1 - fy) * (1 - fz) +
              img[(size_t)iz * NY + iy + 1] * fy * (1 - fz) +
              img[(size_t)(iz + 1) * NY + iy] * (1 - fy) * fz +
              img[(size_t)(iz + 1) * NY + iy + 1] * fy * fz;
        }
        I += v * x->frac;
        A += x->areaWeight;
      }
      if (semantics == 0) {
        if (I != 0 && normalize == 1) I /= A;
        if (flags && flags[b]) I = NAN;
      } else {
        if (A > 1e-9) {
          if (normalize) I /= A;
        } else {
          I = 0.0;
        }
      }
      cake[b] = I;
      if (area) area[b] = A;
    }
  }
}

/* In-bounds area per bin, .cu:817-899. */
void orc_area_inbounds(const orc_entry *E, const int32_t *pairs, int nBins,
                       int NY, int NZ, double *area) {
  for (int b = 0; b < nBins; b++) {
    double A = 0;
    for (int q = 0; q < pairs[2 * b]; q++) {
      const orc_entry *x = &E[pairs[2 * b + 1] + q];
      if (x->y < 0 || x->y >= NY || x->z < 0 || x->z >= NZ) continue;
      A += x->areaWeight;
    }
    area[b] = A;
  }
}

/* Lineouts: IntegratorZarrOMP.c:1829-1843, .cu:1162-1254, .cu:1256-1334. */
void orc_lineouts(const double *cake, const double *ain, int nR, int nEta,
                  double *lo_cpu, double *lo_area, double *lo_mean) {
  for (int r = 0; r < nR; r++) {
    double s = 0, sa = 0, sia = 0, sm = 0;
    int c = 0, cm = 0;
    for (int e = 0; e < nEta; e++) {
      double v = cake[(long long)r * nEta + e];
      if (v != 0 && !isnan(v)) { s += v; c++; }
      double a = ain ? ain[(long long)r * nEta + e] : 0;
      /* NaN (mask) bins are missing values in every definition; the
       * reference GPU path never produces them. */
      if (a > 1e-9 && !isnan(v)) { sia += v * a; sa += a; sm += v; cm++; }
    }
    if (lo_cpu) lo_cpu[r] = c > 0 ? s / c : 0;
    if (lo_area) lo_area[r] = sa > 1e-9 ? sia / sa : 0;
    if (lo_mean) lo_mean[r] = cm > 0 ? sm / cm : 0;
  }
}

/* REtaMap rows R, 2theta, eta, area, Q: IntegratorZarrOMP.c:1754-1763. */
void orc_reta_map(const double *RLo, const double *RHi, const double *ELo,
                  const double *EHi, int nR, int nEta, const double *area,
                  double px, double Lsd, double Lam, double *out) {
  long long nb = (long long)nR * nEta;
  for (int r = 0; r < nR; r++) {
    double Rm = (RLo[r] + RHi[r]) / 2;
    for (int e = 0; e < nEta; e++) {
      long long b = (long long)r * nEta + e;
      out[b] = Rm;
      out[nb + b] = ORC_R2D * atan(Rm * px / Lsd);
      out[2 * nb + b] = (ELo[e] + EHi[e]) / 2;
      out[3 * nb + b] = area[b];
      double t = atan(Rm * px / Lsd);
      out[4 * nb + b] = (Lam > 0) ? (4.0 * M_PI / Lam) * sin(t / 2.0) : 0.0;
    }
  }
}

/* ------------------------------------------------------------------ */
/* Peak models: PeakFit.c:17-90 (TCH), :882-920 (height-normalised pV) */
/* ------------------------------------------------------------------ */
void orc_tch_eta_fwhm(double sig, double gam, double *fwhm, double *eta) {
  double fg = sqrt(fmax(8.0 * log(2.0) * fmax(sig, 1e-12), 0.0)) / 100.0;
  double fl = fmax(gam, 1e-6) / 100.0;
  double g2 = fg * fg, g3 = g2 * fg, g4 = g3 * fg, g5 = g4 * fg;
  double l2 = fl * fl, l3 = l2 * fl, l4 = l3 * fl, l5 = l4 * fl;
  double F = pow(g5 + 2.69269 * g4 * fl + 2.42843 * g3 * l2 +
                     4.47163 * g2 * l3 + 0.07842 * fg * l4 + l5,
                 0.2);
  if (F < 1e-15) F = 1e-15;
  double q = fl / F;
  double e = 1.36603 * q - 0.47719 * q * q + 0.11116 * q * q * q;
  if (e < 0.0) e = 0.0;
  if (e > 1.0) e = 1.0;
  *fwhm = F;
  *eta = e;
}

void orc_tch_model(int nPeaks, const double *p, int n, const double *x,
                   double xlo, double xhi, double *out) {
  double bg0 = p[nPeaks * 4], bg1 = p[nPeaks * 4 + 1];
  double xr = fmax(xhi - xlo, 1e-12);
  for (int i = 0; i < n; i++) {
    double xn = 2.0 * (x[i] - xlo) / xr - 1.0;
    out[i] = bg0 + bg1 * xn;
  }
  for (int k = 0; k < nPeaks; k++) {
    double A = p[k * 4], c = p[k * 4 + 1];
    double sig = fmax(p[k * 4 + 2], 1e-12), gam = fmax(p[k * 4 + 3], 1e-6);
    double F, eta;
    orc_tch_eta_fwhm(sig, gam, &F, &eta);
    if (F < 1e-15) F = 1e-15;
    double sg = F / (2.0 * sqrt(2.0 * log(2.0)));
    double gn = 1.0 / (sg * sqrt(2.0 * M_PI));
    double hf = F / 2.0;
    double ln = hf / M_PI;
    for (int i = 0; i < n; i++) {
      double dx = x[i] - c, d2 = dx * dx;
      double G = gn * exp(-0.5 * d2 / (sg * sg));
      double Lz = ln / (d2 + hf * hf);
      out[i] += A * (eta * Lz + (1.0 - eta) * G);
    }
  }
}

void orc_pv_model(const double *p, int n, const double *x, double *out) {
  double C0 = 4.0 * log(2.0);
  double inv = 1.0 / (p[2] * p[2]);
  for (int i = 0; i < n; i++) {
    double dr = x[i] - p[0], d2 = dr * dr;
    double L = 1.0 / (1.0 + 4.0 * d2 * inv);
    double G = exp(-C0 * d2 * inv);
    out[i] = p[4] + p[3] * (p[1] * L + (1.0 - p[1]) * G);
  }
}

/* Initial values and bounds: pV from PeakFit.c:922-943, TCH single peak
 * from PeakFit.c:356-469 (one peak, ROI = the whole window). */
void orc_window_init(int model, int n, const double *x, const double *y,
                     double rstep, double rmean, double *p, double *lo,
                     double *hi) {
  if (model == 0) {
    double bg = (y[0] + y[n - 1]) / 2;
    if (bg < 0) bg = 0;
    double mx = -1e20;
    for (int i = 0; i < n; i++) if (y[i] > mx) mx = y[i];
    p[0] = rmean; lo[0] = x[0]; hi[0] = x[n - 1];
    p[1] = 0.5; lo[1] = 0; hi[1] = 1;
    p[2] = rstep * 3; lo[2] = rstep / 2; hi[2] = rstep * n / 2;
    p[3] = mx - bg; lo[3] = (mx - bg) / 100; hi[3] = mx * 1.5;
    p[4] = bg; lo[4] = 0; hi[4] = (bg > 0) ? bg * 1.5 : mx * 0.5;
    return;
  }
  /* TCH: nearest bin to rmean is the peak index. */
  int pk = 0;
  double best = 1e10;
  for (int i = 0; i < n; i++)
    if (fabs(x[i] - rmean) < best) { best = fabs(x[i] - rmean); pk = i; }
  double mx = -1e30, mn = 1e30;
  for (int i = 0; i < n; i++) {
    if (y[i] > mx) mx = y[i];
    if (y[i] < mn) mn = y[i];
  }
  double range = fmax(mx - mn, 1.0);
  /* pf_estimate_initial_params, PeakFit.c:215-243: background = mean of
   * min(5, n/4) points at each end, amplitude = peak - background (or the
   * peak value when that is not positive), FWHM = half-maximum crossing
   * distance in bins, at least 2. */
  int bw = (int)fmin(5, n / 4);
  if (bw < 1) bw = 1;
  double bs = 0.0;
  for (int i = 0; i < bw; ++i) {
    bs += y[i];
    bs += y[n - 1 - i];
  }
  double bgg = bs / (2.0 * bw);
  double amp = y[pk] - bgg;
  if (amp <= 0) amp = y[pk];
  double half = bgg + (amp / 2.0);
  int l = pk, r = pk;
  while (l > 0 && y[l] > half) l--;
  while (r < n - 1 && y[r] > half) r++;
  double fw_bins = (double)(r - l);
  if (!(fw_bins > 1.0)) fw_bins = 2.0;
  double fw = fw_bins * rstep;
  double fgi = fw * 0.5;
  double sig0 = (fgi * 100.0) * (fgi * 100.0) / (8.0 * log(2.0));
  if (sig0 < 1e-6) sig0 = 1.0;
  double gam0 = fw * 0.5 * 100.0;
  if (gam0 < 0.01) gam0 = 1.0;
  double a0 = amp * fw * 1.5;
  if (a0 <= 0) a0 = range * rstep;
  double crange = fmax(fw, (double)n * rstep * 0.5);
  double ahi = range * (double)n * rstep * 2.0;
  if (ahi < 1.0) ahi = 1.0;
  double shi = (rstep * 200.0 * 100.0) * (rstep * 200.0 * 100.0) / (8.0 * log(2.0));
  double ghi = rstep * 200.0 * 100.0;
  p[0] = fmax(fmin(a0, ahi * 0.9), 0.01); lo[0] = 0; hi[0] = ahi;
  p[1] = x[pk]; lo[1] = x[pk] - crange; hi[1] = x[pk] + crange;
  p[2] = fmax(fmin(sig0, shi * 0.5), 1e-6 * 2.0); lo[2] = 1e-6; hi[2] = shi;
  p[3] = fmax(fmin(gam0, ghi * 0.5), 0.01 * 2.0); lo[3] = 0.01; hi[3] = ghi;
  for (int k = 0; k < 4; k++) {
    if (p[k] <= lo[k]) p[k] = lo[k] + (hi[k] - lo[k]) * 0.01;
    if (p[k] >= hi[k]) p[k] = hi[k] - (hi[k] - lo[k]) * 0.01;
  }
  p[4] = bgg; lo[4] = fmin(mn - range * 0.5, -1.0); hi[4] = mx + range;
  if (p[4] <= lo[4]) p[4] = lo[4] + 0.01;
  if (p[4] >= hi[4]) p[4] = hi[4] - 0.01;
  p[5] = 0.0; lo[5] = -range; hi[5] = range;
}

/* fitPeaks' start values and bounds for a job of nPeaks peaks on the
 * window (PeakFit.c:370-445): per peak, pf_estimate_initial_params at the
 * sample nearest its location (background = mean of min(5, n/4) samples at
 * each end, amplitude = peak - background or the peak value, FWHM = the
 * half-maximum crossing distance in bins, at least 2), area / centre / sig
 * / gam start and bounds; the background [bg0, bg1] from the first peak's
 * estimate and the window's range.  Parameters 4 per peak, then bg0, bg1. */
void orc_tchn_init(int nPeaks, int n, const double *x, const double *y, double rstep,
                   const double *locs, double *p, double *lo, double *hi) {
  double mx = -1e30, mn = 1e30;
  for (int i = 0; i < n; i++) {
    if (y[i] > mx) mx = y[i];
    if (y[i] < mn) mn = y[i];
  }
  double range = fmax(mx - mn, 1.0);
  int bw = (int)fmin(5, n / 4);
  if (bw < 1) bw = 1;
  double bs = 0.0;
  for (int i = 0; i < bw; ++i) bs += y[i] + y[n - 1 - i];
  double bgg = bs / (2.0 * bw);
  for (int k = 0; k < nPeaks; k++) {
    int pk = 0;
    double best = 1e10;
    for (int i = 0; i < n; i++)
      if (fabs(x[i] - locs[k]) < best) { best = fabs(x[i] - locs[k]); pk = i; }
    double amp = y[pk] - bgg;
    if (amp <= 0) amp = y[pk];
    double half = bgg + amp / 2.0;
    int l = pk, r = pk;
    while (l > 0 && y[l] > half) l--;
    while (r < n - 1 && y[r] > half) r++;
    double fw_bins = (double)(r - l);
    if (!(fw_bins > 1.0)) fw_bins = 2.0;
    double fw = fw_bins * rstep, fgi = fw * 0.5;
    double sig0 = (fgi * 100.0) * (fgi * 100.0) / (8.0 * log(2.0));
    if (sig0 < 1e-6) sig0 = 1.0;
    double gam0 = fw * 0.5 * 100.0;
    if (gam0 < 0.01) gam0 = 1.0;
    double a0 = amp * fw * 1.5;
    if (a0 <= 0) a0 = range * rstep;
    double crange = fmax(fw, (double)n * rstep * 0.5);
    double ahi = range * (double)n * rstep * 2.0;
    if (ahi < 1.0) ahi = 1.0;
    double shi = (rstep * 200.0 * 100.0) * (rstep * 200.0 * 100.0) / (8.0 * log(2.0));
    double ghi = rstep * 200.0 * 100.0;
    double *q = p + 4 * k, *ql = lo + 4 * k, *qh = hi + 4 * k;
    q[0] = fmax(fmin(a0, ahi * 0.9), 0.01); ql[0] = 0; qh[0] = ahi;
    q[1] = x[pk]; ql[1] = x[pk] - crange; qh[1] = x[pk] + crange;
    q[2] = fmax(fmin(sig0, shi * 0.5), 1e-6 * 2.0); ql[2] = 1e-6; qh[2] = shi;
    q[3] = fmax(fmin(gam0, ghi * 0.5), 0.01 * 2.0); ql[3] = 0.01; qh[3] = ghi;
    for (int j = 0; j < 4; j++) {
      if (q[j] <= ql[j]) q[j] = ql[j] + (qh[j] - ql[j]) * 0.01;
      if (q[j] >= qh[j]) q[j] = qh[j] - (qh[j] - ql[j]) * 0.01;
    }
  }
  double *b = p + 4 * nPeaks, *bl = lo + 4 * nPeaks, *bh = hi + 4 * nPeaks;
  b[0] = bgg; bl[0] = fmin(mn - range * 0.5, -1.0); bh[0] = mx + range;
  if (b[0] <= bl[0]) b[0] = bl[0] + 0.01;
  if (b[0] >= bh[0]) b[0] = bh[0] - 0.01;
  b[1] = 0.0; bl[1] = -range; bh[1] = range;
}

/* ------------------------------------------------------------------ */
/* FP64 Levenberg-Marquardt (the build's CPU fit, same objective as the */
/* reference: sum of squared residuals, box bounds by projection).      */
/* ------------------------------------------------------------------ */
/* Peaks of a TCH-family model code: 1 = one peak, 3..9 = 2..8 peaks (a
 * merged fitPeaks job, PeakFit.c:321-351); 0 for pV. */
static int orc_tch_peaks(int model) {
  return model == 1 ? 1 : (model >= 3 && model <= 9) ? model - 1 : 0;
}
static int orc_np(int model) { return model == 0 ? 5 : 4 * orc_tch_peaks(model) + 2; }

static void orc_model_jac(int model, int n, const double *x, const double *p,
                          double xlo, double xhi, double *f, double *J) {
  int np = orc_np(model), npk = orc_tch_peaks(model);
  /* Central differences in FP64 with a relative step: the checker only
   * needs the converged optimum, not the GPU's analytic Jacobian. */
  if (model == 0) orc_pv_model(p, n, x, f);
  else orc_tch_model(npk, p, n, x, xlo, xhi, f);
  double q[34], fp[512], fm[512];
  for (int k = 0; k < np; k++) {
    for (int j = 0; j < np; j++) q[j] = p[j];
    double h = fmax(fabs(p[k]), 1e-3) * 1e-6;
    q[k] = p[k] + h;
    if (model == 0) orc_pv_model(q, n, x, fp);
    else orc_tch_model(npk, q, n, x, xlo, xhi, fp);
    q[k] = p[k] - h;
    if (model == 0) orc_pv_model(q, n, x, fm);
    else orc_tch_model(npk, q, n, x, xlo, xhi, fm);
    for (int i = 0; i < n; i++) J[i * np + k] = (fp[i] - fm[i]) / (2 * h);
  }
}

static int orc_chol_solve(int n, double *A, double *b) {
  for (int j = 0; j < n; j++) {
    double s = A[j * n + j];
    for (int k = 0; k < j; k++) s -= A[j * n + k] * A[j * n + k];
    if (s <= 0) return -1;
    double d = sqrt(s);
    A[j * n + j] = d;
    for (int i = j + 1; i < n; i++) {
      double t = A[i * n + j];
      for (int k = 0; k < j; k++) t -= A[i * n + k] * A[j * n + k];
      A[i * n + j] = t / d;
    }
  }
  for (int i = 0; i < n; i++) {
    double t = b[i];
    for (int k = 0; k < i; k++) t -= A[i * n + k] * b[k];
    b[i] = t / A[i * n + i];
  }
  for (int i = n - 1; i >= 0; i--) {
    double t = b[i];
    for (int k = i + 1; k < n; k++) t -= A[k * n + i] * b[k];
    b[i] = t / A[i * n + i];
  }
  return 0;
}

/* The TCH models are fitted in u = sqrt(sig) in place of sig (the Gaussian
 * FWHM is linear in u): dF_G/dsig grows without bound as sig -> 0, which
 * stalls the damped steps once a path passes near a Lorentzian shape.  Same
 * model, same bounds (sqrt of the sig bounds), results returned in sig. */
static void orc_to_ext(int npk, int np, const double *u, double *p) {
  for (int a = 0; a < np; a++) p[a] = u[a];
  for (int k = 0; k < npk; k++) p[4 * k + 2] = u[4 * k + 2] * u[4 * k + 2];
}

static void orc_eval(int model, int npk, int np, int n, const double *x, const double *u,
                     double xlo, double xhi, double *f, double *J) {
  double p[34];
  orc_to_ext(npk, np, u, p);
  if (J) {
    orc_model_jac(model, n, x, p, xlo, xhi, f, J);
    for (int k = 0; k < npk; k++)
      for (int i = 0; i < n; i++) J[i * np + 4 * k + 2] *= 2.0 * u[4 * k + 2];
  } else if (model == 0) {
    orc_pv_model(p, n, x, f);
  } else {
    orc_tch_model(npk, p, n, x, xlo, xhi, f);
  }
}

int orc_lm_fit(int model, int n, const double *x, const double *y, double *p,
               const double *lo_in, const double *hi_in, int maxIter,
               double *chi2_out) {
  int np = orc_np(model), npk = orc_tch_peaks(model);
  if (n > 512 || (model != 0 && npk == 0)) return -1;
  double xlo = x[0], xhi = x[n - 1];
  static __thread double J[512 * 34];
  double f[512], ft[512], u[34], lo[34], hi[34];
  for (int a = 0; a < np; a++) { u[a] = p[a]; lo[a] = lo_in[a]; hi[a] = hi_in[a]; }
  for (int k = 0; k < npk; k++) {
    u[4 * k + 2] = sqrt(fmax(p[4 * k + 2], 0.0));
    lo[4 * k + 2] = sqrt(fmax(lo_in[4 * k + 2], 0.0));
    hi[4 * k + 2] = sqrt(fmax(hi_in[4 * k + 2], 0.0));
  }
  double lam = 1e-3;
  orc_eval(model, npk, np, n, x, u, xlo, xhi, f, J);
  double chi = 0;
  for (int i = 0; i < n; i++) chi += (f[i] - y[i]) * (f[i] - y[i]);
  int it, done = 0;
  for (it = 0; it < maxIter && !done; it++) {
    double JtJ[34 * 34], g[34];
    for (int a = 0; a < np; a++) {
      g[a] = 0;
      for (int i = 0; i < n; i++) g[a] += J[i * np + a] * (y[i] - f[i]);
      for (int b = 0; b < np; b++) {
        double s = 0;
        for (int i = 0; i < n; i++) s += J[i * np + a] * J[i * np + b];
        JtJ[a * np + b] = s;
      }
    }
    int accepted = 0;
    for (int tries = 0; tries < 12 && !accepted; tries++) {
      double M[34 * 34], d[34], q[34];
      for (int a = 0; a < np * np; a++) M[a] = JtJ[a];
      for (int a = 0; a < np; a++) {
        M[a * np + a] += lam * fmax(JtJ[a * np + a], 1e-30);
        d[a] = g[a];
      }
      /* active set: a parameter on a bound whose descent direction points
       * out of the box is held there for this step */
      for (int a = 0; a < np; a++) {
        if ((u[a] <= lo[a] && g[a] < 0) || (u[a] >= hi[a] && g[a] > 0)) {
          for (int b = 0; b < np; b++) M[a * np + b] = M[b * np + a] = 0;
          M[a * np + a] = 1;
          d[a] = 0;
        }
      }
      if (orc_chol_solve(np, M, d) != 0) { lam *= 10; continue; }
      for (int a = 0; a < np; a++) {
        q[a] = u[a] + d[a];
        if (q[a] < lo[a]) q[a] = lo[a];
        if (q[a] > hi[a]) q[a] = hi[a];
      }
      orc_eval(model, npk, np, n, x, q, xlo, xhi, ft, NULL);
      double c2 = 0;
      for (int i = 0; i < n; i++) c2 += (ft[i] - y[i]) * (ft[i] - y[i]);
      if (c2 <= chi) {
        double step = 0, mag = 0;
        for (int a = 0; a < np; a++) {
          step += (q[a] - u[a]) * (q[a] - u[a]);
          mag += u[a] * u[a];
          u[a] = q[a];
        }
        double dchi = chi - c2;
        chi = c2;
        lam = fmax(lam / 10, 1e-12);
        accepted = 1;
        orc_eval(model, npk, np, n, x, u, xlo, xhi, f, J);
        if (step <= 1e-24 * (mag + 1e-24) || dchi <= 1e-15 * (chi + 1e-300)) done = 1;
      } else {
        lam *= 10;
      }
    }
    if (!accepted) break;
  }
  orc_to_ext(npk, np, u, p);
  if (chi2_out) *chi2_out = chi;
  return it;
}
