/*
 * midas_oracle.h -- FP64 CPU restatement of the MIDAS far-field caking path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the
 * smoke() entry of __graft_entry__.py and the cpu_baseline leg of bench.py
 * may load it; nothing in the product (midas_amd/) links, imports or calls
 * it.  Each function cites the reference file:line it restates.
 *
 * Parity status of each part (what pins it):
 *   orc_parse_mapper_params   -- DetectorMapper.c:116-375 key semantics;
 *                                pinned by tests/test_params.py fixtures.
 *   orc_bin_edges             -- closed form + compiled reference driver.
 *   orc_pixel_to_reta         -- compiled dg_pixel_to_REta_corr vectors
 *                                (oracle/_ref) + flat-detector closed form.
 *   orc_map_build             -- compiled reference mapper_build_map
 *                                (oracle/_ref) byte-for-byte, golden digests.
 *   orc_integrate             -- compiled integration_apply_map + brute force.
 *   orc_lineouts              -- closed-form reductions (tests).
 *   orc_pv / orc_tch model    -- closed forms (area by quadrature, limits).
 *   orc_lm_fit                -- parity unpinned against the reference
 *                                optimiser (NLopt, unpinned `master`, absent);
 *                                pinned by exact recovery on noise-free data.
 */
#ifndef MIDAS_ORACLE_H
#define MIDAS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int NrPixelsY, NrPixelsZ;
  double pxY, pxZ, yCen, zCen, Lsd, RhoD, tx, ty, tz;
  double p[15];
  double RMin, RMax, RBinSize, EtaMin, EtaMax, EtaBinSize;
  double QMin, QMax, QBinSize, Wavelength;
  int NrTransOpt;
  int TransOpt[10];
  int SubPixelLevel;
  double SubPixelCardinalWidth;
  double Parallax;
  int SolidAngleCorrection, PolarizationCorrection;
  double PolarizationFraction;
  int NPanelsY, NPanelsZ, PanelSizeY, PanelSizeZ;
  int PanelGapsY[64], PanelGapsZ[64];
  int hasMask, hasDistortion, hasResidual;
  char MaskFile[1024], DistortionFile[1024], ResidualCorrectionMap[1024],
      PanelShiftsFile[1024];
} orc_params;

typedef struct {
  int id, yMin, yMax, zMin, zMax;
  double dY, dZ, dTheta, dLsd, dP2, centerY, centerZ;
} orc_panel;

/* One Map.bin record (24 bytes, MapperCore.h:19-25). */
#pragma pack(push, 4)
typedef struct {
  float y, z;
  double frac;
  float deltaR, areaWeight;
} orc_entry;
#pragma pack(pop)

void orc_params_default(orc_params *P);
int orc_parse_mapper_params(const char *path, orc_params *P);
/* Edge arrays sized by orc_bin_counts; returns 0/negative. */
int orc_bin_counts(const orc_params *P, int *nR, int *nEta);
int orc_bin_edges(orc_params *P, double *RLo, double *RHi, double *ELo,
                  double *EHi);
void orc_tilt_matrix(double tx, double ty, double tz, double T[9]);
void orc_pixel_to_reta(const orc_params *P, const double T[9], double Y,
                       double Z, double dLsd, double dP2,
                       const double *residual, double *R, double *Eta);
int orc_generate_panels(const orc_params *P, orc_panel *out, int maxPanels);
void orc_image_transform(int nOpt, const int *opt, double *img, int NY,
                         int NZ);
void orc_header_string(const orc_params *P, int qMode, char *buf, int len);

/* Map build: returns an opaque handle; query sizes then copy out. */
typedef struct orc_map orc_map;
orc_map *orc_map_build(const orc_params *P, const double *mask,
                       const double *distY, const double *distZ,
                       const double *residual, const orc_panel *panels,
                       int nPanels, int nThreads);
int64_t orc_map_nnz(const orc_map *m);
int orc_map_nbins(const orc_map *m);
void orc_map_copy(const orc_map *m, orc_entry *entries, int32_t *nmap_pairs,
                  int32_t *mask_flags);
void orc_map_counters(const orc_map *m, int64_t out[3]);
void orc_map_free(orc_map *m);

/* Integration on an already dark-subtracted FP64 image.
 * semantics 0: IntegratorZarrOMP.c:1710-1778 (bilinear, I!=0 normalise,
 *              NaN for flagged bins);
 * semantics 1: IntegratorFitPeaksGPUStream.cu:975-1050 (float offsets in
 *              floor mode, A>1e-9 threshold, no NaN). */
void orc_integrate(const orc_entry *entries, const int32_t *nmap_pairs,
                   const int32_t *mask_flags, int nR, int nEta,
                   const double *image, int NY, int NZ, double BCy,
                   double BCz, int gradient, int normalize, int semantics,
                   int nThreads, double *cake, double *area);
/* Three lineout definitions: cpu mean of nonzero & non-NaN, area-weighted,
 * simple mean over A>1e-9. area_inb = in-bounds area per bin. */
void orc_lineouts(const double *cake, const double *area_inb, int nR,
                  int nEta, double *lo_cpu, double *lo_area, double *lo_mean);
void orc_reta_map(const double *RLo, const double *RHi, const double *ELo,
                  const double *EHi, int nR, int nEta, const double *area,
                  double px, double Lsd, double Lam, double *out5);
void orc_area_inbounds(const orc_entry *entries, const int32_t *nmap_pairs,
                       int nBins, int NY, int NZ, double *area);

/* Peak models (PeakFit.c:17-90, 882-920). */
void orc_tch_eta_fwhm(double sig, double gam, double *fwhm, double *eta);
void orc_tch_model(int nPeaks, const double *params, int n, const double *x,
                   double xlo, double xhi, double *out);
void orc_pv_model(const double *p5, int n, const double *x, double *out);
/* FP64 Levenberg-Marquardt with box bounds; model 0 = pV-5
 * [Rcen, Mu, Gamma, Imax, BG], model 1 = TCH single peak
 * [area, center, sig, gam, bg0, bg1], models 3..9 = TCH with 2..8 peaks
 * (4 params each, then bg0, bg1). Returns iterations. */
int orc_lm_fit(int model, int n, const double *x, const double *y,
               double *params, const double *lo, const double *hi,
               int maxIter, double *chi2_out);
void orc_window_init(int model, int n, const double *x, const double *y,
                     double rstep, double rmean, double *p, double *lo,
                     double *hi);
void orc_tchn_init(int nPeaks, int n, const double *x, const double *y, double rstep,
                   const double *locs, double *p, double *lo, double *hi);

#ifdef __cplusplus
}
#endif
#endif
