/*
 * ref_driver.c -- a thin driver around the reference's own mapping and
 * integration sources, compiled in place from /root/reference by
 * oracle/Makefile into oracle/_ref/libmidasref.so.  TEST INFRASTRUCTURE
 * ONLY: it exists to pin the C restatement in midas_oracle.c and to write
 * the golden fixtures under tests/golden/.
 *
 * What the driver does, in the order DetectorMapper.c does it:
 *   - builds bin edges (dg_build_bin_edges, or the Q-mode formula of
 *     DetectorMapper.c:714-746, which lives in a file that cannot be built
 *     here because it needs blosc2/libzip/libtiff);
 *   - applies midas_image_transform to the mask and distortion maps
 *     (DetectorMapper.c:640-676);
 *   - calls mapper_build_map single-threaded so the within-bin order is the
 *     deterministic single-thread order;
 *   - flattens bin-major like DetectorMapper.c:813-834.
 * and integrates with integration_apply_map + the normalise/NaN rules of
 * IntegratorZarrOMP.c:1748-1767.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "DetectorGeometry.h"
#include "ImageUtils.h"
#include "IntegrationCore.h"
#include "MapperCore.h"
#include "Panel.h"
#include "midas_oracle.h"

typedef struct {
  int64_t nnz;
  int nR, nEta;
  struct MapPixelData *entries;
  int32_t *pairs;
  int32_t *flags;
} ref_map;

ref_map *ref_build_map(const orc_params *Pin, const double *mask_raw,
                       const double *distY_raw, const double *distZ_raw,
                       const double *residual, const double *panel_shifts,
                       int nShiftRows) {
  orc_params P = *Pin;
  const int NY = P.NrPixelsY, NZ = P.NrPixelsZ;
  const size_t npx = (size_t)NY * NZ;
  int nEta = (int)ceil((P.EtaMax - P.EtaMin) / P.EtaBinSize);
  int qMode = (P.QBinSize > 0 && P.Wavelength > 0);
  int nR;
  if (qMode) {
    if (P.QMin > 0)
      P.RMin = (P.Lsd / P.pxY) * tan(2.0 * asin(P.QMin * P.Wavelength / (4.0 * M_PI)));
    if (P.QMax > 0)
      P.RMax = (P.Lsd / P.pxY) * tan(2.0 * asin(P.QMax * P.Wavelength / (4.0 * M_PI)));
    nR = (int)ceil((P.QMax - P.QMin) / P.QBinSize);
  } else {
    nR = (int)ceil((P.RMax - P.RMin) / P.RBinSize);
  }
  double *ELo = malloc(nEta * sizeof(double)), *EHi = malloc(nEta * sizeof(double));
  double *RLo = malloc(nR * sizeof(double)), *RHi = malloc(nR * sizeof(double));
  if (qMode) {
    for (int q = 0; q < nR; q++) {
      double ql = P.QMin + P.QBinSize * q, qh = P.QMin + P.QBinSize * (q + 1);
      RLo[q] = (P.Lsd / P.pxY) * tan(2.0 * asin(ql * P.Wavelength / (4.0 * M_PI)));
      RHi[q] = (P.Lsd / P.pxY) * tan(2.0 * asin(qh * P.Wavelength / (4.0 * M_PI)));
    }
    for (int e = 0; e < nEta; e++) {
      ELo[e] = P.EtaBinSize * e + P.EtaMin;
      EHi[e] = P.EtaBinSize * (e + 1) + P.EtaMin;
    }
  } else {
    dg_build_bin_edges(P.RMin, P.EtaMin, nR, nEta, P.RBinSize, P.EtaBinSize,
                       RLo, RHi, ELo, EHi);
  }
  Panel *panels = NULL;
  int nPanels = 0;
  if (P.NPanelsY > 0 && P.NPanelsZ > 0) {
    GeneratePanels(P.NPanelsY, P.NPanelsZ, P.PanelSizeY, P.PanelSizeZ,
                   P.PanelGapsY, P.PanelGapsZ, &panels, &nPanels);
    for (int r = 0; r < nShiftRows; r++) {
      int id = (int)panel_shifts[r * 6];
      if (id >= 0 && id < nPanels) {
        panels[id].dY = panel_shifts[r * 6 + 1];
        panels[id].dZ = panel_shifts[r * 6 + 2];
        panels[id].dTheta = panel_shifts[r * 6 + 3];
        panels[id].dLsd = panel_shifts[r * 6 + 4];
        panels[id].dP2 = panel_shifts[r * 6 + 5];
      }
    }
  }
  double *mask = NULL;
  if (mask_raw) {
    mask = malloc(npx * sizeof(double));
    midas_image_transform(P.NrTransOpt, P.TransOpt, (double *)mask_raw, mask, NY, NZ);
  }
  double *dY = calloc(npx, sizeof(double)), *dZ = calloc(npx, sizeof(double));
  if (distY_raw) midas_image_transform(P.NrTransOpt, P.TransOpt, (double *)distY_raw, dY, NY, NZ);
  if (distZ_raw) midas_image_transform(P.NrTransOpt, P.TransOpt, (double *)distZ_raw, dZ, NY, NZ);
  DGResidualCorr rc = {residual, NY, NZ};

  struct MapPixelData ***pxList = malloc(nR * sizeof(*pxList));
  int **nPx = malloc(nR * sizeof(*nPx)), **maxPx = malloc(nR * sizeof(*maxPx));
  for (int r = 0; r < nR; r++) {
    pxList[r] = calloc(nEta, sizeof(**pxList));
    nPx[r] = calloc(nEta, sizeof(int));
    maxPx[r] = calloc(nEta, sizeof(int));
  }
  long long nb = (long long)nR * nEta;
  int *flags = calloc(nb, sizeof(int));
  omp_lock_t *locks = malloc(MAPPER_N_BIN_LOCKS * sizeof(omp_lock_t));
  for (int l = 0; l < MAPPER_N_BIN_LOCKS; l++) omp_init_lock(&locks[l]);
  omp_set_num_threads(1);
  long long tot = mapper_build_map(
      P.tx, P.ty, P.tz, NY, NZ, P.pxY, P.pxZ, P.yCen, P.zCen, P.Lsd, P.RhoD,
      P.p[0], P.p[1], P.p[2], P.p[3], P.p[4], P.p[5], P.p[6], P.p[7], P.p[8],
      P.p[9], P.p[10], P.p[11], P.p[12], P.p[13], P.p[14], ELo, EHi, RLo, RHi,
      nR, nEta, pxList, nPx, maxPx, mask, flags, P.NrTransOpt, P.TransOpt,
      locks, P.SubPixelLevel, P.SubPixelCardinalWidth, P.Parallax,
      P.SolidAngleCorrection, P.PolarizationCorrection, P.PolarizationFraction,
      dY, dZ, panels, nPanels, residual ? &rc : NULL);
  for (int l = 0; l < MAPPER_N_BIN_LOCKS; l++) omp_destroy_lock(&locks[l]);
  free(locks);

  ref_map *m = calloc(1, sizeof(ref_map));
  m->nnz = tot;
  m->nR = nR;
  m->nEta = nEta;
  m->entries = malloc((tot ? tot : 1) * sizeof(struct MapPixelData));
  m->pairs = malloc(2 * nb * sizeof(int32_t));
  m->flags = malloc(nb * sizeof(int32_t));
  long long cur = 0;
  for (int r = 0; r < nR; r++)
    for (int e = 0; e < nEta; e++) {
      long long b = (long long)r * nEta + e;
      m->pairs[2 * b] = nPx[r][e];
      m->pairs[2 * b + 1] = (int32_t)cur;
      for (int k = 0; k < nPx[r][e]; k++) m->entries[cur + k] = pxList[r][e][k];
      cur += nPx[r][e];
      m->flags[b] = flags[b];
      free(pxList[r][e]);
    }
  for (int r = 0; r < nR; r++) {
    free(pxList[r]);
    free(nPx[r]);
    free(maxPx[r]);
  }
  free(pxList);
  free(nPx);
  free(maxPx);
  free(flags);
  free(mask);
  free(dY);
  free(dZ);
  free(panels);
  free(RLo);
  free(RHi);
  free(ELo);
  free(EHi);
  return m;
}

int64_t ref_map_nnz(const ref_map *m) { return m->nnz; }
int ref_map_nbins(const ref_map *m) { return m->nR * m->nEta; }

void ref_map_copy(const ref_map *m, void *entries, int32_t *pairs,
                  int32_t *flags) {
  long long nb = (long long)m->nR * m->nEta;
  if (entries) memcpy(entries, m->entries, m->nnz * sizeof(struct MapPixelData));
  if (pairs) memcpy(pairs, m->pairs, 2 * nb * sizeof(int32_t));
  if (flags) memcpy(flags, m->flags, nb * sizeof(int32_t));
}

void ref_map_free(ref_map *m) {
  if (!m) return;
  free(m->entries);
  free(m->pairs);
  free(m->flags);
  free(m);
}

int ref_entry_size(void) { return (int)sizeof(struct MapPixelData); }

/* integration_apply_map over a flat map, then IntegratorZarrOMP's rules:
 * Intensity /= totArea when Intensity != 0 and Normalize (:1748-1752),
 * NaN when the bin is flagged (:1766-1767). */
void ref_integrate(const void *entries, const int32_t *pairs,
                   const int32_t *flags, int nR, int nEta, const double *image,
                   int NY, int NZ, double BCy, double BCz, int gradient,
                   int normalize, double *cake, double *area) {
  const struct MapPixelData *E = (const struct MapPixelData *)entries;
  struct MapPixelData ***pl = malloc(nR * sizeof(*pl));
  int **np = malloc(nR * sizeof(*np));
  for (int r = 0; r < nR; r++) {
    pl[r] = malloc(nEta * sizeof(**pl));
    np[r] = malloc(nEta * sizeof(int));
    for (int e = 0; e < nEta; e++) {
      long long b = (long long)r * nEta + e;
      np[r][e] = pairs[2 * b];
      pl[r][e] = (struct MapPixelData *)(E + pairs[2 * b + 1]);
    }
  }
  integration_apply_map((struct MapPixelData **const *)pl, (int *const *)np,
                        nR, nEta, image, NY, NZ, NULL, BCy, BCz, gradient,
                        cake, area);
  long long nb = (long long)nR * nEta;
  for (long long b = 0; b < nb; b++) {
    if (cake[b] != 0 && normalize == 1) cake[b] /= area[b];
    if (flags && flags[b]) cake[b] = NAN;
  }
  for (int r = 0; r < nR; r++) {
    free(pl[r]);
    free(np[r]);
  }
  free(pl);
  free(np);
}

/* dg_pixel_to_REta_corr on n points (no panel offsets). */
void ref_pixel_to_reta(const orc_params *P, int n, const double *Y,
                       const double *Z, double *R, double *Eta) {
  double T[3][3];
  dg_build_tilt_matrix(P->tx, P->ty, P->tz, T);
  for (int i = 0; i < n; i++)
    dg_pixel_to_REta_corr(Y[i], Z[i], P->yCen, P->zCen, T, P->Lsd, P->RhoD,
                          P->p[0], P->p[1], P->p[2], P->p[3], P->p[4], P->p[5],
                          P->p[6], P->p[7], P->p[8], P->p[9], P->p[10],
                          P->p[11], P->p[12], P->p[13], P->p[14], P->pxY, 0, 0,
                          P->Parallax, NULL, &R[i], &Eta[i], NULL);
}
