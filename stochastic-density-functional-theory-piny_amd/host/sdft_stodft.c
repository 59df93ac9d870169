/*
 * sdft_stodft.c -- the reference's filter-path entry points over the C ABI
 * (declarations and field meanings in include/sdft_stodft.h).
 */
#include "sdft_stodft.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void sdft_die(const char *where, int rc) {
  printf("@@@@@@@@@@@@@@@@@@@@_ERROR_@@@@@@@@@@@@@@@@@@@@\n");
  printf("%s failed (status %d): %s\n", where, rc, sdft_hip_last_error());
  printf("@@@@@@@@@@@@@@@@@@@@_ERROR_@@@@@@@@@@@@@@@@@@@@\n");
  fflush(stdout);
  exit(1);
}

#define SDFT_CALL(where, expr)          \
  do {                                  \
    int rc_ = (expr);                   \
    if (rc_ != 0) sdft_die(where, rc_); \
  } while (0)

/* Internal calls go through these static helpers, never through an exported
 * name, so the library's own call graph cannot be rebound by another
 * definition of the same symbol in the process. */
static void coef_force(SDFT_STODFT *st) {
  SDFT_CALL("calcCoefForceWrapSCF",
            sdft_hip_hpsi(st->ctx, st->numStateUpProc, &st->cre_up[1], &st->cim_up[1],
                          &st->fcre_up[1], &st->fcim_up[1], st->prec));
}

static void cheby_coeff(SDFT_STODFT *st, double mu, double *chebyCoeffs, int funFlag) {
  SDFT_CALL("calcChebyCoeff",
            sdft_hip_cheby_coeffs(st->ctx, st->polynormLength, st->energyMin, st->energyMax, mu,
                                  st->beta, funFlag, chebyCoeffs));
}

static void cheby_moments(SDFT_STODFT *st) {
  SDFT_CALL("calcChebyMoments",
            sdft_hip_cheby_moments(st->ctx, st->numStateUpProc, st->polynormLength, st->scale,
                                   st->energyMean, &st->cre_up[1], &st->cim_up[1],
                                   st->chebyMomentsUp, st->prec));
}

SDFT_API void sdft_calcCoefForceWrapSCF(SDFT_STODFT *st) { coef_force(st); }

/* normh.c:182-344 */
SDFT_API void sdft_normHCheby(SDFT_STODFT *st, int iPoly, int cpyFlag) {
  const int nc = st->numCoeff;
  const int tot = st->numStateUpProc * nc;
  double s1 = -st->scale * 0.25;
  double s2 = -st->scale * 0.5;
  double p = -st->scale * st->energyMean;
  double *cre = st->cre_up, *cim = st->cim_up;
  double *fre = st->fcre_up, *fim = st->fcim_up;
  double *wre = st->wfUpRe1, *wim = st->wfUpIm1;
  int is, i;
  coef_force(st);
  if (iPoly == 1) {
    for (is = 0; is < st->numStateUpProc; is++) {
      const int off = is * nc;
      for (i = 1; i < nc; i++) {
        fre[off + i] = s1 * fre[off + i] + p * cre[off + i];
        fim[off + i] = s1 * fim[off + i] + p * cim[off + i];
      }
      fre[off + nc] = s2 * fre[off + nc] + p * cre[off + nc];
    }
    if (cpyFlag == 1) {
      memcpy(&cre[1], &fre[1], tot * sizeof(double));
      memcpy(&cim[1], &fim[1], tot * sizeof(double));
    }
  } else {
    s1 *= 2.0;
    s2 *= 2.0;
    p *= 2.0;
    for (is = 0; is < st->numStateUpProc; is++) {
      const int off = is * nc;
      for (i = 1; i < nc; i++) {
        fre[off + i] = s1 * fre[off + i] + p * cre[off + i] - wre[off + i];
        fim[off + i] = s1 * fim[off + i] + p * cim[off + i] - wim[off + i];
      }
      fre[off + nc] = s2 * fre[off + nc] + p * cre[off + nc] - wre[off + nc];
    }
    if (cpyFlag == 1) {
      memcpy(&wre[1], &cre[1], tot * sizeof(double));
      memcpy(&wim[1], &cim[1], tot * sizeof(double));
      memcpy(&cre[1], &fre[1], tot * sizeof(double));
      memcpy(&cim[1], &fim[1], tot * sizeof(double));
    }
  }
}

/* filter.c:318-706, the whole loop on the device */
SDFT_API void sdft_filterChebyPolyHerm(SDFT_STODFT *st) {
  const int tot = st->numStateUpProc * st->numCoeff;
  SDFT_CALL("filterChebyPolyHerm",
            sdft_hip_cheby_filter(st->ctx, st->numStateUpProc, st->polynormLength, st->expanCoeff,
                                  st->scale, st->energyMean, &st->cre_up[1], &st->cim_up[1],
                                  &st->stoWfUpRe[0][1], &st->stoWfUpIm[0][1], NULL, st->prec));
  if (st->polynormLength > 1) {
    /* normHCheby with cpyFlag 1 leaves fcre = cre = T_{L-1} */
    memcpy(&st->fcre_up[1], &st->cre_up[1], tot * sizeof(double));
    memcpy(&st->fcim_up[1], &st->cim_up[1], tot * sizeof(double));
  }
}

SDFT_API void sdft_calcChebyCoeff(SDFT_STODFT *st, double *chemPot, double *chebyCoeffs,
                                  int funFlag) {
  cheby_coeff(st, chemPot[0], chebyCoeffs, funFlag);
}

SDFT_API void sdft_calcChebyMoments(SDFT_STODFT *st) { cheby_moments(st); }

/* calc-chempot-chebyshev.c:918-972 (moments already normalised by N_sto) */
SDFT_API double sdft_calcNumElecCheby(SDFT_STODFT *st, double chemPot, double *chebyCoeffs) {
  double ne = 0.0;
  int k;
  cheby_coeff(st, chemPot, chebyCoeffs, 1);
  for (k = 0; k < st->polynormLength; k++) ne += chebyCoeffs[k] * st->chebyMomentsUp[k];
  return ne * st->occNumber;
}

/* calc-chempot-chebyshev.c:30-315 */
SDFT_API void sdft_calcChemPotCheby(SDFT_STODFT *st) {
  double mu = 0.0, ne = 0.0;
  int k;
  cheby_moments(st);
  for (k = 0; k <= st->polynormLength; k++) st->chebyMomentsUp[k] /= (double)st->numStateStoUp;
  SDFT_CALL("calcChemPotCheby",
            sdft_hip_chem_pot(st->ctx, st->polynormLength, st->energyMin, st->energyMax, st->beta,
                              st->chebyMomentsUp, 1.0, st->numElecTrue, st->chemPotInit,
                              st->gapInit, st->occNumber, 1.0e-11 * st->numElecTrue, &mu, &ne));
  st->chemPot[0] = mu;
  printf("Correct Chemical Potential is %.16lg\n", mu);
}

SDFT_API void sdft_calcKSPot(SDFT_STODFT *st) {
  double en[4];
  SDFT_CALL("calcKSPot", sdft_hip_ks_potential(st->ctx, st->rhoUp, SDFT_RHO_FFTW, &st->vextr[1],
                                               &st->vexti[1], st->volume, st->vxcType, en,
                                               st->vKsUp));
  st->ehart = en[0];
  st->eext = en[1];
  st->exc = en[2];
  st->muxc = en[3];
}

SDFT_API void sdft_calcRhoStoHybridCheby(SDFT_STODFT *st) {
  SDFT_CALL("calcRhoStoHybridCheby",
            sdft_hip_density(st->ctx, st->numStateUpProc, &st->stoWfUpRe[0][1],
                             &st->stoWfUpIm[0][1], st->occNumber / (double)st->numStateStoUp,
                             st->rhoUp, SDFT_RHO_FFTW, 0));
}

/* filter.c:33-312: one device call for the whole Newton recurrence; the
 * nmu filtered blocks come back contiguous and are scattered to
 * stoWfUpRe/Im[imu] */
SDFT_API void sdft_filterNewtonPolyHerm(SDFT_STODFT *st) {
  const int tot = st->numStateUpProc * st->numCoeff;
  const int nmu = st->numChemPot > 0 ? st->numChemPot : 1;
  double *ore = (double *)malloc(sizeof(double) * (size_t)tot * nmu);
  double *oim = (double *)malloc(sizeof(double) * (size_t)tot * nmu);
  int imu;
  if (!ore || !oim) sdft_die("filterNewtonPolyHerm", SDFT_ENOMEM);
  SDFT_CALL("filterNewtonPolyHerm",
            sdft_hip_newton_filter(st->ctx, st->numStateUpProc, st->polynormLength, nmu,
                                   st->expanCoeff, st->sampPoint, st->newtonScale,
                                   st->energyMean, &st->cre_up[1], &st->cim_up[1], ore, oim,
                                   st->prec));
  for (imu = 0; imu < nmu; imu++) {
    memcpy(&st->stoWfUpRe[imu][1], ore + (size_t)imu * tot, tot * sizeof(double));
    memcpy(&st->stoWfUpIm[imu][1], oim + (size_t)imu * tot, tot * sizeof(double));
  }
  if (st->polynormLength > 1) {
    memcpy(&st->fcre_up[1], &st->cre_up[1], tot * sizeof(double));
    memcpy(&st->fcim_up[1], &st->cim_up[1], tot * sizeof(double));
  }
  free(ore);
  free(oim);
}

static void spectral(SDFT_STODFT *st, int which, double *re, double *im, const char *name) {
  double e = 0.0;
  int it = 0;
  SDFT_CALL(name, sdft_hip_spectral_bound(st->ctx, which, st->energyMax, 1.0e-5, 1000000, re, im,
                                          &e, &it, st->prec));
  if (which == 0)
    st->energyMax = e + 0.1;
  else
    st->energyMin = e - 0.1;
  st->energyMean = 0.5 * (st->energyMin + st->energyMax);
  st->energyDiff = st->energyMax - st->energyMin;
}

SDFT_API void sdft_genEnergyMax(SDFT_STODFT *st) { spectral(st, 0, st->testWfMaxRe, st->testWfMaxIm, "genEnergyMax"); }

SDFT_API void sdft_genEnergyMin(SDFT_STODFT *st) { spectral(st, 1, st->testWfMinRe, st->testWfMinIm, "genEnergyMin"); }
