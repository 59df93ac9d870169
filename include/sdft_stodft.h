/*
 * sdft_stodft.h -- host C shim with the reference's stochastic-DFT filter-path
 * entry points (stodft/proto_stodft_local.h) over the sdft_hip C ABI.
 *
 * The reference functions take PINY's CP / CLASS / GENERAL_DATA structs; this
 * shim gathers exactly the fields they read into one struct, SDFT_STODFT,
 * keeps the reference's array conventions (1-based `cmalloc(n)-1` arrays,
 * state-major re/im coefficient arrays with G=0 at index ncoef) and its error
 * behaviour (print an @@@..._ERROR_@@@ banner, exit(1)).
 *
 * Every entry point carries the reference function's name with an `sdft_`
 * prefix (sdft_filterChebyPolyHerm for filterChebyPolyHerm, ...), so the
 * library links into a PINY process whose own stodft/ objects still define
 * the unprefixed names with their CP / CLASS / GENERAL_DATA signatures: the
 * reference function keeps its prototype (stodft/proto_stodft_local.h) and
 * its body fills SDFT_STODFT from its structs and calls the prefixed entry
 * point.  libsdft_host.so exports only `sdft_*` symbols and binds its own
 * calls internally.
 */
#ifndef SDFT_STODFT_H
#define SDFT_STODFT_H

#include "sdft_hip.h"

#ifndef SDFT_API
#define SDFT_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct SDFT_STODFT {
  sdft_ctx *ctx;            /* from sdft_hip_init + set_vks + set_projectors */
  int prec;                 /* SDFT_PREC_FP64 / SDFT_PREC_FP32 */
  /* CPCOEFFS_INFO / STODFTINFO / CHEBYSHEVINFO scalars */
  int numStateUpProc;       /* nstate_up_proc */
  int numCoeff;             /* ncoef */
  int polynormLength;       /* Chebyshev length L */
  int numStateStoUp;        /* N_sto (all ranks) */
  double energyMin, energyMax, energyMean, energyDiff;
  double scale;             /* chebyshevInfo->scale = 2 / energyDiff */
  double beta;              /* 1 / kT */
  double occNumber;         /* 2 for unpolarised */
  double numElecTrue;       /* N_e */
  double chemPotInit, gapInit;
  /* CPCOEFFS_POS / STODFTCOEFPOS arrays (1-based, numStateUpProc*numCoeff) */
  double *cre_up, *cim_up;
  double *fcre_up, *fcim_up;
  double *wfUpRe1, *wfUpIm1;   /* T_{k-1} between normHCheby calls */
  double **stoWfUpRe, **stoWfUpIm; /* [0] = filtered orbitals */
  double *expanCoeff;       /* [polynormLength] (numChemPot = 1), 0-based */
  double *chebyMomentsUp;   /* [polynormLength + 1], 0-based */
  double *chemPot;          /* [1] */
  double *rhoUp;            /* [n1*n2*n3] density output, 0-based, FFTW order */
  /* Newton-Hermite filter (NEWTONINFO): numChemPot columns of expanCoeff
   * ([polynormLength * numChemPot], iPoly-major), sample points
   * [polynormLength - 1], the Newton map's scale; stoWfUpRe/Im[imu] */
  int numChemPot;
  double newtonScale;
  double *sampPoint;
  /* spectral range start / result vectors (STODFTCOEFPOS testWfMax/Min,
   * 0-based [numCoeff]) */
  double *testWfMaxRe, *testWfMaxIm, *testWfMinRe, *testWfMinIm;
  /* Kohn-Sham potential (cpGetVksStodft): the context's density sphere set
   * by sdft_hip_ks_init; vextr / vexti the local external potential on it
   * (CPSCR_LOC vextr_loc / vexti_loc, 1-based [ncoef_l]), the cell volume,
   * the LDA form (SDFT_XC_*), and on return the STAT_AVG energies and
   * (nullable, 0-based [n1*n2*n3], FFTW order) v_ks_up */
  double *vextr, *vexti;
  double volume;
  int vxcType;
  double ehart, eext, exc, muxc;
  double *vKsUp;
} SDFT_STODFT;

/* stodft/filter.c:318-706: acc = sum_k c_k T_k chi; cre_up left = T_{L-1} chi. */
SDFT_API void sdft_filterChebyPolyHerm(SDFT_STODFT *st);
/* stodft/normh.c:182-344: one H application + recurrence step iPoly on the
 * host arrays (cpyFlag = 1 rotates cre <- T_k, wfUpRe1 <- T_{k-1}). */
SDFT_API void sdft_normHCheby(SDFT_STODFT *st, int iPoly, int cpyFlag);
/* stodft/reduced-energy/energy-wrapper-scf.c:1066-1197: fcre/fcim = raw force. */
SDFT_API void sdft_calcCoefForceWrapSCF(SDFT_STODFT *st);
/* stodft/coeff.c:1093-1223 (funFlag 1 = f, 2 = sqrt f). */
SDFT_API void sdft_calcChebyCoeff(SDFT_STODFT *st, double *chemPot, double *chebyCoeffs, int funFlag);
/* stodft/calc-chempot-chebyshev.c:596-913 (cre_up = chi, unchanged). */
SDFT_API void sdft_calcChebyMoments(SDFT_STODFT *st);
/* stodft/calc-chempot-chebyshev.c:918-972. */
SDFT_API double sdft_calcNumElecCheby(SDFT_STODFT *st, double chemPot, double *chebyCoeffs);
/* stodft/calc-chempot-chebyshev.c:30-315: moments, then chemPot[0] = mu. */
SDFT_API void sdft_calcChemPotCheby(SDFT_STODFT *st);
/* stodft/density.c:1276-1520 (real-space part): rhoUp = occ/N_sto sum psi^2. */
SDFT_API void sdft_calcRhoStoHybridCheby(SDFT_STODFT *st);
/* stodft/filter.c:33-312 with normHNewtonHerm (normh.c:35-176):
 * stoWfUpRe/Im[imu] = sum_n expanCoeff[n*numChemPot+imu] P_n chi;
 * cre_up left = P_{L-1} chi (fcre_up likewise, as normHNewtonHerm leaves it). */
SDFT_API void sdft_filterNewtonPolyHerm(SDFT_STODFT *st);
/* stodft/calc-spectral-range.c:33-333 / :339-566: power iteration from
 * testWfMax / testWfMin (written back), energyMax = E + 0.1 /
 * energyMin = E - 0.1 (eigValMin's E), energyConv 1e-5. */
SDFT_API void sdft_genEnergyMax(SDFT_STODFT *st);
SDFT_API void sdft_genEnergyMin(SDFT_STODFT *st);
/* stodft/reduced-energy/energy-wrapper-scf.c:584-745 (calcKSPot ->
 * cpGetVksStodft, iperd = 3, no dual grid, LDA) from rhoUp (as
 * calcRhoStoHybridCheby leaves it): the context's v_KS is replaced on the
 * device; ehart / eext / exc / muxc and vKsUp (if set) are written. */
SDFT_API void sdft_calcKSPot(SDFT_STODFT *st);

#ifdef __cplusplus
}
#endif
#endif /* SDFT_STODFT_H */
