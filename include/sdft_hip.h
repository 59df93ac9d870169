/*
 * sdft_hip.h -- C ABI of the MI355X (gfx950) stochastic-DFT filter library.
 *
 * Drop-in boundary for PINY's sDFT Chebyshev filter path.  Every entry point
 * takes plain pointers and sizes; arrays are 0-based (the reference's
 * 1-based `cmalloc(n)-1` arrays are passed as `&arr[1]`).  Orbital blocks use
 * the reference's storage: separate real / imaginary FP64 arrays, state-major,
 * `ncoef` coefficients per state with G=0 stored last (entry ncoef-1), Gamma
 * symmetry c(-G) = c(G)*.  The real-space grid is the FFTW (a,b,c) row-major
 * grid with c fastest.
 *
 * Every function returns 0 on success and a negative SDFT_E* code on error;
 * sdft_hip_last_error() returns a message for the last error on the calling
 * thread.  Calls on one context are not thread-safe; use one context per
 * host thread (the reference's one MPI rank per state group).
 */
#ifndef SDFT_HIP_H
#define SDFT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define SDFT_OK 0
#define SDFT_EINVAL (-1)   /* bad argument / unsupported shape */
#define SDFT_EHIP (-2)     /* HIP runtime error */
#define SDFT_ENOMEM (-3)   /* device allocation failed */
#define SDFT_ESTATE (-4)   /* call out of order (e.g. no v_KS set) */

#define SDFT_PREC_FP64 0   /* plain form: FP64 throughout */
#define SDFT_PREC_FP32 1   /* precision form: FP32 FFT / v.psi / KB GEMM (MFMA),
                              FP64 recurrence, accumulators and reductions */

#define SDFT_RHO_FFTW 0    /* density in FFTW order (a,b,c), c fastest */
#define SDFT_RHO_GENERIC 1 /* density in the generic PINY order, a fastest */

#define SDFT_XC_PZ_LDA 0   /* Perdew-Zunger LDA (vxc_typ "pz_lda") */
#define SDFT_XC_PW_LDA 1   /* Perdew-Wang 1992 LDA ("pw_lda") */
#define SDFT_XC_PADE_LDA 2 /* Goedecker-Teter-Hutter Pade LDA ("pade_lda") */

typedef struct sdft_ctx sdft_ctx;     /* one system: grid, G-set, v_KS, KB table */
typedef struct sdft_block sdft_block; /* a device-resident block of orbitals */

/* Grid and plane-wave set.  Replaces the fields of PARA_FFT_PKG3D that the
 * filter reads (include/typ_defs/typedefs_gen.h:383-516: nkf1..3, ncoef,
 * mapFFTWFilter, mapConFFTWFilter built at mathlib/fft_create_package.c:237-248)
 * and CPEWALD::ak2Kinetic (energy/cp/cp_energy_eext.c:2259-2304). */
typedef struct sdft_grid_desc {
  int n1, n2, n3;            /* FFT grid (nkf1, nkf2, nkf3); cubic n1=n2=n3
                                in {24,32,36,48,64,96,128,192} */
  int ncoef;                 /* coefficients per state incl. G=0 (last) */
  const int *map;            /* [ncoef] 0-based grid index of +G (mapFFTWFilter) */
  const int *map_conj;       /* [ncoef] 0-based grid index of -G (mapConFFTWFilter) */
  const double *ak2_kinetic; /* [ncoef] |G|^2 used by the kinetic term */
} sdft_grid_desc;

/* Create a context on the current HIP device.  Replaces the FFTW plan and map
 * set-up of fft_create_package.c:1627-1689 for the filter's grid. */
int sdft_hip_init(const sdft_grid_desc *desc, sdft_ctx **out);
int sdft_hip_finalize(sdft_ctx *ctx);
const char *sdft_hip_last_error(void);

/* Kohn-Sham potential v_KS(r), n1*n2*n3 doubles in FFTW order (the
 * cpscr_rho.v_ks_up array after the transpose of
 * energy/cp/cp_energy_ee_rho.c:1944-1970). */
int sdft_hip_set_vks(sdft_ctx *ctx, const double *v_ks);

/* Kleinman-Bylander projectors on the coefficient set: Q_p(G) = proj_re +
 * i proj_im ([nproj][ncoef], G=0 last) and E_p (KB energy / volume).  The
 * nonlocal energy of a state is sum_p E_p Z_p^2 with Z_p = sum_{all G} c(G)
 * Q_p(G).  Replaces the per-call tables of the G-space KB path
 * (stodft/reduced-energy/cp-energy-eext-stodft.c:701-1168, getnl_fcoef
 * energy/cp/cp_energy_eext_nonloc.c:1669+).  nproj = 0 disables the term. */
int sdft_hip_set_projectors(sdft_ctx *ctx, int nproj, const double *proj_re,
                            const double *proj_im, const double *proj_energy);

/* Real-space KB projectors (nlppKBRealEnergyFilter,
 * energy/cp/cp_energy_eext_nonloc_real_filter.c:110-450, called from
 * controlPotentialEnergyFilter :36-105 when nlpp_real_on).  Projector p takes
 * the values values[offsets[p] .. offsets[p+1]) at the FFTW linear grid
 * indices grid_index[...] (the atom's gridNlppMap points with one
 * vnlPhiAtomGridRe/Im row per (l, radial, m) channel; indices distinct within
 * a projector); energy[p] is its KB energy (vpsNormList, doubled for the m > 0
 * Re/Im rows as at :318-404); vol_elem = V/N (volElem).  The reference adds
 *   V psi(r) = sum_p energy_p beta_p(r) vol_elem sum_r' beta_p(r') psi(r')
 * to v_KS psi on the grid before the forward FFT.  The library applies
 * that operator in this per-atom form on the real-space grid (see
 * sdft_hip_set_nl_form); restricted to the plane-wave set it is exactly a KB
 * operator with table Q_p(G) = N^(-1/2) sum_r beta_p(r) e^{iG.r} and energies
 * energy_p vol_elem, which the library builds on the device when the force
 * path or nl_form 1 needs it.  Replaces any G-space table. */
int sdft_hip_set_projectors_real(sdft_ctx *ctx, int nproj, const long long *offsets,
                                 const int *grid_index, const double *values,
                                 const double *energy, double vol_elem);

/* Gradient tables of the real-space projectors for the Hellmann-Feynman NL
 * force in its per-atom form (nlppKBRealEnergyForce,
 * energy/cp/cp_energy_eext_nonloc_real_force.c:36-240): dx/dy/dz hold
 * d beta_p / d r_j at the same points and in the same order as the values
 * of the last sdft_hip_set_projectors_real call (offsets[nproj] entries each;
 * PINY's vnlPhiDx/Dy/DzAtomGridRe/Im rows, the derivative with respect to the
 * grid point r - R_a).  With them set, sdft_hip_nl_forces computes
 *   F_a = (occ/nsto) sum_s sum_{p in a} 2 energy_p vol_elem Z_ps Zd_ps / N
 * with Z = sum_r beta_p psi_s and Zd = sum_r (d beta_p / d r) psi_s on the
 * real-space grid (psi_s the unnormalised G->r transform), without the
 * G-space table; its gvec argument is then unused.  Replaced by the next
 * sdft_hip_set_projectors_real / sdft_hip_set_projectors call. */
int sdft_hip_set_projector_gradients(sdft_ctx *ctx, int nproj, const double *dx,
                                     const double *dy, const double *dz);

/* How real-space projectors are applied (no reference counterpart: a choice
 * of this library).  form 0 (default): in their per-atom form, as the
 * reference does -- each H application transforms the pair grids to real
 * space, forms Z_p = sum_r beta_p(r) psi(r) over each projector's grid points
 * (one [channels x points] . [points x pairs] product per atom), adds
 * sum_p energy_p vol_elem beta_p(r) Z_p to v_KS psi at every grid point and
 * transforms back; the cost grows with the projectors' grid points, not with
 * P x ncoef.  form 1: through the equivalent G-space table (built on the
 * device when first needed), applied by the MFMA products of
 * sdft_hip_set_projectors.  sdft_hip_nl_forces uses the table (built on its
 * first call) unless gradient tables are set; fragment batches need form 1. */
int sdft_hip_set_nl_form(sdft_ctx *ctx, int form);

/* Raw coefficient force f = -dE/dc* of H applied to nstate orbitals
 * (-4 Hc for G != 0, -2 Hc at G = 0): calcCoefForceWrapSCF,
 * stodft/reduced-energy/energy-wrapper-scf.c:1066-1197 (zero f, G-space KB,
 * pair-packed FFT v.psi, kinetic; coefForceCalcHybridSCF
 * cp-energy-ee-rho-stodft.c:38-355). */
int sdft_hip_hpsi(sdft_ctx *ctx, int nstate, const double *cre, const double *cim,
                  double *fre, double *fim, int prec);

/* Newton-Hermite filter: filterNewtonPolyHerm (stodft/filter.c:33-312) with
 * normHNewtonHerm (stodft/normh.c:35-176), PINY's default poly_type.  With
 * H^ = scale (H - emean) (scale = the Newton map's newtonInfo->scale):
 * P_0 = chi, P_n = (H^ - z_{n-1}) P_{n-1} for n < L (z = samp_point, L - 1
 * values), and for each of the nmu chemical potentials
 *   out[mu] = sum_{n<L} coeffs[n * nmu + mu] P_n
 * (expanCoeff[iPoly * numChemPot + imu]).  cre/cim hold chi on entry and
 * P_{L-1} on return; out_re/out_im are nmu consecutive blocks of
 * nstate * ncoef values (stoWfUpRe/Im[imu]). */
int sdft_hip_newton_filter(sdft_ctx *ctx, int nstate, int L, int nmu, const double *coeffs,
                           const double *samp_point, double scale, double emean, double *cre,
                           double *cim, double *out_re, double *out_im, int prec);

/* Spectral range: genEnergyMax (which = 0) / genEnergyMin (which = 1),
 * stodft/calc-spectral-range.c:33-333 / :339-566, on the device.  Power
 * iteration of H (which 0) or of (emax - H) (which 1) on one orbital that
 * starts from vre/vim (reference order, normalised first): per iteration
 * E = <phi|H|phi> (Gamma dot), then phi <- H phi (or emax phi - H phi),
 * normalised; it stops when |E_old - E| <= conv (the reference's energyConv
 * 1e-5, starting from E_old = 1e7, E = 1e5) or after max_iter iterations.
 * *energy = the last E (the reference sets energyMax = E + 0.1, energyMin =
 * E - 0.1); vre/vim receive the final vector (testWfMax / testWfMin);
 * *iters (nullable) the iteration count. */
int sdft_hip_spectral_bound(sdft_ctx *ctx, int which, double emax, double conv, int max_iter,
                            double *vre, double *vim, double *energy, int *iters, int prec);

/* Chebyshev expansion coefficients c_0..c_{L-1} of sqrt(f_FD) (fun_flag 2) or
 * f_FD (fun_flag 1) on [emin, emax]: calcChebyCoeff, stodft/coeff.c:1093-1223
 * (M = roundFFT(32 L) cosine grid, coeff.c:1556-1600). FP64 on the device. */
int sdft_hip_cheby_coeffs(sdft_ctx *ctx, int L, double emin, double emax, double mu,
                          double beta, int fun_flag, double *coeffs);

/* Chebyshev filter acc = sum_{k<L} c_k T_k(H^) chi over nstate orbitals with
 * H^ = scale (H - emean): filterChebyPolyHerm, stodft/filter.c:318-706 with
 * normHCheby stodft/normh.c:182-344.  cre/cim hold chi on entry and
 * T_{L-1} chi on return (the reference's side effect on cre_up/cim_up).
 * moments (nullable, L doubles) receives sum_s <chi_s|T_k chi_s> as the
 * storeChebyMomentsFlag branch (filter.c:458-510,572-594). */
int sdft_hip_cheby_filter(sdft_ctx *ctx, int nstate, int L, const double *coeffs,
                          double scale, double emean, double *cre, double *cim,
                          double *out_re, double *out_im, double *moments, int prec);

/* Chebyshev moments mu_0..mu_L (L+1 doubles, summed over the nstate
 * orbitals, not normalised): calcChebyMoments,
 * stodft/calc-chempot-chebyshev.c:596-913 (L/2 H applications, doubling
 * identities).  cre/cim are not modified. */
int sdft_hip_cheby_moments(sdft_ctx *ctx, int nstate, int L, double scale, double emean,
                           const double *cre, const double *cim, double *moments, int prec);

/* Chemical potential from the moments: calcChemPotCheby + calcNumElecCheby,
 * stodft/calc-chempot-chebyshev.c:30-315,918-972.  moments are divided by
 * nsto; N(mu) = occ sum_{k<L} c_k(mu) mu_k; bracket [mu_init -+ gap/2],
 * widen, one interpolation step, bisection until |N - ne| <= tol.  The
 * search runs in FP64 on the device. */
int sdft_hip_chem_pot(sdft_ctx *ctx, int L, double emin, double emax, double beta,
                      const double *moments, double nsto, double ne_true, double mu_init,
                      double gap, double occ, double tol, double *mu_out, double *ne_out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops such as the mu
 * search inside an SCF step, harnesses that forbid hipMalloc): the same
 * search as sdft_hip_chem_pot, same arguments, results and errors, but its
 * scratch (cosine table, h_j, moments, result: 2 roundFFT(32 L) + L + 4
 * doubles) lives in a fixed per-device array of 4.3 MB that the library
 * holds from load to process end, and the call allocates nothing.  L must
 * not exceed 8192 (SDFT_EINVAL beyond).  All calls on one device share that
 * array: calls that overlap on one device (several host threads, or several
 * contexts' streams) race and must be serialised by the caller. */
int sdft_hip_chem_pot_noalloc(sdft_ctx *ctx, int L, double emin, double emax, double beta,
                              const double *moments, double nsto, double ne_true,
                              double mu_init, double gap, double occ, double tol,
                              double *mu_out, double *ne_out);

/* Stochastic density rho(r) = occ_over_nsto * sum_s psi_s(r)^2 (psi from the
 * unnormalised G->r transform): rhoCalcRealStoHybrid stodft/density.c:43-274
 * and the scale of calcRhoStoHybridCheby :1478.  rho: n1*n2*n3 doubles in
 * the requested layout (SDFT_RHO_FFTW or SDFT_RHO_GENERIC); accumulate != 0
 * adds into rho instead of overwriting. */
int sdft_hip_density(sdft_ctx *ctx, int nstate, const double *re, const double *im,
                     double occ_over_nsto, double *rho, int layout, int accumulate);

/* Kinetic and nonlocal energies of nstate orbitals (calcEnergyChemPot,
 * stodft/calc-energy.c:133-191): eke = occ/2 sum 2|G|^2|c|^2 / nsto,
 * enl = occ sum_s sum_p E_p Z_p^2 / nsto.  FP64 reductions on the device. */
int sdft_hip_energies(sdft_ctx *ctx, int nstate, const double *re, const double *im,
                      double occ, double nsto, double *eke, double *enl);

/* Hellmann-Feynman nonlocal forces on the atoms (Ha/bohr) from nstate
 * orbitals: F_a = -(occ/nsto) sum_s sum_{p in a} 2 E_p Z_ps dZ_ps/dR_a with
 * dQ_p/dR_a = i G Q_p (calcEnergyForce's G-space NL branch,
 * stodft/calc-nuclei-force.c:302-489, scaled as :462-478).  gvec: [ncoef][3]
 * cartesian G (G=0 last); proj_atom: [nproj] atom of each projector;
 * forces: [natm][3].  The projections and their derivatives are FP64 MFMA
 * GEMMs on the device. */
int sdft_hip_nl_forces(sdft_ctx *ctx, int nstate, const double *re, const double *im,
                       const double *gvec, const int *proj_atom, int natm, double occ,
                       double nsto, double *forces);

/* Stochastic noise orbitals on the device: genNoiseOrbitalReal,
 * stodft/gen-noise.c:162-374 -- real-space values +-sqrt(N) (:186-195)
 * transformed to the coefficient set, c(G) = (1/N) sum_r chi(r) e^{-iG.r},
 * G=0 real.  The sign of grid point r (FFTW linear index) of orbital
 * first + s is the top bit of splitmix64(seed << 44 | (first + s) << 23 | r)
 * (a counter-based stream replacing the reference's gaussran2 seeding;
 * seed < 2^20, orbitals < 2^21, n^3 < 2^23).  Output in the reference order. */
int sdft_hip_noise(sdft_ctx *ctx, int nstate, long long seed, long long first, double *cre,
                   double *cim);

/* ---- device-resident blocks (orbitals stay in HBM across calls) ---------- */

/* Allocate device storage for nstate orbitals (T_k, T_{k-1}, accumulator,
 * noise copy) in the library's internal coefficient order. */
int sdft_hip_block_create(sdft_ctx *ctx, int nstate, int prec, sdft_block **out);
int sdft_hip_block_destroy(sdft_block *blk);
/* Host -> device copy of chi (reference order, FP64 re/im arrays). */
int sdft_hip_block_set(sdft_block *blk, const double *cre, const double *cim);
/* chi generated on the device (as sdft_hip_noise), no host transfer. */
int sdft_hip_block_noise(sdft_block *blk, long long seed, long long first);
/* Filter on resident data (no host transfers); moments nullable (L doubles). */
int sdft_hip_block_filter(sdft_block *blk, int L, const double *coeffs, double scale,
                          double emean, double *moments);
/* Device -> host copy: which = 0 filtered accumulator, 1 T_{L-1} (current). */
int sdft_hip_block_get(sdft_block *blk, int which, double *re, double *im);

/* ---- fragment batches --------------------------------------------------------
 * fragScf (stodft/fragment/frag-scf.c:37-349, per-fragment loop :211-248)
 * solves the fragments one after another, each a complete system with its own
 * grid (interface-frag/init-coord-hmat-fft.c, reinitFFT.c; FFTW3D forced at
 * parse-frag.c:211-213).  A batch takes nfrag resident blocks (one per
 * fragment, each on its own context, all of one precision, any mix of the
 * supported grid sizes) and runs the same Chebyshev filter on all of them
 * together: one launch per kernel and grid size per Chebyshev step, the steps
 * k >= 2 replayed from a captured HIP graph.  The blocks must outlive the
 * batch and are not to be used through other calls while a batch call runs. */
typedef struct sdft_batch sdft_batch;
int sdft_hip_batch_create(int nfrag, sdft_block *const *blocks, sdft_batch **out);
int sdft_hip_batch_destroy(sdft_batch *bat);
/* filterChebyPolyHerm (stodft/filter.c:318-706) on every fragment: block f's
 * noise copy chi is filtered with coeffs[f*L .. f*L+L) and its own scale /
 * emean (scale[f] = 2/(E_max - E_min), emean[f] = (E_min + E_max)/2); the
 * result is in each block's accumulator (sdft_hip_block_get, which = 0).
 * use_graph = 0 launches every step directly (same kernels and order); 1
 * replays steps 2 .. L-1 from a captured graph of two steps (captured once per
 * launch configuration); 2 captures the whole filter (chi copy, step counter,
 * all L-1 steps) as one graph, once per configuration and L, and runs it as
 * a single graph launch.  All three give identical bits.  L >= 2. */
int sdft_hip_batch_filter(sdft_batch *bat, int L, const double *coeffs, const double *scale,
                          const double *emean, int use_graph);
/* Autotuner: times `steps` Chebyshev steps (after one warm-up step) of each
 * candidate launch configuration -- the plane kernel's b-line tile (0: base
 * sub-chunk, 1: twice as wide, where it fits LDS), whether the KB products
 * overlap the FFT kernels on a second stream, the split of the KB
 * projection's reduction (1, 2, 4 or 8 slices) and, where every fragment
 * allows it, whether the contract kernel writes the next step's expansion
 * (fused expand) or an expand kernel runs per step -- on the batch's current
 * data and keeps the fastest.  config_out (nullable, 4 ints) receives split,
 * overlap, plane tile, fused expand; cand_ms (nullable) the per-candidate
 * step times in ms (at most max_cand; loop order tile, overlap, split, fused
 * expand, the last fastest).  Returns the number of candidates timed (32, or
 * 16 when the expand cannot be fused).  Leaves the blocks' T_k buffers
 * modified (the accumulators are untouched). */
int sdft_hip_batch_autotune(sdft_batch *bat, int steps, int *config_out, int max_cand,
                            double *cand_ms);
/* Set the launch configuration by hand (split in {1,2,4,8}, overlap 0/1,
 * plane tile 0/1, fused expand 0/1; 1 needs every fragment's sphere columns
 * to be ka intervals, else SDFT_EINVAL). */
int sdft_hip_batch_config(sdft_batch *bat, int split, int overlap, int plane_tile,
                          int fused_expand);

/* Launch-configuration tuner for one context (no reference counterpart):
 * times `steps` Chebyshev filter steps of an nstate-orbital block (device
 * noise, zero coefficients) in precision prec for each candidate --
 * plane-kernel variant x KB projection workgroup target (256, 1024) x fused
 * expand (0: an expand kernel per step, 1: the contract kernel writes the
 * next step's a-expansion; offered only when every sphere column is one ka
 * interval) -- and keeps the fastest for this context.  Plane variants: 0
 * the base b-line tile, 1 twice as wide (where it fits LDS); at n = 64 also
 * 2 radix-4 b-lines (4x4x4 instead of 8x8) on the base tile, 3-5 512-thread
 * workgroups with radix-4 c-lines (3: wide tile, radix-8 b-lines; 4: wide
 * tile, radix-4 b-lines; 5: base tile, radix-4 b-lines).  config_out
 * (nullable, 3 ints): variant, workgroup target, fused expand; cand_ms
 * (nullable): per-candidate ms per filter step of the block, loop order
 * variant, target, fused expand.  Returns the number of candidates (2 x 2 x
 * 2 per variant pair: 24 at n = 64, 8 elsewhere; half without the fused
 * form).  sdft_hip_set_config sets the three knobs by hand (variant in
 * [0, 6) at n = 64, [0, 2) elsewhere). */
int sdft_hip_tune(sdft_ctx *ctx, int nstate, int prec, int steps, int *config_out, int max_cand,
                  double *cand_ms);
int sdft_hip_set_config(sdft_ctx *ctx, int plane_tile, int kb_wg_target, int fused_expand);

/* Per-kernel timing of the last block_filter call when profiling is on:
 * names[i] / ms[i] / launches[i] for up to max entries; returns the count. */
int sdft_hip_set_profiling(sdft_ctx *ctx, int on);
int sdft_hip_kernel_times(sdft_ctx *ctx, int max, const char **names, double *ms,
                          int *launches);

/* ---- self-test -------------------------------------------------------------
 * Runs the library's LDS FFT on nlines interleaved complex lines of length n
 * (dir +1: e^{+2 pi i jk/n} as FFTW_BACKWARD, -1: forward), FP64 (prec 0) or
 * FP32 (prec 1) arithmetic; in/out are 2*n*nlines doubles.  No reference
 * counterpart: a diagnostic for the transform the filter path uses in place
 * of mathlib/fft_package.c's FFTW calls. */
int sdft_hip_selftest_fft(int n, int nlines, int dir, int prec, const double *in, double *out);

/* Copies a device buffer of nbytes (rounded down to doubles) reps times with
 * a plain streaming kernel (one read + one write per element) on the current
 * device: the known-size reference against which hardware byte counters of
 * the filter kernels are scaled.  No reference counterpart. */
int sdft_hip_selftest_copy(long long nbytes, int reps);

/* ---- Kohn-Sham potential from the density (closing the SCF loop) ------
 * sdft_hip_ks_init: the density plane-wave half sphere (PINY's large density
 * grid package cp_para_fft_pkg3d_lg: |G|^2 <= 4 x the wave-function cutoff,
 * G = 0 last) on the context's grid; desc->map / map_conj / ak2_kinetic are
 * its +G / -G FFTW indices and |G|^2, n1..n3 must equal the context's.
 * sdft_hip_ks_potential: cpGetVksStodft
 * (stodft/reduced-energy/cp-energy-ee-rho-stodft.c:644-900) after
 * calcRhoStoRecipFullg (stodft/density.c:699-870) for a periodic cell (iperd
 * = 3) without the dual grid: rho (n1 n2 n3 doubles in `layout`, as
 * sdft_hip_density returns it: sum rho / N = N_e) is transformed to the
 * sphere, v(G) = vext(G) + 4 pi rho(G) / (|G|^2 V) (vext: ncoef doubles re /
 * im in the sphere's order, plane-wave amplitudes; G = 0: vext(0)), v(G) ->
 * v(r), and the LDA potential `xc` (SDFT_XC_*) of rho / V is added.  The
 * result becomes the context's v_KS (as sdft_hip_set_vks, without a host
 * copy); energies (nullable, 4 doubles): E_Hartree, E_ext, E_xc and the
 * muxc sum as cpGetVksStodft / excpot_* accumulate them; vks_out (nullable):
 * v_KS in `layout`. */
int sdft_hip_ks_init(sdft_ctx *ctx, const sdft_grid_desc *dens);
/* Local pseudopotential Hellmann-Feynman forces from the density of the last
 * sdft_hip_ks_potential call: the local part of calcEnergyForce
 * (stodft/calc-nuclei-force.c:254 -> calcLocExtPostScf,
 * stodft/reduced-energy/energy-wrapper-post-scf.c:140-220 -> control_ewd_loc),
 * without the ion-ion Ewald term.  With v_ext(G) = sum_a f_type(a)(G)
 * e^{-iG.R_a}: F_a = -dE_ext/dR_a.  pos [natm][3] (bohr), atom_type [natm]
 * in [0, ntype), vloc [ntype][ncoef] the real form factors f(G) and gvec
 * [ncoef][3] the cartesian G, both in the density sphere's order; forces
 * [natm][3] (Ha/bohr). */
int sdft_hip_local_forces(sdft_ctx *ctx, int natm, const double *pos, const int *atom_type,
                          int ntype, const double *vloc, const double *gvec, double *forces);
int sdft_hip_ks_potential(sdft_ctx *ctx, const double *rho, int layout, const double *vext_re,
                          const double *vext_im, double volume, int xc, double *energies,
                          double *vks_out);

#ifdef __cplusplus
}
#endif
#endif /* SDFT_HIP_H */
