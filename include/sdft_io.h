/*
 * sdft_io.h -- the reference's on-disk text formats around the stochastic-DFT
 * SCF loop, written and read by host C (libsdft_host.so), so that densities
 * and orbitals produced on the device interoperate with PINY runs:
 *
 *   density checkpoint   stodft/checkpointIO.c:32-211 (checkpointOutput),
 *                        :218-505 (checkpointInput), :534-650 / :653-866
 *                        (the per-rank "Dist" forms: same layout, one file
 *                        per rank holding that rank's grid share)
 *   density log / final  stodft/density.c:1656-1745 (outputDensity),
 *                        stodft/init.c:862-868 (log truncated at start-up)
 *   fragment orbitals    stodft/fragment/frag-scf.c:254-265 (write),
 *                        :155-176 / :268-292 (read)
 *
 * All arrays are 0-based.  Grid arrays hold the reference's real-space
 * density order (rho_up[1..] of CPSCR_RHO: sdft_hip_density with
 * SDFT_RHO_GENERIC, a fastest); the files store whatever order the caller
 * passes.  Orbital arrays hold PINY's state-major coefficient order (G = 0
 * last), the order every sdft_hip entry point takes and returns.
 *
 * Numbers are printed with "%.16lg" (checkpoint, final density, orbitals) or
 * "%.10lg" (density log), as the reference prints them: 16 significant
 * digits, so a write/read round trip agrees to about 1e-16 relative, not bit
 * for bit.  Return values: 0 on success, SDFT_IO_E* on failure, with the
 * reason in sdft_io_last_error().
 */
#ifndef SDFT_IO_H
#define SDFT_IO_H

#ifdef __cplusplus
extern "C" {
#endif

#define SDFT_IO_EINVAL (-1)  /* bad arguments */
#define SDFT_IO_EOPEN (-10)  /* the file could not be opened */
#define SDFT_IO_EFORMAT (-11) /* short or malformed file, or inconsistent header */
#define SDFT_IO_EWRITE (-12) /* a write or close failed */

const char *sdft_io_last_error(void);

/* Number of DIIS density-bank entries a checkpoint holds for ndiis_now
 * stored steps and a bank of ndiis: ndiis_now + 1 while the bank is filling,
 * ndiis_now once it is full (checkpointIO.c:133-134). */
int sdft_io_checkpoint_nbank(int ndiis_now, int ndiis);

/* checkpointOutput / checkpointOutputDist (checkpointIO.c:32-211, :534-650):
 * writes (truncating) the text file
 *   "iscf ndiis_now"
 *   rho_up[i] * vol                 i < ngrid
 *   rho_dn[i] * vol                 (lsda only)
 *   bank_up[d][i]                   d < sdft_io_checkpoint_nbank(ndiis_now, ndiis)
 *   err_up[d][i]                    d < ndiis_now
 *   bank_dn[d][i], err_dn[d][i]     (lsda only, same counts)
 * one value per line.  vol is the cell volume the reference multiplies the
 * density by (getdeth(hmat_cp)); the banks are written as they are.
 * rho_dn / bank_dn / err_dn may be NULL when lsda == 0. */
int sdft_io_checkpoint_write(const char *path, int iscf, int ndiis_now, int ndiis,
                             long long ngrid, double vol, int lsda, const double *rho_up,
                             const double *rho_dn, const double *const *bank_up,
                             const double *const *err_up, const double *const *bank_dn,
                             const double *const *err_dn);

/* The header of a checkpoint file: *iscf, *ndiis_now (checkpointIO.c:300-302). */
int sdft_io_checkpoint_header(const char *path, int *iscf, int *ndiis_now);

/* checkpointInput (checkpointIO.c:218-505): reads the layout above back.
 * The density is returned as stored (the reference reads it without
 * dividing by the volume, :311-325).  The bank pointer arrays hold ndiis
 * entries of ngrid doubles each, as the reference allocates them
 * (:397-400); entries past the stored counts are left untouched.  The down-
 * spin blocks go to rho_dn / bank_dn / err_dn (the reference's single-rank
 * branch stores them in rhoDnBank and rhoUpErr, :466-498; this reader keeps
 * each block in its own array).  A file whose ndiis_now exceeds ndiis, or
 * that ends early, gives SDFT_IO_EFORMAT. */
int sdft_io_checkpoint_read(const char *path, int ndiis, long long ngrid, int lsda,
                            int *iscf, int *ndiis_now, double *rho_up, double *rho_dn,
                            double *const *bank_up, double *const *err_up,
                            double *const *bank_dn, double *const *err_dn);

/* outputDensity (density.c:1656-1745): appends rho[i] * log_scale ("%.10lg")
 * to the density log and writes (truncating) rho[i] ("%.16lg") to the final
 * density file; up spin then, with lsda, down spin.  log_scale is 1/vol on
 * the reference's gathered multi-rank branch (:1729) and 1 on its single-rank
 * branch (:1748).  Either path may be NULL to skip that file. */
int sdft_io_density_write(const char *log_path, const char *final_path, long long ngrid,
                          int lsda, const double *rho_up, const double *rho_dn,
                          double log_scale);

/* Truncates (creates) the density log, as stodft/init.c:862-868 does once
 * before the SCF loop. */
int sdft_io_density_log_reset(const char *log_path);

/* Reads a final-density file written above: ngrid values (2 ngrid with
 * lsda) into rho_up (and rho_dn). */
int sdft_io_density_read(const char *final_path, long long ngrid, int lsda, double *rho_up,
                         double *rho_dn);

/* frag-MO-<rank> (frag-scf.c:254-265): for each fragment f < nfrag, state
 * s < nstate[f] and coefficient j < ncoef[f], one line
 * "cre[f][s*ncoef[f]+j] cim[f][s*ncoef[f]+j]". */
int sdft_io_frag_mo_write(const char *path, int nfrag, const int *ncoef, const int *nstate,
                          const double *const *cre, const double *const *cim);

/* Reads the file above back into cre[f] / cim[f] (frag-scf.c:155-176). */
int sdft_io_frag_mo_read(const char *path, int nfrag, const int *ncoef, const int *nstate,
                         double *const *cre, double *const *cim);

#ifdef __cplusplus
}
#endif

#endif /* SDFT_IO_H */
