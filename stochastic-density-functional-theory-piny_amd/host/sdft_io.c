/*
 * sdft_io.c -- the reference's checkpoint, density and fragment-orbital text
 * files (layouts and citations in include/sdft_io.h).
 */
#include "sdft_io.h"

#include <errno.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

static char io_err[512];

const char *sdft_io_last_error(void) { return io_err; }

static int io_fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(io_err, sizeof io_err, fmt, ap);
  va_end(ap);
  return code;
}

int sdft_io_checkpoint_nbank(int ndiis_now, int ndiis) {
  return ndiis_now < ndiis ? ndiis_now + 1 : ndiis_now;
}

/* one value per line, each scaled by `scale` (1 writes the value as is:
 * x * 1.0 == x exactly) */
static int put_block(FILE *f, const char *fmt, const double *x, long long n, double scale) {
  for (long long i = 0; i < n; ++i)
    if (fprintf(f, fmt, x[i] * scale) < 0) return -1;
  return 0;
}

static int get_block(FILE *f, double *x, long long n) {
  for (long long i = 0; i < n; ++i)
    if (fscanf(f, "%lg", &x[i]) != 1) return -1;
  return 0;
}

static int close_written(FILE *f, const char *path, int rc) {
  if (fclose(f) != 0 && rc == 0) rc = io_fail(SDFT_IO_EWRITE, "closing %s failed", path);
  return rc;
}

int sdft_io_checkpoint_write(const char *path, int iscf, int ndiis_now, int ndiis,
                             long long ngrid, double vol, int lsda, const double *rho_up,
                             const double *rho_dn, const double *const *bank_up,
                             const double *const *err_up, const double *const *bank_dn,
                             const double *const *err_dn) {
  if (!path || ngrid < 1 || ndiis < 0 || ndiis_now < 0 || ndiis_now > ndiis || !rho_up ||
      (lsda && !rho_dn))
    return io_fail(SDFT_IO_EINVAL, "bad checkpoint arguments");
  const int nbank = sdft_io_checkpoint_nbank(ndiis_now, ndiis);
  for (int d = 0; d < nbank; ++d)
    if (!bank_up || !bank_up[d] || (lsda && (!bank_dn || !bank_dn[d])))
      return io_fail(SDFT_IO_EINVAL, "checkpoint density bank %d missing", d);
  for (int d = 0; d < ndiis_now; ++d)
    if (!err_up || !err_up[d] || (lsda && (!err_dn || !err_dn[d])))
      return io_fail(SDFT_IO_EINVAL, "checkpoint error bank %d missing", d);
  FILE *f = fopen(path, "w");
  if (!f) return io_fail(SDFT_IO_EOPEN, "cannot open %s for writing: %s", path, strerror(errno));
  const char *fmt = "%.16lg\n";
  int bad = fprintf(f, "%i %i\n", iscf, ndiis_now) < 0;
  bad = bad || put_block(f, fmt, rho_up, ngrid, vol);
  if (lsda) bad = bad || put_block(f, fmt, rho_dn, ngrid, vol);
  for (int d = 0; d < nbank; ++d) bad = bad || put_block(f, fmt, bank_up[d], ngrid, 1.0);
  for (int d = 0; d < ndiis_now; ++d) bad = bad || put_block(f, fmt, err_up[d], ngrid, 1.0);
  if (lsda) {
    for (int d = 0; d < nbank; ++d) bad = bad || put_block(f, fmt, bank_dn[d], ngrid, 1.0);
    for (int d = 0; d < ndiis_now; ++d) bad = bad || put_block(f, fmt, err_dn[d], ngrid, 1.0);
  }
  return close_written(f, path, bad ? io_fail(SDFT_IO_EWRITE, "writing %s failed", path) : 0);
}

static int read_header(FILE *f, const char *path, int *iscf, int *ndiis_now) {
  if (fscanf(f, "%i %i", iscf, ndiis_now) != 2 || *ndiis_now < 0)
    return io_fail(SDFT_IO_EFORMAT, "%s: no \"iScf numDiisNow\" header", path);
  return 0;
}

int sdft_io_checkpoint_header(const char *path, int *iscf, int *ndiis_now) {
  if (!path || !iscf || !ndiis_now) return io_fail(SDFT_IO_EINVAL, "bad checkpoint arguments");
  FILE *f = fopen(path, "r");
  if (!f) return io_fail(SDFT_IO_EOPEN, "cannot open %s: %s", path, strerror(errno));
  const int rc = read_header(f, path, iscf, ndiis_now);
  fclose(f);
  return rc;
}

int sdft_io_checkpoint_read(const char *path, int ndiis, long long ngrid, int lsda,
                            int *iscf, int *ndiis_now, double *rho_up, double *rho_dn,
                            double *const *bank_up, double *const *err_up,
                            double *const *bank_dn, double *const *err_dn) {
  if (!path || ngrid < 1 || ndiis < 0 || !iscf || !ndiis_now || !rho_up || (lsda && !rho_dn))
    return io_fail(SDFT_IO_EINVAL, "bad checkpoint arguments");
  FILE *f = fopen(path, "r");
  if (!f) return io_fail(SDFT_IO_EOPEN, "cannot open %s: %s", path, strerror(errno));
  int it = 0, nnow = 0;
  int rc = read_header(f, path, &it, &nnow);
  if (!rc && nnow > ndiis)
    rc = io_fail(SDFT_IO_EFORMAT, "%s: %d stored DIIS steps exceed the bank size %d", path, nnow,
                 ndiis);
  const int nbank = sdft_io_checkpoint_nbank(nnow, ndiis);
  if (!rc)
    for (int d = 0; d < nbank; ++d)
      if (!bank_up || !bank_up[d] || (lsda && (!bank_dn || !bank_dn[d])))
        rc = io_fail(SDFT_IO_EINVAL, "checkpoint density bank %d missing", d);
  if (!rc)
    for (int d = 0; d < nnow; ++d)
      if (!err_up || !err_up[d] || (lsda && (!err_dn || !err_dn[d])))
        rc = io_fail(SDFT_IO_EINVAL, "checkpoint error bank %d missing", d);
  int bad = 0;
  if (!rc) {
    bad = get_block(f, rho_up, ngrid);
    if (lsda) bad = bad || get_block(f, rho_dn, ngrid);
    for (int d = 0; d < nbank; ++d) bad = bad || get_block(f, bank_up[d], ngrid);
    for (int d = 0; d < nnow; ++d) bad = bad || get_block(f, err_up[d], ngrid);
    if (lsda) {
      for (int d = 0; d < nbank; ++d) bad = bad || get_block(f, bank_dn[d], ngrid);
      for (int d = 0; d < nnow; ++d) bad = bad || get_block(f, err_dn[d], ngrid);
    }
    if (bad) rc = io_fail(SDFT_IO_EFORMAT, "%s ends before the stored blocks", path);
  }
  fclose(f);
  if (!rc) {
    *iscf = it;
    *ndiis_now = nnow;
  }
  return rc;
}

int sdft_io_density_log_reset(const char *log_path) {
  if (!log_path) return io_fail(SDFT_IO_EINVAL, "bad density-log path");
  FILE *f = fopen(log_path, "w");
  if (!f) return io_fail(SDFT_IO_EOPEN, "cannot open %s: %s", log_path, strerror(errno));
  return close_written(f, log_path, 0);
}

int sdft_io_density_write(const char *log_path, const char *final_path, long long ngrid,
                          int lsda, const double *rho_up, const double *rho_dn,
                          double log_scale) {
  if (ngrid < 1 || !rho_up || (lsda && !rho_dn))
    return io_fail(SDFT_IO_EINVAL, "bad density-output arguments");
  if (log_path) {
    FILE *f = fopen(log_path, "a");
    if (!f) return io_fail(SDFT_IO_EOPEN, "cannot open %s: %s", log_path, strerror(errno));
    int bad = put_block(f, "%.10lg\n", rho_up, ngrid, log_scale);
    if (lsda) bad = bad || put_block(f, "%.10lg\n", rho_dn, ngrid, log_scale);
    const int rc =
        close_written(f, log_path, bad ? io_fail(SDFT_IO_EWRITE, "writing %s failed", log_path) : 0);
    if (rc) return rc;
  }
  if (final_path) {
    FILE *f = fopen(final_path, "w");
    if (!f) return io_fail(SDFT_IO_EOPEN, "cannot open %s: %s", final_path, strerror(errno));
    int bad = put_block(f, "%.16lg\n", rho_up, ngrid, 1.0);
    if (lsda) bad = bad || put_block(f, "%.16lg\n", rho_dn, ngrid, 1.0);
    return close_written(f, final_path,
                         bad ? io_fail(SDFT_IO_EWRITE, "writing %s failed", final_path) : 0);
  }
  return 0;
}

int sdft_io_density_read(const char *final_path, long long ngrid, int lsda, double *rho_up,
                         double *rho_dn) {
  if (!final_path || ngrid < 1 || !rho_up || (lsda && !rho_dn))
    return io_fail(SDFT_IO_EINVAL, "bad density-read arguments");
  FILE *f = fopen(final_path, "r");
  if (!f) return io_fail(SDFT_IO_EOPEN, "cannot open %s: %s", final_path, strerror(errno));
  int bad = get_block(f, rho_up, ngrid);
  if (lsda) bad = bad || get_block(f, rho_dn, ngrid);
  fclose(f);
  return bad ? io_fail(SDFT_IO_EFORMAT, "%s holds fewer than the expected values", final_path)
             : 0;
}

static int frag_args_ok(int nfrag, const int *ncoef, const int *nstate, const void *cre,
                        const void *cim) {
  if (nfrag < 0 || (nfrag > 0 && (!ncoef || !nstate || !cre || !cim))) return 0;
  for (int i = 0; i < nfrag; ++i)
    if (ncoef[i] < 1 || nstate[i] < 0) return 0;
  return 1;
}

int sdft_io_frag_mo_write(const char *path, int nfrag, const int *ncoef, const int *nstate,
                          const double *const *cre, const double *const *cim) {
  if (!path || !frag_args_ok(nfrag, ncoef, nstate, cre, cim))
    return io_fail(SDFT_IO_EINVAL, "bad fragment-orbital arguments");
  for (int i = 0; i < nfrag; ++i)
    if (nstate[i] > 0 && (!cre[i] || !cim[i]))
      return io_fail(SDFT_IO_EINVAL, "fragment %d has no coefficient arrays", i);
  FILE *f = fopen(path, "w");
  if (!f) return io_fail(SDFT_IO_EOPEN, "cannot open %s for writing: %s", path, strerror(errno));
  int bad = 0;
  for (int i = 0; i < nfrag && !bad; ++i) {
    const long long n = (long long)nstate[i] * ncoef[i];
    for (long long j = 0; j < n && !bad; ++j)
      bad = fprintf(f, "%.16lg %.16lg\n", cre[i][j], cim[i][j]) < 0;
  }
  return close_written(f, path, bad ? io_fail(SDFT_IO_EWRITE, "writing %s failed", path) : 0);
}

int sdft_io_frag_mo_read(const char *path, int nfrag, const int *ncoef, const int *nstate,
                         double *const *cre, double *const *cim) {
  if (!path || !frag_args_ok(nfrag, ncoef, nstate, cre, cim))
    return io_fail(SDFT_IO_EINVAL, "bad fragment-orbital arguments");
  for (int i = 0; i < nfrag; ++i)
    if (nstate[i] > 0 && (!cre[i] || !cim[i]))
      return io_fail(SDFT_IO_EINVAL, "fragment %d has no coefficient arrays", i);
  FILE *f = fopen(path, "r");
  if (!f) return io_fail(SDFT_IO_EOPEN, "cannot open %s: %s", path, strerror(errno));
  int bad = 0;
  for (int i = 0; i < nfrag && !bad; ++i) {
    const long long n = (long long)nstate[i] * ncoef[i];
    for (long long j = 0; j < n && !bad; ++j)
      bad = fscanf(f, "%lg %lg", &cre[i][j], &cim[i][j]) != 2;
  }
  fclose(f);
  return bad ? io_fail(SDFT_IO_EFORMAT, "%s holds fewer coefficients than the fragments need",
                       path)
             : 0;
}
