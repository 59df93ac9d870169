// sdft_cheby.hip -- FP64 scalar side of the Chebyshev path, on the device.
//
//  * expansion coefficients of f_FD or sqrt(f_FD) (calcChebyCoeff,
//    stodft/coeff.c:1093-1223): c_k = (2/M) sum_j g(x_j) cos(2 pi jk/M),
//    c_0 = (1/M) sum_j g(x_j), x_j = (dE/2) cos(2 pi j/M) + E_mean,
//    M = roundFFT(32 L) (coeff.c:1556-1600).  The reference takes the real
//    part of an M-point complex DFT; here the same cosine sums are evaluated
//    directly, one workgroup per k.
//  * electron count and chemical potential (calcNumElecCheby /
//    calcChemPotCheby, stodft/calc-chempot-chebyshev.c:30-315,918-972):
//    N(mu) = occ sum_{k<L} c_k(mu) m_k = (occ/M) sum_j f(x_j; mu) h_j with
//    h_j = m_0 + 2 sum_{k=1}^{L-1} m_k cos(2 pi jk/M) computed once, so each
//    N(mu) is one M-point reduction; the bracket / interpolation / bisection
//    sequence of the reference runs in a single workgroup.
//  * per-step moment partial reduction and the kinetic / nonlocal energy
//    sums (stodft/calc-energy.c:133-191).
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sdft_hip.h"
#include "sdft_internal.h"

namespace sdft {

int round_fft(int n_in);

__device__ inline double fermi_dev(double x, double mu, double beta) {
  return 1.0 / (1.0 + exp(beta * (x - mu)));
}

__global__ void k_costab(double* __restrict__ tab, int M) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < M) tab[j] = cospi(2.0 * (double)j / (double)M);
}

__global__ void k_cheby_fvals(const double* __restrict__ costab, int M, double dE, double em,
                              double mu, double beta, int fun_flag, double* __restrict__ g) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const double x = dE * costab[j] + em;
  double f = fermi_dev(x, mu, beta);
  if (fun_flag == 2) f = sqrt(f);
  g[j] = f;
}

__device__ inline double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double s = 0.0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; ++w) s += sh[w];
  return s;
}

// one workgroup per k
__global__ void k_cheby_project(const double* __restrict__ g, const double* __restrict__ costab,
                                int M, double* __restrict__ c) {
  __shared__ double sh[16];
  const int k = blockIdx.x;
  double s = 0.0;
  for (int j = threadIdx.x; j < M; j += blockDim.x) {
    const long long idx = ((long long)j * k) % M;
    s += g[j] * costab[idx];
  }
  s = block_sum(s, sh);
  if (threadIdx.x == 0) c[k] = (k == 0 ? 1.0 : 2.0) * s / (double)M;
}

// h_j = (m_0 + 2 sum_{k=1}^{L-1} m_k cos(2 pi jk/M)) * inv_nsto
__global__ void k_cheby_h(const double* __restrict__ m, int L, const double* __restrict__ costab,
                          int M, double inv_nsto, double* __restrict__ h) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  double s = 0.0;
  int idx = 0;
  for (int k = 1; k < L; ++k) {
    idx += j;
    if (idx >= M) idx -= M;
    s += m[k] * costab[idx];
  }
  h[j] = (m[0] + 2.0 * s) * inv_nsto;
}

__device__ double num_elec_dev(const double* costab, const double* h, int M, double dE,
                               double em, double mu, double beta, double occ, double* sh) {
  double s = 0.0;
  for (int j = threadIdx.x; j < M; j += blockDim.x) {
    const double x = dE * costab[j] + em;
    s += fermi_dev(x, mu, beta) * h[j];
  }
  s = block_sum(s, sh);
  return occ * s / (double)M;
}

// Bracket + interpolation + bisection of calcChemPotCheby, one workgroup.
__global__ void k_chem_pot(const double* __restrict__ costab, const double* __restrict__ h, int M,
                           double dE, double em, double beta, double occ, double ne_true,
                           double mu_init, double gap, double tol, int max_iter,
                           double* __restrict__ out) {
  __shared__ double sh[16];
  double lo = mu_init - 0.5 * gap, hi = mu_init + 0.5 * gap;
  double nlo = num_elec_dev(costab, h, M, dE, em, lo, beta, occ, sh);
  double nhi = num_elec_dev(costab, h, M, dE, em, hi, beta, occ, sh);
  int it = 0;
  while (nhi < ne_true && it < max_iter) {
    lo = hi;
    hi += gap;
    nlo = num_elec_dev(costab, h, M, dE, em, lo, beta, occ, sh);
    nhi = num_elec_dev(costab, h, M, dE, em, hi, beta, occ, sh);
    ++it;
  }
  while (nlo > ne_true && it < max_iter) {
    hi = lo;
    lo -= gap;
    nlo = num_elec_dev(costab, h, M, dE, em, lo, beta, occ, sh);
    nhi = num_elec_dev(costab, h, M, dE, em, hi, beta, occ, sh);
    ++it;
  }
  double mu = (ne_true - nlo) * (hi - lo) / (nhi - nlo) + lo;
  double nn = num_elec_dev(costab, h, M, dE, em, mu, beta, occ, sh);
  int ib = 0;
  while (fabs(nn - ne_true) > tol && ib < max_iter) {
    if (nn > ne_true) {
      hi = mu;
      nhi = nn;
    }
    if (nn < ne_true) {
      lo = mu;
      nlo = nn;
    }
    mu = 0.5 * (lo + hi);
    nn = num_elec_dev(costab, h, M, dE, em, mu, beta, occ, sh);
    ++ib;
  }
  if (threadIdx.x == 0) {
    out[0] = mu;
    out[1] = nn;
    out[2] = (double)(it + ib);
  }
}

// out[step*3 + i] = sum_w part[(step*nwg + w)*3 + i]
__global__ void k_mom_reduce(const double* __restrict__ part, int nwg, double* __restrict__ out) {
  __shared__ double sh[16];
  const int step = blockIdx.x;
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int w = threadIdx.x; w < nwg; w += blockDim.x) s += part[((size_t)step * nwg + w) * 3 + i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[step * 3 + i] = s;
  }
}

// per state: sum_{j < ncoef-1} 2 ak2 (re^2 + im^2)
__global__ void k_eke(const double* __restrict__ X, int ldx, int ncp, int ncoef,
                      const double* __restrict__ ak2, double* __restrict__ out) {
  __shared__ double sh[16];
  const int s = blockIdx.x;
  const double* x = X + (size_t)s * ldx;
  double acc = 0.0;
  for (int j = threadIdx.x; j < ncoef - 1; j += blockDim.x)
    acc += 2.0 * ak2[j] * (x[j] * x[j] + x[ncp + j] * x[ncp + j]);
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) out[s] = acc;
}

// sum_{p, s < nstate} E_p Z[p][s]^2
__global__ void k_enl(const double* __restrict__ Z, const double* __restrict__ E, int P, int ns,
                      int nstate, double* __restrict__ out) {
  __shared__ double sh[16];
  double acc = 0.0;
  for (int i = threadIdx.x; i < P * ns; i += blockDim.x) {
    const int p = i / ns, s = i - p * ns;
    if (s < nstate) acc += E[p] * Z[i] * Z[i];
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) out[0] = acc;
}

// One power-iteration update of genEnergyMax / genEnergyMin
// (stodft/calc-spectral-range.c:180-221 / :380-420) on orbital 0 of X with
// raw force F: h = -f/4 (G != 0), -f/2 (G = 0); E = 2 sum_{G!=0} (h.re v.re +
// h.im v.im) + h.re(0) v.re(0); w = h (mode 0) or emax v - h (mode 1);
// v = w / sqrt(2 sum_{G!=0} |w|^2 + w.re(0)^2).  out[0] = E.  One workgroup.
__global__ void k_power_update(double* __restrict__ X, const double* __restrict__ F, int ncp,
                               int ncoef, int mode, double emax, double* __restrict__ out) {
  __shared__ double sh[16];
  __shared__ double len_sh;
  double e = 0.0, l2 = 0.0;
  const int g0 = ncoef - 1;
  for (int j = threadIdx.x; j < ncoef; j += blockDim.x) {
    const double s = (j == g0) ? -0.5 : -0.25;
    const double hr = s * F[j], hi = s * F[ncp + j];
    const double vr = X[j], vi = X[ncp + j];
    const double wr = mode ? emax * vr - hr : hr;
    const double wi = mode ? emax * vi - hi : hi;
    if (j == g0) {
      e += hr * vr;
      l2 += wr * wr;
    } else {
      e += 2.0 * (hr * vr + hi * vi);
      l2 += 2.0 * (wr * wr + wi * wi);
    }
    X[j] = wr;
    X[ncp + j] = wi;
  }
  e = block_sum(e, sh);
  l2 = block_sum(l2, sh);
  if (threadIdx.x == 0) {
    out[0] = e;
    len_sh = sqrt(l2);
  }
  __syncthreads();
  const double inv = 1.0 / len_sh;
  for (int j = threadIdx.x; j < ncoef; j += blockDim.x) {
    X[j] *= inv;
    X[ncp + j] *= inv;
  }
}

// v /= sqrt(<v|v>) (Gamma norm) for orbital 0 of X
__global__ void k_normalize0(double* __restrict__ X, int ncp, int ncoef) {
  __shared__ double sh[16];
  __shared__ double len_sh;
  double l2 = 0.0;
  for (int j = threadIdx.x; j < ncoef; j += blockDim.x)
    l2 += (j == ncoef - 1) ? X[j] * X[j] : 2.0 * (X[j] * X[j] + X[ncp + j] * X[ncp + j]);
  l2 = block_sum(l2, sh);
  if (threadIdx.x == 0) len_sh = sqrt(l2);
  __syncthreads();
  const double inv = 1.0 / len_sh;
  for (int j = threadIdx.x; j < ncoef; j += blockDim.x) {
    X[j] *= inv;
    X[ncp + j] *= inv;
  }
}

// ==========================================================================

static int ensure_red(Ctx& c, size_t need) {
  if (c.red_bytes >= need) return 0;
  if (c.red) SDFT_HIP_CHECK(hipFree(c.red));
  c.red = nullptr;
  c.red_bytes = 0;
  if (hipMalloc(&c.red, need) != hipSuccess) {
    set_error("device allocation failed (reduction scratch)");
    return -3;
  }
  c.red_bytes = need;
  return 0;
}

int cheby_coeffs_device(Ctx& c, int L, double emin, double emax, double mu, double beta,
                        int fun_flag, double* coeffs_host) {
  const int M = round_fft(32 * L);
  int rc = ensure_red(c, sizeof(double) * ((size_t)2 * M + L));
  if (rc) return rc;
  double* costab = c.red;
  double* g = costab + M;
  double* cdev = g + M;
  const double dE = (emax - emin) * 0.5, em = 0.5 * (emax + emin);
  const int nb = (M + 255) / 256;
  hipLaunchKernelGGL(k_costab, dim3(nb), dim3(256), 0, c.stream, costab, M);
  hipLaunchKernelGGL(k_cheby_fvals, dim3(nb), dim3(256), 0, c.stream, costab, M, dE, em, mu, beta,
                     fun_flag, g);
  hipLaunchKernelGGL(k_cheby_project, dim3(L), dim3(256), 0, c.stream, g, costab, M, cdev);
  SDFT_HIP_CHECK(hipGetLastError());
  SDFT_HIP_CHECK(hipMemcpyAsync(coeffs_host, cdev, sizeof(double) * L, hipMemcpyDeviceToHost,
                                c.stream));
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

// The chemical-potential search on a device buffer `buf` of at least
// chem_pot_scratch_doubles(L) doubles: cosine table and h_j (M each), the raw
// moments (L), the result (3).  The moments are copied in unnormalised; the
// 1/nsto factor is folded into h_j.
static size_t chem_pot_scratch_doubles(int L) {
  return (size_t)2 * round_fft(32 * L) + L + 4;
}

static int chem_pot_on(Ctx& c, double* buf, int L, double emin, double emax, double beta,
                       const double* moments, double nsto, double ne_true, double mu_init,
                       double gap, double occ, double tol, double* mu_out, double* ne_out) {
  const int M = round_fft(32 * L);
  double* costab = buf;
  double* h = costab + M;
  double* m = h + M;
  double* out = m + L;
  SDFT_HIP_CHECK(hipMemcpyAsync(m, moments, sizeof(double) * L, hipMemcpyHostToDevice, c.stream));
  const int nb = (M + 255) / 256;
  hipLaunchKernelGGL(k_costab, dim3(nb), dim3(256), 0, c.stream, costab, M);
  hipLaunchKernelGGL(k_cheby_h, dim3(nb), dim3(256), 0, c.stream, m, L, costab, M, 1.0 / nsto, h);
  const double dE = (emax - emin) * 0.5, em = 0.5 * (emax + emin);
  hipLaunchKernelGGL(k_chem_pot, dim3(1), dim3(1024), 0, c.stream, costab, h, M, dE, em, beta, occ,
                     ne_true, mu_init, gap, tol, 400, out);
  SDFT_HIP_CHECK(hipGetLastError());
  double res[3];
  SDFT_HIP_CHECK(hipMemcpyAsync(res, out, sizeof(res), hipMemcpyDeviceToHost, c.stream));
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  *mu_out = res[0];
  *ne_out = res[1];
  return 0;
}

int chem_pot_device(Ctx& c, int L, double emin, double emax, double beta, const double* moments,
                    double nsto, double ne_true, double mu_init, double gap, double occ,
                    double tol, double* mu_out, double* ne_out) {
  int rc = ensure_red(c, sizeof(double) * chem_pot_scratch_doubles(L));
  if (rc) return rc;
  return chem_pot_on(c, c.red, L, emin, emax, beta, moments, nsto, ne_true, mu_init, gap, occ,
                     tol, mu_out, ne_out);
}

// ---- chemical-potential search without device allocation ----------------
// A fixed per-device array holds the search's scratch for every L up to
// kNoAllocMaxL: the largest roundFFT(32 L) there is 279936 (2^7 3^7), so
// 2 * 279936 + 8192 + 4 doubles (4.3 MB).  Each device that loads the
// library holds its own copy until the process ends.  The host never uses
// the array's name as an address: at context creation the device address
// is read once per device with hipGetSymbolAddress and kept in
// g_scratch_dev[device].  Every no-allocation call on a device shares that
// device's array, so such calls that overlap on one device (two host
// threads, or two contexts' streams) race; nothing in it has to start at
// zero (every element a search reads is written first by the same call).
constexpr int kNoAllocMaxL = 8192;
constexpr size_t kNoAllocDoubles = (size_t)2 * 279936 + kNoAllocMaxL + 4;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kNoAllocDoubles];

static std::mutex g_scratch_mu;
static double* g_scratch_dev[kMaxDevices] = {nullptr};

int noalloc_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("device ordinal out of range for the no-allocation scratch table");
    return SDFT_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return 0;
  void* p = nullptr;
  SDFT_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_dev[device] = static_cast<double*>(p);
  return 0;
}

int chem_pot_device_noalloc(Ctx& c, int L, double emin, double emax, double beta,
                            const double* moments, double nsto, double ne_true, double mu_init,
                            double gap, double occ, double tol, double* mu_out, double* ne_out) {
  if (L > kNoAllocMaxL || chem_pot_scratch_doubles(L) > kNoAllocDoubles) {
    set_error("sdft_hip_chem_pot_noalloc: L = " + std::to_string(L) +
              " exceeds the fixed scratch (L <= " + std::to_string(kNoAllocMaxL) + ")");
    return SDFT_EINVAL;
  }
  double* buf = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (c.device >= 0 && c.device < kMaxDevices) buf = g_scratch_dev[c.device];
  }
  if (!buf) {
    set_error("sdft_hip_chem_pot_noalloc: no scratch address for this device");
    return SDFT_ESTATE;
  }
  return chem_pot_on(c, buf, L, emin, emax, beta, moments, nsto, ne_true, mu_init, gap, occ, tol,
                     mu_out, ne_out);
}

int moments_reduce(Ctx& c, const double* mom_part, int nsteps, int nwg, double* out3) {
  int rc = ensure_red(c, sizeof(double) * (size_t)3 * nsteps);
  if (rc) return rc;
  hipLaunchKernelGGL(k_mom_reduce, dim3(nsteps), dim3(256), 0, c.stream, mom_part, nwg, c.red);
  SDFT_HIP_CHECK(hipGetLastError());
  SDFT_HIP_CHECK(hipMemcpyAsync(out3, c.red, sizeof(double) * 3 * nsteps, hipMemcpyDeviceToHost,
                                c.stream));
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

int energy_ke_device(Ctx& c, const double* X, int nstate, double* out) {
  int rc = ensure_red(c, sizeof(double) * (size_t)nstate);
  if (rc) return rc;
  hipLaunchKernelGGL(k_eke, dim3(nstate), dim3(256), 0, c.stream, X, 2 * c.g.ncp, c.g.ncp,
                     c.g.ncoef, c.t.ak2, c.red);
  SDFT_HIP_CHECK(hipGetLastError());
  std::vector<double> h(nstate);
  SDFT_HIP_CHECK(hipMemcpyAsync(h.data(), c.red, sizeof(double) * nstate, hipMemcpyDeviceToHost,
                                c.stream));
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  double s = 0.0;
  for (int i = 0; i < nstate; ++i) s += h[i];
  *out = s;
  return 0;
}

int normalize0_device(Ctx& c, double* X) {
  hipLaunchKernelGGL(k_normalize0, dim3(1), dim3(1024), 0, c.stream, X, c.g.ncp, c.g.ncoef);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

int power_update_device(Ctx& c, double* X, const double* F, int mode, double emax,
                        double* energy_host) {
  int rc = ensure_red(c, sizeof(double));
  if (rc) return rc;
  hipLaunchKernelGGL(k_power_update, dim3(1), dim3(1024), 0, c.stream, X, F, c.g.ncp, c.g.ncoef,
                     mode, emax, c.red);
  SDFT_HIP_CHECK(hipGetLastError());
  SDFT_HIP_CHECK(hipMemcpyAsync(energy_host, c.red, sizeof(double), hipMemcpyDeviceToHost,
                                c.stream));
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

int energy_nl_device(Ctx& c, int ns, int nstate, double* out) {
  int rc = ensure_red(c, sizeof(double));
  if (rc) return rc;
  hipLaunchKernelGGL(k_enl, dim3(1), dim3(1024), 0, c.stream, c.Z, c.E, c.nproj, ns, nstate, c.red);
  SDFT_HIP_CHECK(hipGetLastError());
  SDFT_HIP_CHECK(hipMemcpyAsync(out, c.red, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

}  // namespace sdft
