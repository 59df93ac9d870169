// sdft_ks.hip -- the Kohn-Sham potential from the stochastic density on the
// device: the steps of calcRhoStoRecipFullg (stodft/density.c:699-870) and
// cpGetVksStodft (stodft/reduced-energy/cp-energy-ee-rho-stodft.c:644-900)
// for a periodic cell without the dual grid, LDA exchange-correlation
// (energy/cp/xc_functionals.c).  The G-space transforms reuse the sphere
// kernels on the density plane-wave sphere; these kernels are the pointwise
// parts: the density as a complex field, Hartree + external in G space, and
// the LDA potential in real space.
#include "../../include/sdft_hip.h"
#include "sdft_internal.h"

namespace sdft {

// field[o] = (rho[i], 0) and rho_fftw[o] = rho[i] / vol, with o the FFTW
// (c fastest) index of input point i (input in FFTW or generic a-fastest
// order)
__global__ void k_ks_field(const double* __restrict__ rho, int layout, int n, long long N,
                           double rvol, double2* __restrict__ field,
                           double* __restrict__ rho_fftw) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  long long o = i;
  if (layout == 1) {  // generic: i = (ic n + ib) n + ia
    const long long ic = i / ((long long)n * n);
    const long long rem = i - ic * (long long)n * n;
    const long long ib = rem / n, ia = rem - ib * n;
    o = (ia * n + ib) * n + ic;
  }
  const double r = rho[i];
  field[o] = make_double2(r, 0.0);
  rho_fftw[o] = r * rvol;
}

// Hartree + external on the density sphere (internal coefficient order,
// ldx = 2 ncp, G = 0 at j = ncoef - 1): V[j] = vext[j] + 4 pi Q[j] /
// (|G|^2 vol) for G != 0, vext at G = 0 (cp-energy-ee-rho-stodft.c:763-790,
// iperd = 3, no erf screening); per block the partial sums of the half-sphere
// Hartree energy 4 pi |Q|^2 / (|G|^2 vol) and of the external energy
// (2 Re(vext conj Q) for G != 0, vext(0) Q(0) at G = 0).
__global__ __launch_bounds__(256) void k_ks_hartree(const double* __restrict__ Q,
                                                    double* __restrict__ V,
                                                    const double* __restrict__ ak2, int ncoef,
                                                    int ncp, double vol,
                                                    double* __restrict__ part) {
  __shared__ double sh[2][256];
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  double eh = 0.0, ee = 0.0;
  if (j < ncoef) {
    const double qr = Q[j], qi = Q[ncp + j];
    const double vr = V[j], vi = V[ncp + j];
    if (j < ncoef - 1) {
      const double gh = 4.0 * M_PI / (ak2[j] * vol);
      eh = gh * (qr * qr + qi * qi);
      ee = 2.0 * (vr * qr + vi * qi);
      V[j] = vr + gh * qr;
      V[ncp + j] = vi + gh * qi;
    } else {
      ee = vr * qr;
      V[ncp + j] = 0.0;
    }
  }
  sh[0][threadIdx.x] = eh;
  sh[1][threadIdx.x] = ee;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0][0];
    part[2 * blockIdx.x + 1] = sh[1][0];
  }
}

// LDA exchange-correlation of one point (physical density rho): v_xc, rho
// eps_xc and the muxc term, as excpot_pz_lda (xc_functionals.c:268-362),
// excpot_pw_lda (:368-480) and excpot_pade_lda (:32-101).
__device__ inline void lda_point(double rho, int kind, double* v, double* e, double* m) {
  if (kind == SDFT_XC_PADE_LDA) {
    const double a0 = 0.4581652932831429, a1 = 2.217058676663745, a2 = 0.7405551735357053,
                 a3 = 0.01968227878617998;
    const double b1 = 1.0, b2 = 4.504130959426697, b3 = 1.110667363742916,
                 b4 = 0.02359291751427506;
    const double rs = cbrt(3.0 / (4.0 * M_PI * rho));
    const double num = ((a3 * rs + a2) * rs + a1) * rs + a0;
    const double den = (((b4 * rs + b3) * rs + b2) * rs + b1) * rs;
    const double drs = -rs / (3.0 * rho);
    const double num_p = (3.0 * a3 * rs + 2.0 * a2) * rs + a1;
    const double den_p = ((4.0 * b4 * rs + 3.0 * b3) * rs + 2.0 * b2) * rs + b1;
    const double dfact = ((den * num_p - den_p * num) / (den * den)) * drs;
    *v = -(dfact * rho + num / den);
    *e = -(num / den) * rho;
    *m = 0.0;
    return;
  }
  const double r = rho <= 0.0 ? 1.0e-10 : rho;
  const double rs = cbrt(3.0 / (4.0 * M_PI * r));
  const double sq = sqrt(rs);
  const double vx = -cbrt(3.0 * r / M_PI);
  double ec, mu;
  if (kind == SDFT_XC_PZ_LDA) {
    const double g = -0.14230, b1 = 1.05290, b2 = 0.33340;
    const double A = 0.03110, B = -0.0480, C = 0.00200, D = -0.01160;
    if (rs >= 1.0) {
      const double cf1 = 1.0 + b1 * sq + b2 * rs;
      const double cf2 = 1.0 + (7.0 / 6.0) * b1 * sq + (4.0 / 3.0) * b2 * rs;
      ec = g / cf1;
      mu = ec * (cf2 / cf1);
    } else {
      const double lr = log(rs);
      ec = A * lr + B + C * rs * lr + D * rs;
      mu = A * lr + (B - A / 3.0) + (2.0 / 3.0) * C * rs * lr + (1.0 / 3.0) * (2.0 * D - C) * rs;
    }
  } else {
    const double a = 0.03109070, al1 = 0.213700;
    const double be1 = 7.59570, be2 = 3.58760, be3 = 1.63820, be4 = 0.492940;
    const double cf1 = 2.0 * a * (be1 * sq + be2 * rs + be3 * rs * sq + be4 * rs * rs);
    const double cf2 = 2.0 * a * (0.5 * be1 / sq + be2 + 1.5 * be3 * sq + 2.0 * be4 * rs);
    const double lg = log(1.0 + 1.0 / cf1);
    const double opa = 1.0 + al1 * rs;
    ec = -2.0 * a * opa * lg;
    mu = -2.0 * a * (1.0 + 2.0 * al1 * rs / 3.0) * lg +
         (-2.0 * a) * rs * opa / (3.0 * cf1 * (1.0 + cf1)) * cf2;
  }
  *v = vx + mu;
  *e = (0.75 * vx + ec) * r;
  *m = (vx + mu) * r;
}

// vks[o] = vhe[o] + v_xc(rho[o]) (FFTW order); per block the partial sums of
// rho eps_xc and of the muxc term
__global__ __launch_bounds__(256) void k_ks_xc(const double* __restrict__ vhe,
                                               const double* __restrict__ rho, long long N,
                                               int kind, double* __restrict__ vks,
                                               float* __restrict__ vks32,
                                               double* __restrict__ part) {
  __shared__ double sh[2][256];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0, m = 0.0;
  if (i < N) {
    double v;
    lda_point(rho[i], kind, &v, &e, &m);
    const double out = vhe[i] + v;
    vks[i] = out;
    vks32[i] = (float)out;
  }
  sh[0][threadIdx.x] = e;
  sh[1][threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0][0];
    part[2 * blockIdx.x + 1] = sh[1][0];
  }
}

// vks in the caller's layout (FFTW or generic)
__global__ void k_ks_out(const double* __restrict__ vks, int layout, int n, long long N,
                         double* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  long long o = i;
  if (layout == 1) {
    const long long ia = i / ((long long)n * n);
    const long long rem = i - ia * (long long)n * n;
    const long long ib = rem / n, ic = rem - ib * n;
    o = (ic * n + ib) * n + ia;
  }
  out[o] = vks[i];
}

// Local pseudopotential Hellmann-Feynman force on atom a = blockIdx.x: with
// v_ext(G) = sum_a f_type(a)(G) e^{-iG.R_a} and E_ext = sum_G Re(v_ext conj Q),
// F_a = -dE/dR_a = 2 sum_{G != 0, half sphere} G f(G) (cos(G.R_a) Im Q +
// sin(G.R_a) Re Q) (the local part of calcEnergyForce,
// stodft/calc-nuclei-force.c:254 -> control_ewd_loc).  g: [3][ncp] cartesian
// G, vloc: [ntype][ncp], both in the internal coefficient order.
__global__ __launch_bounds__(256) void k_local_forces(const double* __restrict__ Q, int ncp,
                                                      int ncoef, const double* __restrict__ g,
                                                      const double* __restrict__ vloc,
                                                      const double* __restrict__ pos,
                                                      const int* __restrict__ type,
                                                      double* __restrict__ F) {
  __shared__ double sh[3][256];
  const int a = blockIdx.x;
  const double rx = pos[3 * a], ry = pos[3 * a + 1], rz = pos[3 * a + 2];
  const double* f = vloc + (size_t)type[a] * ncp;
  double acc[3] = {0.0, 0.0, 0.0};
  for (int j = threadIdx.x; j < ncoef - 1; j += 256) {
    const double gx = g[j], gy = g[ncp + j], gz = g[2 * ncp + j];
    double sn, cs;
    sincos(gx * rx + gy * ry + gz * rz, &sn, &cs);
    const double t = f[j] * (cs * Q[ncp + j] + sn * Q[j]);
    acc[0] += t * gx;
    acc[1] += t * gy;
    acc[2] += t * gz;
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) sh[d][threadIdx.x] = acc[d];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int d = 0; d < 3; ++d) sh[d][threadIdx.x] += sh[d][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int d = 0; d < 3; ++d) F[3 * a + d] = 2.0 * sh[d][0];
}

int launch_local_forces(hipStream_t s, const double* Q, int ncp, int ncoef, const double* g,
                        const double* vloc, const double* pos, const int* type, int natm,
                        double* F) {
  hipLaunchKernelGGL(k_local_forces, dim3(natm), dim3(256), 0, s, Q, ncp, ncoef, g, vloc, pos,
                     type, F);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ks_field(hipStream_t s, const double* rho, int layout, int n, long long N,
                    double vol, void* field, double* rho_fftw) {
  hipLaunchKernelGGL(k_ks_field, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, rho, layout,
                     n, N, 1.0 / vol, reinterpret_cast<double2*>(field), rho_fftw);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ks_hartree(hipStream_t s, const double* Q, double* V, const double* ak2, int ncoef,
                      int ncp, double vol, double* part, int* nblocks) {
  *nblocks = (ncoef + 255) / 256;
  hipLaunchKernelGGL(k_ks_hartree, dim3(*nblocks), dim3(256), 0, s, Q, V, ak2, ncoef, ncp, vol,
                     part);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ks_xc(hipStream_t s, const double* vhe, const double* rho, long long N, int kind,
                 double* vks, float* vks32, double* part, int* nblocks) {
  *nblocks = (int)((N + 255) / 256);
  hipLaunchKernelGGL(k_ks_xc, dim3(*nblocks), dim3(256), 0, s, vhe, rho, N, kind, vks, vks32,
                     part);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_ks_out(hipStream_t s, const double* vks, int layout, int n, long long N,
                  double* out) {
  hipLaunchKernelGGL(k_ks_out, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, vks, layout, n,
                     N, out);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace sdft
