// sdft_kernels.hip -- the G->r->G sandwich of H.psi for Gamma-packed orbital
// pairs, as three hand-written gfx950 kernels:
//
//   k_expand   : per (pair, column chunk): Gamma pair pack of the sphere
//                coefficients (dble_pack_coef_fftw3d_filter,
//                mathlib/fft_package.c:3195-3254) straight into LDS columns
//                along a, inverse FFT along a, write W[pair][ra][column].
//   k_plane    : per (pair, a-plane): columns -> (kb, c) rows, inverse FFT
//                along c, then in sub-chunks of rc: inverse FFT along b,
//                z *= v_KS (cp_vpsi_threads, energy/cp/cp_energy_ee_rho.c:
//                2817-2845), forward FFT along b; forward FFT along c; back
//                to columns.  The real-space pair grid exists only in LDS.
//   k_contract : per (pair, column-pair chunk): forward FFT along a, then per
//                stored coefficient the unpack of F(G), F(-G)
//                (dble_upack_coef_sum_fftw3d_filter, fft_package.c:3513-3599),
//                the NL force from the KB GEMM, the kinetic term
//                (cp-energy-ee-rho-stodft.c:305-320), and -- in filter mode --
//                the Chebyshev recurrence of normHCheby (stodft/normh.c:
//                259-338) and the accumulation of filter.c:596-612.
//
// Only the sphere's columns (kb, kc) are transformed along a and only the
// sphere's kb rows along c, so an a-plane holds M x n values between the
// first and the last kernel instead of n^2, and no n^3 grid touches HBM.
#include "fft_lds.h"
#include "sdft_internal.h"
#include "sdft_plane.h"

namespace sdft {


int chunk_cols_for(int n) {
  switch (n) {
    case 24: return Cfg<24>::CH;
    case 32: return Cfg<32>::CH;
    case 36: return Cfg<36>::CH;
    case 48: return Cfg<48>::CH;
    case 64: return Cfg<64>::CH;
    case 96: return Cfg<96>::CH;
    case 128: return Cfg<128>::CH;
    case 192: return Cfg<192>::CH;
    default: return 0;
  }
}


// ---- block reduction of 3 doubles ----------------------------------------
__device__ inline void block_sum3(double* v, double* sh, int tid) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double x = v[i];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    v[i] = x;
  }
  const int lane = tid & 63, wave = tid >> 6;
  if (lane == 0) {
    sh[wave * 3 + 0] = v[0];
    sh[wave * 3 + 1] = v[1];
    sh[wave * 3 + 2] = v[2];
  }
  __syncthreads();
  if (tid == 0) {
    double a = 0, b = 0, c = 0;
    for (int w = 0; w < kThreads / 64; ++w) {
      a += sh[w * 3 + 0];
      b += sh[w * 3 + 1];
      c += sh[w * 3 + 2];
    }
    v[0] = a;
    v[1] = b;
    v[2] = c;
  }
}

// ---- K1: pack + inverse FFT along a --------------------------------------
template <typename T, int N>
__device__ __forceinline__ void expand_body(const ExpandArgs& a, int chunk, int q) {
  constexpr int CH = Cfg<N>::CH, LD = Cfg<N>::LD, ML = Cfg<N>::ML;
  __shared__ Cx<T> buf[CH * LD];
  __shared__ Cx<T> tw[N];
  const double* __restrict__ X = a.X;
  const int ldx = a.ldx, ncp = a.ncp, ncolp = a.ncolp;
  const int* __restrict__ chunk_col = a.chunk_col;
  const int* __restrict__ ent_off = a.ent_off;
  const int* __restrict__ ent_j = a.ent_j;
  const int* __restrict__ ent_b = a.ent_b;
  const Cx<T>* __restrict__ twg = reinterpret_cast<const Cx<T>*>(a.tw);
  Cx<T>* __restrict__ W = reinterpret_cast<Cx<T>*>(a.W);
  const int tid = threadIdx.x;
  const int c0 = chunk_col[chunk];
  const int nc = chunk_col[chunk + 1] - c0;
  for (int i = tid; i < N; i += kThreads) tw[i] = twg[i];
  for (int i = tid; i < nc * LD; i += kThreads) buf[i] = cx(T(0), T(0));
  __syncthreads();
  const double* x1 = X + (size_t)(2 * q) * ldx;
  const double* x2 = x1 + ldx;
  const int e1 = ent_off[c0 + nc];
  for (int e = ent_off[c0] + tid; e < e1; e += kThreads) {
    const int j = ent_j[e];
    const int b = ent_b[e];
    const int d = b >> 10, flag = (b >> 8) & 3, ia = b & 255;
    const double ar = x1[j], ai = x1[ncp + j], br = x2[j], bi = x2[ncp + j];
    double zr, zi;
    if (flag == 0) {        // +G: c1 + i c2
      zr = ar - bi;
      zi = ai + br;
    } else if (flag == 1) { // -G: c1* + i c2*
      zr = ar + bi;
      zi = -ai + br;
    } else {                // G = 0: c1re + i c2re
      zr = ar;
      zi = br;
    }
    buf[(d - c0) * LD + ia] = cx(T(zr), T(zi));
  }
  __syncthreads();
  fft_lines_map<T, N, +1, ML, kThreads, kLineFastA<T>>(buf, nc, LD, tw, tid);
  Cx<T>* Wq = W + (size_t)q * N * ncolp;
  const float rnc = slot_rcp(nc);
  for (int e = tid; e < nc * N; e += kThreads) {
    int dl, ra;
    slot_map<true, 1>(e, nc, rnc, dl, ra);
    Wq[(size_t)ra * ncolp + c0 + dl] = buf[dl * LD + ra];
  }
}

template <typename T, int N>
__global__ __launch_bounds__(kThreads) void k_expand(ExpandArgs a) {
  expand_body<T, N>(a, blockIdx.x, blockIdx.y);
}

template <typename T, int N>
__global__ __launch_bounds__(kThreads) void k_expand_batch(const ExpandArgs* __restrict__ args,
                                                           const int* __restrict__ prefix,
                                                           int nfrag) {
  const int f = find_frag(prefix, nfrag, blockIdx.x);
  const ExpandArgs a = args[f];
  const int local = blockIdx.x - prefix[f];
  expand_body<T, N>(a, local % a.gx, local / a.gx);
}


template <typename T, int N, int V, bool PR, int NT = kThreads, int SC = 0, int SB = 0>
__global__ __launch_bounds__(NT) void k_plane(PlaneArgs a) {
  plane_body<T, N, V, PR, 0, NT, SC, SB>(a, blockIdx.x, blockIdx.y);
}

template <typename T, int N, int V, bool PR>
__global__ __launch_bounds__(kThreads) void k_plane_batch(const PlaneArgs* __restrict__ args,
                                                          const int* __restrict__ prefix,
                                                          int nfrag) {
  const int f = find_frag(prefix, nfrag, blockIdx.x);
  const PlaneArgs a = args[f];
  const int local = blockIdx.x - prefix[f];
  plane_body<T, N, V, PR>(a, local % a.gx, local / a.gx);
}


template <typename T, int N>
__global__ __launch_bounds__(kThreads) void k_plane_split(PlaneArgs a) {
  plane_split_body<T, N>(a, blockIdx.x, blockIdx.y);
}

// ---- noise: real-space Rademacher -> sphere columns ----------------------
// genNoiseOrbitalReal (stodft/gen-noise.c:162-374): real-space values
// +-sqrt(N) (ranValue = sqrt(nfft2), :186-195), transformed to G space.  The
// sign of grid point r of orbital o is the top bit of splitmix64(seed << 44 |
// o << 23 | r) (r the FFTW linear index), the same counter-based stream as the
// host system builder.  Two orbitals share one complex field z = chi1 + i chi2.
__device__ inline unsigned long long splitmix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__device__ inline double rademacher(unsigned long long seed, unsigned long long orb,
                                    unsigned long long r) {
  const unsigned long long h = splitmix64((seed << 44) | (orb << 23) | r);
  return (h >> 63) ? -1.0 : 1.0;
}

template <int N>
size_t noise_lds_bytes(int mc) {
  return sizeof(Cx<double>) * ((size_t)Cfg<N>::ML * Cfg<N>::LD + (size_t)mc * Cfg<N>::LD + N);
}

// per (pair, a-plane): forward FFT along c of the generated rows (keeping
// |kc| <= K), then along b (keeping |kb| <= K), into W's sphere columns.
// With `field` non-null the rows are read from a dense pair grid instead
// (field[q][ra][rb][rc], FFTW order): the real-space projector path.  The
// kc lines are produced mc at a time (mc = M, or fewer when M lines of the
// plane do not fit LDS: the rows are then generated and c-transformed once
// per chunk of kc).
template <int N>
__global__ __launch_bounds__(kThreads) void k_noise_plane(Cx<double>* __restrict__ W, int ncolp,
                                                          const int* __restrict__ rowcol, int M,
                                                          const Cx<double>* __restrict__ twg,
                                                          unsigned long long seed, int first,
                                                          int nstate,
                                                          const Cx<double>* __restrict__ field,
                                                          int mc) {
  typedef double T;
  constexpr int LD = Cfg<N>::LD, ML = Cfg<N>::ML;
  constexpr int P = n_passes(N);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cx<T>* Rb = reinterpret_cast<Cx<T>*>(smem_raw);  // ML rows (rb) of the plane
  Cx<T>* Bk = Rb + ML * LD;                        // mc lines (kc) along b
  Cx<T>* tw = Bk + mc * LD;
  const int tid = threadIdx.x;
  const int ra = blockIdx.x;
  const int q = blockIdx.y;
  const int K = (M - 1) / 2;
  for (int i = tid; i < N; i += kThreads) tw[i] = twg[i];
  __syncthreads();
  const int o1 = 2 * q, o2 = 2 * q + 1;
  const double amp = sqrt((double)N * N * N);
  const bool has2 = o2 < nstate;
  Cx<T>* Wp = W + ((size_t)q * N + ra) * ncolp;
  for (int k0 = 0; k0 < M; k0 += mc) {  // kc lines k0 - K .. k0 + mc - 1 - K
  const int mk = M - k0 < mc ? M - k0 : mc;
  for (int rb0 = 0; rb0 < N; rb0 += ML) {
    const int nl = N - rb0 < ML ? N - rb0 : ML;
    auto gen = [&](int l, int pos) {
      const unsigned long long r = ((unsigned long long)ra * N + rb0 + l) * N + pos;
      if (field) return field[(size_t)q * N * N * N + r];
      const double x1 = amp * rademacher(seed, (unsigned long long)(first + o1), r);
      const double x2 = has2 ? amp * rademacher(seed, (unsigned long long)(first + o2), r) : 0.0;
      return cx(x1, x2);
    };
    auto sk = [&](int l, int pos, Cx<T> x) {
      const int kc = pos <= N / 2 ? pos : pos - N;
      const int i = kc + K - k0;
      if (i >= 0 && i < mk) Bk[i * LD + rb0 + l] = x;
    };
    fft_passes_io<T, N, -1, false, 0, P - 1, ML, kThreads, false>(Rb, LD, nl, tw, tid, gen, sk);
  }
  for (int c0 = k0; c0 < k0 + mk; c0 += ML) {
    const int nl = k0 + mk - c0 < ML ? k0 + mk - c0 : ML;
    Cx<T>* buf = Bk + (c0 - k0) * LD;
    auto lb = [&](int l, int pos) { return buf[l * LD + pos]; };
    auto sw = [&](int l, int pos, Cx<T> x) {
      const int kb = pos <= N / 2 ? pos : pos - N;
      if (kb < -K || kb > K) return;
      const int kc = c0 + l - K;
      const int d = rowcol[(kb + K) * N + (kc < 0 ? kc + N : kc)];
      if (d >= 0) Wp[d] = x;
    };
    fft_passes_io<T, N, -1, false, 0, P - 1, ML, kThreads, false>(buf, LD, nl, tw, tid, lb, sw);
  }
  }
}

// ---- K3: forward FFT along a + unpack + force / recurrence ---------------
// E = true compiles the fused expand (a.Wnext); E = false leaves it out, so
// the kernel without it keeps its lower register count.
template <typename T, int N, bool E>
__device__ __forceinline__ void contract_body(const ContractArgs& a, int chunk, int q) {
  constexpr int CH = Cfg<N>::CH, LD = Cfg<N>::LD, ML = Cfg<N>::ML;
  __shared__ Cx<T> buf[CH * LD];
  __shared__ Cx<T> tw[N];
  __shared__ double red[3 * (kThreads / 64)];
  // W may be the buffer Wnext writes (the fused expand): no __restrict__
  const Cx<T>* W = reinterpret_cast<const Cx<T>*>(a.W);
  const int ncolp = a.ncolp;
  const int* __restrict__ chunk_col = a.chunk_col;
  const int* __restrict__ chunk_coef = a.chunk_coef;
  const int* __restrict__ coef_plus = a.coef_plus;
  const int* __restrict__ coef_minus = a.coef_minus;
  const double* __restrict__ ak2 = a.ak2;
  const Cx<T>* __restrict__ twg = reinterpret_cast<const Cx<T>*>(a.tw);
  const double* __restrict__ Xcur = a.Xcur;
  double* __restrict__ Xout = a.Xout;
  double* __restrict__ acc = a.acc;
  const T* __restrict__ fnl = reinterpret_cast<const T*>(a.fnl);
  const double* __restrict__ chi = a.chi;
  const int ldx = a.ldx, ncp = a.ncp, ncoef = a.ncoef, mode = a.mode;
  const double invN = a.invN;
  double* __restrict__ mom_part = a.mom_part;
  StepParams sp = a.sp;
  if (a.coeffs_dev) {
    const int st = *a.step_dev;
    sp.ck = a.coeffs_dev[st];
    sp.step = st;
  }
  const int tid = threadIdx.x;
  const int c0 = chunk_col[chunk];
  const int nc = chunk_col[chunk + 1] - c0;
  // the unpack's first coefficient indices (and, for the fused expand, each
  // column's sphere half-height) are loaded before the transform, so that
  // their latency overlaps it
  const int j0 = chunk_coef[chunk], j1 = chunk_coef[chunk + 1];
  int pp_first = 0, pm_first = 0;
  if (j0 + tid < j1) {
    pp_first = coef_plus[j0 + tid];
    pm_first = coef_minus[j0 + tid];
  }
  Cx<T>* Wn = reinterpret_cast<Cx<T>*>(a.Wnext);
  const bool emit =
      E && Wn != nullptr && (mode == kModeFilter || mode == kModeMoments || mode == kModeNewton);
  __shared__ int col_h[E ? CH : 1];
  if (E && emit && tid < nc)
    col_h[tid] = (a.ent_off[c0 + tid + 1] - a.ent_off[c0 + tid] - 1) >> 1;
  for (int i = tid; i < N; i += kThreads) tw[i] = twg[i];
  const Cx<T>* Wq = W + (size_t)q * N * ncolp;
  const float rnc = slot_rcp(nc);
  for (int e = tid; e < nc * N; e += kThreads) {
    int dl, ra;
    slot_map<true, 1>(e, nc, rnc, dl, ra);
    buf[dl * LD + ra] = Wq[(size_t)ra * ncolp + c0 + dl];
  }
  __syncthreads();
  fft_lines_map<T, N, -1, ML, kThreads, kLineFastA<T>>(buf, nc, LD, tw, tid);

  double dots[3] = {0.0, 0.0, 0.0};
  // (pp, pm) of coefficient j; the increment loads the next trip's pair
  for (int j = j0 + tid, pp = pp_first, pm = pm_first; j < j1;
       j += kThreads, pp = j < j1 ? coef_plus[j] : 0, pm = j < j1 ? coef_minus[j] : 0) {
    Cx<T>* bp = &buf[((pp >> 8) - c0) * LD + (pp & 255)];
    Cx<T>* bm = &buf[((pm >> 8) - c0) * LD + (pm & 255)];
    const Cx<T> t = *bp;
    const Cx<T> tc = *bm;
    const double tr = double(t.x) * invN, ti = double(t.y) * invN;
    const double tcr = double(tc.x) * invN, tci = double(tc.y) * invN;
    const bool g0 = (j == ncoef - 1);
    double f[2][2];
    if (!g0) {
      f[0][0] = -2.0 * (tr + tcr);
      f[0][1] = -2.0 * (ti - tci);
      f[1][0] = -2.0 * (ti + tci);
      f[1][1] = -2.0 * (-tr + tcr);
    } else {
      f[0][0] = -2.0 * tr;
      f[0][1] = 0.0;
      f[1][0] = -2.0 * ti;
      f[1][1] = 0.0;
    }
    const double k2 = ak2[j];
    if (mode == kModeNoise) {
      // c(G) = (1/N) sum_r chi(r) e^{-iG.r} of the two real fields: minus a
      // quarter (G != 0) / half (G = 0) of the unpack
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        const size_t base = (size_t)(2 * q + o) * ldx;
        const double s = g0 ? -0.5 : -0.25;
        Xout[base + j] = s * f[o][0];
        Xout[base + ncp + j] = s * f[o][1];
        acc[base + j] = s * f[o][0];
        acc[base + ncp + j] = s * f[o][1];
      }
      continue;
    }
    double nv[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
    for (int o = 0; o < 2; ++o) {
      const size_t base = (size_t)(2 * q + o) * ldx;
      const double cr = Xcur[base + j], ci = Xcur[base + ncp + j];
      double fr = f[o][0], fi = f[o][1];
      if (fnl) {
        fr += double(fnl[base + j]);
        fi += double(fnl[base + ncp + j]);
      }
      if (!g0) {
        fr -= 2.0 * k2 * cr;
        fi -= 2.0 * k2 * ci;
      } else {
        fi = 0.0;
      }
      if (mode == kModeForce) {
        Xout[base + j] = fr;
        Xout[base + ncp + j] = fi;
        continue;
      }
      double nr, ni;
      if (mode == kModeNewton) {
        // normHNewtonHerm (normh.c:101-150): scale the force, add prefact c
        nr = (g0 ? sp.s2 : sp.s1) * fr + sp.p * cr;
        ni = g0 ? fi + sp.p * ci : sp.s1 * fi + sp.p * ci;
        Xout[base + j] = nr;
        Xout[base + ncp + j] = ni;
        for (int mu = 0; mu < a.nmu; ++mu) {
          double* am = acc + mu * a.acc_stride;
          const double c1 = a.newton_c[(size_t)sp.step * a.nmu + mu];
          if (sp.first) {
            const double c0 = a.newton_c[mu];
            am[base + j] = c0 * cr + c1 * nr;
            am[base + ncp + j] = c0 * ci + c1 * ni;
          } else {
            am[base + j] += c1 * nr;
            am[base + ncp + j] += c1 * ni;
          }
        }
      } else {
        if (sp.first) {
          nr = (g0 ? sp.s2 : sp.s1) * fr + sp.p * cr;
          ni = g0 ? fi : sp.s1 * fi + sp.p * ci;
        } else {
          const double pr = Xout[base + j], pi = Xout[base + ncp + j];
          nr = (g0 ? sp.s2 : sp.s1) * fr + sp.p * cr - pr;
          ni = g0 ? fi : sp.s1 * fi + sp.p * ci - pi;
        }
        Xout[base + j] = nr;
        Xout[base + ncp + j] = ni;
        if (mode == kModeFilter) {
          if (sp.first) {
            acc[base + j] = sp.c0 * cr + sp.ck * nr;
            acc[base + ncp + j] = sp.c0 * ci + sp.ck * ni;
          } else {
            acc[base + j] += sp.ck * nr;
            acc[base + ncp + j] += sp.ck * ni;
          }
        }
        if (mom_part) {
          // Gamma dot as coded at calc-chempot-chebyshev.c:725-729:
          // 2 sum re*re + 2 sum im*im - re(0)*re(0)
          const double wr = g0 ? 1.0 : 2.0;
          const double xr = chi[base + j], xi = chi[base + ncp + j];
          dots[0] += wr * xr * nr + 2.0 * xi * ni;
          if (sp.first)
            dots[1] += wr * cr * cr + 2.0 * ci * ci;
          else
            dots[1] += wr * nr * cr + 2.0 * ni * ci;
          dots[2] += wr * nr * nr + 2.0 * ni * ni;
        }
      }
      nv[o][0] = nr;
      nv[o][1] = ni;
      if (a.Xout32) {
        a.Xout32[base + j] = float(nr);
        a.Xout32[base + ncp + j] = float(ni);
      }
    }
    if (emit) {
      // the Gamma pair pack of the new pair (as in k_expand) into the two
      // positions this coefficient was read from: no other thread touches them
      const double ar = nv[0][0], ai = nv[0][1], br = nv[1][0], bi = nv[1][1];
      if (g0) {
        *bp = cx(T(ar), T(br));
      } else {
        *bp = cx(T(ar - bi), T(ai + br));
        *bm = cx(T(ar + bi), T(br - ai));
      }
    }
  }
  if (mom_part) {
    block_sum3(dots, red, tid);
    if (tid == 0) {
      const size_t row = (size_t)sp.step * a.gx * a.gy + (size_t)q * a.gx + chunk;
      mom_part[row * 3 + 0] = dots[0];
      mom_part[row * 3 + 1] = dots[1];
      mom_part[row * 3 + 2] = dots[2];
    }
  }
  if (emit) {
    // positions outside the sphere (|ka| > h of the column, h from its entry
    // count) become zero: disjoint from the positions written above
    const Cx<T> zero = cx(T(0), T(0));
    for (int e = tid; e < nc * N; e += kThreads) {
      const int dl = e / N, ia = e - dl * N;
      const int h = col_h[E ? dl : 0];
      const int ka = ia <= N / 2 ? ia : ia - N;
      if (ka > h || ka < -h) buf[dl * LD + ia] = zero;
    }
    __syncthreads();
    fft_lines_map<T, N, +1, ML, kThreads, kLineFastA<T>>(buf, nc, LD, tw, tid);
    Cx<T>* Wq2 = Wn + (size_t)q * N * ncolp;
    for (int e = tid; e < nc * N; e += kThreads) {
      int dl, ra;
      slot_map<true, 1>(e, nc, rnc, dl, ra);
      Wq2[(size_t)ra * ncolp + c0 + dl] = buf[dl * LD + ra];
    }
  }
}

template <typename T, int N, bool E>
__global__ __launch_bounds__(kThreads) void k_contract(ContractArgs a) {
  contract_body<T, N, E>(a, blockIdx.x, blockIdx.y);
}

template <typename T, int N, bool E>
__global__ __launch_bounds__(kThreads) void k_contract_batch(const ContractArgs* __restrict__ args,
                                                             const int* __restrict__ prefix,
                                                             int nfrag) {
  const int f = find_frag(prefix, nfrag, blockIdx.x);
  const ContractArgs a = args[f];
  const int local = blockIdx.x - prefix[f];
  contract_body<T, N, E>(a, local % a.gx, local / a.gx);
}

// ---- density: |psi1|^2 + |psi2|^2 summed over the pairs of a group -------
template <typename T, int N, bool RE>
__global__ __launch_bounds__(kThreads) void k_density(const Cx<T>* __restrict__ W, int ncolp,
                                                      int ncol, const int* __restrict__ col_pos,
                                                      int M, const Cx<T>* __restrict__ twg,
                                                      double* __restrict__ rho_part, int npairs,
                                                      int ngroups) {
  constexpr int LD = Cfg<N>::LD, ML = Cfg<N>::ML, CB = Cfg<N>::CB;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cx<T>* P = reinterpret_cast<Cx<T>*>(smem_raw);
  Cx<T>* Q = P + M * LD;
  Cx<T>* tw = Q + CB * LD;
  const int tid = threadIdx.x;
  const int ra = blockIdx.x;
  const int g = blockIdx.y;
  const int K = (M - 1) / 2;
  double* rp = rho_part + ((size_t)g * N + ra) * N * N;
  for (int i = tid; i < N; i += kThreads) tw[i] = twg[i];
  bool first = true;
  for (int q = g; q < npairs; q += ngroups) {
    for (int i = tid; i < M * LD; i += kThreads) P[i] = cx(T(0), T(0));
    __syncthreads();
    const Cx<T>* Wp = W + ((size_t)q * N + ra) * ncolp;
    for (int d = tid; d < ncol; d += kThreads) P[col_pos[d]] = Wp[d];
    __syncthreads();
    fft_lines_batched<T, N, +1, ML, kThreads>(P, M, LD, tw, tid);
    for (int rc0 = 0; rc0 < N; rc0 += CB) {
      for (int i = tid; i < CB * LD; i += kThreads) Q[i] = cx(T(0), T(0));
      __syncthreads();
      for (int e = tid; e < M * CB; e += kThreads) {
        const int r = e / CB, l = e - r * CB;
        Q[l * LD + wrapi(r - K, N)] = P[r * LD + rc0 + l];
      }
      __syncthreads();
      fft_lines<T, N, +1, ML, kThreads>(Q, CB, LD, tw, tid);
      for (int e = tid; e < CB * N; e += kThreads) {
        const int rb = e / CB, l = e - rb * CB;
        const Cx<T> z = Q[l * LD + rb];
        const size_t idx = (size_t)rb * N + rc0 + l;
        if (RE) {
          rp[idx] = double(z.x);
        } else {
          const double s = double(z.x) * double(z.x) + double(z.y) * double(z.y);
          rp[idx] = first ? s : rp[idx] + s;
        }
      }
      __syncthreads();
    }
    first = false;
  }
  if (first) {
    for (int e = tid; e < N * N; e += kThreads) rp[e] = 0.0;
  }
}

// Split form of k_density for planes whose M x (n+1) buffer exceeds LDS: per
// pair and column part (S parts of n / S columns), the inverse c-transform
// of the M rows (through Q, recomputed per part) keeps the part's columns,
// and its b-lines accumulate |psi1|^2 + |psi2|^2 straight from the last
// pass's outputs.
template <typename T, int N, bool RE, int S = 2>
__global__ __launch_bounds__(kThreads) void k_density_split(
    const Cx<T>* __restrict__ W, int ncolp, const int* __restrict__ rowcol, int M,
    const Cx<T>* __restrict__ twg, double* __restrict__ rho_part, int npairs, int ngroups) {
  constexpr int LD = Cfg<N>::LD, ML = Cfg<N>::ML, CB = PlaneCfg<T, N>::CB;
  constexpr int H = SplitCfg<T, N, S>::H, LDH = SplitCfg<T, N, S>::LDH;
  constexpr int QL = SplitCfg<T, N, S>::QL;
  constexpr int P = n_passes(N);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cx<T>* Ph = reinterpret_cast<Cx<T>*>(smem_raw);
  Cx<T>* Q = Ph + M * LDH;
  Cx<T>* tw = Q + QL * LD;
  const int tid = threadIdx.x;
  const int ra = blockIdx.x;
  const int g = blockIdx.y;
  const int K = (M - 1) / 2;
  double* rp = rho_part + ((size_t)g * N + ra) * N * N;
  for (int i = tid; i < N; i += kThreads) tw[i] = twg[i];
  __syncthreads();
  const Cx<T> zero = cx(T(0), T(0));
  for (int q = g; q < npairs; q += ngroups) {
    const bool first = (q == g);
    const Cx<T>* Wp = W + ((size_t)q * N + ra) * ncolp;
    for (int h = 0; h < S; ++h) {
      const int c_lo = h * H;
      for (int r0 = 0; r0 < M; r0 += ML) {
        const int nl = M - r0 < ML ? M - r0 : ML;
        const int* rc = rowcol + (size_t)r0 * N;
        auto lw = [&](int l, int pos) {
          const int d = rc[l * N + pos];
          const Cx<T> x = Wp[d >= 0 ? d : 0];
          const T m = d >= 0 ? T(1) : T(0);
          return cx(x.x * m, x.y * m);
        };
        auto sh = [&](int l, int pos, Cx<T> x) {
          const int c = pos - c_lo;
          if (c >= 0 && c < H) Ph[(r0 + l) * LDH + c] = x;
        };
        fft_passes_io<T, N, +1, false, 0, P - 1, ML, kThreads, false>(Q, LD, nl, tw, tid, lw, sh);
      }
      for (int rc0 = 0; rc0 < H; rc0 += CB) {
        const int nb = H - rc0 < CB ? H - rc0 : CB;
        auto lp = [&](int l, int pos) {
          const int kb = pos <= N / 2 ? pos : pos - N;
          return (kb >= -K && kb <= K) ? Ph[(kb + K) * LDH + rc0 + l] : zero;
        };
        auto sr = [&](int l, int pos, Cx<T> z) {
          const size_t idx = (size_t)pos * N + c_lo + rc0 + l;
          if (RE) {
            rp[idx] = double(z.x);
          } else {
            const double s2 = double(z.x) * double(z.x) + double(z.y) * double(z.y);
            rp[idx] = first ? s2 : rp[idx] + s2;
          }
        };
        fft_passes_io<T, N, +1, false, 0, P - 1, CB, kThreads, true>(Q, LD, nb, tw, tid, lp, sr);
      }
    }
  }
  if (g >= npairs) {
    for (int e = tid; e < N * N; e += kThreads) rp[e] = 0.0;
  }
}

// rho = scale * sum_g rho_part[g] (+ rho), FFTW or generic (a fastest) order
__global__ void k_rho_finish(const double* __restrict__ part, int ngroups, long long Ntot, int n,
                             double scale, int layout, int accumulate, double* __restrict__ rho) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= Ntot) return;
  double s = 0.0;
  for (int g = 0; g < ngroups; ++g) s += part[(size_t)g * Ntot + i];
  s *= scale;
  long long o = i;
  if (layout == 1) {
    const long long ia = i / ((long long)n * n);
    const long long rem = i - ia * (long long)n * n;
    const long long ib = rem / n, ic = rem - ib * n;
    o = (ic * n + ib) * n + ia;
  }
  rho[o] = accumulate ? rho[o] + s : s;
}

// ---- FFT self-test: lines in global memory ------------------------------
template <typename T, int N, int DIR>
__global__ __launch_bounds__(kThreads) void k_fft_test(const double* __restrict__ in,
                                                       double* __restrict__ out, int nlines,
                                                       const Cx<T>* __restrict__ twg) {
  constexpr int LD = Cfg<N>::LD, ML = Cfg<N>::ML;
  __shared__ Cx<T> buf[ML * LD];
  __shared__ Cx<T> tw[N];
  const int tid = threadIdx.x;
  const int l0 = blockIdx.x * ML;
  const int nl = nlines - l0 < ML ? nlines - l0 : ML;
  for (int i = tid; i < N; i += kThreads) tw[i] = twg[i];
  for (int e = tid; e < nl * N; e += kThreads) {
    const int l = e / N, k = e - l * N;
    const size_t g = ((size_t)(l0 + l) * N + k) * 2;
    buf[l * LD + k] = cx(T(in[g]), T(in[g + 1]));
  }
  __syncthreads();
  fft_lines<T, N, DIR, ML, kThreads>(buf, nl, LD, tw, tid);
  for (int e = tid; e < nl * N; e += kThreads) {
    const int l = e / N, k = e - l * N;
    const size_t g = ((size_t)(l0 + l) * N + k) * 2;
    out[g] = double(buf[l * LD + k].x);
    out[g + 1] = double(buf[l * LD + k].y);
  }
}

// ==========================================================================
// launchers
// ==========================================================================


ExpandArgs make_expand_args(const Ctx& c, int prec, const double* X, void* W) {
  ExpandArgs a;
  a.X = X;
  a.ldx = 2 * c.g.ncp;
  a.ncp = c.g.ncp;
  a.chunk_col = c.t.chunk_col;
  a.ent_off = c.t.ent_off;
  a.ent_j = c.t.ent_j;
  a.ent_b = c.t.ent_b;
  a.tw = prec == 0 ? c.t.tw64 : c.t.tw32;
  a.W = W;
  a.ncolp = c.g.ncolp;
  a.gx = c.g.nchunk;
  return a;
}

PlaneArgs make_plane_args(const Ctx& c, int prec, void* W) {
  PlaneArgs a;
  a.W = W;
  a.Wout = nullptr;
  a.field = nullptr;
  a.ncolp = c.g.ncolp;
  a.rowcol = c.t.rowcol;
  a.M = c.g.M;
  a.v = prec == 0 ? (const void*)c.vks64 : (const void*)c.vks32;
  a.tw = prec == 0 ? c.t.tw64 : c.t.tw32;
  a.gx = c.g.n;
  return a;
}

ContractArgs make_contract_args(const Ctx& c, int prec, int mode, const void* W, int npairs,
                                const double* Xcur, double* Xout, double* acc, const void* fnl,
                                const double* chi, const StepParams& sp, double* mom_part) {
  ContractArgs a;
  a.W = W;
  a.ncolp = c.g.ncolp;
  a.chunk_col = c.t.chunk_col;
  a.chunk_coef = c.t.chunk_coef;
  a.coef_plus = c.t.coef_plus;
  a.coef_minus = c.t.coef_minus;
  a.ak2 = c.t.ak2;
  a.tw = prec == 0 ? c.t.tw64 : c.t.tw32;
  a.Xcur = Xcur;
  a.Xout = Xout;
  a.acc = acc;
  a.fnl = fnl;
  a.chi = chi;
  a.ldx = 2 * c.g.ncp;
  a.ncp = c.g.ncp;
  a.ncoef = c.g.ncoef;
  a.invN = 1.0 / (double)c.g.N;
  a.mode = mode;
  a.sp = sp;
  a.mom_part = mom_part;
  a.coeffs_dev = nullptr;
  a.step_dev = nullptr;
  a.nmu = c.newton_nmu;
  a.acc_stride = c.newton_stride;
  a.newton_c = c.newton_c;
  a.Wnext = nullptr;
  a.ent_off = c.t.ent_off;
  a.Xout32 = nullptr;
  a.gx = c.g.nchunk;
  a.gy = npairs;
  return a;
}

size_t plane_lds_for(int n, int prec, int M, int variant) {
  switch (n) {
#define SDFT_PL(NN)                                                                    \
  case NN:                                                                             \
    if (variant == 1)                                                                  \
      return prec == 0 ? plane2_lds_bytes<double, NN, 1>(M) : plane2_lds_bytes<float, NN, 1>(M); \
    return prec == 0 ? plane2_lds_bytes<double, NN, 0>(M) : plane2_lds_bytes<float, NN, 0>(M);
    SDFT_PL(24) SDFT_PL(32) SDFT_PL(36) SDFT_PL(48) SDFT_PL(64) SDFT_PL(96) SDFT_PL(128)
    SDFT_PL(192)
#undef SDFT_PL
    default:
      return 0;
  }
}

// ---- batched launches: one launch per kernel over the fragments of one grid
// size; `prefix` holds each fragment's first workgroup (size nfrag + 1).
template <typename T, int N, int V, bool PR>
static int plane_batch_launch(const void* args, const int* prefix, int nfrag, int total,
                              size_t lds, hipStream_t s) {
  SDFT_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_plane_batch<T, N, V, PR>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL((k_plane_batch<T, N, V, PR>), dim3(total), dim3(kThreads), lds, s,
                     reinterpret_cast<const PlaneArgs*>(args), prefix, nfrag);
  return 0;
}

template <typename T, int N>
static int batch_launch_impl(int kind, const void* args, const int* prefix, int nfrag, int total,
                             size_t lds, int variant, hipStream_t s) {
  if (kind == 0) {
    hipLaunchKernelGGL((k_expand_batch<T, N>), dim3(total), dim3(kThreads), 0, s,
                       reinterpret_cast<const ExpandArgs*>(args), prefix, nfrag);
  } else if (kind == 1) {  // variant: bit 0 tile, bit 1 pruned passes
    int rc = 0;
    switch (variant & 3) {
      case 0: rc = plane_batch_launch<T, N, 0, false>(args, prefix, nfrag, total, lds, s); break;
      case 1: rc = plane_batch_launch<T, N, 1, false>(args, prefix, nfrag, total, lds, s); break;
      case 2: rc = plane_batch_launch<T, N, 0, true>(args, prefix, nfrag, total, lds, s); break;
      default: rc = plane_batch_launch<T, N, 1, true>(args, prefix, nfrag, total, lds, s); break;
    }
    if (rc) return rc;
  } else if (variant == 1) {  // contract with the fused expand
    hipLaunchKernelGGL((k_contract_batch<T, N, true>), dim3(total), dim3(kThreads), 0, s,
                       reinterpret_cast<const ContractArgs*>(args), prefix, nfrag);
  } else {
    hipLaunchKernelGGL((k_contract_batch<T, N, false>), dim3(total), dim3(kThreads), 0, s,
                       reinterpret_cast<const ContractArgs*>(args), prefix, nfrag);
  }
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_batch(int kind, int n, int prec, const void* args, const int* prefix, int nfrag,
                 int total, size_t lds, int variant, hipStream_t s) {
  if (total <= 0) return 0;
  if (prec == 0) {
    SDFT_SIZE_SWITCH(n, return (batch_launch_impl<double, NN>(kind, args, prefix, nfrag, total,
                                                               lds, variant, s)));
  } else {
    SDFT_SIZE_SWITCH(n, return (batch_launch_impl<float, NN>(kind, args, prefix, nfrag, total, lds,
                                                              variant, s)));
  }
  return 0;
}

size_t w_bytes(const Ctx& c, int npairs, int prec) {
  const size_t elem = prec == 0 ? sizeof(Cx<double>) : sizeof(Cx<float>);
  return (size_t)npairs * c.g.n * c.g.ncolp * elem;
}

template <typename T, int N>
static int expand_impl(Ctx& c, const double* X, int npairs, void* W, const void* tw) {
  (void)tw;
  const ExpandArgs a = make_expand_args(c, prec_of<T>(), X, W);
  dim3 grid(a.gx, npairs);
  hipLaunchKernelGGL((k_expand<T, N>), grid, dim3(kThreads), 0, c.stream, a);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_expand(Ctx& c, int prec, const double* X, int npairs, void* W) {
  if (prec == 0) {
    SDFT_SIZE_SWITCH(c.g.n, return (expand_impl<double, NN>(c, X, npairs, W, c.t.tw64)));
  } else {
    SDFT_SIZE_SWITCH(c.g.n, return (expand_impl<float, NN>(c, X, npairs, W, c.t.tw32)));
  }
  return 0;
}

template <typename T, int N, int V, bool PR, int NT = kThreads, int SC = 0, int SB = 0>
static int plane_launch(Ctx& c, const PlaneArgs& a, size_t lds, int npairs) {
  auto kern = &k_plane<T, N, V, PR, NT, SC, SB>;
  SDFT_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kern, dim3(N, npairs), dim3(NT), lds, c.stream, a);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

// Plane-kernel variants (the tuner's plane tile): b-line tile V, threads per
// workgroup NT and the radix schedules of the c- and b-lines (SC, SB).  At
// n = 64 every thread of a pass has a butterfly when (lines x butterflies
// per line) reaches NT: with radix-4 b-lines (SB = 1) a 256-thread FP64
// workgroup fills its threads from a 16-line b tile, and 512-thread ones keep
// the plane's LDS buffer for twice the waves.
//   0: V 0, 256, 8x8 / 8x8     1: V 1, 256, 8x8 / 8x8   2: V 0, 256, 8x8 / 4x4x4
//   3: V 1, 512, 4x4x4 / 8x8   4: V 1, 512, 4x4x4 / 4x4x4
//   5: V 0, 512, 4x4x4 / 4x4x4
// Variants 2-5 differ from 0 / 1 only where the power of two in n is 2^6
// (64, 192) and are built for
// n = 64; other sizes take variant (pv & 1).
template <typename T, int N, bool PR>
static int plane_variant_launch(Ctx& c, const PlaneArgs& a, int pv, int npairs) {
  const int M = c.g.M;
  const size_t lds0 = plane2_lds_bytes<T, N, 0>(M), lds1 = plane2_lds_bytes<T, N, 1>(M);
  if constexpr (N == 64) {
    switch (pv) {
      case 2: return plane_launch<T, N, 0, PR, kThreads, 0, 1>(c, a, lds0, npairs);
      case 3: if (lds1 <= 160 * 1024) return plane_launch<T, N, 1, PR, 512, 1, 0>(c, a, lds1, npairs);
              break;
      case 4: if (lds1 <= 160 * 1024) return plane_launch<T, N, 1, PR, 512, 1, 1>(c, a, lds1, npairs);
              break;
      case 5: return plane_launch<T, N, 0, PR, 512, 1, 1>(c, a, lds0, npairs);
      default: break;
    }
  }
  if ((pv & 1) && lds1 <= 160 * 1024) return plane_launch<T, N, 1, PR>(c, a, lds1, npairs);
  return plane_launch<T, N, 0, PR>(c, a, lds0, npairs);
}

template <typename T, int N>
static int plane_impl(Ctx& c, void* W, int npairs, void** Wout) {
  const Geometry& g = c.g;
  PlaneArgs a = make_plane_args(c, prec_of<T>(), W);
  dim3 grid(N, npairs);
  const bool pr = plane_prunable(g.K, N);
  const size_t lds = plane2_lds_bytes<T, N, 0>(g.M);
  if (lds <= 160 * 1024) {
    *Wout = W;
    return pr ? plane_variant_launch<T, N, true>(c, a, c.plane_variant, npairs)
              : plane_variant_launch<T, N, false>(c, a, c.plane_variant, npairs);
  }
  const size_t lds2 = plane_split_lds_bytes<T, N, 2, kThreads, true>(g.M);
  if (lds2 > 160 * 1024) {
    set_error("plane buffer exceeds LDS for this grid in this precision");
    return SDFT_EINVAL_CODE;
  }
  const size_t need = w_bytes(c, npairs, prec_of<T>());
  if (c.W2_bytes < need) {
    if (c.W2) SDFT_HIP_CHECK(hipFree(c.W2));
    c.W2 = nullptr;
    c.W2_bytes = 0;
    if (hipMalloc(&c.W2, need) != hipSuccess) {
      set_error("device allocation failed (split-plane output)");
      return -3;
    }
    c.W2_bytes = need;
  }
  a.Wout = c.W2;
  SDFT_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_plane_split<T, N>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2));
  hipLaunchKernelGGL((k_plane_split<T, N>), grid, dim3(kThreads), lds2, c.stream, a);
  SDFT_HIP_CHECK(hipGetLastError());
  *Wout = c.W2;
  return 0;
}

int launch_plane(Ctx& c, int prec, void* W, int npairs, void** Wout) {
  if (prec == 0) {
    SDFT_SIZE_SWITCH(c.g.n, return (plane_impl<double, NN>(c, W, npairs, Wout)));
  } else {
    SDFT_SIZE_SWITCH(c.g.n, return (plane_impl<float, NN>(c, W, npairs, Wout)));
  }
  return 0;
}

template <typename T, int N>
static int contract_impl(Ctx& c, int mode, const void* W, int npairs, const double* Xcur,
                         double* Xout, double* acc, const void* fnl, const double* chi,
                         const StepParams& sp, double* mom_part, void* Wnext, float* Xout32) {
  ContractArgs a = make_contract_args(c, prec_of<T>(), mode, W, npairs, Xcur, Xout, acc, fnl,
                                      chi, sp, mom_part);
  a.Wnext = Wnext;
  a.Xout32 = Xout32;
  dim3 grid(a.gx, npairs);
  if (Wnext)
    hipLaunchKernelGGL((k_contract<T, N, true>), grid, dim3(kThreads), 0, c.stream, a);
  else
    hipLaunchKernelGGL((k_contract<T, N, false>), grid, dim3(kThreads), 0, c.stream, a);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_contract(Ctx& c, int prec, int mode, const void* W, int npairs, const double* Xcur,
                    double* Xout, double* acc, const void* fnl, const double* chi,
                    const StepParams& sp, double* mom_part, void* Wnext, float* Xout32) {
  if (Wnext && !c.g.col_intervals) {
    set_error("internal error: fused expand needs interval columns");
    return SDFT_EINVAL_CODE;
  }
  if (prec == 0) {
    SDFT_SIZE_SWITCH(c.g.n, return (contract_impl<double, NN>(c, mode, W, npairs, Xcur, Xout, acc,
                                                              fnl, chi, sp, mom_part, Wnext, Xout32)));
  } else {
    SDFT_SIZE_SWITCH(c.g.n, return (contract_impl<float, NN>(c, mode, W, npairs, Xcur, Xout, acc,
                                                             fnl, chi, sp, mom_part, Wnext, Xout32)));
  }
  return 0;
}

template <typename T, int N, bool RE>
static int density_impl(Ctx& c, const void* W, int npairs, double* rho_part, int ngroups,
                        const void* tw) {
  const Geometry& g = c.g;
  const size_t lds = plane_lds_bytes<T, N>(g.M);
  if (lds > 160 * 1024) {
    const size_t lds2 = plane_split_lds_bytes<T, N, 2>(g.M);
    const size_t lds4 = plane_split_lds_bytes<T, N, 4>(g.M);
    if (lds2 <= 160 * 1024) {
      SDFT_HIP_CHECK(hipFuncSetAttribute(
          reinterpret_cast<const void*>(&k_density_split<T, N, RE, 2>),
          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds2));
      hipLaunchKernelGGL((k_density_split<T, N, RE, 2>), dim3(N, ngroups), dim3(kThreads), lds2,
                         c.stream,
                         reinterpret_cast<const Cx<T>*>(W), g.ncolp, c.t.rowcol, g.M,
                         reinterpret_cast<const Cx<T>*>(tw), rho_part, npairs, ngroups);
    } else if (N % 4 == 0 && lds4 <= 160 * 1024) {
      SDFT_HIP_CHECK(hipFuncSetAttribute(
          reinterpret_cast<const void*>(&k_density_split<T, N, RE, 4>),
          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds4));
      hipLaunchKernelGGL((k_density_split<T, N, RE, 4>), dim3(N, ngroups), dim3(kThreads), lds4,
                         c.stream,
                         reinterpret_cast<const Cx<T>*>(W), g.ncolp, c.t.rowcol, g.M,
                         reinterpret_cast<const Cx<T>*>(tw), rho_part, npairs, ngroups);
    } else {
      set_error("plane buffer exceeds LDS for this grid in this precision");
      return SDFT_EINVAL_CODE;
    }
    SDFT_HIP_CHECK(hipGetLastError());
    return 0;
  }
  SDFT_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_density<T, N, RE>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  dim3 grid(N, ngroups);
  hipLaunchKernelGGL((k_density<T, N, RE>), grid, dim3(kThreads), lds, c.stream,
                     reinterpret_cast<const Cx<T>*>(W), g.ncolp, g.ncol, c.t.col_pos, g.M,
                     reinterpret_cast<const Cx<T>*>(tw), rho_part, npairs, ngroups);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_density(Ctx& c, const void* W, int npairs, double* rho_part, int ngroups, bool re) {
  if (re) {
    if (npairs != 1 || ngroups != 1) {
      set_error("internal error: real field output takes one pair");
      return SDFT_EINVAL_CODE;
    }
    SDFT_SIZE_SWITCH(c.g.n, return (density_impl<double, NN, true>(c, W, npairs, rho_part,
                                                                   ngroups, c.t.tw64)));
  }
  SDFT_SIZE_SWITCH(c.g.n, return (density_impl<double, NN, false>(c, W, npairs, rho_part,
                                                                  ngroups, c.t.tw64)));
  return 0;
}

template <int N>
static int noise_impl(Ctx& c, void* W, int npairs, unsigned long long seed, int first,
                      int nstate, const void* field) {
  const Geometry& g = c.g;
  // kc lines per chunk: all M when they fit, else the fewest chunks that do
  int mc = g.M;
  for (int parts = 1; noise_lds_bytes<N>(mc) > 160 * 1024; ++parts) {
    mc = (g.M + parts) / (parts + 1);
    if (mc < 1) {
      set_error("noise plane buffer exceeds LDS for this grid");
      return SDFT_EINVAL_CODE;
    }
  }
  const size_t lds = noise_lds_bytes<N>(mc);
  SDFT_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_noise_plane<N>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  dim3 grid(N, npairs);
  hipLaunchKernelGGL((k_noise_plane<N>), grid, dim3(kThreads), lds, c.stream,
                     reinterpret_cast<Cx<double>*>(W), g.ncolp, c.t.rowcol, g.M,
                     reinterpret_cast<const Cx<double>*>(c.t.tw64), seed, first, nstate,
                     reinterpret_cast<const Cx<double>*>(field), mc);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_noise_plane(Ctx& c, void* W, int npairs, unsigned long long seed, int first,
                       int nstate) {
  SDFT_SIZE_SWITCH(c.g.n, return (noise_impl<NN>(c, W, npairs, seed, first, nstate, nullptr)));
  return 0;
}

int launch_grid_plane(Ctx& c, void* W, int npairs, const void* field) {
  SDFT_SIZE_SWITCH(c.g.n, return (noise_impl<NN>(c, W, npairs, 0, 0, 2 * npairs, field)));
  return 0;
}

// field[(p / 2) N + index] .x (p even) / .y (p odd) = value for the entries of
// projectors p0 <= p < p1 (CSR offsets); the caller clears the field first
__global__ void k_scatter_proj(const long long* __restrict__ off, const int* __restrict__ idx,
                               const double* __restrict__ val, int p0, int p1, long long N,
                               double* __restrict__ field) {
  const long long e0 = off[p0], e1 = off[p1];
  for (long long e = e0 + (long long)blockIdx.x * blockDim.x + threadIdx.x; e < e1;
       e += (long long)gridDim.x * blockDim.x) {
    // projector of entry e: binary search over off[p0..p1]
    int lo = p0, hi = p1;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (off[mid] <= e)
        lo = mid;
      else
        hi = mid;
    }
    const int q = (lo - p0) >> 1;
    field[((size_t)q * N + idx[e]) * 2 + ((lo - p0) & 1)] = val[e];
  }
}

int launch_scatter_proj(Ctx& c, const long long* off, const int* idx, const double* val, int p0,
                        int p1, double* field) {
  hipLaunchKernelGGL(k_scatter_proj, dim3(1024), dim3(256), 0, c.stream, off, idx, val, p0, p1,
                     c.g.N, field);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_rho_finish(Ctx& c, const double* rho_part, int ngroups, double scale, int layout,
                      int accumulate, double* rho) {
  const long long Ntot = c.g.N;
  const int nb = (int)((Ntot + 255) / 256);
  hipLaunchKernelGGL(k_rho_finish, dim3(nb), dim3(256), 0, c.stream, rho_part, ngroups, Ntot,
                     c.g.n, scale, layout, accumulate, rho);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

template <typename T, int N>
static int fft_test_impl(int nlines, int dir, const double* in, double* out, const void* tw) {
  constexpr int ML = Cfg<N>::ML;
  const int nb = (nlines + ML - 1) / ML;
  if (dir > 0)
    hipLaunchKernelGGL((k_fft_test<T, N, 1>), dim3(nb), dim3(kThreads), 0, 0, in, out, nlines,
                       reinterpret_cast<const Cx<T>*>(tw));
  else
    hipLaunchKernelGGL((k_fft_test<T, N, -1>), dim3(nb), dim3(kThreads), 0, 0, in, out, nlines,
                       reinterpret_cast<const Cx<T>*>(tw));
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

int selftest_fft_dispatch(int n, int nlines, int dir, int prec, const double* in, double* out,
                          const void* tw) {
  if (prec == 0) {
    SDFT_SIZE_SWITCH(n, return (fft_test_impl<double, NN>(nlines, dir, in, out, tw)));
  } else {
    SDFT_SIZE_SWITCH(n, return (fft_test_impl<float, NN>(nlines, dir, in, out, tw)));
  }
  return 0;
}

}  // namespace sdft
