// sdft_gemm.hip -- Kleinman-Bylander projector contraction as MFMA GEMMs.
//
// The G-space KB path of the reference loops over G x atoms x (l,m) x states
// with scalar arithmetic (controlEwdNonlocFilter / getNlmatFilter,
// stodft/reduced-energy/cp-energy-eext-stodft.c:865-946,1105-1168, and
// getnl_fcoef, energy/cp/cp_energy_eext_nonloc.c:1669+).  With the projector
// table folded into A[p][2 ncp] (Gamma weights included) the same arithmetic
// is two dense products over the orbital block X (2 ncp x nstate):
//   Z  = A X                      (projections, split-K, reduced in FP64)
//   Y  = -2 diag(E) Z
//   F  = A^T Y                    (raw NL force, -4/-2 scaling included)
// Tiles: 64x64 per workgroup of 4 waves, each wave 2x2 MFMA 16x16 blocks,
// K staged through LDS 16 at a time.  FP64 form: v_mfma_f64_16x16x4f64;
// FP32 form: v_mfma_f32_16x16x4f32 with FP32 accumulation, results stored
// and reduced in FP64.
#include "sdft_internal.h"

namespace sdft {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename TI>
struct Mfma;

// Operand layout of both 16x16x4 instructions: lane l holds A[l%16][l/16]
// and B[l/16][l%16].  Accumulator register r of lane l holds column l%16 and
// row `row(l, r)`: rows 4(l/16)+r for the FP32 form, rows l/16+4r for FP64.
template <>
struct Mfma<double> {
  typedef d4_t acc_t;
  __device__ static inline acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

template <>
struct Mfma<float> {
  typedef f4_t acc_t;
  __device__ static inline acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

constexpr int BM = 64, BN = 64, BK = 16;

// C[m][n] (+ split offset) = alpha * sum_{k in split} A(m,k) B(k,n)
// A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]; one of sam/sak and one
// of sbk/sbn is 1 (the contiguous direction, used for coalesced tile loads).
template <typename TA, typename TB, typename TI>
__global__ __launch_bounds__(256) void k_gemm(int M, int Nn, int K, int kchunk,
                                              const TA* __restrict__ A, long sam, long sak,
                                              const TB* __restrict__ B, long sbk, long sbn,
                                              double* __restrict__ C, long scm, long scn,
                                              long c_split, double alpha) {
  __shared__ TI As[BK][BM + 1];
  __shared__ TI Bs[BK][BN + 1];
  typedef typename Mfma<TI>::acc_t acc_t;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.x * BM;
  const int n0 = blockIdx.y * BN;
  const int kbeg = blockIdx.z * kchunk;
  const int kend = min(K, kbeg + kchunk);
  acc_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = TI(0);

  const bool a_kfast = (sak == 1);
  const bool b_nfast = (sbn == 1);
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
#pragma unroll
    for (int i = 0; i < (BM * BK) / 256; ++i) {
      const int e = tid + i * 256;
      int m, k;
      if (a_kfast) {
        k = e % BK;
        m = e / BK;
      } else {
        m = e % BM;
        k = e / BM;
      }
      const int gm = m0 + m, gk = k0 + k;
      TI val = TI(0);
      if (gm < M && gk < kend) val = TI(A[(long)gm * sam + (long)gk * sak]);
      As[k][m] = val;
    }
#pragma unroll
    for (int i = 0; i < (BN * BK) / 256; ++i) {
      const int e = tid + i * 256;
      int n, k;
      if (b_nfast) {
        n = e % BN;
        k = e / BN;
      } else {
        k = e % BK;
        n = e / BK;
      }
      const int gn = n0 + n, gk = k0 + k;
      TI val = TI(0);
      if (gn < Nn && gk < kend) val = TI(B[(long)gk * sbk + (long)gn * sbn]);
      Bs[k][n] = val;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4);
      TI a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[kr][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[kr][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = Mfma<TI>::run(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  double* Cz = C + (long)blockIdx.z * c_split;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm * 32 + i * 16 + Mfma<TI>::row(lane, r);
        const int gn = n0 + wn * 32 + j * 16 + (lane & 15);
        if (gm < M && gn < Nn) Cz[(long)gm * scm + (long)gn * scn] = alpha * double(acc[i][j][r]);
      }
}

// Z[p][s] = sum_z part[z][p][s];  Y[p][s] = -2 E_p Z[p][s]
__global__ void k_nl_reduce(const double* __restrict__ part, int nsplit, int P, int ns,
                            const double* __restrict__ E, double* __restrict__ Z,
                            double* __restrict__ Y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P * ns) return;
  double s = 0.0;
  for (int z = 0; z < nsplit; ++z) s += part[(size_t)z * P * ns + i];
  const int p = i / ns;
  Z[i] = s;
  Y[i] = -2.0 * E[p] * s;
}

template <typename TA, typename TB, typename TI>
static int gemm(Ctx& c, int M, int Nn, int K, int nsplit, const TA* A, long sam, long sak,
                const TB* B, long sbk, long sbn, double* C, long scm, long scn, long c_split) {
  int kchunk = (K + nsplit - 1) / nsplit;
  kchunk = (kchunk + BK - 1) / BK * BK;
  nsplit = (K + kchunk - 1) / kchunk;
  dim3 grid((M + BM - 1) / BM, (Nn + BN - 1) / BN, nsplit);
  hipLaunchKernelGGL((k_gemm<TA, TB, TI>), grid, dim3(256), 0, c.stream, M, Nn, K, kchunk, A, sam,
                     sak, B, sbk, sbn, C, scm, scn, c_split, 1.0);
  SDFT_HIP_CHECK(hipGetLastError());
  return nsplit;
}

static int ensure(void** p, size_t* have, size_t need) {
  if (*have >= need) return 0;
  if (*p) SDFT_HIP_CHECK(hipFree(*p));
  *p = nullptr;
  *have = 0;
  hipError_t e = hipMalloc(p, need);
  if (e != hipSuccess) {
    set_error("device allocation failed (KB scratch)");
    return -3;
  }
  *have = need;
  return 0;
}

int launch_nl_project(Ctx& c, int prec, const double* X, int ns) {
  const int P = c.nproj;
  const int K = 2 * c.g.ncp;
  const int tiles = ((P + BM - 1) / BM) * ((ns + BN - 1) / BN);
  int nsplit = (512 + tiles - 1) / tiles;
  const int kmax = (K + 63) / 64;
  if (nsplit > kmax) nsplit = kmax;
  if (nsplit < 1) nsplit = 1;
  int rc = ensure(reinterpret_cast<void**>(&c.zpart), &c.zpart_bytes,
                  sizeof(double) * (size_t)nsplit * P * ns);
  if (rc) return rc;
  if (c.Y_bytes < sizeof(double) * (size_t)P * ns) {
    if (c.Y) SDFT_HIP_CHECK(hipFree(c.Y));
    if (c.Z) SDFT_HIP_CHECK(hipFree(c.Z));
    c.Y = c.Z = nullptr;
    c.Y_bytes = 0;
    if (hipMalloc(&c.Y, sizeof(double) * (size_t)P * ns) != hipSuccess ||
        hipMalloc(&c.Z, sizeof(double) * (size_t)P * ns) != hipSuccess) {
      set_error("device allocation failed (KB projections)");
      return -3;
    }
    c.Y_bytes = sizeof(double) * (size_t)P * ns;
  }
  int used;
  if (prec == 0)
    used = gemm<double, double, double>(c, P, ns, K, nsplit, c.A64, K, 1, X, 1, K, c.zpart, ns, 1,
                                        (long)P * ns);
  else
    used = gemm<float, double, float>(c, P, ns, K, nsplit, c.A32, K, 1, X, 1, K, c.zpart, ns, 1,
                                      (long)P * ns);
  if (used < 0) return used;
  const int tot = P * ns;
  hipLaunchKernelGGL(k_nl_reduce, dim3((tot + 255) / 256), dim3(256), 0, c.stream, c.zpart, used,
                     P, ns, c.E, c.Z, c.Y);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

int launch_nl_force(Ctx& c, int prec, int ns, double* fnl) {
  const int P = c.nproj;
  const int K2 = 2 * c.g.ncp;
  int used;
  if (prec == 0)
    used = gemm<double, double, double>(c, ns, K2, P, 1, c.Y, 1, ns, c.A64, K2, 1, fnl, K2, 1, 0);
  else
    used = gemm<double, float, float>(c, ns, K2, P, 1, c.Y, 1, ns, c.A32, K2, 1, fnl, K2, 1, 0);
  return used < 0 ? used : 0;
}

}  // namespace sdft
