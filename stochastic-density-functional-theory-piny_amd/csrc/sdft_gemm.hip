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
//
// Both products run on the 16x16x4 MFMA (v_mfma_f64_16x16x4f64 for the FP64
// form, v_mfma_f32_16x16x4f32 for the FP32 form).  The K order of an MFMA
// accumulation is free, so lane l (row or column l%16, group g = l/16) owns
// the 8 consecutive k = k0 + 8g .. k0 + 8g + 7 of a 32-wide K step and loads
// them with vector loads straight into registers: no LDS staging and no
// barriers inside the K loop.  Each wave accumulates a 32x32 tile (2x2 MFMA
// blocks).
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

constexpr int KS = 32;   // K per step: 4 lane groups x 8 consecutive k
constexpr int KPL = 8;   // k per lane per step

// 8 consecutive values src[0..7] in 16-byte vector loads (src is 32-byte
// aligned for float, 64-byte for double: row strides and run starts are
// multiples of 8 elements of 256-byte aligned buffers)
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static constexpr int n = 4;
};
template <>
struct Vec16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  static constexpr int n = 2;
};

template <typename T>
__device__ inline void load_run(const T* __restrict__ src, T (&dst)[KPL]) {
  typedef typename Vec16<T>::type V;
  constexpr int NV = Vec16<T>::n;
  const V* s = reinterpret_cast<const V*>(src);
#pragma unroll
  for (int v = 0; v < KPL / NV; ++v) {
    const V x = s[v];
#pragma unroll
    for (int e = 0; e < NV; ++e) dst[v * NV + e] = x[e];
  }
}

// ---- Z partials: zpart[z][p][s] = sum_{k in split z} A[p][k] X[s][k] -------
// Tiles (ceil(P/32), ceil(ns/32), nsplit); 4 waves per workgroup share the
// 32x32 output tile and interleave the split's K steps; the four partial
// tiles are summed through LDS.  Both operands must have their row count
// padded to a multiple of 32 with readable rows (zeros, or values whose
// products land only in the rows / columns the epilogue does not store):
// the K loop loads without bounds tests.
template <typename TA, typename TX, typename TI>
__device__ __forceinline__ void kb_project_tile(const TA* __restrict__ A,
                                                const TX* __restrict__ X, int P, int ns,
                                                int K, int ksteps_per_split,
                                                double* __restrict__ zpart, int bx, int by, int bz,
                                                double (*red)[32 * 32]) {
  typedef typename Mfma<TI>::acc_t acc_t;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: scalar K loop
  const int p0 = bx * 32, s0 = by * 32;
  const int g = lane >> 4, r16 = lane & 15;
  acc_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = TI(0);
  const int step0 = bz * ksteps_per_split;
  const int nsteps_total = K / KS;
  const int step1 = min(nsteps_total, step0 + ksteps_per_split);
  const TA* arow[2];
  const TX* xrow[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    arow[i] = A + (size_t)(p0 + 16 * i + r16) * K + KPL * g;
    xrow[i] = X + (size_t)(s0 + 16 * i + r16) * K + KPL * g;
  }
  // Operands are loaded in their storage types and converted to TI only at
  // the MFMAs.  The next step's loads are issued under a condition, and the
  // compiler's wait before the current step's MFMAs then covers them too, so
  // a wave does not overlap its loads with its MFMAs.  A counted, branch-free
  // form that does (correct waits, rings of 2-6 steps in flight) ran the
  // products 6 % faster alone at 2 steps and slower at more; beside the
  // plane kernel on the other stream it left the Si64 step unchanged and
  // the fragment batch 1 % slower.  These products are bound by how the
  // memory system serves their operand stream, shared with the plane kernel.
  TA ra0[2][KPL], ra1[2][KPL];
  TX rb0[2][KPL], rb1[2][KPL];
  auto load_raw = [&](int st, TA (&a)[2][KPL], TX (&b)[2][KPL]) {
    const int k0 = st * KS;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      load_run<TA>(arow[i] + k0, a[i]);
      load_run<TX>(xrow[i] + k0, b[i]);
    }
  };
  auto mma_raw = [&](TA (&a)[2][KPL], TX (&b)[2][KPL]) {
#pragma unroll
    for (int t = 0; t < KPL; ++t)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = Mfma<TI>::run(TI(a[i][t]), TI(b[j][t]), acc[i][j]);
  };
  int st = step0 + wave;
  if (st < step1) load_raw(st, ra0, rb0);
  while (st < step1) {
    const int nx = st + 4;
    if (nx < step1) load_raw(nx, ra1, rb1);
    mma_raw(ra0, rb0);
    st = nx;
    if (st >= step1) break;
    const int nx2 = st + 4;
    if (nx2 < step1) load_raw(nx2, ra0, rb0);
    mma_raw(ra1, rb1);
    st = nx2;
  }
  // sum the four waves' tiles
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = 16 * i + Mfma<TI>::row(lane, r);
        const int col = 16 * j + r16;
        red[wave][row * 32 + col] = double(acc[i][j][r]);
      }
  __syncthreads();
  double* zp = zpart + (size_t)bz * P * ns;
  for (int e = tid; e < 32 * 32; e += 256) {
    const int row = e >> 5, col = e & 31;
    const int p = p0 + row, s = s0 + col;
    if (p < P && s < ns) zp[(size_t)p * ns + s] = red[0][e] + red[1][e] + red[2][e] + red[3][e];
  }
}

// ---- F[s][k] = sum_p Y[s][p] At[k][p] ----------------------------------------
// The reduction (over the padded projector index, ppad) is short, so it is
// not split across waves: a workgroup covers 32 rows s x 64 NJ columns k,
// wave w owns columns 16 NJ w .. 16 NJ (w + 1) - 1 of that block (2 x NJ MFMA
// blocks) and runs the whole reduction (two-stage pipeline as above), and the
// accumulators are stored directly -- no LDS, no barrier.  K2 is a multiple
// of 128, so a wave's columns lie either all below K2 or all above (that
// wave returns at once).  NJ = 2 keeps the kernel's registers low enough for
// it to run beside the plane kernel on the other stream (NJ = 4 reuses each
// Y run twice as often but needs about twice the registers).  Y has its rows padded
// to a multiple of 32 with readable memory.  All operands and F are in the
// form's type T (float for FP32: the MFMA's own operand and accumulator type).
constexpr int kForceNJ = kForceCols / 64;

template <typename T>
__device__ __forceinline__ void kb_force_block(const T* __restrict__ Y, const T* __restrict__ At,
                                               int ns, int K2, int ppad, T* __restrict__ F,
                                               int bx, int by) {
  constexpr int NJ = kForceNJ;
  typedef typename Mfma<T>::acc_t acc_t;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int s0 = bx * 32, k0 = by * (64 * NJ) + wave * (16 * NJ);
  if (k0 >= K2) return;
  const int g = lane >> 4, r16 = lane & 15;
  acc_t acc[2][NJ];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = T(0);
  const T* arow[2];
  const T* xrow[NJ];
#pragma unroll
  for (int i = 0; i < 2; ++i) arow[i] = Y + (size_t)(s0 + 16 * i + r16) * ppad + KPL * g;
#pragma unroll
  for (int j = 0; j < NJ; ++j) xrow[j] = At + (size_t)(k0 + 16 * j + r16) * ppad + KPL * g;
  T ra0[2][KPL], ra1[2][KPL];
  T rb0[NJ][KPL], rb1[NJ][KPL];
  auto load_raw = [&](int st, T (&a)[2][KPL], T (&b)[NJ][KPL]) {
    const int kk = st * KS;
#pragma unroll
    for (int i = 0; i < 2; ++i) load_run<T>(arow[i] + kk, a[i]);
#pragma unroll
    for (int j = 0; j < NJ; ++j) load_run<T>(xrow[j] + kk, b[j]);
  };
  auto mma_raw = [&](T (&a)[2][KPL], T (&b)[NJ][KPL]) {
#pragma unroll
    for (int t = 0; t < KPL; ++t)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) acc[i][j] = Mfma<T>::run(a[i][t], b[j][t], acc[i][j]);
  };
  const int nsteps = ppad / KS;
  if constexpr (sizeof(T) == 4) {
    int st = 0;
    load_raw(0, ra0, rb0);
    while (st < nsteps) {
      if (st + 1 < nsteps) load_raw(st + 1, ra1, rb1);
      mma_raw(ra0, rb0);
      if (++st >= nsteps) break;
      if (st + 1 < nsteps) load_raw(st + 1, ra0, rb0);
      mma_raw(ra1, rb1);
      ++st;
    }
  } else {
    // FP64: one operand set in flight (64 fewer VGPRs: 4 waves per SIMD
    // instead of 2 hide the load latency)
    for (int st = 0; st < nsteps; ++st) {
      load_raw(st, ra0, rb0);
      mma_raw(ra0, rb0);
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int s = s0 + 16 * i + Mfma<T>::row(lane, r);
      if (s < ns) {
#pragma unroll
        for (int j = 0; j < NJ; ++j) F[(size_t)s * K2 + k0 + 16 * j + r16] = acc[i][j][r];
      }
    }
}

// 1-D launch over ts * ceil(K2 / kForceCols) blocks, XCD-swizzled, s fastest:
// the blocks sharing one slab of At sit on one XCD.
template <typename T>
__global__ __launch_bounds__(256) void k_kb_force(const T* __restrict__ Y, const T* __restrict__ At,
                                                  int ns, int K2, int ppad, T* __restrict__ F,
                                                  int ts, int total) {
  const int L = xcd_swizzle(blockIdx.x, total);
  kb_force_block<T>(Y, At, ns, K2, ppad, F, L % ts, L / ts);
}

// 1-D launch over tp * ts * nsplit tiles.  Logical tile L (XCD-swizzled) is
// split-major, then column tile, then row tile: the tiles of one K slice sit
// on one XCD, whose L2 then holds that slice of both operands.
template <typename TA, typename TX, typename TI>
__global__ __launch_bounds__(256) void k_kb_project(const TA* __restrict__ A,
                                                    const TX* __restrict__ X, int P, int ns,
                                                    int K, int ksteps_per_split,
                                                    double* __restrict__ zpart, int tp, int ts,
                                                    int total) {
  __shared__ double red[4][32 * 32];
  const int L = xcd_swizzle(blockIdx.x, total);
  const int bx = L % tp, r = L / tp;
  kb_project_tile<TA, TX, TI>(A, X, P, ns, K, ksteps_per_split, zpart, bx, r % ts, r / ts, red);
}

// One launch over the fragments of a batch: workgroup w belongs to fragment
// f with prefix[f] <= w < prefix[f+1]; its tile is (local % tp, local / tp %
// ts, local / (tp ts)).
template <typename TA, typename TI>
__global__ __launch_bounds__(256) void k_kb_project_batch(const KbProjectArgs* __restrict__ args,
                                                          const int* __restrict__ prefix,
                                                          int nfrag) {
  __shared__ double red[4][32 * 32];
  const int L = xcd_swizzle(blockIdx.x, prefix[nfrag]);
  const int f = find_frag(prefix, nfrag, L);
  const KbProjectArgs a = args[f];
  const int local = L - prefix[f];
  const int bx = local % a.tp, r = local / a.tp;
  kb_project_tile<TA, TA, TI>(reinterpret_cast<const TA*>(a.A), reinterpret_cast<const TA*>(a.X),
                              a.P, a.ns, a.K, a.per, a.zpart, bx, r % a.ts, r / a.ts, red);
}

// Z[p][s] = sum_z part[z][p][s];  Y[s][p] = -2 E_p Z[p][s] (transposed, row
// stride ppad, in the form's type TY; the padding columns p >= P are never
// written and stay zero)
template <typename TY>
__device__ __forceinline__ void nl_reduce_elem(const double* __restrict__ part, int nsplit, int P,
                                               int ns, int ppad, const double* __restrict__ E,
                                               double* __restrict__ Z, TY* __restrict__ Y, int i) {
  if (i >= P * ns) return;
  double s = 0.0;
  for (int z = 0; z < nsplit; ++z) s += part[(size_t)z * P * ns + i];
  const int p = i / ns, st = i - p * ns;
  Z[i] = s;
  if (Y) Y[(size_t)st * ppad + p] = TY(-2.0 * E[p] * s);
}

template <typename TY>
__global__ void k_nl_reduce(const double* __restrict__ part, int nsplit, int P, int ns, int ppad,
                            const double* __restrict__ E, double* __restrict__ Z,
                            TY* __restrict__ Y) {
  nl_reduce_elem<TY>(part, nsplit, P, ns, ppad, E, Z, Y, blockIdx.x * blockDim.x + threadIdx.x);
}

template <typename TY>
__global__ __launch_bounds__(256) void k_nl_reduce_batch(const KbReduceArgs* __restrict__ args,
                                                         const int* __restrict__ prefix,
                                                         int nfrag) {
  const int f = find_frag(prefix, nfrag, blockIdx.x);
  const KbReduceArgs a = args[f];
  nl_reduce_elem<TY>(a.zpart, a.nsplit, a.P, a.ns, a.ppad, a.E, a.Z, reinterpret_cast<TY*>(a.Y),
                     (blockIdx.x - prefix[f]) * 256 + threadIdx.x);
}

// ---- batched F[s][k] = sum_p Y[s][p] At[k][p] ------------------------------
// kb_force_block per fragment: ts x tk blocks of 32 states x kForceCols columns.
template <typename T>
__global__ __launch_bounds__(256) void k_kb_force_batch(const KbForceArgs* __restrict__ args,
                                                        const int* __restrict__ prefix,
                                                        int nfrag) {
  const int L = xcd_swizzle(blockIdx.x, prefix[nfrag]);
  const int f = find_frag(prefix, nfrag, L);
  const KbForceArgs a = args[f];
  const int local = L - prefix[f];
  kb_force_block<T>(reinterpret_cast<const T*>(a.Y), reinterpret_cast<const T*>(a.At), a.ns, a.K2,
                    a.ppad, reinterpret_cast<T*>(a.F), local % a.ts, local / a.ts);
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

// K split of the projection GEMM: about c.kb_wg_target workgroups (512 unless
// tuned), at least 16 K steps (four per wave) per split.  Returns the split
// count, *per = steps per split.
int kb_project_split(const Ctx& c, int ns, int* per) {
  const int P = c.nproj;
  const int K = 2 * c.g.ncp;  // multiple of 128
  const int tiles = ((P + 31) / 32) * ((ns + 31) / 32);
  const int nsteps = K / KS;
  int nsplit = (c.kb_wg_target + tiles - 1) / tiles;
  if (nsplit > nsteps / 16) nsplit = nsteps / 16;
  if (nsplit < 1) nsplit = 1;
  *per = (nsteps + nsplit - 1) / nsplit;
  return (nsteps + *per - 1) / *per;
}

int launch_nl_project(Ctx& c, int prec, const void* X, int ns) {
  const int P = c.nproj;
  const int K = 2 * c.g.ncp;
  int per = 0;
  const int nsplit = kb_project_split(c, ns, &per);
  int rc = ensure(reinterpret_cast<void**>(&c.zpart), &c.zpart_bytes,
                  sizeof(double) * (size_t)nsplit * P * ns);
  if (rc) return rc;
  // Y is [ns][ppad] with zero padding columns: (re)allocated and cleared
  // whenever it grows or the padded projector count changes.  Z ([P][ns],
  // written below for every state) is allocated with the same padded row
  // count, so it holds P * ns doubles whenever Y needs no reallocation.
  const int ns_pad = (ns + 31) / 32 * 32;
  const size_t ybytes = sizeof(double) * (size_t)c.ppad * ns_pad;
  if (c.Y_bytes < ybytes || c.Y_ppad != c.ppad) {
    if (c.Y) SDFT_HIP_CHECK(hipFree(c.Y));
    if (c.Y32) SDFT_HIP_CHECK(hipFree(c.Y32));
    if (c.Z) SDFT_HIP_CHECK(hipFree(c.Z));
    c.Y = c.Z = nullptr;
    c.Y32 = nullptr;
    c.Y_bytes = 0;
    if (hipMalloc(&c.Y, ybytes) != hipSuccess || hipMalloc(&c.Y32, ybytes / 2) != hipSuccess ||
        hipMalloc(&c.Z, sizeof(double) * (size_t)c.ppad * ns_pad) != hipSuccess) {
      set_error("device allocation failed (KB projections)");
      return -3;
    }
    SDFT_HIP_CHECK(hipMemsetAsync(c.Y, 0, ybytes, c.stream));
    SDFT_HIP_CHECK(hipMemsetAsync(c.Y32, 0, ybytes / 2, c.stream));
    c.Y_bytes = ybytes;
    c.Y_ppad = c.ppad;
  }
  const int tp = (P + 31) / 32, ts = (ns + 31) / 32, tot = tp * ts * nsplit;
  const int nel = P * ns;
  if (prec == 0) {
    hipLaunchKernelGGL((k_kb_project<double, double, double>), dim3(tot), dim3(256), 0, c.stream,
                       c.A64, static_cast<const double*>(X), P, ns, K, per, c.zpart, tp, ts, tot);
    hipLaunchKernelGGL(k_nl_reduce<double>, dim3((nel + 255) / 256), dim3(256), 0, c.stream,
                       c.zpart, nsplit, P, ns, c.ppad, c.E, c.Z, c.Y);
  } else {
    hipLaunchKernelGGL((k_kb_project<float, float, float>), dim3(tot), dim3(256), 0, c.stream,
                       c.A32, static_cast<const float*>(X), P, ns, K, per, c.zpart, tp, ts, tot);
    hipLaunchKernelGGL(k_nl_reduce<float>, dim3((nel + 255) / 256), dim3(256), 0, c.stream,
                       c.zpart, nsplit, P, ns, c.ppad, c.E, c.Z, c.Y32);
  }
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

// Ad[p][j] = G_d(j) A[p][ncp + j], Ad[p][ncp + j] = -G_d(j) A[p][j]: the table
// of dQ_p/dR_a = i G Q_p in the same Gamma packing (G = 0 gives zero).
__global__ void k_kb_deriv_table(const double* __restrict__ A, const double* __restrict__ gd,
                                 int P, int ncp, double* __restrict__ Ad) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)P * ncp) return;
  const size_t p = i / ncp, j = i - p * ncp;
  const double g = gd[j];
  Ad[p * 2 * ncp + j] = g * A[p * 2 * ncp + ncp + j];
  Ad[p * 2 * ncp + ncp + j] = -g * A[p * 2 * ncp + j];
}

// S[p*3 + d] = sum_{s < nstate} Z[p][s] Zd[d][p][s]
__global__ void k_force_sums(const double* __restrict__ Z, const double* __restrict__ Zd, int P,
                             int ns, int nstate, double* __restrict__ S) {
  __shared__ double sh[3][256];
  const int p = blockIdx.x;
  double a[3] = {0, 0, 0};
  for (int s = threadIdx.x; s < nstate; s += blockDim.x) {
    const double z = Z[(size_t)p * ns + s];
#pragma unroll
    for (int d = 0; d < 3; ++d) a[d] += z * Zd[((size_t)d * P + p) * ns + s];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) sh[d][threadIdx.x] = a[d];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int d = 0; d < 3; ++d) sh[d][threadIdx.x] += sh[d][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int d = 0; d < 3; ++d) S[p * 3 + d] = sh[d][0];
}

// dZ/dR for the three directions (FP64) and the per-projector sums; gint is
// G in internal coefficient order [3][ncp].
int nl_force_sums(Ctx& c, const double* X, int ns, int nstate, const double* gint,
                  double* Ad, double* Zd, double* part, double* S) {
  const int P = c.nproj;
  const int K = 2 * c.g.ncp;
  int rc = launch_nl_project(c, 0, X, ns);  // c.Z = A X
  if (rc) return rc;
  int per = 0;
  const int nsplit = kb_project_split(c, ns, &per);
  for (int d = 0; d < 3; ++d) {
    const size_t tot = (size_t)P * c.g.ncp;
    hipLaunchKernelGGL(k_kb_deriv_table, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       c.stream, c.A64, gint + (size_t)d * c.g.ncp, P, c.g.ncp, Ad);
    const int tp = (P + 31) / 32, ts = (ns + 31) / 32, ntile = tp * ts * nsplit;
    hipLaunchKernelGGL((k_kb_project<double, double, double>), dim3(ntile), dim3(256), 0,
                       c.stream, Ad, X, P, ns, K, per, part, tp, ts, ntile);
    // reduce the split partials into Zd[d] (the E factor is applied later)
    const int n2 = P * ns;
    hipLaunchKernelGGL(k_nl_reduce<double>, dim3((n2 + 255) / 256), dim3(256), 0, c.stream, part,
                       nsplit, P, ns, c.ppad, c.E, Zd + (size_t)d * n2, (double*)nullptr);
    SDFT_HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_force_sums, dim3(P), dim3(256), 0, c.stream, c.Z, Zd, P, ns, nstate, S);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

// F[s][k] = sum_p Y[s][p] At[k][p] (K2 = 2 ncp is a multiple of 128), F in
// the form's type
int launch_nl_force(Ctx& c, int prec, int ns, void* fnl) {
  const int K2 = 2 * c.g.ncp;
  const int ts = (ns + 31) / 32, tot = ts * ((K2 + kForceCols - 1) / kForceCols);
  if (prec == 0)
    hipLaunchKernelGGL((k_kb_force<double>), dim3(tot), dim3(256), 0, c.stream, c.Y, c.At64, ns,
                       K2, c.ppad, static_cast<double*>(fnl), ts, tot);
  else
    hipLaunchKernelGGL((k_kb_force<float>), dim3(tot), dim3(256), 0, c.stream, c.Y32, c.At32, ns,
                       K2, c.ppad, static_cast<float*>(fnl), ts, tot);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

// kind 0: projection partials, 1: split reduction + Y, 2: force product
int launch_kb_batch(int kind, int prec, const void* args, const int* prefix, int nfrag,
                    int total, hipStream_t s) {
  if (total <= 0) return 0;
  if (kind == 0) {
    if (prec == 0)
      hipLaunchKernelGGL((k_kb_project_batch<double, double>), dim3(total), dim3(256), 0, s,
                         reinterpret_cast<const KbProjectArgs*>(args), prefix, nfrag);
    else
      hipLaunchKernelGGL((k_kb_project_batch<float, float>), dim3(total), dim3(256), 0, s,
                         reinterpret_cast<const KbProjectArgs*>(args), prefix, nfrag);
  } else if (kind == 1) {
    if (prec == 0)
      hipLaunchKernelGGL(k_nl_reduce_batch<double>, dim3(total), dim3(256), 0, s,
                         reinterpret_cast<const KbReduceArgs*>(args), prefix, nfrag);
    else
      hipLaunchKernelGGL(k_nl_reduce_batch<float>, dim3(total), dim3(256), 0, s,
                         reinterpret_cast<const KbReduceArgs*>(args), prefix, nfrag);
  } else {
    if (prec == 0)
      hipLaunchKernelGGL((k_kb_force_batch<double>), dim3(total), dim3(256), 0, s,
                         reinterpret_cast<const KbForceArgs*>(args), prefix, nfrag);
    else
      hipLaunchKernelGGL((k_kb_force_batch<float>), dim3(total), dim3(256), 0, s,
                         reinterpret_cast<const KbForceArgs*>(args), prefix, nfrag);
  }
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace sdft
