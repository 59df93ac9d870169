// sdft_plane.h -- per-size FFT configuration and the a-plane stage of the
// G->r->G sandwich (c and b transforms of one (pair, a-plane)), shared by the
// fused plane kernels (sdft_kernels.hip) and the real-space projector path
// (sdft_rs.hip), which runs the same stage split at the real-space grid.
#pragma once
#include "fft_lds.h"
#include "sdft_internal.h"

namespace sdft {

// ---- per-size configuration ----------------------------------------------
constexpr int largest_divisor_le(int n, int cap) {
  for (int d = cap; d >= 1; --d)
    if (n % d == 0) return d;
  return 1;
}

template <int N>
struct Cfg {
  static constexpr int ML = fft_max_lines(N, kThreads);  // lines per FFT call
  static constexpr int CH = ML < 32 ? ML : 32;           // columns per chunk
  static constexpr int CB = largest_divisor_le(N, ML < 16 ? ML : 16);
  static constexpr int LD = N + 1;                       // padded LDS line stride
};


// Thread mapping of the passes over lines that sit next to each other in
// LDS (c rows of a plane, a columns of a chunk): with consecutive threads on
// consecutive butterflies of one line, a first pass stores with a stride of R
// elements, which for 16-byte (double) elements puts a line's butterflies in
// one LDS bank; consecutive threads on consecutive lines (stride n + 1)
// avoid that.  For the plane's c rows in float the line-major order keeps
// the global gathers of the rows denser and measured faster (kLineFast);
// the a columns of the expand / contract chunks are transformed in LDS only,
// and take the line-fast order in both precisions (kLineFastA).
template <typename T>
constexpr bool kLineFast = sizeof(T) == 8;
template <typename T>
constexpr bool kLineFastA = true;

template <typename T, int N>
size_t plane_lds_bytes(int M) {
  return sizeof(Cx<T>) * ((size_t)M * Cfg<N>::LD + (size_t)Cfg<N>::CB * Cfg<N>::LD + N);
}

__device__ inline int wrapi(int k, int n) { return k < 0 ? k + n : k; }

// ---- K2: one a-plane: c and b transforms around v.psi --------------------
// rowcol[row * N + pos] is the sphere column whose (kb, kc) is (row - K, pos)
// (pos = kc mod N), or -1: the c-lines' first pass reads W through it and
// their last pass writes W through it, so the plane is never zero-filled or
// scattered.  The b-lines read and write only the 2K+1 rows |kb| <= K of the
// c-transformed plane; their inverse last pass, the v_KS multiply and the
// forward first pass run fused in registers.
// Tile variants of the b-line sub-chunk CB (columns rc transformed together):
// V = 0 the base width (one radix butterfly per thread and pass at n = 64),
// V = 1 twice it (half the sub-chunks and barriers, two butterflies per
// thread); the tuner picks per context or batch.  PR = true (for 4K < n, as
// on every density grid of the reference, n >= 4K + 1) prunes the first pass
// of the two inverse transforms (c and b: inputs with |k| >= n/4 are zero)
// and the last pass of the two forward ones (outputs with |k| >= n/4 are
// not kept); the c-forward runs in reversed radix order so that its last
// pass is the radix-4/8 one.
template <typename T, int N, int V = 0>
struct PlaneCfg {
  static constexpr int ML = Cfg<N>::ML;
  static constexpr int CB0 = largest_divisor_le(N, ML < (sizeof(T) == 4 ? 32 : 16)
                                                       ? ML
                                                       : (sizeof(T) == 4 ? 32 : 16));
  static constexpr int CB = V == 0 ? CB0 : largest_divisor_le(N, 2 * CB0 < N ? 2 * CB0 : N);
};

// Padded slot layout (lds_slot<true>: one unused slot after every 8
// elements, line stride n + ceil(n / 8)).  Lines in it run line-major
// (consecutive threads on the butterflies of one line: the W gathers and
// scatters of a wave touch a few rows' consecutive sphere columns), and the
// radix-8 first pass stores with a stride of 8 elements, which unpadded puts
// the n / 8 butterflies of a line on 4 bank pairs and padded spreads over
// distinct banks, for 8- and 16-byte elements.  The c rows of the plane
// buffer use it at n = 64 (measured: FP32 74.7 -> 70.9 us, FP64 190.8 ->
// 183.5 us per Si64 plane launch); at 128 its larger rows cost a workgroup
// per CU, and at the fragments' 24-48 the unpadded rows (FP64 line-fast)
// measured faster.  The split form's c-lines in Q use it at every size.
constexpr int padded_ld(int n) { return n + (n + 7) / 8; }
template <typename T, int N>
struct PlaneRows {
  static constexpr bool PAD = N == 64;
  static constexpr int LDC = PAD ? padded_ld(N) : Cfg<N>::LD;
  static constexpr bool LINE_FAST = PAD ? false : kLineFast<T>;
};

template <typename T, int N, int V = 0>
size_t plane2_lds_bytes(int M) {
  return sizeof(Cx<T>) * ((size_t)M * PlaneRows<T, N>::LDC +
                          (size_t)PlaneCfg<T, N, V>::CB * Cfg<N>::LD + N);
}

// MODE 0: the fused stage (inverse c, inverse b, v.psi, forward b, forward
// c).  MODE 1 stops at real space: the inverse b-lines are stored to the pair
// grid a.field[q][ra][rb][rc] (FFTW order).  MODE 2 starts there: the forward
// b-lines are read from a.field (which already holds v.psi plus the real-space
// KB term) and the stage continues as MODE 0.
// SC / SB: the radix schedules (fft_lds.h) of the c-lines and the b-lines.
template <typename T, int N, int V, bool PR, int MODE = 0, int NT = kThreads, int SC = 0,
          int SB = 0>
__device__ __forceinline__ void plane_body(const PlaneArgs& a, int ra, int q) {
  constexpr int LD = Cfg<N>::LD, ML = fft_max_lines(N, NT, SC), CB = PlaneCfg<T, N, V>::CB;
  constexpr bool PAD = PlaneRows<T, N>::PAD;
  constexpr int LDC = PlaneRows<T, N>::LDC;
  constexpr int PC = n_passes(N, SC), P = n_passes(N, SB);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cx<T>* __restrict__ W = reinterpret_cast<Cx<T>*>(a.W);
  const int ncolp = a.ncolp, M = a.M;
  const int* __restrict__ rowcol = a.rowcol;
  const T* __restrict__ v = reinterpret_cast<const T*>(a.v);
  const Cx<T>* __restrict__ twg = reinterpret_cast<const Cx<T>*>(a.tw);
  Cx<T>* Pl = reinterpret_cast<Cx<T>*>(smem_raw);  // M rows, stride LDC
  Cx<T>* Q = Pl + M * LDC;
  Cx<T>* tw = Q + CB * LD;
  const int tid = threadIdx.x;
  const int K = (M - 1) / 2;
  for (int i = tid; i < N; i += NT) tw[i] = twg[i];
  __syncthreads();
  Cx<T>* Wp = W + ((size_t)q * N + ra) * ncolp;
  const Cx<T> zero = cx(T(0), T(0));

  Cx<T>* __restrict__ F = reinterpret_cast<Cx<T>*>(a.field) + ((size_t)q * N + ra) * N * N;

  // inverse along c: rows of the plane, W -> LDS
  if constexpr (MODE != 2)
  for (int r0 = 0; r0 < M; r0 += ML) {
    const int nl = M - r0 < ML ? M - r0 : ML;
    const int* rc = rowcol + (size_t)r0 * N;
    // loads are unconditional (a column outside the sphere reads entry 0)
    // and masked by a multiply afterwards: one vector load per element, all
    // of a butterfly's index and data loads in flight together, no
    // divergent branch (a select would be turned back into a guarded load)
    auto lw = [&](int l, int pos) {
      const int d = rc[l * N + pos];
      const Cx<T> x = Wp[d >= 0 ? d : 0];
      const T m = d >= 0 ? T(1) : T(0);
      return cx(x.x * m, x.y * m);
    };
    Cx<T>* buf = Pl + r0 * LDC;
    auto sb = [&](int l, int pos, Cx<T> x) { buf[l * LDC + lds_slot<PAD>(pos)] = x; };
    // the rows are not read before this pass (W in, rows out: no barrier
    // between the first pass's loads and stores)
    fft_passes_io<T, N, +1, false, 0, PC - 1, ML, NT, PlaneRows<T, N>::LINE_FAST, PR, false,
                  PAD, SC, true, false>(buf, LDC, nl, tw, tid, lw, sb);
  }

  // along b, CB columns rc at a time
  const T* vplane = v + (size_t)ra * N * N;
  for (int rc0 = 0; rc0 < N; rc0 += CB) {
    auto lp = [&](int l, int pos) {
      const int kb = pos <= N / 2 ? pos : pos - N;
      const bool in = kb >= -K && kb <= K;
      const Cx<T> x = Pl[(in ? kb + K : 0) * LDC + lds_slot<PAD>(rc0 + l)];
      return in ? x : zero;
    };
    auto lq = [&](int l, int pos) { return Q[l * LD + pos]; };
    auto sq = [&](int l, int pos, Cx<T> x) { Q[l * LD + pos] = x; };
    auto vf = [&](int l, int pos) { return vplane[pos * N + rc0 + l]; };
    auto sp = [&](int l, int pos, Cx<T> x) {
      const int kb = pos <= N / 2 ? pos : pos - N;
      if (kb >= -K && kb <= K) Pl[(kb + K) * LDC + lds_slot<PAD>(rc0 + l)] = x;
    };
    if constexpr (MODE == 1) {
      // inverse b-lines straight to the real-space grid (consecutive threads
      // on consecutive lines: runs of CB consecutive rc)
      auto sf = [&](int l, int pos, Cx<T> x) { F[(size_t)pos * N + rc0 + l] = x; };
      fft_passes_io<T, N, +1, false, 0, P - 1, CB, NT, true, PR, false, false, SB, true, true>(
          Q, LD, CB, tw, tid, lp, sf);
    } else if constexpr (MODE == 2) {
      auto lf = [&](int l, int pos) { return F[(size_t)pos * N + rc0 + l]; };
      fft_passes_io<T, N, -1, true, 0, P - 1, CB, NT, true, false, PR, false, SB, true, true>(
          Q, LD, CB, tw, tid, lf, sp);
    } else {
      // Q is free here (the previous sub-chunk's last pass read it before
      // its closing barrier) and the plane rows are read-only until the
      // last forward pass: the first and last passes skip their middle
      // barrier
      if constexpr (P >= 2)
        fft_passes_io<T, N, +1, false, 0, P - 2, CB, NT, true, PR, false, false, SB, true,
                      false>(Q, LD, CB, tw, tid, lp, sq);
      fft_fused_vmul<T, N, CB, NT, true, SB>(CB, tw, tid, lq, vf, sq);
      if constexpr (P >= 2)
        fft_passes_io<T, N, -1, true, 1, P - 1, CB, NT, true, false, PR, false, SB, false,
                      true>(Q, LD, CB, tw, tid, lq, sp);
    }
  }

  // forward along c: LDS -> W
  if constexpr (MODE != 1)
  for (int r0 = 0; r0 < M; r0 += ML) {
    const int nl = M - r0 < ML ? M - r0 : ML;
    const int* rc = rowcol + (size_t)r0 * N;
    Cx<T>* buf = Pl + r0 * LDC;
    auto lb = [&](int l, int pos) { return buf[l * LDC + lds_slot<PAD>(pos)]; };
    auto sw = [&](int l, int pos, Cx<T> x) {
      const int d = rc[l * N + pos];
      if (d >= 0) Wp[d] = x;
    };
    fft_passes_io<T, N, -1, PR, 0, PC - 1, ML, NT, PlaneRows<T, N>::LINE_FAST, false, PR,
                  PAD, SC, false, true>(buf, LDC, nl, tw, tid, lb, sw);
  }
}

// ---- K2, split form: for planes whose M x (n+1) buffer exceeds LDS --------
// The plane is processed in two column halves (rc in [h n/2, (h+1) n/2)).
// Per half the inverse c-transform of the M rows is recomputed from W (ML
// rows at a time through Q), only the half's columns are kept in LDS
// (M x (n/2+1)), their b-lines run as in plane_body, and the forward
// c-transform of the rows with the other half's entries zero is stored
// (h = 0) or added (h = 1) to Wout: the forward transform is linear in its
// input, and W (the input of both halves) is left intact.
// S column parts of width H = n / S (the plane kernels use halves; the
// density kernel also quarters, for the density sphere's rows at 192^3)
// V: the b-line tile (PlaneCfg); the 512-thread split kernels take the
// doubled tile, whose lines fit in the rows the padded c-lines reserve.
template <typename T, int N, int S = 2, int NT = kThreads, int V = 0>
struct SplitCfg {
  static constexpr int H = N / S, LDH = H + 1;
  static constexpr int ML = fft_max_lines(N, NT);
  static constexpr int CB = PlaneCfg<T, N, V>::CB;
  static constexpr int QL = ML > CB ? ML : CB;
  // elements of Q: its QL lines of stride n + 1, or (plane split kernels)
  // the c-lines' ML rows in the padded slot layout, whichever is larger
  static constexpr int QE = QL * Cfg<N>::LD;
  static constexpr int QEP = ML * padded_ld(N) > QE ? ML * padded_ld(N) : QE;
};

// PADQ: the plane split kernels, whose c-lines run padded in Q (the density
// split kernel keeps stride n + 1)
template <typename T, int N, int S = 2, int NT = kThreads, bool PADQ = false, int V = 0>
size_t plane_split_lds_bytes(int M) {
  using C = SplitCfg<T, N, S, NT, V>;
  return sizeof(Cx<T>) * ((size_t)M * C::LDH + (size_t)(PADQ ? C::QEP : C::QE) + N);
}

// MODE 1 / 2 as in plane_body: per half, the inverse b-lines go to the
// real-space grid a.field (no forward part), or the forward b-lines come from
// it (no inverse part; W is not read, the output goes to Wout).
template <typename T, int N, int MODE = 0, int NT = kThreads, int V = 0>
__device__ __forceinline__ void plane_split_body(const PlaneArgs& a, int ra, int q) {
  using C = SplitCfg<T, N, 2, NT, V>;
  constexpr int LD = Cfg<N>::LD, ML = C::ML, CB = C::CB;
  constexpr int H = C::H, LDH = C::LDH;
  constexpr int P = n_passes(N);
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const Cx<T>* __restrict__ W = reinterpret_cast<const Cx<T>*>(a.W);
  Cx<T>* __restrict__ Wo = reinterpret_cast<Cx<T>*>(a.Wout);
  const int ncolp = a.ncolp, M = a.M;
  const int* __restrict__ rowcol = a.rowcol;
  const T* __restrict__ v = reinterpret_cast<const T*>(a.v);
  const Cx<T>* __restrict__ twg = reinterpret_cast<const Cx<T>*>(a.tw);
  Cx<T>* Ph = reinterpret_cast<Cx<T>*>(smem_raw);  // M rows x the half's columns
  Cx<T>* Q = Ph + M * LDH;  // QL b-lines of stride n + 1 / ML padded c-lines
  Cx<T>* tw = Q + C::QEP;
  constexpr int LDP = padded_ld(N);
  const int tid = threadIdx.x;
  const int K = (M - 1) / 2;
  for (int i = tid; i < N; i += NT) tw[i] = twg[i];
  __syncthreads();
  const Cx<T>* Wp = W + ((size_t)q * N + ra) * ncolp;
  Cx<T>* Wq = Wo + ((size_t)q * N + ra) * ncolp;
  const T* vplane = v + (size_t)ra * N * N;
  const Cx<T> zero = cx(T(0), T(0));
  Cx<T>* __restrict__ F = reinterpret_cast<Cx<T>*>(a.field) + ((size_t)q * N + ra) * N * N;
  for (int h = 0; h < 2; ++h) {
    const int c_lo = h * H;
    // inverse along c, keeping this half's columns
    if constexpr (MODE != 2)
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
      // W in, the half's rows out: Q is free (closing barrier of its last
      // use) and the rows are not read before the b-lines
      fft_passes_io<T, N, +1, false, 0, P - 1, ML, NT, false, false, false, true, 0, true,
                    true>(Q, LDP, nl, tw, tid, lw, sh);
    }
    // along b, CB columns at a time.  MODE 2: the forward b-lines start from
    // the grid F, and the first pass's inputs of the next sub-chunk are
    // loaded while this sub-chunk's other passes run (they stay in flight
    // across the LDS barriers)
    using P0 = PassShape<T, N, 0, true, CB, NT>;
    typename P0::Regs fpre;
    int rcl = 0;
    auto lfp = [&](int l, int pos) { return F[(size_t)pos * N + c_lo + rcl + l]; };
    if constexpr (MODE == 2)
      fft_pass_load<T, N, 0, true, CB, NT, true>(H < CB ? H : CB, tid, lfp, fpre);
    for (int rc0 = 0; rc0 < H; rc0 += CB) {
      const int nb = H - rc0 < CB ? H - rc0 : CB;
      auto lp = [&](int l, int pos) {
        const int kb = pos <= N / 2 ? pos : pos - N;
        return (kb >= -K && kb <= K) ? Ph[(kb + K) * LDH + rc0 + l] : zero;
      };
      auto lq = [&](int l, int pos) { return Q[l * LD + pos]; };
      auto sq = [&](int l, int pos, Cx<T> x) { Q[l * LD + pos] = x; };
      auto vf = [&](int l, int pos) { return vplane[pos * N + c_lo + rc0 + l]; };
      auto sp = [&](int l, int pos, Cx<T> x) {
        const int kb = pos <= N / 2 ? pos : pos - N;
        if (kb >= -K && kb <= K) Ph[(kb + K) * LDH + rc0 + l] = x;
      };
      // Q is free at each sub-chunk's first pass (closing barrier of its last
      // use) and the half's rows are read by the first pass only (modes 0 /
      // 1) or written by the last pass only (mode 2): the passes that read or
      // write outside Q skip the barrier between their loads and stores
      if constexpr (MODE == 1) {
        auto sf = [&](int l, int pos, Cx<T> x) { F[(size_t)pos * N + c_lo + rc0 + l] = x; };
        fft_passes_io<T, N, +1, false, 0, P - 1, CB, NT, true, false, false, false, 0, true,
                      true>(Q, LD, nb, tw, tid, lp, sf);
      } else if constexpr (MODE == 2) {
        fft_pass_finish<T, N, -1, 0, true, CB, NT, true>(nb, tw, tid, fpre, sq);
        if (rc0 + CB < H) {
          rcl = rc0 + CB;
          fft_pass_load<T, N, 0, true, CB, NT, true>(H - rcl < CB ? H - rcl : CB, tid, lfp,
                                                     fpre);
        }
        fft_passes_io<T, N, -1, true, 1, P - 1, CB, NT, true, false, false, false, 0, false,
                      true>(Q, LD, nb, tw, tid, lq, sp);
      } else {
        if constexpr (P >= 2)
          fft_passes_io<T, N, +1, false, 0, P - 2, CB, NT, true, false, false, false, 0, true,
                        false>(Q, LD, nb, tw, tid, lp, sq);
        fft_fused_vmul<T, N, CB, NT, true>(nb, tw, tid, lq, vf, sq);
        if constexpr (P >= 2)
          fft_passes_io<T, N, -1, true, 1, P - 1, CB, NT, true, false, false, false, 0, false,
                        true>(Q, LD, nb, tw, tid, lq, sp);
      }
    }
    // forward along c of the rows with the other half zero: store / add
    if constexpr (MODE != 1)
    for (int r0 = 0; r0 < M; r0 += ML) {
      const int nl = M - r0 < ML ? M - r0 : ML;
      const int* rc = rowcol + (size_t)r0 * N;
      auto lb = [&](int l, int pos) {
        const int c = pos - c_lo;
        return (c >= 0 && c < H) ? Ph[(r0 + l) * LDH + c] : zero;
      };
      auto sw = [&](int l, int pos, Cx<T> x) {
        const int d = rc[l * N + pos];
        if (d < 0) return;
        if (h == 0)
          Wq[d] = x;
        else
          Wq[d] = Wq[d] + x;
      };
      fft_passes_io<T, N, -1, false, 0, P - 1, ML, NT, false, false, false, true, 0, true,
                    true>(Q, LDP, nl, tw, tid, lb, sw);
    }
  }
}

constexpr int SDFT_EINVAL_CODE = -1;

#define SDFT_SIZE_SWITCH(n, CALL)          \
  switch (n) {                             \
    case 24: { constexpr int NN = 24; CALL; } break;   \
    case 32: { constexpr int NN = 32; CALL; } break;   \
    case 36: { constexpr int NN = 36; CALL; } break;   \
    case 48: { constexpr int NN = 48; CALL; } break;   \
    case 64: { constexpr int NN = 64; CALL; } break;   \
    case 96: { constexpr int NN = 96; CALL; } break;   \
    case 128: { constexpr int NN = 128; CALL; } break; \
    case 192: { constexpr int NN = 192; CALL; } break; \
    default: set_error("unsupported grid size"); return SDFT_EINVAL_CODE; \
  }

template <typename T>
constexpr int prec_of() { return sizeof(T) == 8 ? 0 : 1; }

}  // namespace sdft
