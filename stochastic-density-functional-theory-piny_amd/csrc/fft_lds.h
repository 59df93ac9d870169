// fft_lds.h -- mixed-radix Stockham FFT on lines held in LDS (gfx950).
//
// A workgroup transforms `nlines` lines of compile-time length N that sit in
// LDS (line l starts at buf + l*ld, elements contiguous).  Each pass with
// radix R has N/R butterflies per line; the workgroup's threads first load
// every butterfly's R inputs into registers, barrier, then twiddle, apply the
// radix-R DFT and store in autosorted (Stockham) order, so the transform is
// in place in LDS without a ping-pong buffer.  DIR = +1 computes
// sum_k x_k e^{+2 pi i jk/N} (FFTW_BACKWARD, G->r), DIR = -1 the forward
// transform; neither is normalised.
//
// Radix schedule: powers of two in radix-8 passes (one or two radix-4 passes
// for the remainder) or, by choice, in radix-4 passes; then radix-3 passes.
// Supported N = 2^a 3^b.
#pragma once

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define SDFT_HD __host__ __device__
#define SDFT_DEV __device__
#define SDFT_DEVI __device__ __forceinline__
#else
#define SDFT_HD
#define SDFT_DEV
#define SDFT_DEVI inline
#endif

namespace sdft {

// A complex value, aligned to its size so that loads and stores of one
// element are single 64-bit (float) or 128-bit (double) accesses.
template <typename T>
struct alignas(2 * sizeof(T)) Cx {
  T x, y;
};

template <typename T>
SDFT_HD inline Cx<T> cx(T a, T b) {
  Cx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
SDFT_HD inline Cx<T> operator+(Cx<T> a, Cx<T> b) { return cx(a.x + b.x, a.y + b.y); }
template <typename T>
SDFT_HD inline Cx<T> operator-(Cx<T> a, Cx<T> b) { return cx(a.x - b.x, a.y - b.y); }
template <typename T>
SDFT_HD inline Cx<T> cmul(Cx<T> a, Cx<T> b) {
  return cx(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename T>
SDFT_HD inline Cx<T> cmulc(Cx<T> a, Cx<T> b) {
  return cx(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T>
SDFT_HD inline Cx<T> cscale(Cx<T> a, T s) { return cx(a.x * s, a.y * s); }
// multiply by DIR*i
template <typename T, int DIR>
SDFT_HD inline Cx<T> mul_di(Cx<T> a) {
  return DIR > 0 ? cx(-a.y, a.x) : cx(a.y, -a.x);
}

// Index products below 2^24 (line offsets, twiddle indices) through the
// full-rate 24-bit multiply; a 32-bit v_mul_lo_u32 issues at a quarter of the
// VALU rate.
SDFT_HD inline int mul24(int a, int b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return static_cast<int>(__umul24(static_cast<unsigned>(a), static_cast<unsigned>(b)));
#else
  return a * b;
#endif
}

// ---- radix schedule ------------------------------------------------------
// Schedule S = 0 takes the powers of two in radix-8 passes (one or two
// radix-4 passes for the remainder); S = 1 in radix-4 passes (one radix-8
// pass for an odd power).  Among the supported sizes they differ where the
// power of two in N is 2^6 (64, 192): 8 x 8 against 4 x 4 x 4.  Radix-4 passes have twice the butterflies per
// line, so a workgroup can keep every thread busy on half as many lines
// (half the LDS for the lines in flight).  Radix-3 passes come last.
constexpr int exp2_of(int n) { int e = 0; while (n % 2 == 0) { n /= 2; ++e; } return e; }
constexpr int exp3_of(int n) { int e = 0; while (n % 3 == 0) { n /= 3; ++e; } return e; }
constexpr bool fft_size_ok(int n) {
  int m = n;
  while (m % 2 == 0) m /= 2;
  while (m % 3 == 0) m /= 3;
  return m == 1 && n > 1;
}
constexpr int n_pow2_passes(int N, int S = 0) {
  int e = exp2_of(N);
  if (e == 0) return 0;
  if (S == 1) return e == 1 ? 1 : (e % 2 == 0 ? e / 2 : (e - 3) / 2 + 1);
  if (e % 3 == 0) return e / 3;
  if (e % 3 == 1) return e == 1 ? 1 : (e - 4) / 3 + 2;
  return e / 3 + 1;
}
constexpr int n_passes(int N, int S = 0) { return n_pow2_passes(N, S) + exp3_of(N); }
constexpr int radix_at(int N, int i, int S = 0) {
  int e = exp2_of(N);
  int p2 = n_pow2_passes(N, S);
  if (i >= p2) return 3;
  if (S == 1) {
    if (e == 1) return 2;
    return (e % 2 == 1 && i == 0) ? 8 : 4;
  }
  if (e % 3 == 0) return 8;
  if (e % 3 == 1) {
    if (e == 1) return 2;
    return i < (e - 4) / 3 ? 8 : 4;
  }
  return i < e / 3 ? 8 : 4;
}
constexpr int ns_before(int N, int i, int S = 0) {
  int ns = 1;
  for (int k = 0; k < i; ++k) ns *= radix_at(N, k, S);
  return ns;
}
constexpr int min_radix(int N, int S = 0) {
  int r = 8;
  for (int i = 0; i < n_passes(N, S); ++i) r = radix_at(N, i, S) < r ? radix_at(N, i, S) : r;
  return r;
}
// Lines a workgroup of NT threads transforms per call: enough that the
// smallest-radix pass gives each thread at most two butterflies.
constexpr int fft_max_lines(int N, int NT, int S = 0) {
  int m = 2 * NT * min_radix(N, S) / N;
  return m < 1 ? 1 : m;
}

// ---- small DFTs (in registers), sign DIR -------------------------------
template <typename T, int R, int DIR>
struct Dft;

template <typename T, int DIR>
struct Dft<T, 2, DIR> {
  SDFT_HD static inline void run(Cx<T>* v) {
    Cx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  }
};

template <typename T, int DIR>
struct Dft<T, 3, DIR> {
  SDFT_HD static inline void run(Cx<T>* v) {
    const T c = T(-0.5);
    const T s = T(DIR) * T(0.86602540378443864676372317075294);
    Cx<T> t1 = v[1] + v[2];
    Cx<T> t2 = v[1] - v[2];
    Cx<T> m = cx(v[0].x + c * t1.x, v[0].y + c * t1.y);
    // i*s*t2
    Cx<T> is = cx(-s * t2.y, s * t2.x);
    v[0] = v[0] + t1;
    v[1] = m + is;
    v[2] = m - is;
  }
};

template <typename T, int DIR>
struct Dft<T, 4, DIR> {
  SDFT_HD static inline void run(Cx<T>* v) {
    Cx<T> a = v[0] + v[2];
    Cx<T> b = v[0] - v[2];
    Cx<T> c = v[1] + v[3];
    Cx<T> d = mul_di<T, DIR>(v[1] - v[3]);
    v[0] = a + c;
    v[1] = b + d;
    v[2] = a - c;
    v[3] = b - d;
  }
  // the same with v[1] = v[2] = 0 (not read)
  SDFT_HD static inline void run_zin(Cx<T>* v) {
    const Cx<T> a = v[0], d = mul_di<T, DIR>(v[3]);
    v[0] = a + v[3];
    v[1] = a - d;
    v[2] = a - v[3];
    v[3] = a + d;
  }
};

template <typename T, int DIR>
struct Dft<T, 8, DIR> {
  SDFT_HD static inline void run(Cx<T>* v) {
    Cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    Cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    Dft<T, 4, DIR>::run(e);
    Dft<T, 4, DIR>::run(o);
    combine(v, e, o);
  }
  // the same with v[2..5] = 0 (not read): both radix-4 halves have their
  // middle inputs zero
  SDFT_HD static inline void run_zin(Cx<T>* v) {
    Cx<T> e[4] = {v[0], v[0], v[0], v[6]};
    Cx<T> o[4] = {v[1], v[1], v[1], v[7]};
    Dft<T, 4, DIR>::run_zin(e);
    Dft<T, 4, DIR>::run_zin(o);
    combine(v, e, o);
  }
  SDFT_HD static inline void combine(Cx<T>* v, const Cx<T>* e, const Cx<T>* o) {
    const T h = T(0.70710678118654752440084436210485);
    // w^1 = (1 + DIR i)/sqrt2, w^2 = DIR i, w^3 = (-1 + DIR i)/sqrt2
    Cx<T> o1 = o[1], o2 = o[2], o3 = o[3];
    Cx<T> w1o = cx(h * (o1.x - T(DIR) * o1.y), h * (o1.y + T(DIR) * o1.x));
    Cx<T> w2o = mul_di<T, DIR>(o2);
    Cx<T> w3o = cx(h * (-o3.x - T(DIR) * o3.y), h * (-o3.y + T(DIR) * o3.x));
    v[0] = e[0] + o[0];
    v[4] = e[0] - o[0];
    v[1] = e[1] + w1o;
    v[5] = e[1] - w1o;
    v[2] = e[2] + w2o;
    v[6] = e[2] - w2o;
    v[3] = e[3] + w3o;
    v[7] = e[3] - w3o;
  }
};

// ---- one Stockham pass over lines in LDS ---------------------------------
// Pruning of a line whose nonzero spectrum is |k| < N/4 (the plane-wave
// sphere on a grid with n > 4K): in the first pass (NS = 1) of a radix R
// divisible by 4, butterfly j's inputs j + r N/R with R/4 <= r < 3R/4 lie in
// [N/4, 3N/4) and are zero (ZIN: not loaded, pruned radix-R DFT); in the last
// pass (NS = N/R) its outputs there are not needed (ZOUT: not stored, and the
// compiler drops their arithmetic).
template <int R>
constexpr bool mid_index(int r) {
  return R % 4 == 0 && r >= R / 4 && r < 3 * R / 4;
}

// tw[m] = exp(+2 pi i m / N), m < N (the table for the line length).
template <typename T, int N, int DIR, int PASS, int MAXLINES, int NT>
SDFT_DEVI void fft_pass(Cx<T>* buf, int nlines, int ld, const Cx<T>* tw, int tid) {
  constexpr int R = radix_at(N, PASS);
  constexpr int NB = N / R;
  constexpr int MAXB = (MAXLINES * NB + NT - 1) / NT;
  constexpr int NS = ns_before(N, PASS);
  constexpr int STEP = N / (NS * R);
  const int total = nlines * NB;
  Cx<T> v[MAXB][R];
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int idx = tid + b * NT;
    if (idx < total) {
      const int line = idx / NB;
      const int j = idx - line * NB;
      const Cx<T>* src = buf + line * ld + j;
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = src[r * NB];
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int idx = tid + b * NT;
    if (idx < total) {
      const int line = idx / NB;
      const int j = idx - line * NB;
      const int k = j % NS;
      if (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const Cx<T> w = tw[mul24(r, k * STEP)];
          v[b][r] = DIR > 0 ? cmul(v[b][r], w) : cmulc(v[b][r], w);
        }
      }
      Dft<T, R, DIR>::run(v[b]);
      Cx<T>* dst = buf + line * ld + (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * NS] = v[b][r];
    }
  }
  __syncthreads();
}

// All passes of a length-N transform over nlines <= MAXLINES lines (the
// pass count is a compile-time constant; the if-constexpr chain unrolls it).
template <typename T, int N, int DIR, int MAXLINES, int NT>
SDFT_DEVI void fft_lines(Cx<T>* buf, int nlines, int ld, const Cx<T>* tw, int tid) {
  static_assert(fft_size_ok(N), "FFT length must be 2^a 3^b");
  constexpr int P = n_passes(N);
  static_assert(P >= 1 && P <= 6, "unsupported pass count");
  fft_pass<T, N, DIR, 0, MAXLINES, NT>(buf, nlines, ld, tw, tid);
  if constexpr (P > 1) fft_pass<T, N, DIR, (P > 1 ? 1 : 0), MAXLINES, NT>(buf, nlines, ld, tw, tid);
  if constexpr (P > 2) fft_pass<T, N, DIR, (P > 2 ? 2 : 0), MAXLINES, NT>(buf, nlines, ld, tw, tid);
  if constexpr (P > 3) fft_pass<T, N, DIR, (P > 3 ? 3 : 0), MAXLINES, NT>(buf, nlines, ld, tw, tid);
  if constexpr (P > 4) fft_pass<T, N, DIR, (P > 4 ? 4 : 0), MAXLINES, NT>(buf, nlines, ld, tw, tid);
  if constexpr (P > 5) fft_pass<T, N, DIR, (P > 5 ? 5 : 0), MAXLINES, NT>(buf, nlines, ld, tw, tid);
}

// ---- passes with caller-supplied first loads / last stores ---------------
// Radix order of a transform: standard (radix_at) or reversed.  A forward
// transform in reversed order can start where an inverse transform in
// standard order ends: the inverse's last pass writes positions j + r N/R,
// exactly the inputs of the reversed forward's first butterfly j.
constexpr int radix_sched(int N, int i, bool rev, int S = 0) {
  return rev ? radix_at(N, n_passes(N, S) - 1 - i, S) : radix_at(N, i, S);
}
constexpr int ns_sched(int N, int i, bool rev, int S = 0) {
  int ns = 1;
  for (int k = 0; k < i; ++k) ns *= radix_sched(N, k, rev, S);
  return ns;
}

// Thread slot idx -> (line, butterfly j) of a pass over nlines lines.  For a
// line count known at compile time the division folds; for a runtime count
// it goes through a single-precision reciprocal instead of an integer
// division: (idx + 1/2) / nlines lies at least 1/(2 nlines) from an integer,
// and for idx < 2^16 the rounding of v_rcp_f32 (1 ulp) and of the product
// stays below 2^-6 / nlines, so the truncation is the exact quotient.
SDFT_DEVI float slot_rcp(int nlines) {
  return __builtin_amdgcn_rcpf(static_cast<float>(nlines));
}
template <bool LINE_FAST, int NB>
SDFT_DEVI void slot_map(int idx, int nlines, float rn, int& line, int& j) {
  if constexpr (LINE_FAST) {
    if (__builtin_constant_p(nlines)) {
      line = idx % nlines;
      j = idx / nlines;
    } else {
      j = static_cast<int>((static_cast<float>(idx) + 0.5f) * rn);
      line = idx - mul24(j, nlines);
    }
  } else {
    line = idx / NB;
    j = idx - line * NB;
  }
}

// One Stockham pass over nlines <= MAXLINES lines.  Inputs come from
// lf(line, pos), outputs go to sf(line, pos, value); a barrier separates all
// loads from all stores (SYNC_MID; needed only when lf and sf address the
// same LDS buffer), and one follows the stores.  LINE_FAST maps consecutive
// threads to consecutive lines (same butterfly), otherwise to consecutive
// butterflies of a line.
// Shape of pass PASS: radix, butterflies per line, butterflies per thread
// (MAXB, for MAXLINES lines over NT threads), stride of the outputs, and the
// register type holding a thread's inputs.
template <typename T, int N, int PASS, bool REV, int MAXLINES, int NT, int S = 0>
struct PassShape {
  static constexpr int R = radix_sched(N, PASS, REV, S);
  static constexpr int NB = N / R;
  static constexpr int MAXB = (MAXLINES * NB + NT - 1) / NT;
  static constexpr int NS = ns_sched(N, PASS, REV, S);
  static constexpr int STEP = N / (NS * R);
  typedef Cx<T> Regs[MAXB][R];
};

// The load phase of a pass: each thread's butterfly inputs from lf into v.
// A caller may issue it early (the loads stay in flight across barriers
// until the inputs are first used) and finish the pass later.
template <typename T, int N, int PASS, bool REV, int MAXLINES, int NT, bool LINE_FAST,
          bool ZIN = false, int S = 0, class LF>
SDFT_DEVI void fft_pass_load(int nlines, int tid, LF& lf,
                             typename PassShape<T, N, PASS, REV, MAXLINES, NT, S>::Regs& v) {
  using Sh = PassShape<T, N, PASS, REV, MAXLINES, NT, S>;
  constexpr int R = Sh::R, NB = Sh::NB, MAXB = Sh::MAXB;
  static_assert(!ZIN || (Sh::NS == 1 && R % 4 == 0 && (R == 4 || R == 8)),
                "ZIN: first pass, radix 4/8");
  const int total = nlines * NB;
  const float rn = slot_rcp(nlines);
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int idx = tid + b * NT;
    if (idx < total) {
      int line, j;
      slot_map<LINE_FAST, NB>(idx, nlines, rn, line, j);
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (!(ZIN && mid_index<R>(r))) v[b][r] = lf(line, j + r * NB);
    }
  }
}

// The rest of a pass: twiddle, radix-R DFT, stores through sf, and the
// closing barrier.
template <typename T, int N, int DIR, int PASS, bool REV, int MAXLINES, int NT, bool LINE_FAST,
          bool ZIN = false, bool ZOUT = false, int S = 0, class SF>
SDFT_DEVI void fft_pass_finish(int nlines, const Cx<T>* tw, int tid,
                               typename PassShape<T, N, PASS, REV, MAXLINES, NT, S>::Regs& v,
                               SF& sf) {
  using Sh = PassShape<T, N, PASS, REV, MAXLINES, NT, S>;
  constexpr int R = Sh::R, NB = Sh::NB, MAXB = Sh::MAXB, NS = Sh::NS, STEP = Sh::STEP;
  static_assert(!ZOUT || (NS * R == N && R % 4 == 0), "ZOUT: last pass, radix 4/8");
  const int total = nlines * NB;
  const float rn = slot_rcp(nlines);
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int idx = tid + b * NT;
    if (idx < total) {
      int line, j;
      slot_map<LINE_FAST, NB>(idx, nlines, rn, line, j);
      const int k = j % NS;
      if (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const Cx<T> w = tw[mul24(r, k * STEP)];
          v[b][r] = DIR > 0 ? cmul(v[b][r], w) : cmulc(v[b][r], w);
        }
      }
      if constexpr (ZIN)
        Dft<T, R, DIR>::run_zin(v[b]);
      else
        Dft<T, R, DIR>::run(v[b]);
      const int base = (j / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (!(ZOUT && mid_index<R>(r))) sf(line, base + r * NS, v[b][r]);
    }
  }
  __syncthreads();
}

template <typename T, int N, int DIR, int PASS, bool REV, int MAXLINES, int NT, bool LINE_FAST,
          bool ZIN = false, bool ZOUT = false, int S = 0, bool SYNC_MID = true, class LF,
          class SF>
SDFT_DEVI void fft_pass_io(int nlines, const Cx<T>* tw, int tid, LF& lf, SF& sf) {
  typename PassShape<T, N, PASS, REV, MAXLINES, NT, S>::Regs v;
  fft_pass_load<T, N, PASS, REV, MAXLINES, NT, LINE_FAST, ZIN, S>(nlines, tid, lf, v);
  if constexpr (SYNC_MID) __syncthreads();
  fft_pass_finish<T, N, DIR, PASS, REV, MAXLINES, NT, LINE_FAST, ZIN, ZOUT, S>(nlines, tw, tid,
                                                                               v, sf);
}

// LDS slot of element pos of a line in the padded layout: one unused slot
// after every 8 elements.  A radix-8 first pass over lines mapped line-major
// (8 consecutive threads on the 8 butterflies of a line) stores element
// 8 j + r from thread j; unpadded, those 8 threads hit 2 bank pairs (4-way
// conflicts for 8-byte elements); padded (slot 9 j + r, line stride 72
// slots) the threads of 2 (or 4) consecutive lines, 16 (or 32) threads, hit
// distinct bank pairs.
template <bool PAD8>
SDFT_HD inline int lds_slot(int pos) {
  return PAD8 ? pos + (pos >> 3) : pos;
}

// Passes FIRST..LAST (inclusive) of one transform: the first loads through
// lf0, the last stores through sf1, the ones between go through the LDS lines
// buf + line * ld (padded slots with PAD8).  ZIN / ZOUT prune the first /
// last pass (see fft_pass_io).  SEP_IN: lf0 reads memory other than the lines
// (global memory, another LDS buffer), so the first pass stores without a
// barrier after its loads -- the caller guarantees that the lines are no
// longer read (a barrier since their last use); SEP_OUT: sf1 writes memory
// other than the lines, so the last pass does the same.
template <typename T, int N, int DIR, bool REV, int FIRST, int LAST, int MAXLINES, int NT,
          bool LINE_FAST, bool ZIN = false, bool ZOUT = false, bool PAD8 = false, int S = 0,
          bool SEP_IN = false, bool SEP_OUT = false, class LF, class SF>
SDFT_DEVI void fft_passes_io(Cx<T>* buf, int ld, int nlines, const Cx<T>* tw, int tid,
                                   LF& lf0, SF& sf1) {
  auto lb = [&](int line, int pos) { return buf[mul24(line, ld) + lds_slot<PAD8>(pos)]; };
  auto sb = [&](int line, int pos, Cx<T> x) { buf[mul24(line, ld) + lds_slot<PAD8>(pos)] = x; };
  if constexpr (FIRST == LAST) {
    // one pass from lf0 to sf1: they cannot alias when exactly one of them
    // addresses the lines
    fft_pass_io<T, N, DIR, FIRST, REV, MAXLINES, NT, LINE_FAST, ZIN, ZOUT, S,
                !(SEP_IN != SEP_OUT)>(nlines, tw, tid, lf0, sf1);
  } else {
    fft_pass_io<T, N, DIR, FIRST, REV, MAXLINES, NT, LINE_FAST, ZIN, false, S, !SEP_IN>(
        nlines, tw, tid, lf0, sb);
    if constexpr (FIRST + 1 < LAST) {
      fft_passes_io<T, N, DIR, REV, FIRST + 1, LAST - 1, MAXLINES, NT, LINE_FAST, false, false,
                    PAD8, S>(buf, ld, nlines, tw, tid, lb, sb);
    }
    fft_pass_io<T, N, DIR, LAST, REV, MAXLINES, NT, LINE_FAST, false, ZOUT, S, !SEP_OUT>(
        nlines, tw, tid, lb, sf1);
  }
}

// Last pass of an inverse transform (standard order), a real multiplier
// vf(line, pos) at every output position, and the first pass of the forward
// transform in reversed order, in registers.  Outputs land at j*R + r.
template <typename T, int N, int MAXLINES, int NT, bool LINE_FAST, int S = 0, class LF, class VF,
          class SF>
SDFT_DEVI void fft_fused_vmul(int nlines, const Cx<T>* tw, int tid, LF& lf, VF& vf,
                                    SF& sf) {
  constexpr int P = n_passes(N, S);
  constexpr int R = radix_at(N, P - 1, S);
  constexpr int NB = N / R;
  constexpr int MAXB = (MAXLINES * NB + NT - 1) / NT;
  const int total = nlines * NB;
  const float rn = slot_rcp(nlines);
  Cx<T> v[MAXB][R];
  // the multipliers (global memory) are loaded with the inputs, before the
  // barrier, so that their latency is not exposed after it
  T sv[MAXB][R];
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int idx = tid + b * NT;
    if (idx < total) {
      int line, j;
      slot_map<LINE_FAST, NB>(idx, nlines, rn, line, j);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        v[b][r] = lf(line, j + r * NB);
        sv[b][r] = vf(line, j + r * NB);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < MAXB; ++b) {
    const int idx = tid + b * NT;
    if (idx < total) {
      int line, j;
      slot_map<LINE_FAST, NB>(idx, nlines, rn, line, j);
      if (P > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], tw[mul24(r, j)]);
      }
      Dft<T, R, +1>::run(v[b]);
#pragma unroll
      for (int r = 0; r < R; ++r) v[b][r] = cx(v[b][r].x * sv[b][r], v[b][r].y * sv[b][r]);
      Dft<T, R, -1>::run(v[b]);
#pragma unroll
      for (int r = 0; r < R; ++r) sf(line, j * R + r, v[b][r]);
    }
  }
  __syncthreads();
}

// fft_lines with a choice of thread mapping (LINE_FAST: consecutive threads
// on consecutive lines; otherwise on consecutive butterflies of one line).
template <typename T, int N, int DIR, int MAXLINES, int NT, bool LINE_FAST>
SDFT_DEVI void fft_lines_map(Cx<T>* buf, int nlines, int ld, const Cx<T>* tw, int tid) {
  if constexpr (!LINE_FAST) {
    fft_lines<T, N, DIR, MAXLINES, NT>(buf, nlines, ld, tw, tid);
  } else {
    constexpr int P = n_passes(N);
    auto lb = [&](int line, int pos) { return buf[mul24(line, ld) + pos]; };
    auto sb = [&](int line, int pos, Cx<T> x) { buf[mul24(line, ld) + pos] = x; };
    fft_passes_io<T, N, DIR, false, 0, P - 1, MAXLINES, NT, true>(buf, ld, nlines, tw, tid, lb,
                                                                  sb);
  }
}

// Any number of lines: batches of MAXLINES.
template <typename T, int N, int DIR, int MAXLINES, int NT>
SDFT_DEVI void fft_lines_batched(Cx<T>* buf, int nlines, int ld, const Cx<T>* tw, int tid) {
  for (int l0 = 0; l0 < nlines; l0 += MAXLINES) {
    const int nl = nlines - l0 < MAXLINES ? nlines - l0 : MAXLINES;
    fft_lines<T, N, DIR, MAXLINES, NT>(buf + l0 * ld, nl, ld, tw, tid);
  }
}

}  // namespace sdft
