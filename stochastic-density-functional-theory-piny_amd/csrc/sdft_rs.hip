// sdft_rs.hip -- the real-space Kleinman-Bylander term in its per-atom form:
// nlppKBRealEnergyFilter (energy/cp/cp_energy_eext_nonloc_real_filter.c:
// 110-450), which controlPotentialEnergyFilter (:36-105) adds to v.psi on the
// real-space grid before the forward FFT.
//
// Per H application on npairs Gamma-packed pairs z = psi1 + i psi2:
//   rs_to_grid   (plane stage, MODE 1): W -> z(r) on the pair grids
//                F[pair][ra][rb][rc] (inverse c and b transforms);
//   rs_project   k_rs_project: Z_p(pair) = sum_k beta_p(r_k) z(r_k) over the
//                projector's grid points -- the gather of :246-258 and the
//                per-channel dots of :318-404 -- as one [nch x K_a] .
//                [K_a x pairs] product per atom (the channels of an atom
//                share its point list, gridNlppMap);
//   rs_apply     k_rs_apply: z(r) <- v(r) z(r) + sum_p w_p(r) Z_p(pair) with
//                w_p = E_p vol_elem beta_p, at every grid point through the
//                point -> projector inverse lists -- the daxpy into
//                forceTemp and the scatter-add of :404-441 plus z v (:78, :97);
//   rs_from_grid (plane stage, MODE 2): forward b and c transforms back to W.
// beta is real, so the complex z carries both orbitals of a pair through the
// arithmetic the reference applies to the real and imaginary parts in turn.
// Cost per H application: O(sum_a K_a) for the gather and scatter, against
// O(P ncoef) for the equivalent G-space table.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sdft_hip.h"
#include "sdft_plane.h"

namespace sdft {

constexpr int kRsThreads = 256;
// The apply kernel's point blocks, one point per thread: a 16 x 16 tile of
// (rb, rc) in one a-plane where n is a multiple of 16 (the projectors a tile
// touches are those of the few atoms near it), else a run of 256 consecutive
// grid points.  At most kRsMaxSlots projectors per block are staged in LDS,
// and at most kRsStageBytes of their Z rows at a time; a point's first
// kRsRegEntries list entries are held in registers across the pairs, and
// kRsAccPairs pairs are accumulated at a time.  The apply kernel streams the
// pair grids and is bound by memory latency: 8 + 8 (94 VGPRs in FP32, 124 in
// FP64) keeps 4-5 waves per SIMD in flight where 16 + 16 (159 / 213 VGPRs)
// kept 2-3, and measured 19 % (FP32) / 26 % (FP64) faster at 192^3.
constexpr int kRsApplyPoints = kRsThreads;
constexpr int kRsTile = 16;
constexpr int kRsMaxSlots = 256;
constexpr size_t kRsStageBytes = 64 * 1024;
constexpr int kRsRegEntries = 8;
constexpr int kRsAccPairs = 8;

// grid point of thread t of apply block b.  Tiled: blocks of kRsApplyPoints
// points, tb x tc in (rb, rc) with tc the run along rc (tb = points / tc),
// n / tb x n / tc tiles per a-plane; else runs of kRsApplyPoints points.
__host__ __device__ inline long long rs_block_point(int b, int t, int n, int tc) {
  if (tc == 0) return (long long)b * kRsApplyPoints + t;
  const int tb = kRsApplyPoints / tc;
  const int nb = n / tb, nc = n / tc;
  const int ra = b / (nb * nc), rem = b - ra * nb * nc;
  const int ib = rem / nc, ic = rem - ib * nc;
  const int rb = ib * tb + t / tc, rc = ic * tc + t % tc;
  return ((long long)ra * n + rb) * n + rc;
}

struct RsKB {
  int nproj = 0;
  int ngroup = 0;
  long long nent = 0;      // entries over all projectors
  long long nent_g = 0;    // entries over the groups (one per shared point)
  // groups: consecutive projectors (at most 4) with one point list
  long long* goff = nullptr;  // [ngroup + 1]
  int* gidx = nullptr;        // [nent_g] FFTW linear grid index
  double* gval64 = nullptr;   // [nent_g][4] beta of the group's channels (0 past nch)
  float* gval32 = nullptr;
  int* gp0 = nullptr;         // [ngroup] first projector
  int* gnch = nullptr;        // [ngroup] channels
  // inverse lists: grid point -> (projector, E_p vol_elem beta_p(r))
  long long* ptoff = nullptr; // [N + 1]
  int* ptslot = nullptr;      // [nent] slot of the entry's projector in its point block's
                              // staged set, or -1 - p when the set is full
  double* ptw64 = nullptr;    // [nent]
  float* ptw32 = nullptr;
  // per block of kRsApplyPoints consecutive grid points: the projectors its
  // points touch (staged in LDS by the apply kernel)
  int* bpoff = nullptr;       // [nblock + 1]
  int* bpp = nullptr;         // projector ids
  int nblock = 0;
  int max_slots = 0;          // largest staged set
  bool tiled = false;         // (rb, rc) tiles of 256 points (else runs of 256 points)
  int tile_c = 0;             // the tiles' extent along rc (0: runs)
  double* escale = nullptr;   // [nproj] E_p vol_elem / N (energies)
  // per-call buffers
  void* F = nullptr;          // pair grids, complex in the form's type
  size_t F_bytes = 0;
  void* Z = nullptr;          // [nproj][npairs] complex in the form's type
  size_t Z_bytes = 0;
  double* epart = nullptr;    // energy partials
  size_t epart_bytes = 0;
  // gradient tables d beta_p / d r (forces), per group entry [nent_g][4][3]
  double* gdev = nullptr;
  double* fpart = nullptr;    // force partials
  size_t fpart_bytes = 0;
  std::vector<long long> goff_h;  // the group structure (host)
  std::vector<int> gp0_h, gnch_h;
  // the caller's lists (host), for the equivalent G-space table on demand
  std::vector<long long> off;
  std::vector<int> idx;
  std::vector<double> val, energy;
  double vol_elem = 0.0;
};

constexpr int SDFT_ENOMEM_CODE = -3;

template <typename P>
static int rs_alloc(P** p, size_t bytes) {
  *p = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(p), bytes ? bytes : 16) != hipSuccess) {
    *p = nullptr;
    set_error("device allocation failed (real-space projectors, " + std::to_string(bytes) +
              " bytes)");
    return SDFT_ENOMEM_CODE;
  }
  return 0;
}

template <typename V>
static int rs_upload(V** dst, const std::vector<V>& src, hipStream_t s) {
  int rc = rs_alloc(dst, sizeof(V) * src.size());
  if (rc) return rc;
  if (!src.empty())
    SDFT_HIP_CHECK(hipMemcpyAsync(*dst, src.data(), sizeof(V) * src.size(),
                                  hipMemcpyHostToDevice, s));
  return 0;
}

static int rs_grow(void** p, size_t* have, size_t need) {
  if (*have >= need) return 0;
  if (*p) SDFT_HIP_CHECK(hipFree(*p));
  *p = nullptr;
  *have = 0;
  int rc = rs_alloc(p, need);
  if (rc) return rc;
  *have = need;
  return 0;
}

static void rs_release(RsKB* r) {
  void* ptrs[] = {r->goff,   r->gidx,  r->gval64, r->gval32, r->gp0,   r->gnch,
                  r->ptoff,  r->ptslot, r->ptw64, r->ptw32, r->bpoff, r->bpp,
                  r->escale, r->F,     r->Z,     r->epart,  r->gdev,  r->fpart};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete r;
}

void rs_free(Ctx& c) {
  if (c.rs) {
    (void)hipStreamSynchronize(c.stream);
    rs_release(c.rs);
  }
  c.rs = nullptr;
}

bool rs_active(const Ctx& c) { return c.rs != nullptr && c.nl_form == 0; }

void rs_host_lists(const Ctx& c, int* nproj, const long long** off, const int** idx,
                   const double** val, const double** energy, double* vol_elem) {
  const RsKB* r = c.rs;
  *nproj = r ? r->nproj : 0;
  *off = r ? r->off.data() : nullptr;
  *idx = r ? r->idx.data() : nullptr;
  *val = r ? r->val.data() : nullptr;
  *energy = r ? r->energy.data() : nullptr;
  *vol_elem = r ? r->vol_elem : 0.0;
}

bool rs_has_gradients(const Ctx& c) { return c.rs != nullptr && c.rs->gdev != nullptr; }

int rs_set_gradients(Ctx& c, int nproj, const double* dx, const double* dy, const double* dz) {
  RsKB* r = c.rs;
  if (!r || nproj != r->nproj) {
    set_error("projector gradients: set the real-space projectors first, with the same count");
    return SDFT_ESTATE;
  }
  const double* dd[3] = {dx, dy, dz};
  for (int j = 0; j < 3; ++j)
    for (long long e = 0; e < r->nent; ++e)
      if (!std::isfinite(dd[j][e])) {
        set_error("projector gradient value is not finite");
        return SDFT_EINVAL;
      }
  std::vector<double> gd((size_t)r->nent_g * 12, 0.0);
  for (int g = 0; g < r->ngroup; ++g) {
    const int p0 = r->gp0_h[(size_t)g], nch = r->gnch_h[(size_t)g];
    const long long k0 = r->goff_h[(size_t)g], kn = r->goff_h[(size_t)g + 1] - k0;
    for (int ch = 0; ch < nch; ++ch)
      for (long long k = 0; k < kn; ++k)
        for (int j = 0; j < 3; ++j)
          gd[(size_t)(k0 + k) * 12 + ch * 3 + j] = dd[j][r->off[(size_t)p0 + ch] + k];
  }
  if (r->gdev) SDFT_HIP_CHECK(hipFree(r->gdev));
  r->gdev = nullptr;
  int rc = rs_upload(&r->gdev, gd, c.stream);
  if (!rc) SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  return rc;
}

int rs_set(Ctx& c, int nproj, const long long* off, const int* idx, const double* val,
           const double* energy, double vol_elem) {
  rs_free(c);
  if (nproj == 0) return 0;
  const long long N = c.g.N;
  if (off[0] != 0) {
    set_error("real-space projector offsets must start at 0");
    return SDFT_EINVAL;
  }
  for (int p = 0; p < nproj; ++p) {
    if (off[p + 1] < off[p]) {
      set_error("real-space projector offsets must be non-decreasing");
      return SDFT_EINVAL;
    }
    std::vector<int> ix(idx + off[p], idx + off[p + 1]);
    for (int v : ix)
      if (v < 0 || v >= N) {
        set_error("real-space projector grid index out of range");
        return SDFT_EINVAL;
      }
    std::sort(ix.begin(), ix.end());
    if (std::adjacent_find(ix.begin(), ix.end()) != ix.end()) {
      set_error("a real-space projector lists a grid point twice");
      return SDFT_EINVAL;
    }
    if (!std::isfinite(energy[p])) {
      set_error("real-space projector energy is not finite");
      return SDFT_EINVAL;
    }
  }
  const long long nent = off[nproj];
  for (long long e = 0; e < nent; ++e)
    if (!std::isfinite(val[e])) {
      set_error("real-space projector value is not finite");
      return SDFT_EINVAL;
    }
  if (nent >= (1LL << 31)) {
    set_error("real-space projectors: more than 2^31 entries");
    return SDFT_EINVAL;
  }
  RsKB* r = new RsKB();
  r->nproj = nproj;
  r->nent = nent;
  r->vol_elem = vol_elem;
  r->off.assign(off, off + nproj + 1);
  r->idx.assign(idx, idx + nent);
  r->val.assign(val, val + nent);
  r->energy.assign(energy, energy + nproj);
  // groups of consecutive projectors with identical point lists (an atom's
  // channels), at most 4 per group
  std::vector<long long> goff{0};
  std::vector<int> gidx, gp0, gnch;
  std::vector<double> gval;
  for (int p = 0; p < nproj;) {
    const long long k0 = off[p], kn = off[p + 1] - off[p];
    int nch = 1;
    while (nch < 4 && p + nch < nproj && off[p + nch + 1] - off[p + nch] == kn &&
           std::equal(idx + k0, idx + k0 + kn, idx + off[p + nch]))
      ++nch;
    gp0.push_back(p);
    gnch.push_back(nch);
    for (long long k = 0; k < kn; ++k) {
      gidx.push_back(idx[k0 + k]);
      for (int ch = 0; ch < 4; ++ch) gval.push_back(ch < nch ? val[off[p + ch] + k] : 0.0);
    }
    goff.push_back(goff.back() + kn);
    p += nch;
  }
  r->ngroup = (int)gp0.size();
  r->nent_g = goff.back();
  r->goff_h = goff;
  r->gp0_h = gp0;
  r->gnch_h = gnch;
  // inverse lists by grid point, projectors ascending within a point
  std::vector<long long> ptoff((size_t)N + 1, 0);
  for (long long e = 0; e < nent; ++e) ++ptoff[(size_t)idx[e] + 1];
  for (long long i = 0; i < N; ++i) ptoff[(size_t)i + 1] += ptoff[(size_t)i];
  std::vector<long long> fill(ptoff.begin(), ptoff.end() - 1);
  std::vector<int> ptp((size_t)nent);
  std::vector<double> ptw((size_t)nent);
  for (int p = 0; p < nproj; ++p) {
    const double s = energy[p] * vol_elem;
    for (long long e = off[p]; e < off[p + 1]; ++e) {
      const long long d = fill[(size_t)idx[e]]++;
      ptp[(size_t)d] = p;
      ptw[(size_t)d] = s * val[e];
    }
  }
  // per point block: the distinct projectors (ascending), each entry's slot
  const int n = c.g.n;
  const int tcw = kRsTile;
  const int tbw = kRsApplyPoints / tcw;
  const bool tiled = n % tcw == 0 && n % tbw == 0;
  const int tc = tiled ? tcw : 0;
  const int nblock = tiled ? n * (n / tbw) * (n / tcw)
                           : (int)((N + kRsApplyPoints - 1) / kRsApplyPoints);
  std::vector<int> bpoff{0}, bpp, ptslot((size_t)nent);
  std::vector<int> mark(nproj, -1);
  int max_slots = 0;
  for (int b = 0; b < nblock; ++b) {
    std::vector<int> set;
    for (int t = 0; t < kRsApplyPoints; ++t) {
      const long long pt = rs_block_point(b, t, n, tc);
      if (pt >= N) break;
      for (long long e = ptoff[(size_t)pt]; e < ptoff[(size_t)pt + 1]; ++e) {
        const int p = ptp[(size_t)e];
        if (mark[p] != b) {
          mark[p] = b;
          set.push_back(p);
        }
      }
    }
    std::sort(set.begin(), set.end());
    if ((int)set.size() > kRsMaxSlots) set.resize(kRsMaxSlots);
    for (int t = 0; t < kRsApplyPoints; ++t) {
      const long long pt = rs_block_point(b, t, n, tc);
      if (pt >= N) break;
      for (long long e = ptoff[(size_t)pt]; e < ptoff[(size_t)pt + 1]; ++e) {
        const int p = ptp[(size_t)e];
        const auto it = std::lower_bound(set.begin(), set.end(), p);
        ptslot[(size_t)e] = (it != set.end() && *it == p) ? (int)(it - set.begin()) : -1 - p;
      }
    }
    max_slots = std::max(max_slots, (int)set.size());
    bpp.insert(bpp.end(), set.begin(), set.end());
    bpoff.push_back((int)bpp.size());
  }
  r->tiled = tiled;
  r->tile_c = tc;
  r->nblock = nblock;
  r->max_slots = max_slots;
  std::vector<float> gval32(gval.begin(), gval.end());
  std::vector<float> ptw32(ptw.begin(), ptw.end());
  std::vector<double> esc(nproj);
  for (int p = 0; p < nproj; ++p) esc[p] = energy[p] * vol_elem / (double)N;
  int rc = 0;
  if (!rc) rc = rs_upload(&r->goff, goff, c.stream);
  if (!rc) rc = rs_upload(&r->gidx, gidx, c.stream);
  if (!rc) rc = rs_upload(&r->gval64, gval, c.stream);
  if (!rc) rc = rs_upload(&r->gval32, gval32, c.stream);
  if (!rc) rc = rs_upload(&r->gp0, gp0, c.stream);
  if (!rc) rc = rs_upload(&r->gnch, gnch, c.stream);
  if (!rc) rc = rs_upload(&r->ptoff, ptoff, c.stream);
  if (!rc) rc = rs_upload(&r->ptslot, ptslot, c.stream);
  if (!rc) rc = rs_upload(&r->bpoff, bpoff, c.stream);
  if (!rc) rc = rs_upload(&r->bpp, bpp, c.stream);
  if (!rc) rc = rs_upload(&r->ptw64, ptw, c.stream);
  if (!rc) rc = rs_upload(&r->ptw32, ptw32, c.stream);
  if (!rc) rc = rs_upload(&r->escale, esc, c.stream);
  if (!rc && hipStreamSynchronize(c.stream) != hipSuccess) {
    set_error("real-space projector upload failed");
    rc = SDFT_EHIP;
  }
  if (rc) {
    rs_release(r);
    return rc;
  }
  c.rs = r;
  return 0;
}

// ---- kernels ----------------------------------------------------------------

// NT = 512 (with the doubled b-line tile, V = 1) puts eight waves behind one
// workgroup's LDS buffer where only one workgroup fits a CU
template <typename T, int N, bool PR, int MODE, int NT>
__global__ __launch_bounds__(NT) void k_plane_field(PlaneArgs a) {
  plane_body<T, N, NT == kThreads ? 0 : 1, PR, MODE, NT>(a, blockIdx.x, blockIdx.y);
}

// FP32 512-thread form: at most 128 VGPRs, so that two workgroups (16 waves)
// share a CU where their LDS allows it
template <typename T, int N, int MODE, int NT>
__global__ __launch_bounds__(NT)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && NT == 512 ? 4 : 1))) void
k_plane_split_field(PlaneArgs a) {
  plane_split_body<T, N, MODE, NT, NT == kThreads ? 0 : 1>(a, blockIdx.x, blockIdx.y);
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Z[p][q] = sum_k beta_p[k] z_q(r_k) for the channels p of group g = blockIdx.x
// and the pairs q0 <= q < q0 + QC (q0 = blockIdx.y QC).  Each thread keeps
// QC x 4 complex partial sums over its share of the group's points; the wave
// sums, then the workgroup's four waves, give Z.
// Workgroup w runs group g's pair chunk j with w = ((g / 8) nchunk + j) 8 +
// g % 8: workgroups are dealt to the 8 XCDs in turn (w mod 8), so all chunks
// of a group run on one XCD at about the same time and re-read its beta
// values and indices from that XCD's L2.
template <typename T, int QC, int U = 1>
__global__ __launch_bounds__(kRsThreads) void k_rs_project(
    const long long* __restrict__ goff, const int* __restrict__ gidx, const T* __restrict__ gval,
    const int* __restrict__ gp0, const int* __restrict__ gnch, const Cx<T>* __restrict__ F,
    long long Ngrid, int npairs, int ngroup, int nchunk, Cx<T>* __restrict__ Z) {
  __shared__ T red[kRsThreads / 64][QC * 8];
  const int w = blockIdx.x;
  const int rest = w >> 3;
  const int g = (rest / nchunk) * 8 + (w & 7);
  if (g >= ngroup) return;
  const int q0 = (rest % nchunk) * QC;
  const int nq = npairs - q0 < QC ? npairs - q0 : QC;
  const long long e0 = goff[g], e1 = goff[g + 1];
  T ar[QC][4], ai[QC][4];
#pragma unroll
  for (int qq = 0; qq < QC; ++qq)
#pragma unroll
    for (int ch = 0; ch < 4; ++ch) {
      ar[qq][ch] = T(0);
      ai[qq][ch] = T(0);
    }
  const Cx<T>* Fq = F + (size_t)q0 * Ngrid;
  // U points per thread and trip: their index loads, then all their grid
  // loads, are in flight together (a point past the list reads point 0 of
  // the grid with zero weights)
  for (long long e = e0 + threadIdx.x; e < e1; e += U * kRsThreads) {
    size_t r[U];
    T b[U][4];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const long long eu = e + u * kRsThreads;
      const bool ok = eu < e1;
      r[u] = ok ? (size_t)gidx[eu] : 0;
#pragma unroll
      for (int ch = 0; ch < 4; ++ch) b[u][ch] = ok ? gval[eu * 4 + ch] : T(0);
    }
#pragma unroll
    for (int qq = 0; qq < QC; ++qq) {
      if (qq < nq) {
        Cx<T> z[U];
#pragma unroll
        for (int u = 0; u < U; ++u) z[u] = Fq[(size_t)qq * Ngrid + r[u]];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int ch = 0; ch < 4; ++ch) {
            ar[qq][ch] += b[u][ch] * z[u].x;
            ai[qq][ch] += b[u][ch] * z[u].y;
          }
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int qq = 0; qq < QC; ++qq)
#pragma unroll
    for (int ch = 0; ch < 4; ++ch) {
      const T sr = wave_sum(ar[qq][ch]);
      const T si = wave_sum(ai[qq][ch]);
      if (lane == 0) {
        red[wave][(qq * 4 + ch) * 2 + 0] = sr;
        red[wave][(qq * 4 + ch) * 2 + 1] = si;
      }
    }
  __syncthreads();
  const int nch = gnch[g], p0 = gp0[g];
  for (int i = threadIdx.x; i < QC * 4; i += kRsThreads) {
    const int qq = i >> 2, ch = i & 3;
    if (qq < nq && ch < nch) {
      T sr = T(0), si = T(0);
#pragma unroll
      for (int w = 0; w < kRsThreads / 64; ++w) {
        sr += red[w][i * 2 + 0];
        si += red[w][i * 2 + 1];
      }
      Z[(size_t)(p0 + ch) * npairs + q0 + qq] = cx(sr, si);
    }
  }
}

// z_q(r) <- v(r) z_q(r) + sum_{e in point r} w_e Z[p_e][q] over all pairs,
// one grid point per thread, one point block per workgroup.  The Z rows of
// the projectors the block's points touch are staged in LDS, qz pairs at a
// time, so each row is read from memory once per block instead of once per
// point and entry; an entry whose projector did not fit the staged set
// (slot < 0) reads Z directly.  A point's first kRsRegEntries entries stay in
// registers across the pairs; the rest (crowded points) are re-read.
template <typename T, int QA = kRsAccPairs, int ER = kRsRegEntries>
__global__ __launch_bounds__(kRsThreads) void k_rs_apply(
    const long long* __restrict__ ptoff, const int* __restrict__ ptslot, const T* __restrict__ ptw,
    const int* __restrict__ bpoff, const int* __restrict__ bpp, const T* __restrict__ v,
    const Cx<T>* __restrict__ Z, long long Ngrid, int n, int tc, int npairs, int qz,
    Cx<T>* __restrict__ F) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Cx<T>* Zl = reinterpret_cast<Cx<T>*>(smem_raw);  // [slots][qz + 1]
  const int ldz = qz + 1;
  const int b = blockIdx.x;
  const long long r = rs_block_point(b, threadIdx.x, n, tc);
  const bool live = r < Ngrid;
  const int s0 = bpoff[b], ns = bpoff[b + 1] - s0;
  const T vr = live ? v[r] : T(0);
  const long long e0 = live ? ptoff[r] : 0, e1 = live ? ptoff[r + 1] : 0;
  const int ne = (int)(e1 - e0);
  int sl_r[ER];
  T w_r[ER];
#pragma unroll
  for (int i = 0; i < ER; ++i) {
    sl_r[i] = i < ne ? ptslot[e0 + i] : 0;
    w_r[i] = i < ne ? ptw[e0 + i] : T(0);
  }
  const Cx<T> zero = cx(T(0), T(0));
  for (int q0 = 0; q0 < npairs; q0 += qz) {
    const int nqs = npairs - q0 < qz ? npairs - q0 : qz;
    __syncthreads();  // the previous pass's readers are done with Zl
    for (int i = threadIdx.x; i < ns * qz; i += kRsThreads) {
      const int sl = i / qz, qq = i - sl * qz;
      Zl[sl * ldz + qq] = qq < nqs ? Z[(size_t)bpp[s0 + sl] * npairs + q0 + qq] : zero;
    }
    __syncthreads();
    if (!live) continue;
    // software pipeline over the accumulator blocks: the next block's grid
    // values are loaded while this block's entries are summed (other pairs,
    // so no overlap with this block's stores).  The kernel is bound by memory
    // latency; this doubles the loads in flight per wave at +1 VGPR (FP32):
    // 2.51 -> 2.02 ms per 192^3 FP32 step, 3.79 -> 3.59 ms in FP64.
    Cx<T> zc[QA];
    {
      const int nq0 = nqs < QA ? nqs : QA;
      const Cx<T>* F0 = F + (size_t)q0 * Ngrid + r;
#pragma unroll
      for (int qq = 0; qq < QA; ++qq) zc[qq] = qq < nq0 ? F0[(size_t)qq * Ngrid] : zero;
    }
    for (int q1 = 0; q1 < nqs; q1 += QA) {
      const int nq = nqs - q1 < QA ? nqs - q1 : QA;
      Cx<T> acc[QA];
      Cx<T>* Fq = F + (size_t)(q0 + q1) * Ngrid + r;
      {
        Cx<T> zn[QA];
        const int q2 = q1 + QA;
        const int nq2 = nqs - q2 < QA ? nqs - q2 : QA;
#pragma unroll
        for (int qq = 0; qq < QA; ++qq)
          zn[qq] = qq < nq2 ? Fq[(size_t)(QA + qq) * Ngrid] : zero;
#pragma unroll
        for (int qq = 0; qq < QA; ++qq) {
          acc[qq] = cx(vr * zc[qq].x, vr * zc[qq].y);
          zc[qq] = zn[qq];
        }
      }
      // the LDS and the direct reads stay in separate branches with their
      // own index expressions (one merged pointer would be a flat address)
#define SDFT_RS_ADD(SL, W)                                                  \
  do {                                                                      \
    const int sl_ = (SL);                                                   \
    const T w_ = (W);                                                       \
    if (sl_ >= 0) {                                                         \
      _Pragma("unroll") for (int qq = 0; qq < QA; ++qq) {                   \
        if (qq < nq) {                                                      \
          const Cx<T> zz = Zl[sl_ * ldz + q1 + qq];                         \
          acc[qq].x += w_ * zz.x;                                           \
          acc[qq].y += w_ * zz.y;                                           \
        }                                                                   \
      }                                                                     \
    } else {                                                                \
      _Pragma("unroll") for (int qq = 0; qq < QA; ++qq) {                   \
        if (qq < nq) {                                                      \
          const Cx<T> zz = Z[(size_t)(-1 - sl_) * npairs + q0 + q1 + qq];   \
          acc[qq].x += w_ * zz.x;                                           \
          acc[qq].y += w_ * zz.y;                                           \
        }                                                                   \
      }                                                                     \
    }                                                                       \
  } while (0)
#pragma unroll
      for (int i = 0; i < ER; ++i)
        if (i < ne) SDFT_RS_ADD(sl_r[i], w_r[i]);
      for (long long e = e0 + ER; e < e1; ++e) SDFT_RS_ADD(ptslot[e], ptw[e]);
#undef SDFT_RS_ADD
#pragma unroll
      for (int qq = 0; qq < QA; ++qq)
        if (qq < nq) Fq[(size_t)qq * Ngrid] = acc[qq];
    }
  }
}

// Force sums of group g = blockIdx-mapped as in k_rs_project (FP64): over the
// group's points Z_ch(q) = sum_k beta_ch z_q and Zd_ch,j(q) = sum_k (d beta_ch
// / d r_j) z_q for the QC pairs of the chunk, then per channel and direction
// S = sum_q (Re Z Re Zd + Im Z Im Zd) -- the two orbitals of each pair --
// written to part[(g nchunk + chunk) 12 + 3 ch + j].
template <int QC>
__global__ __launch_bounds__(kRsThreads) void k_rs_force_sums(
    const long long* __restrict__ goff, const int* __restrict__ gidx,
    const double* __restrict__ gval, const double* __restrict__ gdev,
    const Cx<double>* __restrict__ F, long long Ngrid, int npairs, int ngroup, int nchunk,
    double* __restrict__ part) {
  constexpr int NV = QC * 32;  // Z (4 x 2) + Zd (12 x 2) per pair
  __shared__ double red[kRsThreads / 64][NV];
  const int w = blockIdx.x;
  const int rest = w >> 3;
  const int g = (rest / nchunk) * 8 + (w & 7);
  if (g >= ngroup) return;
  const int chunk = rest % nchunk;
  const int q0 = chunk * QC;
  const int nq = npairs - q0 < QC ? npairs - q0 : QC;
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = 0.0;
  const Cx<double>* Fq = F + (size_t)q0 * Ngrid;
  for (long long e = goff[g] + threadIdx.x; e < goff[g + 1]; e += kRsThreads) {
    const size_t r = (size_t)gidx[e];
    double b[4], d[12];
#pragma unroll
    for (int ch = 0; ch < 4; ++ch) b[ch] = gval[e * 4 + ch];
#pragma unroll
    for (int i = 0; i < 12; ++i) d[i] = gdev[e * 12 + i];
#pragma unroll
    for (int qq = 0; qq < QC; ++qq) {
      if (qq < nq) {
        const Cx<double> z = Fq[(size_t)qq * Ngrid + r];
#pragma unroll
        for (int ch = 0; ch < 4; ++ch) {
          acc[qq * 32 + ch * 2 + 0] += b[ch] * z.x;
          acc[qq * 32 + ch * 2 + 1] += b[ch] * z.y;
        }
#pragma unroll
        for (int i = 0; i < 12; ++i) {
          acc[qq * 32 + 8 + i * 2 + 0] += d[i] * z.x;
          acc[qq * 32 + 8 + i * 2 + 1] += d[i] * z.y;
        }
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    const double v = wave_sum(acc[i]);
    if (lane == 0) red[wave][i] = v;
  }
  __syncthreads();
  if (threadIdx.x < 12) {
    const int ch = threadIdx.x / 3, j = threadIdx.x % 3;
    double sum = 0.0;
    for (int qq = 0; qq < nq; ++qq) {
      double zr = 0.0, zi = 0.0, dr = 0.0, di = 0.0;
      for (int wv = 0; wv < kRsThreads / 64; ++wv) {
        zr += red[wv][qq * 32 + ch * 2 + 0];
        zi += red[wv][qq * 32 + ch * 2 + 1];
        dr += red[wv][qq * 32 + 8 + (ch * 3 + j) * 2 + 0];
        di += red[wv][qq * 32 + 8 + (ch * 3 + j) * 2 + 1];
      }
      sum += zr * dr + zi * di;
    }
    part[((size_t)g * nchunk + chunk) * 12 + threadIdx.x] = sum;
  }
}

// part[b] = sum over (p, q) of escale[p] (Re Z^2 + Im Z^2), the imaginary
// part only where orbital 2q + 1 exists
__global__ void k_rs_energy(const Cx<double>* __restrict__ Z, const double* __restrict__ esc,
                            int nproj, int npairs, int nstate, double* __restrict__ part) {
  __shared__ double sh[kRsThreads / 64];
  double acc = 0.0;
  const long long tot = (long long)nproj * npairs;
  for (long long i = (long long)blockIdx.x * kRsThreads + threadIdx.x; i < tot;
       i += (long long)gridDim.x * kRsThreads) {
    const int p = (int)(i / npairs), q = (int)(i - (long long)p * npairs);
    const Cx<double> z = Z[i];
    acc += esc[p] * (z.x * z.x + (2 * q + 1 < nstate ? z.y * z.y : 0.0));
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kRsThreads / 64; ++w) s += sh[w];
    part[blockIdx.x] = s;
  }
}

// ---- launchers ----------------------------------------------------------------

template <typename T, int N, bool PR, int MODE, int NT>
static int launch_plane_field(Ctx& c, const PlaneArgs& a, size_t lds, int npairs) {
  SDFT_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_plane_field<T, N, PR, MODE, NT>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL((k_plane_field<T, N, PR, MODE, NT>), dim3(N, npairs), dim3(NT), lds, c.stream,
                     a);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

template <typename T, int N, int MODE, int NT>
static int launch_plane_split_field(Ctx& c, const PlaneArgs& a, size_t lds, int npairs) {
  SDFT_HIP_CHECK(hipFuncSetAttribute(
      reinterpret_cast<const void*>(&k_plane_split_field<T, N, MODE, NT>),
      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL((k_plane_split_field<T, N, MODE, NT>), dim3(N, npairs), dim3(NT), lds,
                     c.stream, a);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

// The plane stage's launch form, by waves per CU (LDS-limited): the whole
// plane with 256 threads, or with 512 threads and the doubled b-line tile,
// or -- where it at least doubles the waves, or where nothing else fits --
// the split form (column halves, c-transforms recomputed per half) with 256
// or 512 threads.
template <typename T, int N, int MODE>
static int plane_field_impl(Ctx& c, void* W, void* F, int npairs) {
  const Geometry& g = c.g;
  PlaneArgs a = make_plane_args(c, prec_of<T>(), W);
  a.field = F;
  constexpr size_t kLds = 160 * 1024;
  auto waves = [&](size_t lds, int nt) { return lds <= kLds ? (int)(kLds / lds) * (nt / 64) : 0; };
  const size_t l0 = plane2_lds_bytes<T, N, 0>(g.M), l1 = plane2_lds_bytes<T, N, 1>(g.M);
  const size_t s0 = plane_split_lds_bytes<T, N, 2, kThreads, true>(g.M);
  const size_t s1 = plane_split_lds_bytes<T, N, 2, 512, true, 1>(g.M);
  const int w0 = waves(l0, kThreads), w1 = waves(l1, 512);
  const int ws0 = waves(s0, kThreads), ws1 = waves(s1, 512);
  const int wun = std::max(w0, w1), wsp = std::max(ws0, ws1);
  const bool pr = plane_prunable(g.K, N);
  if (wun > 0 && wsp < 2 * wun) {
    if (w0 >= w1)
      return pr ? launch_plane_field<T, N, true, MODE, kThreads>(c, a, l0, npairs)
                : launch_plane_field<T, N, false, MODE, kThreads>(c, a, l0, npairs);
    return pr ? launch_plane_field<T, N, true, MODE, 512>(c, a, l1, npairs)
              : launch_plane_field<T, N, false, MODE, 512>(c, a, l1, npairs);
  }
  if (wsp == 0) {
    set_error("plane buffer exceeds LDS for this grid in this precision");
    return SDFT_EINVAL_CODE;
  }
  a.Wout = W;  // MODE 2 writes the result in place of its (unread) input
  if (ws0 >= ws1) return launch_plane_split_field<T, N, MODE, kThreads>(c, a, s0, npairs);
  return launch_plane_split_field<T, N, MODE, 512>(c, a, s1, npairs);
}

template <int MODE>
static int plane_field(Ctx& c, int prec, void* W, int npairs) {
  RsKB& r = *c.rs;
  const size_t need = (prec == 0 ? sizeof(Cx<double>) : sizeof(Cx<float>)) * (size_t)npairs *
                      (size_t)c.g.N;
  if (MODE == 1) {
    int rc = rs_grow(&r.F, &r.F_bytes, need);
    if (rc) return rc;
  } else if (r.F_bytes < need) {
    set_error("internal error: real-space grid not produced before its forward transform");
    return SDFT_EINVAL_CODE;
  }
  if (prec == 0) {
    SDFT_SIZE_SWITCH(c.g.n, return (plane_field_impl<double, NN, MODE>(c, W, r.F, npairs)));
  } else {
    SDFT_SIZE_SWITCH(c.g.n, return (plane_field_impl<float, NN, MODE>(c, W, r.F, npairs)));
  }
  return 0;
}

int rs_to_grid(Ctx& c, int prec, void* W, int npairs) { return plane_field<1>(c, prec, W, npairs); }

int rs_from_grid(Ctx& c, int prec, void* W, int npairs) {
  return plane_field<2>(c, prec, W, npairs);
}

template <typename T, int QC, int U>
static int project_launch(Ctx& c, int npairs) {
  RsKB& r = *c.rs;
  int rc = rs_grow(&r.Z, &r.Z_bytes, sizeof(Cx<T>) * (size_t)r.nproj * npairs);
  if (rc) return rc;
  const int nchunk = (npairs + QC - 1) / QC;
  const dim3 grid((unsigned)((r.ngroup + 7) / 8 * 8 * nchunk));
  const T* gval = sizeof(T) == 8 ? reinterpret_cast<const T*>(r.gval64)
                                 : reinterpret_cast<const T*>(r.gval32);
  hipLaunchKernelGGL((k_rs_project<T, QC, U>), grid, dim3(kRsThreads), 0, c.stream, r.goff,
                     r.gidx, gval, r.gp0, r.gnch, reinterpret_cast<const Cx<T>*>(r.F),
                     (long long)c.g.N, npairs, r.ngroup, nchunk, reinterpret_cast<Cx<T>*>(r.Z));
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

// 4 pairs per product workgroup; FP64 with two points per thread and trip.
// Measured at 128^3 / 192^3 against 8 pairs (FP32) and one point per trip:
// FP32 0.675 -> 0.591 / 2.00 -> 1.95 ms, FP64 1.035 -> 0.981 / 3.22 -> 3.19 ms;
// 16 pairs (FP32) or 8 (FP64) were 1.2-1.5x slower (144-150 VGPRs).
template <typename T>
static int project_impl(Ctx& c, int npairs) {
  return project_launch<T, 4, sizeof(T) == 8 ? 2 : 1>(c, npairs);
}

template <typename T, int QA, int ER>
static int apply_launch(Ctx& c, int npairs, size_t stage_bytes) {
  RsKB& r = *c.rs;
  const long long N = c.g.N;
  const int slots = std::max(1, r.max_slots);
  // pairs staged per pass: as many as stage_bytes hold for the largest set,
  // a multiple of the accumulator width, at most all pairs
  int qz = (int)(stage_bytes / (sizeof(Cx<T>) * slots));
  qz = std::max(QA, qz / QA * QA);
  qz = std::min(qz, (npairs + QA - 1) / QA * QA);
  // row stride qz + 1: the lanes of a wave read the rows of different
  // projectors at one pair index, and with a stride of qz (a multiple of 16
  // elements) every row would start in the same LDS bank
  const size_t lds = sizeof(Cx<T>) * (size_t)slots * (qz + 1);
  const T* ptw = sizeof(T) == 8 ? reinterpret_cast<const T*>(r.ptw64)
                                : reinterpret_cast<const T*>(r.ptw32);
  const T* v = sizeof(T) == 8 ? reinterpret_cast<const T*>(c.vks64)
                              : reinterpret_cast<const T*>(c.vks32);
  auto kern = &k_rs_apply<T, QA, ER>;
  SDFT_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(kern, dim3(r.nblock), dim3(kRsThreads), lds, c.stream, r.ptoff, r.ptslot,
                     ptw, r.bpoff, r.bpp, v, reinterpret_cast<const Cx<T>*>(r.Z), N, c.g.n,
                     r.tile_c, npairs, qz, reinterpret_cast<Cx<T>*>(r.F));
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

template <typename T>
static int apply_impl(Ctx& c, int npairs) {
  return apply_launch<T, kRsAccPairs, kRsRegEntries>(c, npairs, kRsStageBytes);
}

int rs_project(Ctx& c, int prec, int npairs) {
  return prec == 0 ? project_impl<double>(c, npairs) : project_impl<float>(c, npairs);
}

int rs_apply(Ctx& c, int prec, int npairs) {
  return prec == 0 ? apply_impl<double>(c, npairs) : apply_impl<float>(c, npairs);
}

// per projector and direction S_p,j = sum over the pairs' orbitals of
// Z_p dZ_p/d r_j (real-space grids of npairs FP64 pairs in r.F); out
// [nproj][3]
int rs_force_sums(Ctx& c, int npairs, double* out) {
  RsKB& r = *c.rs;
  if (!r.gdev) {
    set_error("internal error: no projector gradients");
    return SDFT_EINVAL_CODE;
  }
  constexpr int QC = 2;
  const int nchunk = (npairs + QC - 1) / QC;
  const size_t nparts = (size_t)r.ngroup * nchunk * 12;
  int rc = rs_grow(reinterpret_cast<void**>(&r.fpart), &r.fpart_bytes, sizeof(double) * nparts);
  if (rc) return rc;
  const dim3 grid((unsigned)((r.ngroup + 7) / 8 * 8 * nchunk));
  hipLaunchKernelGGL((k_rs_force_sums<QC>), grid, dim3(kRsThreads), 0, c.stream, r.goff, r.gidx,
                     r.gval64, r.gdev, reinterpret_cast<const Cx<double>*>(r.F),
                     (long long)c.g.N, npairs, r.ngroup, nchunk, r.fpart);
  SDFT_HIP_CHECK(hipGetLastError());
  std::vector<double> h(nparts);
  SDFT_HIP_CHECK(hipMemcpyAsync(h.data(), r.fpart, sizeof(double) * nparts,
                                hipMemcpyDeviceToHost, c.stream));
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  for (int p = 0; p < r.nproj * 3; ++p) out[p] = 0.0;
  for (int g = 0; g < r.ngroup; ++g)
    for (int ch = 0; ch < r.gnch_h[(size_t)g]; ++ch)
      for (int j = 0; j < 3; ++j) {
        double sum = 0.0;
        for (int k = 0; k < nchunk; ++k) sum += h[((size_t)g * nchunk + k) * 12 + ch * 3 + j];
        out[(size_t)(r.gp0_h[(size_t)g] + ch) * 3 + j] = sum;
      }
  return 0;
}

// after rs_project in FP64 on the same pairs
int rs_energy(Ctx& c, int npairs, int nstate, double* out) {
  RsKB& r = *c.rs;
  const long long tot = (long long)r.nproj * npairs;
  const int nb = (int)std::max<long long>(1, std::min<long long>(256, (tot + kRsThreads - 1) /
                                                                          kRsThreads));
  int rc = rs_grow(reinterpret_cast<void**>(&r.epart), &r.epart_bytes, sizeof(double) * nb);
  if (rc) return rc;
  hipLaunchKernelGGL(k_rs_energy, dim3(nb), dim3(kRsThreads), 0, c.stream,
                     reinterpret_cast<const Cx<double>*>(r.Z), r.escale, r.nproj, npairs, nstate,
                     r.epart);
  SDFT_HIP_CHECK(hipGetLastError());
  std::vector<double> h(nb);
  SDFT_HIP_CHECK(hipMemcpyAsync(h.data(), r.epart, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                c.stream));
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  double s = 0.0;
  for (int i = 0; i < nb; ++i) s += h[i];
  *out = s;
  return 0;
}

}  // namespace sdft
