// sdft_batch.hip -- fragment batches: many independent systems (each with its
// own grid, G-set, v_KS, KB table and orbital block) filtered together.
//
// The reference solves fragments one after another (fragScf,
// stodft/fragment/frag-scf.c:211-248).  Here one Chebyshev step of the whole
// batch is a fixed sequence of launches:
//   KB projection partials, split reduction, KB force  (one launch each over
//     every fragment with projectors; on a second stream when overlapping)
//   expand, plane, contract                             (one launch each per
//     grid size over the fragments of that size)
//   step-counter advance                                (one thread)
// Each launch walks a table of per-fragment argument blocks: workgroup w
// belongs to fragment f with prefix[f] <= w < prefix[f + 1].  The Chebyshev
// coefficient of step k is read on the device from coeffs[f][k] at the step
// counter, so every step k >= 2 is the same launch sequence up to the T_k
// ping-pong parity; two consecutive steps are captured once in a HIP graph
// and replayed.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sdft_hip.h"
#include "sdft_internal.h"

namespace sdft {

namespace {

__global__ void k_counter_set(int* c, int v) {
  if (threadIdx.x == 0) c[0] = v;
}

__global__ void k_counter_advance(int* c) {
  if (threadIdx.x == 0) c[0] = c[0] + 1;
}

// T_0 := chi for every fragment (one launch; sizes from a table)
struct CopyArgs {
  const double* src;
  double* dst;
  long long n;
  float* dst32;  // FP32 form with projectors: the float copy as well
};

__global__ void k_copy_batch(const CopyArgs* __restrict__ args, const int* __restrict__ prefix,
                             int nfrag) {
  const int f = find_frag(prefix, nfrag, blockIdx.x);
  const CopyArgs a = args[f];
  const long long i = (long long)(blockIdx.x - prefix[f]) * 256 + threadIdx.x;
  if (i < a.n) {
    const double x = a.src[i];
    a.dst[i] = x;
    if (a.dst32) a.dst32[i] = float(x);
  }
}

constexpr int kVariants = 3;  // 0: first step (X0 -> X1), 1: even k (X1 -> X0), 2: odd k >= 3

struct Frag {
  Block* b = nullptr;
  Ctx* c = nullptr;
  int npairs = 0;
  // projector count and padded count of the context when the batch was
  // created: zpart, Y and Z are sized for them, and the KB launch lists
  // include the fragment only when P > 0
  int P = 0, ppad = 0;
  void* W = nullptr;
  double* zpart = nullptr;
  double* Y = nullptr;
  double* Z = nullptr;
};

// one grid size
struct Group {
  int n = 0;
  std::vector<int> frags;
  size_t lds = 0;    // plane kernel dynamic LDS, tile variant 0
  size_t lds1 = 0;   // tile variant 1 (0 when it does not fit)
  bool prune = true;  // every fragment's sphere inside |k| < n/4 (pruned plane passes)
  int tot_exp = 0, tot_plane = 0;
  int* pre_exp = nullptr;    // expand and contract share the (chunk, pair) tiling
  int* pre_plane = nullptr;
  ExpandArgs* exp[2] = {nullptr, nullptr};  // cur = X0 / X1
  PlaneArgs* plane = nullptr;
  ContractArgs* con[kVariants] = {nullptr, nullptr, nullptr};
};

template <typename P>
int dmalloc(P** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    set_error("device allocation failed (batch)");
    return -3;
  }
  return 0;
}

template <typename V>
int upload_vec(V* dst, const std::vector<V>& v, hipStream_t s) {
  if (v.empty()) return 0;
  SDFT_HIP_CHECK(
      hipMemcpyAsync(dst, v.data(), sizeof(V) * v.size(), hipMemcpyHostToDevice, s));
  return 0;
}

}  // namespace

struct Batch {
  int device = 0;
  int prec = 0;
  std::vector<Frag> frags;
  std::vector<Group> groups;
  hipStream_t stream = nullptr, stream_nl = nullptr;
  hipEvent_t ev_fork[2] = {nullptr, nullptr}, ev_join[2] = {nullptr, nullptr};
  // KB launches (fragments with projectors)
  std::vector<int> nl_frags;
  int split = 1;       // target slices of the projection's reduction
  bool overlap = true;
  // fused_ok: every fragment's columns are ka intervals, so the contract can
  // write the next step's expansion; fused: it does (only the first step then
  // launches the expand kernel), chosen by the autotuner or batch_config
  bool fused_ok = true;
  bool fused = true;
  int plane_variant = 0;  // plane-kernel tile variant (falls back to 0 per group)
  int tot_proj = 0, tot_red = 0, tot_force = 0;
  int* pre_proj = nullptr;
  int* pre_red = nullptr;
  int* pre_force = nullptr;
  KbProjectArgs* proj[2] = {nullptr, nullptr};  // X = X0 / X1
  KbReduceArgs* red = nullptr;
  KbForceArgs* force = nullptr;
  // copy chi -> X0
  int tot_copy = 0;
  int* pre_copy = nullptr;
  CopyArgs* copy = nullptr;
  // Chebyshev coefficients [nfrag][L] and the step counter
  double* coeffs = nullptr;
  size_t coeffs_cap = 0;
  int* counter = nullptr;
  // captured pair of steps (even, odd)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool graph_valid = false;
  // the whole filter of full_L terms (chi copy, counter reset, steps 1 ..
  // L-1) as one graph; full_L = 0 when absent or out of date
  hipGraph_t full_graph = nullptr;
  hipGraphExec_t full_exec = nullptr;
  int full_L = 0;
  std::vector<int> split_of;  // per fragment slices at the current `split`
  std::vector<int> per_of;    // K steps per slice
  size_t max_zpart_slices = 8;
};

namespace {

void free_batch(Batch& B) {
  if (B.graph_exec) (void)hipGraphExecDestroy(B.graph_exec);
  if (B.graph) (void)hipGraphDestroy(B.graph);
  if (B.full_exec) (void)hipGraphExecDestroy(B.full_exec);
  if (B.full_graph) (void)hipGraphDestroy(B.full_graph);
  for (Frag& f : B.frags) {
    void* ptrs[] = {f.W, f.zpart, f.Y, f.Z};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
  }
  for (Group& g : B.groups) {
    void* ptrs[] = {g.pre_exp, g.pre_plane, g.exp[0], g.exp[1], g.plane,
                    g.con[0],  g.con[1],    g.con[2]};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
  }
  void* ptrs[] = {B.pre_proj, B.pre_red, B.pre_force, B.proj[0], B.proj[1], B.red,
                  B.force,    B.pre_copy, B.copy,     B.coeffs,  B.counter};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (int i = 0; i < 2; ++i) {
    if (B.ev_fork[i]) (void)hipEventDestroy(B.ev_fork[i]);
    if (B.ev_join[i]) (void)hipEventDestroy(B.ev_join[i]);
  }
  if (B.stream) (void)hipStreamDestroy(B.stream);
  if (B.stream_nl) (void)hipStreamDestroy(B.stream_nl);
}

// Split of the projection's K steps into `want` slices (at least 16 K steps
// per slice, as for single systems).
void split_for(const Frag& f, int want, int* nsplit, int* per) {
  const int nsteps = (2 * f.c->g.ncp) / 32;
  int s = want;
  if (s > nsteps / 16) s = nsteps / 16;
  if (s < 1) s = 1;
  *per = (nsteps + s - 1) / s;
  *nsplit = (nsteps + *per - 1) / *per;
}

int build_kb_tables(Batch& B) {
  const int nf = (int)B.frags.size();
  B.split_of.assign(nf, 1);
  B.per_of.assign(nf, 1);
  std::vector<int> pp{0}, pr{0}, pf{0};
  for (int fi : B.nl_frags) {
    Frag& f = B.frags[fi];
    int ns = 0, per = 0;
    split_for(f, B.split, &ns, &per);
    B.split_of[fi] = ns;
    B.per_of[fi] = per;
    const int P = f.c->nproj, S = f.b->nstate_pad, K2 = 2 * f.c->g.ncp;
    pp.push_back(pp.back() + ((P + 31) / 32) * ((S + 31) / 32) * ns);
    pr.push_back(pr.back() + (P * S + 255) / 256);
    pf.push_back(pf.back() + ((S + 31) / 32) * ((K2 + kForceCols - 1) / kForceCols));
  }
  B.tot_proj = pp.back();
  B.tot_red = pr.back();
  B.tot_force = pf.back();
  int rc;
  if ((rc = upload_vec(B.pre_proj, pp, B.stream))) return rc;
  if ((rc = upload_vec(B.pre_red, pr, B.stream))) return rc;
  if ((rc = upload_vec(B.pre_force, pf, B.stream))) return rc;
  SDFT_HIP_CHECK(hipStreamSynchronize(B.stream));
  return 0;
}

// Per-fragment argument blocks for this filter call.
int build_args(Batch& B, int L, const double* coeffs_host, const double* scale,
               const double* emean) {
  const int nf = (int)B.frags.size();
  // coefficients first: the contract blocks point into them
  const size_t need = sizeof(double) * (size_t)nf * L;
  if (B.coeffs_cap < need) {
    if (B.coeffs) SDFT_HIP_CHECK(hipFree(B.coeffs));
    B.coeffs = nullptr;
    B.coeffs_cap = 0;
    int rc0 = dmalloc(&B.coeffs, need);
    if (rc0) return rc0;
    B.coeffs_cap = need;
  }
  for (int fi = 0; fi < nf; ++fi) {
    const Frag& f = B.frags[fi];
    if (rs_active(*f.c)) {
      set_error("fragment " + std::to_string(fi) + ": the context switched to the per-atom "
                "real-space KB form after sdft_hip_batch_create");
      return SDFT_ESTATE;
    }
    if (f.c->nproj != f.P || f.c->ppad != f.ppad) {
      set_error("fragment " + std::to_string(fi) + ": the context's projectors changed after "
                "sdft_hip_batch_create (the batch's KB buffers are sized for " +
                std::to_string(f.P) + " projectors, the context now has " +
                std::to_string(f.c->nproj) + "); destroy and re-create the batch");
      return SDFT_ESTATE;
    }
  }
  SDFT_HIP_CHECK(hipMemcpyAsync(B.coeffs, coeffs_host, need, hipMemcpyHostToDevice, B.stream));
  for (Group& g : B.groups) {
    std::vector<ExpandArgs> ex[2];
    std::vector<PlaneArgs> pl;
    std::vector<ContractArgs> co[kVariants];
    for (int fi : g.frags) {
      Frag& f = B.frags[fi];
      Block& b = *f.b;
      Ctx& c = *f.c;
      double* X[2] = {b.X0, b.X1};
      for (int v = 0; v < 2; ++v) ex[v].push_back(make_expand_args(c, B.prec, X[v], f.W));
      pl.push_back(make_plane_args(c, B.prec, f.W));
      const double sc = scale[fi], em = emean[fi];
      for (int v = 0; v < kVariants; ++v) {
        StepParams sp;
        const double m = v == 0 ? 1.0 : 2.0;
        sp.s1 = -sc * 0.25 * m;
        sp.s2 = -sc * 0.5 * m;
        sp.p = -sc * em * m;
        sp.ck = 0.0;
        sp.c0 = coeffs_host[(size_t)fi * L];
        sp.first = v == 0;
        sp.step = 0;
        const double* cur = v == 1 ? b.X1 : b.X0;
        double* out = v == 1 ? b.X0 : b.X1;
        ContractArgs a = make_contract_args(c, B.prec, kModeFilter, f.W, f.npairs, cur, out,
                                            b.acc, c.nproj > 0 ? b.fnl : nullptr, b.chi, sp,
                                            nullptr);
        a.Wnext = B.fused ? f.W : nullptr;
        a.Xout32 = (B.prec == 1 && c.nproj > 0) ? b.X32 : nullptr;
        a.coeffs_dev = B.coeffs + (size_t)fi * L;
        a.step_dev = B.counter;
        co[v].push_back(a);
      }
    }
    int rc;
    for (int v = 0; v < 2; ++v)
      if ((rc = upload_vec(g.exp[v], ex[v], B.stream))) return rc;
    if ((rc = upload_vec(g.plane, pl, B.stream))) return rc;
    for (int v = 0; v < kVariants; ++v)
      if ((rc = upload_vec(g.con[v], co[v], B.stream))) return rc;
    SDFT_HIP_CHECK(hipStreamSynchronize(B.stream));
  }
  std::vector<KbProjectArgs> pj[2];
  std::vector<KbReduceArgs> rd;
  std::vector<KbForceArgs> fo;
  for (int fi : B.nl_frags) {
    Frag& f = B.frags[fi];
    Block& b = *f.b;
    Ctx& c = *f.c;
    const int P = c.nproj, S = b.nstate_pad, K2 = 2 * c.g.ncp;
    const void* A = B.prec == 0 ? (const void*)c.A64 : (const void*)c.A32;
    double* X[2] = {b.X0, b.X1};
    for (int v = 0; v < 2; ++v) {
      KbProjectArgs a;
      a.A = A;
      a.X = B.prec == 1 ? (const void*)b.X32 : (const void*)X[v];
      a.P = P;
      a.ns = S;
      a.K = K2;
      a.per = B.per_of[fi];
      a.nsplit = B.split_of[fi];
      a.tp = (P + 31) / 32;
      a.ts = (S + 31) / 32;
      a.zpart = f.zpart;
      pj[v].push_back(a);
    }
    KbReduceArgs r;
    r.zpart = f.zpart;
    r.nsplit = B.split_of[fi];
    r.P = P;
    r.ns = S;
    r.ppad = c.ppad;
    r.E = c.E;
    r.Z = f.Z;
    r.Y = f.Y;
    rd.push_back(r);
    KbForceArgs q;
    q.Y = f.Y;
    q.At = B.prec == 0 ? (const void*)c.At64 : (const void*)c.At32;
    q.ns = S;
    q.K2 = K2;
    q.ppad = c.ppad;
    q.ts = (S + 31) / 32;
    q.tk = (K2 + kForceCols - 1) / kForceCols;
    q.F = b.fnl;
    fo.push_back(q);
  }
  int rc;
  for (int v = 0; v < 2; ++v)
    if ((rc = upload_vec(B.proj[v], pj[v], B.stream))) return rc;
  if ((rc = upload_vec(B.red, rd, B.stream))) return rc;
  if ((rc = upload_vec(B.force, fo, B.stream))) return rc;
  SDFT_HIP_CHECK(hipStreamSynchronize(B.stream));
  return 0;
}

// One Chebyshev step of every fragment; variant selects the T_k parity (and
// the first-step recurrence), slot the fork/join event pair.
int enqueue_step(Batch& B, int variant, int slot) {
  const int xv = variant == 1 ? 1 : 0;  // which buffer is T_k
  const bool nl = !B.nl_frags.empty();
  const int nnl = (int)B.nl_frags.size();
  hipStream_t s = B.stream;
  hipStream_t snl = B.overlap ? B.stream_nl : B.stream;
  int rc;
  if (nl) {
    if (B.overlap) {
      SDFT_HIP_CHECK(hipEventRecord(B.ev_fork[slot], s));
      SDFT_HIP_CHECK(hipStreamWaitEvent(snl, B.ev_fork[slot], 0));
    }
    if ((rc = launch_kb_batch(0, B.prec, B.proj[xv], B.pre_proj, nnl, B.tot_proj, snl))) return rc;
    if ((rc = launch_kb_batch(1, B.prec, B.red, B.pre_red, nnl, B.tot_red, snl))) return rc;
    if ((rc = launch_kb_batch(2, B.prec, B.force, B.pre_force, nnl, B.tot_force, snl))) return rc;
    if (B.overlap) SDFT_HIP_CHECK(hipEventRecord(B.ev_join[slot], snl));
  }
  if (variant == 0 || !B.fused)
    for (Group& g : B.groups)
      if ((rc = launch_batch(0, g.n, B.prec, g.exp[xv], g.pre_exp, (int)g.frags.size(),
                             g.tot_exp, 0, 0, s)))
        return rc;
  for (Group& g : B.groups) {
    const int v = (B.plane_variant == 1 && g.lds1 > 0) ? 1 : 0;
    if ((rc = launch_batch(1, g.n, B.prec, g.plane, g.pre_plane, (int)g.frags.size(),
                           g.tot_plane, v ? g.lds1 : g.lds, v | (g.prune ? 2 : 0), s)))
      return rc;
  }
  if (nl && B.overlap) SDFT_HIP_CHECK(hipStreamWaitEvent(s, B.ev_join[slot], 0));
  for (Group& g : B.groups)
    if ((rc = launch_batch(2, g.n, B.prec, g.con[variant], g.pre_exp, (int)g.frags.size(),
                           g.tot_exp, 0, B.fused ? 1 : 0, s)))
      return rc;
  hipLaunchKernelGGL(k_counter_advance, dim3(1), dim3(64), 0, s, B.counter);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

int capture_graph(Batch& B) {
  if (B.graph_exec) SDFT_HIP_CHECK(hipGraphExecDestroy(B.graph_exec));
  if (B.graph) SDFT_HIP_CHECK(hipGraphDestroy(B.graph));
  B.graph_exec = nullptr;
  B.graph = nullptr;
  B.graph_valid = false;
  SDFT_HIP_CHECK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeRelaxed));
  int rc = enqueue_step(B, 1, 0);
  if (!rc) rc = enqueue_step(B, 2, 1);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(B.stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
  B.graph = g;
  SDFT_HIP_CHECK(hipGraphInstantiate(&B.graph_exec, g, nullptr, nullptr, 0));
  B.graph_valid = true;
  return 0;
}

int set_counter(Batch& B, int v) {
  hipLaunchKernelGGL(k_counter_set, dim3(1), dim3(64), 0, B.stream, B.counter, v);
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

int copy_chi(Batch& B) {
  hipLaunchKernelGGL(k_copy_batch, dim3(B.tot_copy), dim3(256), 0, B.stream, B.copy, B.pre_copy,
                     (int)B.frags.size());
  SDFT_HIP_CHECK(hipGetLastError());
  return 0;
}

// The whole filter of L terms: chi copy, step counter, steps 1 .. L-1.
int enqueue_filter(Batch& B, int L) {
  int rc;
  if ((rc = copy_chi(B))) return rc;
  if ((rc = set_counter(B, 1))) return rc;
  if ((rc = enqueue_step(B, 0, 0))) return rc;  // k = 1
  for (int k = 2; k < L; ++k)
    if ((rc = enqueue_step(B, (k & 1) ? 2 : 1, k & 1))) return rc;
  return 0;
}

int capture_full_graph(Batch& B, int L) {
  if (B.full_exec) SDFT_HIP_CHECK(hipGraphExecDestroy(B.full_exec));
  if (B.full_graph) SDFT_HIP_CHECK(hipGraphDestroy(B.full_graph));
  B.full_exec = nullptr;
  B.full_graph = nullptr;
  B.full_L = 0;
  SDFT_HIP_CHECK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeRelaxed));
  const int rc = enqueue_filter(B, L);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(B.stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
  B.full_graph = g;
  SDFT_HIP_CHECK(hipGraphInstantiate(&B.full_exec, g, nullptr, nullptr, 0));
  B.full_L = L;
  return 0;
}

// Steps 1 .. L-1 of the recurrence (filter.c:560-633) on every fragment.
// use_graph 0: every kernel launched directly; 1: steps 2 .. replayed from
// the captured pair of steps; 2: the whole filter replayed as one graph.
int run_filter(Batch& B, int L, int use_graph) {
  int rc;
  if (use_graph == 2) {
    if (B.full_L != L && (rc = capture_full_graph(B, L))) return rc;
    SDFT_HIP_CHECK(hipGraphLaunch(B.full_exec, B.stream));
  } else {
    if ((rc = copy_chi(B))) return rc;
    if ((rc = set_counter(B, 1))) return rc;
    if ((rc = enqueue_step(B, 0, 0))) return rc;  // k = 1
    int k = 2;
    if (use_graph == 1 && L - 2 >= 2) {
      if (!B.graph_valid && (rc = capture_graph(B))) return rc;
      for (; k + 1 < L; k += 2) SDFT_HIP_CHECK(hipGraphLaunch(B.graph_exec, B.stream));
    }
    for (; k < L; ++k)
      if ((rc = enqueue_step(B, (k & 1) ? 2 : 1, k & 1))) return rc;
  }
  SDFT_HIP_CHECK(hipStreamSynchronize(B.stream));
  // T_{L-1} sits in X[(L-1) & 1]
  for (Frag& f : B.frags) f.b->cur = (L - 1) & 1;
  return 0;
}

}  // namespace

}  // namespace sdft

struct sdft_batch {
  sdft::Batch B;
};

using namespace sdft;

extern "C" {

int sdft_hip_batch_create(int nfrag, sdft_block* const* blocks, sdft_batch** out) {
  if (nfrag < 1 || !blocks || !out) {
    set_error("bad batch arguments");
    return SDFT_EINVAL;
  }
  *out = nullptr;
  const int prec = blocks[0] ? blocks[0]->b.prec : 0;
  int dev = 0;
  SDFT_HIP_CHECK(hipGetDevice(&dev));
  for (int i = 0; i < nfrag; ++i) {
    if (!blocks[i] || !blocks[i]->b.ctx) {
      set_error("null block in batch");
      return SDFT_EINVAL;
    }
    const Block& b = blocks[i]->b;
    if (b.prec != prec) {
      set_error("all blocks of a batch must have the same precision");
      return SDFT_EINVAL;
    }
    if (b.ctx->device != dev) {
      set_error("all blocks of a batch must live on the current device");
      return SDFT_EINVAL;
    }
    if (!b.ctx->have_vks) {
      set_error("v_KS not set on a fragment context (call sdft_hip_set_vks first)");
      return SDFT_ESTATE;
    }
    if (rs_active(*b.ctx)) {
      set_error("fragment context with real-space projectors in their per-atom form: the batch "
                "applies the G-space table (sdft_hip_set_nl_form(ctx, 1) first)");
      return SDFT_EINVAL;
    }
    for (int j = 0; j < i; ++j)
      if (blocks[j] == blocks[i]) {
        set_error("a block appears twice in the batch");
        return SDFT_EINVAL;
      }
    const int n = b.ctx->g.n;
    if (n != 24 && n != 32 && n != 36 && n != 48 && n != 64 && n != 96 && n != 128 &&
        n != 192) {
      set_error("unsupported grid size in batch");
      return SDFT_EINVAL;
    }
  }
  sdft_batch* h = new sdft_batch();
  Batch& B = h->B;
  B.device = dev;
  B.prec = prec;
  int rc = 0;
  auto fail = [&](int code) {
    free_batch(B);
    delete h;
    return code;
  };
  hipError_t e = hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&B.stream_nl, hipStreamNonBlocking);
  for (int i = 0; i < 2 && e == hipSuccess; ++i) {
    e = hipEventCreateWithFlags(&B.ev_fork[i], kSyncEventFlags);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&B.ev_join[i], kSyncEventFlags);
  }
  if (e != hipSuccess) return fail(hip_fail(e, "batch stream/event", __FILE__, __LINE__));

  B.frags.resize(nfrag);
  std::vector<int> pc{0};
  std::vector<CopyArgs> ca;
  for (int i = 0; i < nfrag; ++i) {
    Frag& f = B.frags[i];
    f.b = &blocks[i]->b;
    f.c = f.b->ctx;
    f.npairs = f.b->nstate_pad / 2;
    if (!f.c->g.col_intervals) B.fused_ok = B.fused = false;
    if ((rc = dmalloc(&f.W, w_bytes(*f.c, f.npairs, prec)))) return fail(rc);
    const int P = f.c->nproj, S = f.b->nstate_pad;
    f.P = P;
    f.ppad = f.c->ppad;
    if (P > 0) {
      B.nl_frags.push_back(i);
      if ((rc = dmalloc(&f.zpart, sizeof(double) * B.max_zpart_slices * P * S))) return fail(rc);
      const size_t yb = sizeof(double) * (size_t)f.c->ppad * ((S + 31) / 32 * 32);
      if ((rc = dmalloc(&f.Y, yb))) return fail(rc);
      e = hipMemsetAsync(f.Y, 0, yb, B.stream);  // padding columns stay zero
      if (e != hipSuccess) return fail(hip_fail(e, "hipMemsetAsync", __FILE__, __LINE__));
      if ((rc = dmalloc(&f.Z, sizeof(double) * (size_t)P * S))) return fail(rc);
    }
    const long long nelem = (long long)S * 2 * f.c->g.ncp;
    ca.push_back(CopyArgs{f.b->chi, f.b->X0, nelem,
                          (prec == 1 && f.c->nproj > 0) ? f.b->X32 : nullptr});
    pc.push_back(pc.back() + (int)((nelem + 255) / 256));
    // group by grid size
    const int n = f.c->g.n;
    auto it = std::find_if(B.groups.begin(), B.groups.end(), [n](const Group& g) { return g.n == n; });
    if (it == B.groups.end()) {
      B.groups.emplace_back();
      B.groups.back().n = n;
      it = B.groups.end() - 1;
    }
    it->frags.push_back(i);
  }
  B.tot_copy = pc.back();
  if ((rc = dmalloc(&B.pre_copy, sizeof(int) * pc.size()))) return fail(rc);
  if ((rc = dmalloc(&B.copy, sizeof(CopyArgs) * ca.size()))) return fail(rc);
  if ((rc = upload_vec(B.pre_copy, pc, B.stream))) return fail(rc);
  if ((rc = upload_vec(B.copy, ca, B.stream))) return fail(rc);
  for (Group& g : B.groups) {
    std::vector<int> pe{0}, pp{0};
    size_t lds = 0, lds1 = 0;
    for (int fi : g.frags) {
      const Frag& f = B.frags[fi];
      pe.push_back(pe.back() + f.c->g.nchunk * f.npairs);
      pp.push_back(pp.back() + f.c->g.n * f.npairs);
      lds = std::max(lds, plane_lds_for(g.n, prec, f.c->g.M, 0));
      lds1 = std::max(lds1, plane_lds_for(g.n, prec, f.c->g.M, 1));
      if (!plane_prunable(f.c->g.K, g.n)) g.prune = false;
    }
    if (lds > 160 * 1024) {
      set_error("plane buffer exceeds LDS for a fragment grid in this precision");
      return fail(SDFT_EINVAL);
    }
    g.lds = lds;
    g.lds1 = lds1 <= 160 * 1024 ? lds1 : 0;
    g.tot_exp = pe.back();
    g.tot_plane = pp.back();
    const size_t nf = g.frags.size();
    if ((rc = dmalloc(&g.pre_exp, sizeof(int) * pe.size()))) return fail(rc);
    if ((rc = dmalloc(&g.pre_plane, sizeof(int) * pp.size()))) return fail(rc);
    for (int v = 0; v < 2; ++v)
      if ((rc = dmalloc(&g.exp[v], sizeof(ExpandArgs) * nf))) return fail(rc);
    if ((rc = dmalloc(&g.plane, sizeof(PlaneArgs) * nf))) return fail(rc);
    for (int v = 0; v < kVariants; ++v)
      if ((rc = dmalloc(&g.con[v], sizeof(ContractArgs) * nf))) return fail(rc);
    if ((rc = upload_vec(g.pre_exp, pe, B.stream))) return fail(rc);
    if ((rc = upload_vec(g.pre_plane, pp, B.stream))) return fail(rc);
  }
  const size_t nnl = B.nl_frags.size() + 1;
  if ((rc = dmalloc(&B.pre_proj, sizeof(int) * nnl))) return fail(rc);
  if ((rc = dmalloc(&B.pre_red, sizeof(int) * nnl))) return fail(rc);
  if ((rc = dmalloc(&B.pre_force, sizeof(int) * nnl))) return fail(rc);
  for (int v = 0; v < 2; ++v)
    if ((rc = dmalloc(&B.proj[v], sizeof(KbProjectArgs) * nnl))) return fail(rc);
  if ((rc = dmalloc(&B.red, sizeof(KbReduceArgs) * nnl))) return fail(rc);
  if ((rc = dmalloc(&B.force, sizeof(KbForceArgs) * nnl))) return fail(rc);
  if ((rc = dmalloc(&B.counter, 256))) return fail(rc);
  if ((rc = build_kb_tables(B))) return fail(rc);
  *out = h;
  return 0;
}

int sdft_hip_batch_destroy(sdft_batch* h) {
  if (!h) return 0;
  if (h->B.stream) (void)hipStreamSynchronize(h->B.stream);
  free_batch(h->B);
  delete h;
  return 0;
}

int sdft_hip_batch_config(sdft_batch* h, int split, int overlap, int plane_tile,
                          int fused_expand) {
  if (!h || (split != 1 && split != 2 && split != 4 && split != 8) ||
      (overlap != 0 && overlap != 1) || (plane_tile != 0 && plane_tile != 1) ||
      (fused_expand != 0 && fused_expand != 1)) {
    set_error("bad batch configuration");
    return SDFT_EINVAL;
  }
  Batch& B = h->B;
  if (fused_expand && !B.fused_ok) {
    set_error("fused expand needs every fragment's sphere columns to be ka intervals");
    return SDFT_EINVAL;
  }
  B.split = split;
  B.overlap = overlap != 0;
  B.plane_variant = plane_tile;
  B.fused = fused_expand != 0;
  B.graph_valid = false;
  B.full_L = 0;
  return build_kb_tables(B);
}

int sdft_hip_batch_filter(sdft_batch* h, int L, const double* coeffs, const double* scale,
                          const double* emean, int use_graph) {
  if (!h || !coeffs || !scale || !emean) {
    set_error("null argument");
    return SDFT_EINVAL;
  }
  if (L < 2) {
    set_error("batch filter needs L >= 2");
    return SDFT_EINVAL;
  }
  Batch& B = h->B;
  int rc = build_args(B, L, coeffs, scale, emean);
  if (rc) return rc;
  if (use_graph < 0 || use_graph > 2) {
    set_error("use_graph must be 0, 1 or 2");
    return SDFT_EINVAL;
  }
  return run_filter(B, L, use_graph);
}

int sdft_hip_batch_autotune(sdft_batch* h, int steps, int* config_out, int max_cand,
                            double* cand_ms) {
  if (!h || steps < 1) {
    set_error("bad autotune arguments");
    return SDFT_EINVAL;
  }
  Batch& B = h->B;
  const int nf = (int)B.frags.size();
  // Timed steps k >= 2 with c_k = 0 leave the accumulators as they are; a
  // tiny scale and zero shift keep the recurrence x_{k+1} = A x_k - x_{k-1}
  // bounded (|A| << 2), so T_k stays finite however many steps run.
  const int L = steps + 4;
  std::vector<double> co((size_t)nf * L, 0.0), sc(nf, 1e-4), em(nf, 0.0);
  hipEvent_t t0 = nullptr, t1 = nullptr;
  SDFT_HIP_CHECK(hipEventCreate(&t0));
  SDFT_HIP_CHECK(hipEventCreate(&t1));
  const int splits[4] = {1, 2, 4, 8};
  const int nfu = B.fused_ok ? 2 : 1;
  double best = 1e300;
  int best_split = B.split, best_ov = B.overlap ? 1 : 0, best_pv = B.plane_variant;
  bool best_fu = B.fused;
  int ncand = 0;
  int rc = 0;
  if ((rc = copy_chi(B))) return rc;
  for (int pv = 0; pv < 2 && !rc; ++pv)
  for (int ov = 0; ov < 2 && !rc; ++ov)
  for (int si = 0; si < 4 && !rc; ++si) {
    for (int fu = 0; fu < nfu && !rc; ++fu) {
      B.split = splits[si];
      B.overlap = ov != 0;
      B.plane_variant = pv;
      B.fused = fu != 0;
      if ((rc = build_kb_tables(B))) break;
      if ((rc = build_args(B, L, co.data(), sc.data(), em.data()))) break;
      // warm-up: the first step (which expands chi) and one more, so that the
      // timed steps start from the state a filter has at k = 3
      if ((rc = set_counter(B, 1))) break;
      if ((rc = enqueue_step(B, 0, 0))) break;
      if ((rc = enqueue_step(B, 1, 1))) break;
      hipError_t e = hipEventRecord(t0, B.stream);
      for (int k = 0; k < steps && e == hipSuccess && !rc; ++k)
        rc = enqueue_step(B, (k & 1) ? 1 : 2, k & 1);
      if (rc) break;
      if (e == hipSuccess) e = hipEventRecord(t1, B.stream);
      if (e == hipSuccess) e = hipEventSynchronize(t1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
      if (e != hipSuccess) {
        rc = hip_fail(e, "autotune timing", __FILE__, __LINE__);
        break;
      }
      const double per = ms / steps;
      if (cand_ms && ncand < max_cand) cand_ms[ncand] = per;
      ++ncand;
      if (per < best) {
        best = per;
        best_split = splits[si];
        best_ov = ov;
        best_pv = pv;
        best_fu = fu != 0;
      }
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  B.split = best_split;
  B.overlap = best_ov != 0;
  B.plane_variant = best_pv;
  B.fused = best_fu;
  B.graph_valid = false;
  B.full_L = 0;
  int rc2 = build_kb_tables(B);
  if (rc) return rc;
  if (rc2) return rc2;
  if (config_out) {
    config_out[0] = best_split;
    config_out[1] = best_ov;
    config_out[2] = best_pv;
    config_out[3] = best_fu ? 1 : 0;
  }
  return ncand;
}

}  // extern "C"
