// sdft_api.hip -- the C ABI (include/sdft_hip.h): context set-up, host<->device
// transfers and the orchestration of the filter path on one HIP stream.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/sdft_hip.h"
#include "sdft_internal.h"


namespace sdft {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  char buf[512];
  snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  g_err = buf;
  return SDFT_EHIP;
}

// roundFFT, stodft/coeff.c:1556-1600: closest 2^n 3^m to n_in by the
// reference's search order.
int round_fft(int n_in) {
  const double ln2 = std::log(2.0), ln3 = std::log(3.0);
  const double lnn = std::log((double)n_in);
  const int nmax = (int)(lnn / ln2) + 1;
  double best = std::fabs(std::pow(2.0, nmax) - n_in);
  int n = nmax, m = 0;
  for (int i = 0; i < nmax; ++i) {
    const double x = std::fmax((lnn - i * ln2) / ln3, 0.0);
    const int mnow = (int)x + 1;
    const double diff = std::fabs(std::pow(2.0, i) * std::pow(3.0, mnow) - n_in);
    if (diff < best) {
      m = mnow;
      n = i;
      best = diff;
    }
  }
  int out = 1;
  for (int i = 0; i < n; ++i) out *= 2;
  for (int i = 0; i < m; ++i) out *= 3;
  return out;
}

template <typename P>
static int dalloc(P** p, size_t bytes) {
  void* v = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(&v, bytes);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return SDFT_ENOMEM;
  }
  *p = reinterpret_cast<P*>(v);
  return 0;
}

template <typename V>
static int upload(V** dst, const std::vector<V>& src) {
  int rc = dalloc(dst, sizeof(V) * src.size());
  if (rc) return rc;
  SDFT_HIP_CHECK(hipMemcpy(*dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice));
  return 0;
}

static int signed_k(int i, int n) { return i <= n / 2 ? i : i - n; }

// Build the column / chunk / entry tables from the caller's maps.
static int build_geometry(Ctx& c, const sdft_grid_desc* d) {
  Geometry& g = c.g;
  if (d->n1 != d->n2 || d->n1 != d->n3) {
    set_error("only cubic grids n1 = n2 = n3 are supported");
    return SDFT_EINVAL;
  }
  const int n = d->n1;
  if (chunk_cols_for(n) == 0) {
    set_error("grid size must be one of 24,32,36,48,64,96,128,192");
    return SDFT_EINVAL;
  }
  if (d->ncoef < 2 || !d->map || !d->map_conj || !d->ak2_kinetic) {
    set_error("ncoef < 2 or null map / ak2 pointer");
    return SDFT_EINVAL;
  }
  g.n = n;
  g.N = (long long)n * n * n;
  g.ncoef = d->ncoef;
  g.ncp = (g.ncoef + 63) / 64 * 64;
  const int nc = g.ncoef;
  if (d->map[nc - 1] != 0) {
    set_error("the last coefficient must be G=0 (map[ncoef-1] == 0)");
    return SDFT_EINVAL;
  }
  std::vector<int> ka(nc), kb(nc), kc(nc);
  int K = 0;
  for (int i = 0; i < nc; ++i) {
    const long long m = d->map[i];
    if (m < 0 || m >= g.N) {
      set_error("map index out of the grid");
      return SDFT_EINVAL;
    }
    const int ia = (int)(m / ((long long)n * n)), ib = (int)((m / n) % n), ic = (int)(m % n);
    ka[i] = signed_k(ia, n);
    kb[i] = signed_k(ib, n);
    kc[i] = signed_k(ic, n);
    if (i < nc - 1) {
      if (ka[i] == 0 && kb[i] == 0 && kc[i] == 0) {
        set_error("G=0 appears before the last coefficient");
        return SDFT_EINVAL;
      }
      const long long mc = d->map_conj[i];
      const long long want = ((long long)((-ka[i] + n) % n) * n + (-kb[i] + n) % n) * n +
                             (-kc[i] + n) % n;
      if (mc != want) {
        set_error("map_conj is not the index of -G");
        return SDFT_EINVAL;
      }
    }
    K = std::max(K, std::max(std::abs(ka[i]), std::max(std::abs(kb[i]), std::abs(kc[i]))));
  }
  if (2 * K >= n) {
    set_error("plane-wave sphere too large for the grid (need 2K+1 <= n)");
    return SDFT_EINVAL;
  }
  g.K = K;
  g.M = 2 * K + 1;

  // columns (kb, kc) of the full sphere
  std::map<std::pair<int, int>, std::vector<std::tuple<int, int, int>>> cols;  // -> (ka, i, flag)
  for (int i = 0; i < nc - 1; ++i) {
    cols[{kb[i], kc[i]}].push_back(std::make_tuple(ka[i], i, 0));
    cols[{-kb[i], -kc[i]}].push_back(std::make_tuple(-ka[i], i, 1));
  }
  cols[{0, 0}].push_back(std::make_tuple(0, nc - 1, 2));
  // column order: pairs (key, -key) with key positive, then (0,0) last
  std::vector<std::pair<int, int>> order;
  for (auto& kv : cols) {
    const int b = kv.first.first, cc = kv.first.second;
    const bool pos = (b > 0) || (b == 0 && cc > 0);
    if (!pos) continue;
    order.push_back(kv.first);
    if (!cols.count({-b, -cc})) {
      set_error("sphere is not inversion symmetric");
      return SDFT_EINVAL;
    }
    order.push_back({-b, -cc});
  }
  order.push_back({0, 0});
  g.ncol = (int)order.size();
  g.col_intervals = true;
  for (auto& key : order) {
    std::vector<int> kas;
    for (auto& e : cols[key]) kas.push_back(std::get<0>(e));
    std::sort(kas.begin(), kas.end());
    const int h = ((int)kas.size() - 1) / 2;
    for (int t = 0; t < (int)kas.size(); ++t)
      if (kas[t] != t - h) g.col_intervals = false;
  }
  g.ncolp = (g.ncol + 3) / 4 * 4;
  std::map<std::pair<int, int>, int> colid;
  for (int dd = 0; dd < g.ncol; ++dd) colid[order[dd]] = dd;

  // chunks: column pairs packed up to CH columns; (0,0) joins the last chunk
  const int CH = chunk_cols_for(n);
  std::vector<int> chunk_col{0};
  int cur = 0;
  for (int dd = 0; dd + 1 < g.ncol; dd += 2) {
    if (cur + 2 > CH) {
      chunk_col.push_back(dd);
      cur = 0;
    }
    cur += 2;
  }
  if (cur + 1 > CH) {
    chunk_col.push_back(g.ncol - 1);
    cur = 0;
  }
  chunk_col.push_back(g.ncol);
  g.nchunk = (int)chunk_col.size() - 1;
  g.chunk_max = CH;

  // internal coefficient order: chunk by chunk, column by column, +G entries
  // sorted by ka; G=0 last
  g.perm.assign(nc, -1);
  g.iperm.assign(nc, -1);
  std::vector<int> chunk_coef{0};
  int j = 0;
  for (int ch = 0; ch < g.nchunk; ++ch) {
    for (int dd = chunk_col[ch]; dd < chunk_col[ch + 1]; ++dd) {
      auto& ents = cols[order[dd]];
      std::sort(ents.begin(), ents.end());
      for (auto& e : ents) {
        if (std::get<2>(e) != 0) continue;
        const int i = std::get<1>(e);
        g.perm[j] = i;
        g.iperm[i] = j;
        ++j;
      }
    }
    if (ch == g.nchunk - 1) {
      g.perm[j] = nc - 1;
      g.iperm[nc - 1] = j;
      ++j;
    }
    chunk_coef.push_back(j);
  }
  if (j != nc) {
    set_error("internal error: coefficient permutation incomplete");
    return SDFT_EINVAL;
  }
  // entries per column (for the pack), coefficient positions (for the unpack)
  std::vector<int> ent_off{0}, ent_j, ent_b, col_pos(g.ncol);
  std::vector<int> rowcol((size_t)g.M * n, -1);
  std::vector<int> coef_plus(nc), coef_minus(nc);
  const int LD = n + 1;
  for (int dd = 0; dd < g.ncol; ++dd) {
    const int b = order[dd].first, cc = order[dd].second;
    col_pos[dd] = (b + K) * LD + ((cc % n) + n) % n;
    rowcol[(size_t)(b + K) * n + ((cc % n) + n) % n] = dd;
    for (auto& e : cols[order[dd]]) {
      const int a = std::get<0>(e), i = std::get<1>(e), flag = std::get<2>(e);
      const int ia = ((a % n) + n) % n;
      ent_j.push_back(g.iperm[i]);
      ent_b.push_back((dd << 10) | (flag << 8) | ia);
      if (flag == 0 || flag == 2) coef_plus[g.iperm[i]] = (dd << 8) | ia;
      if (flag == 1 || flag == 2) coef_minus[g.iperm[i]] = (dd << 8) | ia;
    }
    ent_off.push_back((int)ent_j.size());
  }
  g.ak2.assign(g.ncp, 0.0);
  for (int jj = 0; jj < nc; ++jj) g.ak2[jj] = d->ak2_kinetic[g.perm[jj]];

  std::vector<double> tw64(2 * n);
  std::vector<float> tw32(2 * n);
  for (int m = 0; m < n; ++m) {
    const long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)m /
                            (long double)n;
    tw64[2 * m] = (double)cosl(ang);
    tw64[2 * m + 1] = (double)sinl(ang);
    tw32[2 * m] = (float)cosl(ang);
    tw32[2 * m + 1] = (float)sinl(ang);
  }
  int rc;
  if ((rc = upload(&c.t.col_pos, col_pos))) return rc;
  if ((rc = upload(&c.t.rowcol, rowcol))) return rc;
  if ((rc = upload(&c.t.ent_off, ent_off))) return rc;
  if ((rc = upload(&c.t.ent_j, ent_j))) return rc;
  if ((rc = upload(&c.t.ent_b, ent_b))) return rc;
  if ((rc = upload(&c.t.chunk_col, chunk_col))) return rc;
  if ((rc = upload(&c.t.chunk_coef, chunk_coef))) return rc;
  if ((rc = upload(&c.t.coef_plus, coef_plus))) return rc;
  if ((rc = upload(&c.t.coef_minus, coef_minus))) return rc;
  if ((rc = upload(&c.t.ak2, g.ak2))) return rc;
  double* t64 = nullptr;
  float* t32 = nullptr;
  if ((rc = upload(&t64, tw64))) return rc;
  if ((rc = upload(&t32, tw32))) return rc;
  c.t.tw64 = t64;
  c.t.tw32 = t32;
  return 0;
}

static void free_ctx(Ctx& c) {
  rs_free(c);
  if (c.dens) {
    free_ctx(*c.dens);
    delete c.dens;
    c.dens = nullptr;
  }
  void* ptrs[] = {c.ks_field, c.ks_rho, c.ks_vhe, c.ks_part, c.ks_io, c.ks_q,c.t.col_pos,  c.t.rowcol,     c.t.ent_off,   c.t.ent_j,      c.t.ent_b,
                  c.t.chunk_col,
                  c.t.chunk_coef, c.t.coef_plus, c.t.coef_minus, c.t.ak2,       c.t.tw64,
                  c.t.tw32,     c.vks64,        c.vks32,       c.A64,          c.A32,
                  c.At64,       c.At32,
                  c.E,          c.W,            c.W2,          c.zpart,        c.Y,           c.Y32,
                  c.Z,          c.red};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c.ev_cur) (void)hipEventDestroy(c.ev_cur);
  if (c.ev_nl) (void)hipEventDestroy(c.ev_nl);
  if (c.stream_nl) (void)hipStreamDestroy(c.stream_nl);
  if (c.stream) (void)hipStreamDestroy(c.stream);
}

// ---- profiling ------------------------------------------------------------
struct ProfRec {
  int cls;
  hipEvent_t a, b;
};
static thread_local std::vector<ProfRec> g_prof;
static thread_local std::vector<hipEvent_t> g_evpool;

static hipEvent_t ev_get() {
  if (!g_evpool.empty()) {
    hipEvent_t e = g_evpool.back();
    g_evpool.pop_back();
    return e;
  }
  hipEvent_t e;
  (void)hipEventCreate(&e);
  return e;
}

static int prof_class(Ctx& c, const char* name) {
  for (size_t i = 0; i < c.prof_names.size(); ++i)
    if (c.prof_names[i] == name) return (int)i;
  c.prof_names.push_back(name);
  c.prof_ms.push_back(0.0);
  c.prof_launch.push_back(0);
  return (int)c.prof_names.size() - 1;
}

struct ProfScope {
  Ctx& c;
  int cls = -1;
  hipEvent_t a{}, b{};
  ProfScope(Ctx& cc, const char* name) : c(cc) {
    if (!c.profiling) return;
    cls = prof_class(c, name);
    a = ev_get();
    b = ev_get();
    (void)hipEventRecord(a, c.stream);
  }
  ~ProfScope() {
    if (cls < 0) return;
    (void)hipEventRecord(b, c.stream);
    g_prof.push_back({cls, a, b});
  }
};

static void prof_reset(Ctx& c) {
  for (size_t i = 0; i < c.prof_ms.size(); ++i) {
    c.prof_ms[i] = 0.0;
    c.prof_launch[i] = 0;
  }
}

static int prof_collect(Ctx& c) {
  if (g_prof.empty()) return 0;
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  for (auto& r : g_prof) {
    float ms = 0.f;
    SDFT_HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
    c.prof_ms[r.cls] += ms;
    c.prof_launch[r.cls] += 1;
    g_evpool.push_back(r.a);
    g_evpool.push_back(r.b);
  }
  g_prof.clear();
  return 0;
}

// ---- scratch --------------------------------------------------------------
static int ensure_W(Ctx& c, size_t bytes) {
  if (c.W_bytes >= bytes) return 0;
  if (c.W) SDFT_HIP_CHECK(hipFree(c.W));
  c.W = nullptr;
  c.W_bytes = 0;
  int rc = dalloc(&c.W, bytes);
  if (rc) return rc;
  c.W_bytes = bytes;
  return 0;
}

// ---- block helpers -------------------------------------------------------
static int block_alloc(Ctx& c, int nstate, int prec, Block& b) {
  b.ctx = &c;
  b.nstate = nstate;
  b.nstate_pad = nstate + (nstate & 1);
  b.prec = prec;
  const size_t bytes = sizeof(double) * (size_t)b.nstate_pad * 2 * c.g.ncp;
  // the T buffers are the KB projection's column operand: rows padded to a
  // multiple of 32 (zeros, never written) so that its K loop needs no bounds
  const size_t xbytes = sizeof(double) * (size_t)((b.nstate_pad + 31) / 32 * 32) * 2 * c.g.ncp;
  int rc;
  if ((rc = dalloc(&b.X0, xbytes))) return rc;
  if ((rc = dalloc(&b.X1, xbytes))) return rc;
  if ((rc = dalloc(&b.acc, bytes))) return rc;
  if ((rc = dalloc(&b.chi, bytes))) return rc;
  if ((rc = dalloc(&b.fnl, bytes))) return rc;
  SDFT_HIP_CHECK(hipMemsetAsync(b.X0, 0, xbytes, c.stream));
  SDFT_HIP_CHECK(hipMemsetAsync(b.X1, 0, xbytes, c.stream));
  SDFT_HIP_CHECK(hipMemsetAsync(b.acc, 0, bytes, c.stream));
  SDFT_HIP_CHECK(hipMemsetAsync(b.chi, 0, bytes, c.stream));
  SDFT_HIP_CHECK(hipMemsetAsync(b.fnl, 0, bytes, c.stream));
  if (prec == 1) {
    if ((rc = dalloc(&b.X32, xbytes / 2))) return rc;
    SDFT_HIP_CHECK(hipMemsetAsync(b.X32, 0, xbytes / 2, c.stream));
  }
  b.cur = 0;
  return 0;
}

static void block_free(Block& b) {
  void* ptrs[] = {b.X0, b.X1, b.acc, b.chi, b.fnl, b.mom_part, b.X32};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  b = Block();
}

// host (reference order, re/im arrays) -> device internal layout
static int block_put(Block& b, double* dst, const double* re, const double* im) {
  Ctx& c = *b.ctx;
  const int nc = c.g.ncoef, ncp = c.g.ncp;
  std::vector<double> h((size_t)b.nstate_pad * 2 * ncp, 0.0);
  for (int s = 0; s < b.nstate; ++s) {
    double* row = h.data() + (size_t)s * 2 * ncp;
    const double* r = re + (size_t)s * nc;
    const double* i = im + (size_t)s * nc;
    for (int j = 0; j < nc; ++j) {
      row[j] = r[c.g.perm[j]];
      row[ncp + j] = i[c.g.perm[j]];
    }
  }
  SDFT_HIP_CHECK(hipMemcpyAsync(dst, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice,
                                c.stream));
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

static int block_fetch(Block& b, const double* src, double* re, double* im) {
  Ctx& c = *b.ctx;
  const int nc = c.g.ncoef, ncp = c.g.ncp;
  std::vector<double> h((size_t)b.nstate_pad * 2 * ncp);
  SDFT_HIP_CHECK(hipMemcpyAsync(h.data(), src, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                                c.stream));
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  for (int s = 0; s < b.nstate; ++s) {
    const double* row = h.data() + (size_t)s * 2 * ncp;
    double* r = re + (size_t)s * nc;
    double* i = im + (size_t)s * nc;
    for (int j = 0; j < nc; ++j) {
      r[c.g.perm[j]] = row[j];
      i[c.g.perm[j]] = row[ncp + j];
    }
  }
  return 0;
}

__global__ void k_to_f32(const double* __restrict__ x, float* __restrict__ y, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = float(x[i]);
}

// apply_step flags: state a previous step of the same loop left behind, and
// what this step leaves for the next one.
enum StepFlags {
  kHaveW = 1,    // W holds the a-expansion of `cur` (no expand kernel)
  kEmitW = 2,    // the contract writes the a-expansion of `out` into W
  kHaveX32 = 4,  // FP32 form: X32 holds `cur` in float (no conversion)
  kEmitX32 = 8   // FP32 form: the contract writes `out` in float into X32
};

// One application of the H sandwich to `cur`; `mode` selects the epilogue,
// `flags` (StepFlags) the state carried between the steps of a loop.
static int apply_step(Block& b, int mode, const double* cur, double* out, const StepParams& sp,
                      double* mom_part, int flags = 0) {
  Ctx& c = *b.ctx;
  const int npairs = b.nstate_pad / 2;
  int rc;
  if ((rc = ensure_W(c, w_bytes(c, npairs, b.prec)))) return rc;
  const void* fnl = nullptr;
  // real-space KB in its per-atom form: applied on the real-space grid
  // between the inverse and forward plane stages, no G-space products
  const bool rs = rs_active(c);
  const bool table = c.nproj > 0 && !rs;
  // the FP32 form's KB products read T in float
  const bool x32 = b.prec == 1 && table;
  if (x32 && !(flags & kHaveX32)) {
    const size_t n = (size_t)b.nstate_pad * 2 * c.g.ncp;
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, cur,
                       b.X32, n);
    SDFT_HIP_CHECK(hipGetLastError());
  }
  // The KB chain runs on its own stream, overlapping the FFT kernels; with
  // profiling on, each kernel's events are recorded on the stream it runs on.
  const bool overlap = table;
  if (table) {
    hipStream_t main = c.stream;
    if (overlap) {
      SDFT_HIP_CHECK(hipEventRecord(c.ev_cur, c.stream));
      SDFT_HIP_CHECK(hipStreamWaitEvent(c.stream_nl, c.ev_cur, 0));
      c.stream = c.stream_nl;
    }
    {
      ProfScope ps(c, "kb_project");
      rc = launch_nl_project(c, b.prec, x32 ? (const void*)b.X32 : (const void*)cur,
                             b.nstate_pad);
    }
    if (!rc) {
      ProfScope ps(c, "kb_force");
      rc = launch_nl_force(c, b.prec, b.nstate_pad, b.fnl);
    }
    if (overlap) {
      c.stream = main;
      SDFT_HIP_CHECK(hipEventRecord(c.ev_nl, c.stream_nl));
    }
    if (rc) return rc;
    fnl = b.fnl;
  }
  if (!(flags & kHaveW)) {
    ProfScope ps(c, "expand_a");
    if ((rc = launch_expand(c, b.prec, cur, npairs, c.W))) return rc;
  }
  void* Wres = c.W;
  if (rs) {
    {
      ProfScope ps(c, "plane_bc_to_r");
      if ((rc = rs_to_grid(c, b.prec, c.W, npairs))) return rc;
    }
    {
      ProfScope ps(c, "rs_project");
      if ((rc = rs_project(c, b.prec, npairs))) return rc;
    }
    {
      ProfScope ps(c, "rs_apply_vpsi");
      if ((rc = rs_apply(c, b.prec, npairs))) return rc;
    }
    {
      ProfScope ps(c, "plane_r_to_bc");
      if ((rc = rs_from_grid(c, b.prec, c.W, npairs))) return rc;
    }
  } else {
    ProfScope ps(c, "plane_bc_vpsi");
    if ((rc = launch_plane(c, b.prec, c.W, npairs, &Wres))) return rc;
  }
  if (overlap) SDFT_HIP_CHECK(hipStreamWaitEvent(c.stream, c.ev_nl, 0));
  {
    ProfScope ps(c, "contract_a");
    if ((rc = launch_contract(c, b.prec, mode, Wres, npairs, cur, out, b.acc, fnl, b.chi, sp,
                              mom_part, (flags & kEmitW) ? c.W : nullptr,
                              (x32 && (flags & kEmitX32)) ? b.X32 : nullptr)))
      return rc;
  }
  return 0;
}

// apply_step flags for step k of a loop over 1 <= k < L: from the second
// step on, W (when the contract may re-expand: interval columns, and the
// tuner has not switched it off) and X32 hold what the previous contract
// wrote, and every step but the last writes them for its output.
static int fuse_flags(const Ctx& c, int k, int L) {
  int f = (k > 1 ? kHaveX32 : 0) | (k + 1 < L ? kEmitX32 : 0);
  if (c.g.col_intervals && c.fuse_expand) f |= (k > 1 ? kHaveW : 0) | (k + 1 < L ? kEmitW : 0);
  return f;
}

static int check_vks(Ctx& c) {
  if (!c.have_vks) {
    set_error("v_KS not set (call sdft_hip_set_vks first)");
    return SDFT_ESTATE;
  }
  return 0;
}

static int ensure_mom(Block& b, int nsteps) {
  Ctx& c = *b.ctx;
  const size_t nwg = (size_t)c.g.nchunk * (b.nstate_pad / 2);
  const size_t need = sizeof(double) * 3 * nwg * (size_t)nsteps;
  if (b.mom_part_bytes >= need) return 0;
  if (b.mom_part) SDFT_HIP_CHECK(hipFree(b.mom_part));
  b.mom_part = nullptr;
  b.mom_part_bytes = 0;
  int rc = dalloc(&b.mom_part, need);
  if (rc) return rc;
  b.mom_part_bytes = need;
  return 0;
}

__global__ void k_scale_copy(const double* __restrict__ x, double a, double* __restrict__ y,
                             size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i];
}

// genNoiseOrbitalReal on the device: block X0 and chi receive the noise.
static int block_noise_impl(Block& b, long long seed, long long first) {
  Ctx& c = *b.ctx;
  if (seed < 0 || seed >= (1LL << 20) || first < 0 || first + b.nstate_pad > (1LL << 21) ||
      c.g.N > (1LL << 23)) {
    set_error("noise key out of range (seed < 2^20, orbitals < 2^21, grid < 2^23)");
    return SDFT_EINVAL;
  }
  const int npairs = b.nstate_pad / 2;
  int rc = ensure_W(c, w_bytes(c, npairs, 0));
  if (!rc) rc = launch_noise_plane(c, c.W, npairs, (unsigned long long)seed, (int)first, b.nstate);
  StepParams sp{};
  if (!rc)
    rc = launch_contract(c, 0, kModeNoise, c.W, npairs, b.X0, b.X0, b.chi, nullptr, nullptr, sp,
                         nullptr);
  if (!rc) {
    b.cur = 0;
    SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  }
  return rc;
}

// The Chebyshev filter on resident data (filter.c:318-706).
static int block_filter_impl(Block& b, int L, const double* coeffs, double scale, double emean,
                             double* moments) {
  Ctx& c = *b.ctx;
  int rc;
  if ((rc = check_vks(c))) return rc;
  if (L < 1) {
    set_error("polynomial length must be >= 1");
    return SDFT_EINVAL;
  }
  if (moments && L < 2) {
    set_error("moments need L >= 2");
    return SDFT_EINVAL;
  }
  double* X[2] = {b.X0, b.X1};
  double* T0 = X[b.cur];
  double* T1 = X[1 - b.cur];
  const size_t n = (size_t)b.nstate_pad * 2 * c.g.ncp;
  if (L == 1) {
    hipLaunchKernelGGL(k_scale_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, T0,
                       coeffs[0], b.acc, n);
    SDFT_HIP_CHECK(hipGetLastError());
    SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
    return 0;
  }
  if (moments && (rc = ensure_mom(b, L))) return rc;
  prof_reset(c);
  double* bufs[2] = {T0, T1};
  for (int k = 1; k < L; ++k) {
    StepParams sp;
    const double m = (k == 1) ? 1.0 : 2.0;
    sp.s1 = -scale * 0.25 * m;
    sp.s2 = -scale * 0.5 * m;
    sp.p = -scale * emean * m;
    sp.ck = coeffs[k];
    sp.c0 = coeffs[0];
    sp.first = (k == 1);
    sp.step = k;
    double* cur = bufs[(k - 1) & 1];
    double* out = bufs[k & 1];
    if ((rc = apply_step(b, kModeFilter, cur, out, sp, moments ? b.mom_part : nullptr,
                         fuse_flags(c, k, L))))
      return rc;
  }
  // T_{L-1} sits in bufs[(L-1) & 1]
  b.cur = (bufs[(L - 1) & 1] == b.X0) ? 0 : 1;
  if ((rc = prof_collect(c))) return rc;
  if (moments) {
    const int nwg = c.g.nchunk * (b.nstate_pad / 2);
    std::vector<double> red((size_t)3 * L);
    if ((rc = moments_reduce(c, b.mom_part, L, nwg, red.data()))) return rc;
    moments[0] = red[3 * 1 + 1];
    for (int k = 1; k < L; ++k) moments[k] = red[3 * k + 0];
  }
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  return 0;
}

// G-space KB table from Q_p(G) in the caller's order (see sdft_hip_set_projectors).
static int set_projectors_impl(Ctx& c, int nproj, const double* pre, const double* pim,
                               const double* pe) {
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream_nl));
  void* old[] = {c.A64, c.A32, c.At64, c.At32, c.E};
  for (void* p : old)
    if (p) (void)hipFree(p);
  c.A64 = nullptr;
  c.A32 = nullptr;
  c.At64 = nullptr;
  c.At32 = nullptr;
  c.E = nullptr;
  c.nproj = 0;
  c.ppad = 0;
  if (nproj == 0) return 0;
  const int nc = c.g.ncoef, ncp = c.g.ncp;
  const size_t K = 2 * (size_t)ncp;
  const int ppad = (nproj + 31) / 32 * 32;  // zero rows up to ppad (GEMM row padding)
  std::vector<double> A((size_t)ppad * K, 0.0);
  std::vector<float> A32((size_t)ppad * K, 0.f);
  for (int p = 0; p < nproj; ++p) {
    for (int j = 0; j < nc; ++j) {
      const int i = c.g.perm[j];
      const bool g0 = (j == nc - 1);
      const double w = g0 ? 1.0 : 2.0;
      const double qr = pre[(size_t)p * nc + i], qi = pim[(size_t)p * nc + i];
      A[(size_t)p * K + j] = w * qr;
      A[(size_t)p * K + ncp + j] = g0 ? 0.0 : -w * qi;
    }
  }
  for (size_t i = 0; i < A.size(); ++i) A32[i] = (float)A[i];
  std::vector<double> At(K * ppad, 0.0);
  std::vector<float> At32(K * ppad, 0.f);
  for (int p = 0; p < nproj; ++p)
    for (size_t k = 0; k < K; ++k) {
      At[k * ppad + p] = A[(size_t)p * K + k];
      At32[k * ppad + p] = A32[(size_t)p * K + k];
    }
  int rc;
  if ((rc = upload(&c.A64, A))) return rc;
  if ((rc = upload(&c.A32, A32))) return rc;
  if ((rc = upload(&c.At64, At))) return rc;
  if ((rc = upload(&c.At32, At32))) return rc;
  c.ppad = ppad;
  std::vector<double> E(pe, pe + nproj);
  if ((rc = upload(&c.E, E))) return rc;
  c.E_host = E;
  c.nproj = nproj;
  return 0;
}


// Real-space KB projectors -> the G-space table they restrict to on the
// plane-wave set: Q_p(G) = N^(-1/2) sum_r beta_p(r) e^{iG.r}, energies
// energy_p vol_elem.  Pairs of projectors are scattered onto dense complex
// grids (re = p even, im = p odd) in batches, transformed with the noise
// path's pruned forward transform (c(G) = (1/N) sum_r f(r) e^{-iG.r} for both
// real fields) and conjugated: Q = sqrt(N) conj(c).
static int real_table_impl(Ctx& c) {
  int nproj;
  const long long* off;
  const int* idx;
  const double *val, *energy;
  double vol_elem;
  rs_host_lists(c, &nproj, &off, &idx, &val, &energy, &vol_elem);
  if (nproj == 0) return set_projectors_impl(c, 0, nullptr, nullptr, nullptr);
  const long long N = c.g.N;
  const long long nent = off[nproj];
  int rc = 0;
  long long* doff = nullptr;
  int* didx = nullptr;
  double* dval = nullptr;
  double* field = nullptr;
  const int pairs = (nproj + 1) / 2;
  const size_t pair_bytes = sizeof(double) * 2 * (size_t)N;
  int bq = (int)std::max<size_t>(1, std::min<size_t>((size_t)pairs, (1ull << 30) / pair_bytes));
  std::vector<double> qre((size_t)nproj * c.g.ncoef), qim((size_t)nproj * c.g.ncoef);
  std::vector<double> bre, bim;
  Block tmp;
  const double sq = std::sqrt((double)N);
  do {
    if ((rc = dalloc(&doff, sizeof(long long) * (nproj + 1)))) break;
    if ((rc = dalloc(&didx, sizeof(int) * std::max<long long>(1, nent)))) break;
    if ((rc = dalloc(&dval, sizeof(double) * std::max<long long>(1, nent)))) break;
    if ((rc = dalloc(&field, pair_bytes * bq))) break;
    SDFT_HIP_TRY(rc, hipMemcpyAsync(doff, off, sizeof(long long) * (nproj + 1),
                                    hipMemcpyHostToDevice, c.stream));
    if (rc) break;
    if (nent > 0) {
      SDFT_HIP_TRY(rc, hipMemcpyAsync(didx, idx, sizeof(int) * nent, hipMemcpyHostToDevice,
                                      c.stream));
      if (rc) break;
      SDFT_HIP_TRY(rc, hipMemcpyAsync(dval, val, sizeof(double) * nent, hipMemcpyHostToDevice,
                                      c.stream));
      if (rc) break;
    }
    if ((rc = block_alloc(c, 2 * bq, 0, tmp))) break;
    if ((rc = ensure_W(c, w_bytes(c, bq, 0)))) break;
    bre.resize((size_t)2 * bq * c.g.ncoef);
    bim.resize((size_t)2 * bq * c.g.ncoef);
    for (int q0 = 0; q0 < pairs && !rc; q0 += bq) {
      const int nq = std::min(bq, pairs - q0);
      const int p0 = 2 * q0, p1 = std::min(nproj, 2 * (q0 + nq));
      SDFT_HIP_TRY(rc, hipMemsetAsync(field, 0, pair_bytes * nq, c.stream));
      if (rc) break;
      if ((rc = launch_scatter_proj(c, doff, didx, dval, p0, p1, field))) break;
      if ((rc = launch_grid_plane(c, c.W, nq, field))) break;
      StepParams sp{};
      if ((rc = launch_contract(c, 0, kModeNoise, c.W, nq, tmp.X0, tmp.X0, tmp.chi, nullptr,
                                nullptr, sp, nullptr)))
        break;
      if ((rc = block_fetch(tmp, tmp.X0, bre.data(), bim.data()))) break;
      for (int p = p0; p < p1; ++p) {
        const double* r = bre.data() + (size_t)(p - p0) * c.g.ncoef;
        const double* i = bim.data() + (size_t)(p - p0) * c.g.ncoef;
        double* qr = qre.data() + (size_t)p * c.g.ncoef;
        double* qi = qim.data() + (size_t)p * c.g.ncoef;
        for (int j = 0; j < c.g.ncoef; ++j) {
          qr[j] = sq * r[j];
          qi[j] = -sq * i[j];
        }
      }
    }
  } while (0);
  block_free(tmp);
  void* ptrs[] = {doff, didx, dval, field};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (rc) return rc;
  std::vector<double> e(nproj);
  for (int p = 0; p < nproj; ++p) e[p] = energy[p] * vol_elem;
  return set_projectors_impl(c, nproj, qre.data(), qim.data(), e.data());
}

// The G-space table of the current real-space projectors, if not built yet
// (the force path and nl_form 1 use it).
static int ensure_real_table(Ctx& c) {
  if (!c.rs || c.nproj > 0) return 0;
  return real_table_impl(c);
}

// Real-space projectors: validated and kept in their per-atom form; any
// G-space table is dropped (rebuilt from these lists on demand).
static int set_projectors_real_impl(Ctx& c, int nproj, const long long* off, const int* idx,
                                    const double* val, const double* energy, double vol_elem) {
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream_nl));
  int rc = set_projectors_impl(c, 0, nullptr, nullptr, nullptr);
  if (!rc) rc = rs_set(c, nproj, off, idx, val, energy, vol_elem);
  if (!rc && c.nl_form == 1) rc = ensure_real_table(c);
  return rc;
}

// Real-space KB forces in the per-atom form (nlppKBRealEnergyForce): the
// orbitals' grids, then per projector S_p,j = sum_s Z_ps Zd_ps,j with
// Zd = sum_r (d beta_p / d r_j) psi_s; F_a,j = (occ/nsto) sum_{p in a}
// 2 E_p dV S_p,j / N (d beta / d R_a = -d beta / d r).
static int nl_forces_real(Ctx& c, int nstate, const double* re, const double* im,
                          const int* proj_atom, int natm, double occ, double nsto,
                          double* forces) {
  int nproj;
  const long long* off;
  const int* idx;
  const double *val, *energy;
  double vol_elem;
  rs_host_lists(c, &nproj, &off, &idx, &val, &energy, &vol_elem);
  for (int p = 0; p < nproj; ++p)
    if (proj_atom[p] < 0 || proj_atom[p] >= natm) {
      set_error("projector atom index out of range");
      return SDFT_EINVAL;
    }
  Block b;
  int rc = block_alloc(c, nstate, 0, b);
  if (!rc) rc = block_put(b, b.X0, re, im);
  const int npairs = b.nstate_pad / 2;
  if (!rc) rc = ensure_W(c, w_bytes(c, npairs, 0));
  if (!rc) rc = launch_expand(c, 0, b.X0, npairs, c.W);
  if (!rc) rc = rs_to_grid(c, 0, c.W, npairs);
  std::vector<double> S((size_t)nproj * 3);
  if (!rc) rc = rs_force_sums(c, npairs, S.data());
  if (!rc) {
    const double N = (double)c.g.N;
    for (int p = 0; p < nproj; ++p)
      for (int j = 0; j < 3; ++j)
        forces[proj_atom[p] * 3 + j] +=
            2.0 * energy[p] * vol_elem * S[(size_t)p * 3 + j] / N * occ / nsto;
  }
  block_free(b);
  return rc;
}

}  // namespace sdft

using namespace sdft;

extern "C" {

const char* sdft_hip_last_error(void) { return g_err.c_str(); }

int sdft_hip_init(const sdft_grid_desc* desc, sdft_ctx** out) {
  if (!desc || !out) {
    set_error("null argument");
    return SDFT_EINVAL;
  }
  *out = nullptr;
  sdft_ctx* h = new sdft_ctx();
  Ctx& c = h->c;
  hipError_t e = hipGetDevice(&c.device);
  if (e != hipSuccess) {
    delete h;
    return hip_fail(e, "hipGetDevice", __FILE__, __LINE__);
  }
  e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c.stream_nl, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c.ev_cur, kSyncEventFlags);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c.ev_nl, kSyncEventFlags);
  if (e != hipSuccess) {
    free_ctx(c);
    delete h;
    return hip_fail(e, "hipStreamCreate", __FILE__, __LINE__);
  }
  int rc = build_geometry(c, desc);
  if (!rc) rc = noalloc_scratch_setup(c.device);
  if (rc) {
    free_ctx(c);
    delete h;
    return rc;
  }
  *out = h;
  return 0;
}

int sdft_hip_finalize(sdft_ctx* h) {
  if (!h) return 0;
  (void)hipStreamSynchronize(h->c.stream);
  free_ctx(h->c);
  delete h;
  return 0;
}

int sdft_hip_set_vks(sdft_ctx* h, const double* v) {
  if (!h || !v) {
    set_error("null argument");
    return SDFT_EINVAL;
  }
  Ctx& c = h->c;
  const size_t N = (size_t)c.g.N;
  for (size_t i = 0; i < N; ++i)
    if (!std::isfinite(v[i])) {
      set_error("v_KS contains a non-finite value");
      return SDFT_EINVAL;
    }
  int rc;
  if (!c.vks64 && (rc = dalloc(&c.vks64, sizeof(double) * N))) return rc;
  if (!c.vks32 && (rc = dalloc(&c.vks32, sizeof(float) * N))) return rc;
  std::vector<float> v32(N);
  for (size_t i = 0; i < N; ++i) v32[i] = (float)v[i];
  SDFT_HIP_CHECK(hipMemcpy(c.vks64, v, sizeof(double) * N, hipMemcpyHostToDevice));
  SDFT_HIP_CHECK(hipMemcpy(c.vks32, v32.data(), sizeof(float) * N, hipMemcpyHostToDevice));
  c.have_vks = true;
  return 0;
}

int sdft_hip_set_projectors(sdft_ctx* h, int nproj, const double* pre, const double* pim,
                            const double* pe) {
  if (!h || nproj < 0 || (nproj > 0 && (!pre || !pim || !pe))) {
    set_error("bad projector arguments");
    return SDFT_EINVAL;
  }
  rs_free(h->c);  // a G-space table replaces any real-space projectors
  return set_projectors_impl(h->c, nproj, pre, pim, pe);
}

int sdft_hip_set_nl_form(sdft_ctx* h, int form) {
  if (!h || (form != 0 && form != 1)) {
    set_error("bad nonlocal form (0 per-atom real space, 1 G-space table)");
    return SDFT_EINVAL;
  }
  Ctx& c = h->c;
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  c.nl_form = form;
  return form == 1 ? ensure_real_table(c) : 0;
}

int sdft_hip_set_projectors_real(sdft_ctx* h, int nproj, const long long* offsets,
                                 const int* grid_index, const double* values,
                                 const double* energy, double vol_elem) {
  if (!h || nproj < 0 || (nproj > 0 && (!offsets || !grid_index || !values || !energy)) ||
      !(vol_elem > 0)) {
    set_error("bad real-space projector arguments");
    return SDFT_EINVAL;
  }
  return set_projectors_real_impl(h->c, nproj, offsets, grid_index, values, energy, vol_elem);
}
int sdft_hip_set_projector_gradients(sdft_ctx* h, int nproj, const double* dx, const double* dy,
                                     const double* dz) {
  if (!h || nproj < 1 || !dx || !dy || !dz) {
    set_error("bad projector gradient arguments");
    return SDFT_EINVAL;
  }
  return rs_set_gradients(h->c, nproj, dx, dy, dz);
}

int sdft_hip_nl_forces(sdft_ctx* h, int nstate, const double* re, const double* im,
                       const double* gvec, const int* proj_atom, int natm, double occ,
                       double nsto, double* forces) {
  if (!h || nstate < 1 || !re || !im || !gvec || !proj_atom || natm < 1 || !forces ||
      !(nsto > 0)) {
    set_error("bad force arguments");
    return SDFT_EINVAL;
  }
  Ctx& c = h->c;
  for (int a = 0; a < natm * 3; ++a) forces[a] = 0.0;
  if (rs_has_gradients(c)) return nl_forces_real(c, nstate, re, im, proj_atom, natm, occ, nsto,
                                                 forces);
  // real-space projectors without gradient tables: the forces use the
  // G-space table they restrict to
  int rc0 = ensure_real_table(c);
  if (rc0) return rc0;
  if (c.nproj == 0) return 0;
  const int P = c.nproj;
  for (int p = 0; p < P; ++p)
    if (proj_atom[p] < 0 || proj_atom[p] >= natm) {
      set_error("projector atom index out of range");
      return SDFT_EINVAL;
    }
  Block b;
  int rc = block_alloc(c, nstate, 0, b);
  if (!rc) rc = block_put(b, b.X0, re, im);
  const int ns = b.nstate_pad, ncp = c.g.ncp;
  std::vector<double> gint((size_t)3 * ncp, 0.0);
  for (int j = 0; j < c.g.ncoef; ++j)
    for (int d = 0; d < 3; ++d) gint[(size_t)d * ncp + j] = gvec[(size_t)c.g.perm[j] * 3 + d];
  double *dg = nullptr, *Ad = nullptr, *Zd = nullptr, *part = nullptr, *S = nullptr;
  int per = 0;
  const int nsplit = kb_project_split(c, ns, &per);
  if (!rc) rc = upload(&dg, gint);
  const size_t adbytes = sizeof(double) * (size_t)c.ppad * 2 * ncp;  // zero padding rows
  if (!rc) rc = dalloc(&Ad, adbytes);
  if (!rc && hipMemsetAsync(Ad, 0, adbytes, c.stream) != hipSuccess) {
    set_error("hipMemsetAsync failed");
    rc = SDFT_EHIP;
  }
  if (!rc) rc = dalloc(&Zd, sizeof(double) * (size_t)3 * P * ns);
  if (!rc) rc = dalloc(&part, sizeof(double) * (size_t)nsplit * P * ns);
  if (!rc) rc = dalloc(&S, sizeof(double) * (size_t)3 * P);
  if (!rc) rc = nl_force_sums(c, b.X0, ns, nstate, dg, Ad, Zd, part, S);
  std::vector<double> hs((size_t)3 * P);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(hs.data(), S, sizeof(double) * 3 * P, hipMemcpyDeviceToHost,
                                  c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) rc = hip_fail(e, "force download", __FILE__, __LINE__);
  }
  if (!rc) {
    // F_a = -(occ/nsto) sum_{p in a} 2 E_p sum_s Z dZ/dR (calc-nuclei-force.c:462-478 scaling)
    for (int p = 0; p < P; ++p)
      for (int d = 0; d < 3; ++d)
        forces[proj_atom[p] * 3 + d] -= 2.0 * c.E_host[p] * hs[(size_t)p * 3 + d] * occ / nsto;
  }
  void* ptrs[] = {dg, Ad, Zd, part, S};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  block_free(b);
  return rc;
}

int sdft_hip_hpsi(sdft_ctx* h, int nstate, const double* cre, const double* cim, double* fre,
                  double* fim, int prec) {
  if (!h || nstate < 1 || !cre || !cim || !fre || !fim || (prec != 0 && prec != 1)) {
    set_error("bad hpsi arguments");
    return SDFT_EINVAL;
  }
  Ctx& c = h->c;
  int rc;
  if ((rc = check_vks(c))) return rc;
  Block b;
  if ((rc = block_alloc(c, nstate, prec, b))) {
    block_free(b);
    return rc;
  }
  StepParams sp{};
  if (!(rc = block_put(b, b.X0, cre, cim)) && !(rc = apply_step(b, kModeForce, b.X0, b.X1, sp,
                                                                  nullptr)))
    rc = block_fetch(b, b.X1, fre, fim);
  if (!rc) rc = prof_collect(c);
  block_free(b);
  return rc;
}

int sdft_hip_tune(sdft_ctx* h, int nstate, int prec, int steps, int* config_out, int max_cand,
                  double* cand_ms) {
  if (!h || nstate < 1 || steps < 1 || (prec != 0 && prec != 1)) {
    set_error("bad tune arguments");
    return SDFT_EINVAL;
  }
  Ctx& c = h->c;
  int rc;
  if ((rc = check_vks(c))) return rc;
  Block b;
  if ((rc = block_alloc(c, nstate, prec, b))) {
    block_free(b);
    return rc;
  }
  rc = block_noise_impl(b, 1, 0);
  const bool prof = c.profiling;
  c.profiling = false;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (!rc && (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess)) {
    set_error("hipEventCreate failed");
    rc = SDFT_EHIP;
  }
  // candidates: plane variant x KB workgroup target x fused expand (the last
  // only where the columns allow it), each timed on filter steps k >= 3 of
  // a loop as block_filter_impl runs it (zero coefficients: the accumulator
  // is left as it is; a tiny scale keeps T_k bounded)
  const int targets[2] = {256, 1024};
  const int nfuse = c.g.col_intervals ? 2 : 1;
  const int pv0 = c.plane_variant, kt0 = c.kb_wg_target;
  const bool fu0 = c.fuse_expand;
  int best_pv = pv0, best_kt = kt0, ncand = 0;
  bool best_fu = fu0;
  double best = 1e300;
  const int L = steps + 4;
  double* bufs[2] = {b.X0, b.X1};
  auto step_params = [](int k) {
    StepParams sp{};
    const double m = (k == 1) ? 1.0 : 2.0, scale = 1e-4;
    sp.s1 = -scale * 0.25 * m;
    sp.s2 = -scale * 0.5 * m;
    sp.first = (k == 1);
    sp.step = k;
    return sp;
  };
  const int npv = plane_variant_count(c.g.n);
  for (int pv = 0; pv < npv && !rc; ++pv)
    for (int ti = 0; ti < 2 && !rc; ++ti)
      for (int fu = 0; fu < nfuse && !rc; ++fu) {
        c.plane_variant = pv;
        c.kb_wg_target = targets[ti];
        c.fuse_expand = fu != 0;
        for (int k = 1; k <= 2 && !rc; ++k)  // warm-up: the first two steps
          rc = apply_step(b, kModeFilter, bufs[(k - 1) & 1], bufs[k & 1], step_params(k),
                          nullptr, fuse_flags(c, k, L));
        if (rc) break;
        if (hipEventRecord(t0, c.stream) != hipSuccess) {
          rc = SDFT_EHIP;
          break;
        }
        for (int k = 3; k < steps + 3 && !rc; ++k)
          rc = apply_step(b, kModeFilter, bufs[(k - 1) & 1], bufs[k & 1], step_params(k),
                          nullptr, fuse_flags(c, k, L));
        if (rc) break;
        float ms = 0.f;
        if (hipEventRecord(t1, c.stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
            hipEventElapsedTime(&ms, t0, t1) != hipSuccess) {
          set_error("tune timing failed");
          rc = SDFT_EHIP;
          break;
        }
        const double per = ms / steps;
        if (cand_ms && ncand < max_cand) cand_ms[ncand] = per;
        ++ncand;
        if (per < best) {
          best = per;
          best_pv = pv;
          best_kt = targets[ti];
          best_fu = fu != 0;
        }
      }
  c.profiling = prof;
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  (void)hipStreamSynchronize(c.stream);
  block_free(b);
  if (rc) {
    c.plane_variant = pv0;
    c.kb_wg_target = kt0;
    c.fuse_expand = fu0;
    return rc;
  }
  c.plane_variant = best_pv;
  c.kb_wg_target = best_kt;
  c.fuse_expand = best_fu;
  if (config_out) {
    config_out[0] = best_pv;
    config_out[1] = best_kt;
    config_out[2] = best_fu ? 1 : 0;
  }
  return ncand;
}

int sdft_hip_set_config(sdft_ctx* h, int plane_tile, int kb_wg_target, int fused_expand) {
  if (!h || plane_tile < 0 || plane_tile >= plane_variant_count(h->c.g.n) || kb_wg_target < 1 ||
      (fused_expand != 0 && fused_expand != 1)) {
    set_error("bad configuration");
    return SDFT_EINVAL;
  }
  h->c.plane_variant = plane_tile;
  h->c.kb_wg_target = kb_wg_target;
  h->c.fuse_expand = fused_expand != 0;
  return 0;
}

int sdft_hip_newton_filter(sdft_ctx* h, int nstate, int L, int nmu, const double* coeffs,
                           const double* samp_point, double scale, double emean, double* cre,
                           double* cim, double* out_re, double* out_im, int prec) {
  if (!h || nstate < 1 || L < 1 || nmu < 1 || !coeffs || (L > 1 && !samp_point) || !cre ||
      !cim || !out_re || !out_im || (prec != 0 && prec != 1)) {
    set_error("bad Newton filter arguments");
    return SDFT_EINVAL;
  }
  Ctx& c = h->c;
  int rc;
  if ((rc = check_vks(c))) return rc;
  Block b;
  if ((rc = block_alloc(c, nstate, prec, b))) {
    block_free(b);
    return rc;
  }
  const long long stride = (long long)b.nstate_pad * 2 * c.g.ncp;
  double* accs = nullptr;
  double* cdev = nullptr;
  do {
    if ((rc = dalloc(&accs, sizeof(double) * (size_t)stride * nmu))) break;
    if ((rc = dalloc(&cdev, sizeof(double) * (size_t)L * nmu))) break;
    SDFT_HIP_TRY(rc, hipMemcpyAsync(cdev, coeffs, sizeof(double) * (size_t)L * nmu,
                                    hipMemcpyHostToDevice, c.stream));
    if (rc) break;
    if ((rc = block_put(b, b.X0, cre, cim))) break;
    if (L == 1) {
      for (int mu = 0; mu < nmu; ++mu) {
        hipLaunchKernelGGL(k_scale_copy, dim3((unsigned)((stride + 255) / 256)), dim3(256), 0,
                           c.stream, b.X0, coeffs[mu], accs + (size_t)mu * stride,
                           (size_t)stride);
        SDFT_HIP_TRY(rc, hipGetLastError());
      }
      if (rc) break;
    }
    // the contract kernel writes accumulator mu at acc + mu * stride: point
    // the block's accumulator at the nmu-accumulator buffer for the loop
    c.newton_nmu = nmu;
    c.newton_stride = stride;
    c.newton_c = cdev;
    double* own_acc = b.acc;
    b.acc = accs;
    prof_reset(c);
    double* X[2] = {b.X0, b.X1};
    // filter.c:151-208: P_n = (H^ - z_{n-1}) P_{n-1}, acc_mu += a[n][mu] P_n
    for (int n = 1; n < L && !rc; ++n) {
      StepParams sp;
      sp.s1 = -scale * 0.25;
      sp.s2 = -scale * 0.5;
      sp.p = -scale * emean - samp_point[n - 1];
      sp.ck = 0.0;
      sp.c0 = 0.0;
      sp.first = (n == 1);
      sp.step = n;
      rc = apply_step(b, kModeNewton, X[(n - 1) & 1], X[n & 1], sp, nullptr, fuse_flags(c, n, L));
    }
    b.acc = own_acc;
    c.newton_nmu = 0;
    c.newton_stride = 0;
    c.newton_c = nullptr;
    if (rc) break;
    if ((rc = prof_collect(c))) break;
    const size_t per = (size_t)nstate * c.g.ncoef;
    for (int mu = 0; mu < nmu && !rc; ++mu)
      rc = block_fetch(b, accs + (size_t)mu * stride, out_re + mu * per, out_im + mu * per);
    if (!rc) rc = block_fetch(b, X[(L - 1) & 1], cre, cim);
  } while (0);
  c.newton_nmu = 0;
  c.newton_stride = 0;
  c.newton_c = nullptr;
  void* ptrs[] = {accs, cdev};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  block_free(b);
  return rc;
}

int sdft_hip_spectral_bound(sdft_ctx* h, int which, double emax, double conv, int max_iter,
                            double* vre, double* vim, double* energy, int* iters, int prec) {
  if (!h || (which != 0 && which != 1) || max_iter < 1 || !vre || !vim || !energy ||
      (prec != 0 && prec != 1) || !(conv >= 0.0)) {
    set_error("bad spectral-bound arguments");
    return SDFT_EINVAL;
  }
  Ctx& c = h->c;
  int rc;
  if ((rc = check_vks(c))) return rc;
  Block b;
  if ((rc = block_alloc(c, 1, prec, b))) {
    block_free(b);
    return rc;
  }
  // the reference starts from energyOld = 1e7, energy = 1e5 (:168-170) and
  // iterates while |energyOld - energy| > energyConv
  double e_old = 1.0e7, e = 1.0e5;
  int it = 0;
  StepParams sp{};
  if (!(rc = block_put(b, b.X0, vre, vim))) rc = normalize0_device(c, b.X0);
  while (!rc && std::fabs(e_old - e) > conv && it < max_iter) {
    if ((rc = apply_step(b, kModeForce, b.X0, b.X1, sp, nullptr))) break;
    e_old = e;
    if ((rc = power_update_device(c, b.X0, b.X1, which, emax, &e))) break;
    ++it;
  }
  if (!rc) rc = block_fetch(b, b.X0, vre, vim);
  if (!rc) rc = prof_collect(c);
  block_free(b);
  if (rc) return rc;
  *energy = e;
  if (iters) *iters = it;
  return 0;
}

int sdft_hip_cheby_coeffs(sdft_ctx* h, int L, double emin, double emax, double mu, double beta,
                          int fun_flag, double* coeffs) {
  if (!h || L < 1 || !coeffs || !(emax > emin) || (fun_flag != 1 && fun_flag != 2)) {
    set_error("bad coefficient arguments");
    return SDFT_EINVAL;
  }
  return cheby_coeffs_device(h->c, L, emin, emax, mu, beta, fun_flag, coeffs);
}

int sdft_hip_cheby_filter(sdft_ctx* h, int nstate, int L, const double* coeffs, double scale,
                          double emean, double* cre, double* cim, double* ore, double* oim,
                          double* moments, int prec) {
  if (!h || nstate < 1 || L < 1 || !coeffs || !cre || !cim || !ore || !oim ||
      (prec != 0 && prec != 1)) {
    set_error("bad filter arguments");
    return SDFT_EINVAL;
  }
  Ctx& c = h->c;
  Block b;
  int rc = block_alloc(c, nstate, prec, b);
  if (!rc) rc = block_put(b, b.X0, cre, cim);
  if (!rc && moments) {
    const size_t bytes = sizeof(double) * (size_t)b.nstate_pad * 2 * c.g.ncp;
    SDFT_HIP_TRY(rc, hipMemcpyAsync(b.chi, b.X0, bytes, hipMemcpyDeviceToDevice, c.stream));
  }
  if (!rc) rc = block_filter_impl(b, L, coeffs, scale, emean, moments);
  if (!rc) rc = block_fetch(b, b.acc, ore, oim);
  if (!rc) rc = block_fetch(b, b.cur ? b.X1 : b.X0, cre, cim);
  block_free(b);
  return rc;
}

int sdft_hip_cheby_moments(sdft_ctx* h, int nstate, int L, double scale, double emean,
                           const double* cre, const double* cim, double* moments, int prec) {
  if (!h || nstate < 1 || L < 2 || !cre || !cim || !moments || (prec != 0 && prec != 1)) {
    set_error("bad moment arguments");
    return SDFT_EINVAL;
  }
  Ctx& c = h->c;
  int rc;
  if ((rc = check_vks(c))) return rc;
  Block b;
  rc = block_alloc(c, nstate, prec, b);
  if (!rc) rc = block_put(b, b.X0, cre, cim);
  const int nmom = (L % 2 == 0) ? L / 2 + 1 : (L + 1) / 2;
  if (!rc) rc = ensure_mom(b, nmom);
  if (!rc) {
    const size_t bytes = sizeof(double) * (size_t)b.nstate_pad * 2 * c.g.ncp;
    SDFT_HIP_TRY(rc, hipMemcpyAsync(b.chi, b.X0, bytes, hipMemcpyDeviceToDevice, c.stream));
    prof_reset(c);
    double* bufs[2] = {b.X0, b.X1};
    for (int k = 1; k < nmom && !rc; ++k) {
      StepParams sp;
      const double m = (k == 1) ? 1.0 : 2.0;
      sp.s1 = -scale * 0.25 * m;
      sp.s2 = -scale * 0.5 * m;
      sp.p = -scale * emean * m;
      sp.ck = 0.0;
      sp.c0 = 0.0;
      sp.first = (k == 1);
      sp.step = k;
      rc = apply_step(b, kModeMoments, bufs[(k - 1) & 1], bufs[k & 1], sp, b.mom_part,
                      fuse_flags(c, k, nmom));
    }
    if (!rc) rc = prof_collect(c);
  }
  if (!rc) {
    const int nwg = c.g.nchunk * (b.nstate_pad / 2);
    std::vector<double> red((size_t)3 * nmom);
    rc = moments_reduce(c, b.mom_part, nmom, nwg, red.data());
    if (!rc) {
      // calc-chempot-chebyshev.c:722-897 write order (direct mu_k wins)
      for (int k = 0; k <= L; ++k) moments[k] = 0.0;
      moments[0] = red[3 * 1 + 1];
      moments[1] = red[3 * 1 + 0];
      for (int k = 2; k < nmom; ++k) {
        moments[k] = red[3 * k + 0];
        moments[2 * k - 1] = 2.0 * red[3 * k + 1] - moments[1];
        moments[2 * k] = 2.0 * red[3 * k + 2] - moments[0];
      }
    }
  }
  block_free(b);
  return rc;
}

int sdft_hip_chem_pot(sdft_ctx* h, int L, double emin, double emax, double beta,
                      const double* moments, double nsto, double ne_true, double mu_init,
                      double gap, double occ, double tol, double* mu_out, double* ne_out) {
  if (!h || L < 1 || !moments || !mu_out || !ne_out || !(emax > emin) || !(nsto > 0) ||
      !(gap > 0) || !(tol > 0)) {
    set_error("bad chemical-potential arguments");
    return SDFT_EINVAL;
  }
  return chem_pot_device(h->c, L, emin, emax, beta, moments, nsto, ne_true, mu_init, gap, occ, tol,
                         mu_out, ne_out);
}

int sdft_hip_chem_pot_noalloc(sdft_ctx* h, int L, double emin, double emax, double beta,
                              const double* moments, double nsto, double ne_true, double mu_init,
                              double gap, double occ, double tol, double* mu_out,
                              double* ne_out) {
  if (!h || L < 1 || !moments || !mu_out || !ne_out || !(emax > emin) || !(nsto > 0) ||
      !(gap > 0) || !(tol > 0)) {
    set_error("bad chemical-potential arguments");
    return SDFT_EINVAL;
  }
  return chem_pot_device_noalloc(h->c, L, emin, emax, beta, moments, nsto, ne_true, mu_init, gap,
                                 occ, tol, mu_out, ne_out);
}


int sdft_hip_density(sdft_ctx* h, int nstate, const double* re, const double* im,
                     double occ_over_nsto, double* rho, int layout, int accumulate) {
  if (!h || nstate < 1 || !re || !im || !rho || (layout != 0 && layout != 1)) {
    set_error("bad density arguments");
    return SDFT_EINVAL;
  }
  Ctx& c = h->c;
  Block b;
  int rc = block_alloc(c, nstate, 0, b);
  if (!rc) rc = block_put(b, b.X0, re, im);
  const int npairs = b.nstate_pad / 2;
  const int ngroups = std::max(1, std::min(npairs, 2048 / c.g.n));
  double* part = nullptr;
  double* drho = nullptr;
  if (!rc) rc = ensure_W(c, w_bytes(c, npairs, 0));
  if (!rc) rc = dalloc(&part, sizeof(double) * (size_t)ngroups * c.g.N);
  if (!rc) rc = dalloc(&drho, sizeof(double) * (size_t)c.g.N);
  if (!rc && accumulate) {
    hipError_t e = hipMemcpy(drho, rho, sizeof(double) * c.g.N, hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = hip_fail(e, "hipMemcpy rho", __FILE__, __LINE__);
  }
  if (!rc) rc = launch_expand(c, 0, b.X0, npairs, c.W);
  if (!rc) rc = launch_density(c, c.W, npairs, part, ngroups);
  if (!rc) rc = launch_rho_finish(c, part, ngroups, occ_over_nsto, layout, accumulate, drho);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(rho, drho, sizeof(double) * c.g.N, hipMemcpyDeviceToHost,
                                  c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) rc = hip_fail(e, "density download", __FILE__, __LINE__);
  }
  if (part) (void)hipFree(part);
  if (drho) (void)hipFree(drho);
  block_free(b);
  return rc;
}

int sdft_hip_energies(sdft_ctx* h, int nstate, const double* re, const double* im, double occ,
                      double nsto, double* eke, double* enl) {
  if (!h || nstate < 1 || !re || !im || !eke || !enl || !(nsto > 0)) {
    set_error("bad energy arguments");
    return SDFT_EINVAL;
  }
  Ctx& c = h->c;
  Block b;
  int rc = block_alloc(c, nstate, 0, b);
  if (!rc) rc = block_put(b, b.X0, re, im);
  double ke = 0.0, nl = 0.0;
  if (!rc) rc = energy_ke_device(c, b.X0, nstate, &ke);
  if (!rc && rs_active(c)) {
    // per-atom form: Z_p from the real-space grids, E = sum E_p dV |Z_p|^2 / N
    const int npairs = b.nstate_pad / 2;
    rc = ensure_W(c, w_bytes(c, npairs, 0));
    if (!rc) rc = launch_expand(c, 0, b.X0, npairs, c.W);
    if (!rc) rc = rs_to_grid(c, 0, c.W, npairs);
    if (!rc) rc = rs_project(c, 0, npairs);
    if (!rc) rc = rs_energy(c, npairs, nstate, &nl);
  } else if (!rc && c.nproj > 0) {
    rc = launch_nl_project(c, 0, b.X0, b.nstate_pad);
    if (!rc) rc = energy_nl_device(c, b.nstate_pad, nstate, &nl);
  }
  if (!rc) {
    *eke = occ * 0.5 * ke / nsto;
    *enl = occ * nl / nsto;
  }
  block_free(b);
  return rc;
}

int sdft_hip_block_create(sdft_ctx* h, int nstate, int prec, sdft_block** out) {
  if (!h || nstate < 1 || !out || (prec != 0 && prec != 1)) {
    set_error("bad block arguments");
    return SDFT_EINVAL;
  }
  sdft_block* bb = new sdft_block();
  int rc = block_alloc(h->c, nstate, prec, bb->b);
  if (rc) {
    block_free(bb->b);
    delete bb;
    return rc;
  }
  *out = bb;
  return 0;
}

int sdft_hip_block_destroy(sdft_block* bb) {
  if (!bb) return 0;
  if (bb->b.ctx) (void)hipStreamSynchronize(bb->b.ctx->stream);
  block_free(bb->b);
  delete bb;
  return 0;
}

int sdft_hip_block_set(sdft_block* bb, const double* cre, const double* cim) {
  if (!bb || !cre || !cim) {
    set_error("null argument");
    return SDFT_EINVAL;
  }
  Block& b = bb->b;
  b.cur = 0;
  int rc = block_put(b, b.X0, cre, cim);
  if (rc) return rc;
  const size_t bytes = sizeof(double) * (size_t)b.nstate_pad * 2 * b.ctx->g.ncp;
  SDFT_HIP_CHECK(hipMemcpyAsync(b.chi, b.X0, bytes, hipMemcpyDeviceToDevice, b.ctx->stream));
  SDFT_HIP_CHECK(hipStreamSynchronize(b.ctx->stream));
  return 0;
}

int sdft_hip_noise(sdft_ctx* h, int nstate, long long seed, long long first, double* cre,
                   double* cim) {
  if (!h || nstate < 1 || !cre || !cim) {
    set_error("bad noise arguments");
    return SDFT_EINVAL;
  }
  Block b;
  int rc = block_alloc(h->c, nstate, 0, b);
  if (!rc) rc = block_noise_impl(b, seed, first);
  if (!rc) rc = block_fetch(b, b.X0, cre, cim);
  block_free(b);
  return rc;
}

int sdft_hip_block_noise(sdft_block* bb, long long seed, long long first) {
  if (!bb) {
    set_error("null argument");
    return SDFT_EINVAL;
  }
  return block_noise_impl(bb->b, seed, first);
}

int sdft_hip_block_filter(sdft_block* bb, int L, const double* coeffs, double scale, double emean,
                          double* moments) {
  if (!bb || !coeffs) {
    set_error("null argument");
    return SDFT_EINVAL;
  }
  Block& b = bb->b;
  // every call filters chi (the reference re-filters the same noise)
  const size_t bytes = sizeof(double) * (size_t)b.nstate_pad * 2 * b.ctx->g.ncp;
  b.cur = 0;
  SDFT_HIP_CHECK(hipMemcpyAsync(b.X0, b.chi, bytes, hipMemcpyDeviceToDevice, b.ctx->stream));
  return block_filter_impl(b, L, coeffs, scale, emean, moments);
}

int sdft_hip_block_get(sdft_block* bb, int which, double* re, double* im) {
  if (!bb || !re || !im || (which != 0 && which != 1)) {
    set_error("bad block_get arguments");
    return SDFT_EINVAL;
  }
  Block& b = bb->b;
  const double* src = which == 0 ? b.acc : (b.cur ? b.X1 : b.X0);
  return block_fetch(b, src, re, im);
}

int sdft_hip_set_profiling(sdft_ctx* h, int on) {
  if (!h) {
    set_error("null argument");
    return SDFT_EINVAL;
  }
  h->c.profiling = on != 0;
  return 0;
}

int sdft_hip_kernel_times(sdft_ctx* h, int max, const char** names, double* ms, int* launches) {
  if (!h) {
    set_error("null argument");
    return SDFT_EINVAL;
  }
  Ctx& c = h->c;
  const int n = (int)c.prof_names.size();
  for (int i = 0; i < n && i < max; ++i) {
    if (names) names[i] = c.prof_names[i].c_str();
    if (ms) ms[i] = c.prof_ms[i];
    if (launches) launches[i] = c.prof_launch[i];
  }
  return n;
}

// ---- Kohn-Sham potential from the density ----------------------------------
int sdft_hip_ks_init(sdft_ctx* h, const sdft_grid_desc* d) {
  if (!h || !d) {
    set_error("null argument");
    return SDFT_EINVAL;
  }
  Ctx& c = h->c;
  if (d->n1 != c.g.n || d->n2 != c.g.n || d->n3 != c.g.n) {
    set_error("the density sphere must be on the context's grid");
    return SDFT_EINVAL;
  }
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  if (c.dens) {
    free_ctx(*c.dens);
    delete c.dens;
    c.dens = nullptr;
  }
  Ctx* dc = new Ctx();
  dc->device = c.device;
  hipError_t e = hipStreamCreateWithFlags(&dc->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete dc;
    return hip_fail(e, "hipStreamCreate", __FILE__, __LINE__);
  }
  int rc = build_geometry(*dc, d);
  if (rc) {
    free_ctx(*dc);
    delete dc;
    return rc;
  }
  c.dens = dc;
  const size_t N = (size_t)c.g.N;
  const size_t nb = std::max((N + 255) / 256, (size_t)(dc->g.ncoef + 255) / 256);
  if (!c.ks_field && (rc = dalloc(&c.ks_field, 2 * sizeof(double) * N))) return rc;
  if (!c.ks_rho && (rc = dalloc(&c.ks_rho, sizeof(double) * N))) return rc;
  if (!c.ks_vhe && (rc = dalloc(&c.ks_vhe, sizeof(double) * N))) return rc;
  if (!c.ks_io && (rc = dalloc(&c.ks_io, sizeof(double) * N))) return rc;
  if (c.ks_q) SDFT_HIP_CHECK(hipFree(c.ks_q));
  c.ks_q = nullptr;
  if ((rc = dalloc(&c.ks_q, sizeof(double) * 2 * dc->g.ncp))) return rc;
  if (c.ks_part_n < 4 * nb) {
    if (c.ks_part) SDFT_HIP_CHECK(hipFree(c.ks_part));
    c.ks_part = nullptr;
    c.ks_part_n = 0;
    if ((rc = dalloc(&c.ks_part, sizeof(double) * 4 * nb))) return rc;
    c.ks_part_n = 4 * nb;
  }
  return 0;
}

int sdft_hip_ks_potential(sdft_ctx* h, const double* rho, int layout, const double* vext_re,
                          const double* vext_im, double volume, int xc, double* energies,
                          double* vks_out) {
  if (!h || !rho || !vext_re || !vext_im || !(volume > 0.0) || (layout != SDFT_RHO_FFTW &&
      layout != SDFT_RHO_GENERIC) || xc < SDFT_XC_PZ_LDA || xc > SDFT_XC_PADE_LDA) {
    set_error("bad Kohn-Sham potential arguments");
    return SDFT_EINVAL;
  }
  Ctx& c = h->c;
  if (!c.dens) {
    set_error("density sphere not set (call sdft_hip_ks_init first)");
    return SDFT_ESTATE;
  }
  Ctx& dc = *c.dens;
  const long long N = c.g.N;
  const int n = c.g.n;
  int rc;
  // the context's v_KS is rewritten below: its queued filter work comes first
  SDFT_HIP_CHECK(hipStreamSynchronize(c.stream));
  if (!c.vks64 && (rc = dalloc(&c.vks64, sizeof(double) * N))) return rc;
  if (!c.vks32 && (rc = dalloc(&c.vks32, sizeof(float) * N))) return rc;
  hipStream_t s = dc.stream;
  // rho -> the sphere: Q(G) = (1/N) sum_r rho e^{-iG.r} (row 0 of X0)
  SDFT_HIP_CHECK(hipMemcpyAsync(c.ks_io, rho, sizeof(double) * N, hipMemcpyHostToDevice, s));
  if ((rc = launch_ks_field(s, c.ks_io, layout, n, N, volume, c.ks_field, c.ks_rho))) return rc;
  if ((rc = ensure_W(dc, w_bytes(dc, 1, 0)))) return rc;
  Block b;
  rc = block_alloc(dc, 1, 0, b);
  if (!rc) rc = launch_grid_plane(dc, dc.W, 1, c.ks_field);
  StepParams sp{};
  if (!rc)
    rc = launch_contract(dc, 0, kModeNoise, dc.W, 1, b.X0, b.X0, b.chi, nullptr, nullptr, sp,
                         nullptr);
  // rho(G) kept for sdft_hip_local_forces
  if (!rc)
    SDFT_HIP_CHECK(hipMemcpyAsync(c.ks_q, b.X0, sizeof(double) * 2 * dc.g.ncp,
                                  hipMemcpyDeviceToDevice, s));
  // v(G) = vext + Hartree (row 0 of X1; row 1 stays zero)
  if (!rc) rc = block_put(b, b.X1, vext_re, vext_im);
  int nb1 = 0, nb2 = 0;
  if (!rc)
    rc = launch_ks_hartree(s, b.X0, b.X1, dc.t.ak2, dc.g.ncoef, dc.g.ncp, volume, c.ks_part,
                           &nb1);
  // v(G) -> v(r): expand, then the c / b transforms with the real part kept
  if (!rc) rc = launch_expand(dc, 0, b.X1, 1, dc.W);
  if (!rc) rc = launch_density(dc, dc.W, 1, c.ks_vhe, 1, true);
  // + LDA exchange-correlation -> the context's v_KS
  if (!rc)
    rc = launch_ks_xc(s, c.ks_vhe, c.ks_rho, N, xc, c.vks64, c.vks32, c.ks_part + 2 * nb1, &nb2);
  std::vector<double> part;
  if (!rc) {
    part.resize(2 * (size_t)(nb1 + nb2));
    SDFT_HIP_CHECK(hipMemcpyAsync(part.data(), c.ks_part, sizeof(double) * part.size(),
                                  hipMemcpyDeviceToHost, s));
    if (vks_out) {
      rc = launch_ks_out(s, c.vks64, layout, n, N, c.ks_io);
      if (!rc)
        SDFT_HIP_CHECK(hipMemcpyAsync(vks_out, c.ks_io, sizeof(double) * N,
                                      hipMemcpyDeviceToHost, s));
    }
  }
  (void)hipStreamSynchronize(s);
  block_free(b);
  if (rc) return rc;
  c.have_vks = true;
  c.ks_have_q = true;
  if (energies) {
    double eh = 0.0, ee = 0.0, ex = 0.0, mx = 0.0;
    for (int i = 0; i < nb1; ++i) {
      eh += part[2 * i];
      ee += part[2 * i + 1];
    }
    for (int i = nb1; i < nb1 + nb2; ++i) {
      ex += part[2 * i];
      mx += part[2 * i + 1];
    }
    const double vscale = volume / (double)N;
    energies[0] = eh;
    energies[1] = ee;
    energies[2] = ex * vscale;
    energies[3] = mx * vscale;
  }
  return 0;
}

int sdft_hip_local_forces(sdft_ctx* h, int natm, const double* pos, const int* atom_type,
                          int ntype, const double* vloc, const double* gvec, double* forces) {
  if (!h || natm < 1 || ntype < 1 || !pos || !atom_type || !vloc || !gvec || !forces) {
    set_error("bad local force arguments");
    return SDFT_EINVAL;
  }
  Ctx& c = h->c;
  if (!c.dens || !c.ks_have_q) {
    set_error("no density on the device (call sdft_hip_ks_potential first)");
    return SDFT_ESTATE;
  }
  for (int a = 0; a < natm; ++a)
    if (atom_type[a] < 0 || atom_type[a] >= ntype) {
      set_error("atom type out of range");
      return SDFT_EINVAL;
    }
  Ctx& dc = *c.dens;
  const int nc = dc.g.ncoef, ncp = dc.g.ncp;
  std::vector<double> g((size_t)3 * ncp, 0.0), f((size_t)ntype * ncp, 0.0);
  for (int j = 0; j < nc; ++j) {
    const int src = dc.g.perm[j];
    for (int d = 0; d < 3; ++d) g[(size_t)d * ncp + j] = gvec[(size_t)src * 3 + d];
    for (int t = 0; t < ntype; ++t) f[(size_t)t * ncp + j] = vloc[(size_t)t * nc + src];
  }
  double *dg = nullptr, *df = nullptr, *dp = nullptr, *dF = nullptr;
  int* dt = nullptr;
  int rc = dalloc(&dg, sizeof(double) * g.size());
  if (!rc) rc = dalloc(&df, sizeof(double) * f.size());
  if (!rc) rc = dalloc(&dp, sizeof(double) * 3 * natm);
  if (!rc) rc = dalloc(&dF, sizeof(double) * 3 * natm);
  if (!rc) rc = dalloc(&dt, sizeof(int) * natm);
  hipStream_t s = dc.stream;
  hipError_t e = hipSuccess;
  if (!rc) {
    e = hipMemcpyAsync(dg, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(df, f.data(), sizeof(double) * f.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dp, pos, sizeof(double) * 3 * natm, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dt, atom_type, sizeof(int) * natm, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) rc = hip_fail(e, "local force upload", __FILE__, __LINE__);
  }
  if (!rc) rc = launch_local_forces(s, c.ks_q, ncp, nc, dg, df, dp, dt, natm, dF);
  if (!rc) {
    e = hipMemcpyAsync(forces, dF, sizeof(double) * 3 * natm, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "local force download", __FILE__, __LINE__);
  }
  (void)hipStreamSynchronize(s);
  void* ptrs[] = {dg, df, dp, dF, dt};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  return rc;
}

int sdft_hip_selftest_copy(long long nbytes, int reps) {
  if (nbytes < 8 || reps < 1) {
    set_error("bad copy selftest arguments");
    return SDFT_EINVAL;
  }
  const size_t n = (size_t)nbytes / sizeof(double);
  double *a = nullptr, *b = nullptr;
  int rc = dalloc(&a, n * sizeof(double));
  if (!rc) rc = dalloc(&b, n * sizeof(double));
  if (!rc) {
    hipError_t e = hipMemset(a, 0, n * sizeof(double));
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
      hipLaunchKernelGGL(k_scale_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, a, 1.0,
                         b, n);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) rc = hip_fail(e, "copy selftest", __FILE__, __LINE__);
  }
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

int sdft_hip_selftest_fft(int n, int nlines, int dir, int prec, const double* in, double* out) {
  if (nlines < 1 || !in || !out || (dir != 1 && dir != -1) || (prec != 0 && prec != 1) ||
      chunk_cols_for(n) == 0) {
    set_error("bad selftest arguments");
    return SDFT_EINVAL;
  }
  std::vector<double> tw64(2 * n);
  std::vector<float> tw32(2 * n);
  for (int m = 0; m < n; ++m) {
    const long double ang = 2.0L * 3.14159265358979323846264338327950288L * m / (long double)n;
    tw64[2 * m] = (double)cosl(ang);
    tw64[2 * m + 1] = (double)sinl(ang);
    tw32[2 * m] = (float)cosl(ang);
    tw32[2 * m + 1] = (float)sinl(ang);
  }
  const size_t bytes = sizeof(double) * 2 * (size_t)n * nlines;
  double *din = nullptr, *dout = nullptr;
  void* dtw = nullptr;
  int rc = dalloc(&din, bytes);
  if (!rc) rc = dalloc(&dout, bytes);
  if (!rc) rc = dalloc(&dtw, sizeof(double) * 2 * n);
  if (!rc) {
    hipError_t e = hipMemcpy(din, in, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(dtw, prec == 0 ? (const void*)tw64.data() : (const void*)tw32.data(),
                    prec == 0 ? sizeof(double) * 2 * n : sizeof(float) * 2 * n,
                    hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = hip_fail(e, "selftest upload", __FILE__, __LINE__);
  }
  if (!rc) rc = selftest_fft_dispatch(n, nlines, dir, prec, din, dout, dtw);
  if (!rc) {
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "selftest run", __FILE__, __LINE__);
  }
  if (din) (void)hipFree(din);
  if (dout) (void)hipFree(dout);
  if (dtw) (void)hipFree(dtw);
  return rc;
}

}  // extern "C"
