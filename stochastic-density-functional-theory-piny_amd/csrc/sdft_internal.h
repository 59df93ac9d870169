// sdft_internal.h -- device context and kernel launchers shared by the HIP
// translation units of libsdft_hip.
//
// Internal data layout in HBM (all device arrays owned by the context/block):
//  * coefficient order: the caller's half-sphere entries are permuted so that
//    coefficients whose +G lie in the same column chunk are contiguous
//    (columns run along the a axis and are indexed by (kb, kc)); G=0 stays
//    last.  The permutation is applied on upload and undone on download.
//  * an orbital block X holds nstate_pad (even) states, each 2*ncp doubles:
//    real parts [0, ncoef), zero pad to ncp, imaginary parts [ncp, ncp+ncoef),
//    zero pad.  As a matrix it is (2 ncp) x nstate, column-major: the
//    operand of the Kleinman-Bylander GEMMs.
//  * W: partially transformed pair grids, [pair][ra][ncolp] complex (T),
//    i.e. for every a-plane the values of all sphere columns (kb, kc).  It is
//    the only intermediate of the FFT sandwich; no full n^3 grid is stored.
//  * v_KS in FFTW order [ra][rb][rc], FP64 and FP32 copies.
//  * KB table A[p][2 ncp]: A = w Re Q | -w Im Q with w = 2 (G != 0), 1 (G=0),
//    so that Z = A X and the NL force is -2 A^T diag(E) Z.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace sdft {

constexpr int kThreads = 256;       // threads per workgroup of the FFT kernels

// Error reporting (thread-local message, negative status codes).
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define SDFT_HIP_CHECK(expr)                                               \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess) return ::sdft::hip_fail(e_, #expr, __FILE__, __LINE__); \
  } while (0)

// The same check without the early return: rc receives the status so that a
// caller's shared cleanup still runs (use inside do { ... } while (0) bodies
// as `SDFT_HIP_TRY(rc, expr); if (rc) break;`).
#define SDFT_HIP_TRY(rc, expr)                                              \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess) (rc) = ::sdft::hip_fail(e_, #expr, __FILE__, __LINE__); \
  } while (0)

struct Geometry {
  int n = 0;          // cubic grid extent
  long long N = 0;    // n^3
  int ncoef = 0;      // coefficients incl. G=0
  int ncp = 0;        // padded coefficient stride (multiple of 64)
  int K = 0;          // max |k| over the sphere (all axes)
  int M = 0;          // rows of the (kb) plane buffer: 2K+1
  int ncol = 0;       // sphere columns (kb, kc)
  int ncolp = 0;      // padded column stride of W
  int nchunk = 0;     // column chunks (column-pair aligned)
  int chunk_max = 0;  // max columns in a chunk
  // every column's ka set is one interval [-h, h] (h from its entry count):
  // the contract kernel can then re-expand its own columns for the next step
  bool col_intervals = false;
  // host copies
  std::vector<int> perm;        // internal j -> caller index
  std::vector<int> iperm;       // caller index -> internal j
  std::vector<double> ak2;      // internal order, ncp
};

struct DeviceTables {
  int* col_pos = nullptr;     // [ncol] LDS offset (kb+K)*(n+1) + kc mod n in the plane buffer
  int* rowcol = nullptr;      // [M][n] column d with (kb, kc mod n) = (row - K, pos), else -1
  int* ent_off = nullptr;     // [ncol+1] entry offsets per column
  int* ent_j = nullptr;       // [nent] internal coefficient index
  int* ent_b = nullptr;       // [nent] (d << 10) | (flag << 8) | ia ; flag 0 +G, 1 -G, 2 G=0
  int* chunk_col = nullptr;   // [nchunk+1]
  int* chunk_coef = nullptr;  // [nchunk+1]
  int* coef_plus = nullptr;   // [ncoef] (d << 8) | ia of +G (internal order)
  int* coef_minus = nullptr;  // [ncoef] (d << 8) | ia of -G
  double* ak2 = nullptr;      // [ncp] internal order
  void* tw64 = nullptr;       // [n] double2 exp(+2 pi i m/n)
  void* tw32 = nullptr;       // [n] float2
};

// Events that only order work between the library's own streams on one
// device: no timing, and no system-scope fence (nothing is handed to the
// host or another device through them), so a fork or join costs no
// system-wide cache writeback.
constexpr unsigned kSyncEventFlags = hipEventDisableTiming | hipEventDisableSystemFence;

// Per-atom real-space KB data (sdft_hip_set_projectors_real), see sdft_rs.hip.
struct RsKB;

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // the KB GEMM chain of a step runs on stream_nl, concurrently with the
  // expand / plane kernels; ev_cur orders it after the previous step, ev_nl
  // orders the contract kernel after it
  hipStream_t stream_nl = nullptr;
  hipEvent_t ev_cur = nullptr;
  hipEvent_t ev_nl = nullptr;
  Geometry g;
  DeviceTables t;
  double* vks64 = nullptr;   // [N]
  float* vks32 = nullptr;    // [N]
  bool have_vks = false;
  int nproj = 0;
  double* A64 = nullptr;     // [nproj][2 ncp]
  float* A32 = nullptr;      // [nproj][2 ncp]
  // A transposed, projectors padded with zeros to ppad (multiple of 32):
  // the force product's operand, read in runs of consecutive projectors
  double* At64 = nullptr;    // [2 ncp][ppad]
  float* At32 = nullptr;     // [2 ncp][ppad]
  int ppad = 0;
  double* E = nullptr;       // [nproj]
  std::vector<double> E_host;
  // scratch (grown on demand)
  void* W = nullptr;
  size_t W_bytes = 0;
  void* W2 = nullptr;        // output of the split-plane form
  size_t W2_bytes = 0;
  double* zpart = nullptr;   // split-K partials of Z
  size_t zpart_bytes = 0;
  double* Y = nullptr;       // [nstate_pad][ppad] = -2 E Z, transposed, zero padded
  float* Y32 = nullptr;      // the same for the FP32 form (its force operand)
  double* Z = nullptr;       // [nproj][nstate_pad]
  size_t Y_bytes = 0;
  int Y_ppad = 0;
  double* red = nullptr;     // reduction scratch
  size_t red_bytes = 0;
  // launch configuration (the tuner's pick): plane-kernel tile variant (0
  // base sub-chunk, 1 double) and the KB projection's target workgroup count
  int plane_variant = 0;  // plane-kernel variant, 0 .. plane_variant_count(n) - 1
  int kb_wg_target = 512;
  // filter loops let the contract kernel write the next step's a-expansion
  // (when the columns allow it, Geometry::col_intervals); the tuner decides
  bool fuse_expand = true;
  // Newton-form accumulators for the contract kernel (set for the duration
  // of a Newton filter call only)
  int newton_nmu = 0;
  long long newton_stride = 0;
  const double* newton_c = nullptr;
  // Kohn-Sham potential from the density: the density-sphere geometry (its
  // own context on the same grid, owned) and its device buffers
  Ctx* dens = nullptr;
  void* ks_field = nullptr;     // [N] complex: the density as a field
  double* ks_rho = nullptr;     // [N] physical density, FFTW order
  double* ks_vhe = nullptr;     // [N] v_H + v_ext in real space
  double* ks_part = nullptr;    // per-block energy partials
  double* ks_io = nullptr;      // [N] the caller's density / v_KS staging
  double* ks_q = nullptr;       // rho(G) of the last call (density sphere, internal order)
  bool ks_have_q = false;
  size_t ks_part_n = 0;
  // real-space KB projectors in their per-atom form (null: none set);
  // nl_form 0 applies them per atom on the real-space grid, 1 through the
  // G-space table they restrict to (built on first use)
  RsKB* rs = nullptr;
  int nl_form = 0;
  // profiling
  bool profiling = false;
  std::vector<std::string> prof_names;
  std::vector<double> prof_ms;
  std::vector<int> prof_launch;
};

struct Block {
  Ctx* ctx = nullptr;
  int nstate = 0;
  int nstate_pad = 0;
  int prec = 0;
  double* X0 = nullptr;    // T ping
  double* X1 = nullptr;    // T pong
  double* acc = nullptr;   // filtered accumulator
  double* chi = nullptr;   // copy of the input noise (moments)
  // NL raw force of the current step, stored in the form's precision
  // (double for FP64, float for FP32: the MFMA accumulates in that type)
  void* fnl = nullptr;
  // FP32 form: float copy of the T buffer the KB projection reads (its MFMA
  // operand), rows padded like X0 / X1
  float* X32 = nullptr;
  double* mom_part = nullptr;  // per-step moment partials
  size_t mom_part_bytes = 0;
  int cur = 0;             // which of X0/X1 holds the current T
};

// ---- recurrence parameters passed to the contract kernel ----------------
struct StepParams {
  double s1, s2, p;   // -scale/4, -scale/2, -scale*emean (x2 after step 1)
  double ck;          // coefficient of the produced T_k
  double c0;          // coefficient of T_0 (step 1 initialises acc)
  int first;          // 1 on the first step (T_1 = s1 f + p T_0)
  int step;           // step index (moment partial row)
};

// ---- kernel argument blocks ----------------------------------------------
// The FFT-sandwich kernels take one of these by value (single system) or an
// array of them plus a workgroup prefix table (one entry per fragment of a
// batch).  gx / gy are the per-system grid extents.
struct ExpandArgs {
  const double* X;
  int ldx, ncp;
  const int* chunk_col;
  const int* ent_off;
  const int* ent_j;
  const int* ent_b;
  const void* tw;
  void* W;
  int ncolp;
  int gx;  // column chunks; gy = pairs
};

struct PlaneArgs {
  void* W;
  void* Wout;  // split-plane form only: the output buffer (W stays the input)
  void* field; // real-space KB path: the pair grids [pair][ra][rb][rc]
  int ncolp;
  const int* rowcol;
  int M;
  const void* v;
  const void* tw;
  int gx;  // = n (a-planes); gy = pairs
};

struct ContractArgs {
  const void* W;
  int ncolp;
  const int* chunk_col;
  const int* chunk_coef;
  const int* coef_plus;
  const int* coef_minus;
  const double* ak2;
  const void* tw;
  const double* Xcur;
  double* Xout;
  double* acc;
  const void* fnl;  // in the kernel's precision T
  const double* chi;
  int ldx, ncp, ncoef;
  double invN;
  int mode;
  StepParams sp;
  double* mom_part;
  // when non-null, c_k of the produced T_k is coeffs_dev[*step_dev] (the
  // step counter of a replayed graph); otherwise sp.ck
  const double* coeffs_dev;
  const int* step_dev;
  // Newton form: nmu accumulators acc + mu * acc_stride and the coefficient
  // table newton_c[n * nmu + mu] (device)
  int nmu;
  long long acc_stride;
  const double* newton_c;
  // when non-null (filter / moments / Newton modes), the kernel also packs
  // the new vector Xout of its column pairs into its LDS columns, inverse
  // transforms them along a and writes them to Wnext: the expand of the
  // next step (needs Geometry::col_intervals)
  void* Wnext;
  const int* ent_off;
  // FP32 form: when non-null, the new vector is also stored here in float
  // (the next KB projection's operand)
  float* Xout32;
  int gx, gy;  // column chunks, pairs
};

// kModeNoise writes the transformed noise to Xout and to acc (the caller
// passes the block's chi buffer as acc).  kModeNewton: P_{n+1} = (H^ - z_n) P_n
// (sp.p = -scale emean - z_n) and acc_mu += a[n+1][mu] P_{n+1}.
enum ContractMode {
  kModeForce = 0,
  kModeFilter = 1,
  kModeMoments = 2,
  kModeNoise = 3,
  kModeNewton = 4
};

// KB GEMM arguments of one fragment of a batch (tile counts: tp over
// projectors, ts over states (32 each); tk over the 2 ncp coefficient rows in
// kForceCols-row blocks).
struct KbProjectArgs {
  const void* A;  // [P][K] (double or float)
  const void* X;  // [ns][K] in the same type
  int P, ns, K, per, nsplit, tp, ts;
  double* zpart;
};

struct KbReduceArgs {
  const double* zpart;
  int nsplit, P, ns, ppad;
  const double* E;
  double* Z;
  void* Y;  // [ns][ppad], double (FP64 form) or float (FP32 form)
};

// columns k of F per workgroup of the force product (four waves of
// kForceCols / 4)
constexpr int kForceCols = 128;

struct KbForceArgs {
  const void* Y;   // [ns][ppad]   (all three in the form's type)
  const void* At;  // [K2][ppad]
  int ns, K2, ppad, ts, tk;
  void* F;         // [ns][K2]
};

// Workgroups of a launch are dispatched to the eight XCDs round-robin (XCD =
// index mod 8), each XCD with its own L2.  This maps the hardware index w of
// a 1-D launch of `total` workgroups to a logical index so that every XCD
// receives one contiguous range of logical indices (sizes differing by at
// most one): workgroups that share operands, numbered next to each other,
// then read them through the same L2.  Bijective on [0, total).
__device__ inline int xcd_swizzle(int w, int total) {
  constexpr int X = 8;
  const int q = total / X, r = total - q * X;
  const int x = w % X, j = w / X;
  return x * q + (x < r ? x : r) + j;
}

// workgroup -> fragment of a batch launch: prefix[f] <= wg < prefix[f + 1]
__device__ inline int find_frag(const int* __restrict__ prefix, int nfrag, int wg) {
  int lo = 0, hi = nfrag;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= wg)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// ---- launchers (sdft_kernels.hip) ----------------------------------------
size_t w_bytes(const Ctx& c, int npairs, int prec);
int launch_expand(Ctx& c, int prec, const double* X, int npairs, void* W);
// *Wout receives the buffer holding the result: W itself, or the context's
// second buffer W2 when the plane runs in the split form (M x (n+1) plane
// buffer above the LDS size)
int launch_plane(Ctx& c, int prec, void* W, int npairs, void** Wout);
// Plane-kernel variants offered at grid size n (b-line tile, threads per
// workgroup, radix schedules): 6 at n = 64, 2 elsewhere.
inline int plane_variant_count(int n) { return n == 64 ? 6 : 2; }
int launch_contract(Ctx& c, int prec, int mode, const void* W, int npairs,
                    const double* Xcur, double* Xout, double* acc, const void* fnl,
                    const double* chi, const StepParams& sp, double* mom_part,
                    void* Wnext = nullptr, float* Xout32 = nullptr);
// re = true: a single pair (ngroups = 1) whose real part is written to
// rho_part (the real field sum_G c1(G) e^{iG.r}) instead of |z|^2
int launch_density(Ctx& c, const void* W, int npairs, double* rho_part, int ngroups,
                   bool re = false);
// Kohn-Sham potential kernels (sdft_ks.hip)
int launch_ks_field(hipStream_t s, const double* rho, int layout, int n, long long N,
                    double vol, void* field, double* rho_fftw);
int launch_ks_hartree(hipStream_t s, const double* Q, double* V, const double* ak2, int ncoef,
                      int ncp, double vol, double* part, int* nblocks);
int launch_ks_xc(hipStream_t s, const double* vhe, const double* rho, long long N, int kind,
                 double* vks, float* vks32, double* part, int* nblocks);
int launch_ks_out(hipStream_t s, const double* vks, int layout, int n, long long N,
                  double* out);
int launch_local_forces(hipStream_t s, const double* Q, int ncp, int ncoef, const double* g,
                        const double* vloc, const double* pos, const int* type, int natm,
                        double* F);
int launch_noise_plane(Ctx& c, void* W, int npairs, unsigned long long seed, int first,
                       int nstate);
int launch_grid_plane(Ctx& c, void* W, int npairs, const void* field);
int launch_scatter_proj(Ctx& c, const long long* off, const int* idx, const double* val, int p0,
                        int p1, double* field);
int launch_rho_finish(Ctx& c, const double* rho_part, int ngroups, double scale,
                      int layout, int accumulate, double* rho);
int selftest_fft_dispatch(int n, int nlines, int dir, int prec, const double* in, double* out,
                          const void* tw);
ExpandArgs make_expand_args(const Ctx& c, int prec, const double* X, void* W);
PlaneArgs make_plane_args(const Ctx& c, int prec, void* W);
ContractArgs make_contract_args(const Ctx& c, int prec, int mode, const void* W, int npairs,
                                const double* Xcur, double* Xout, double* acc, const void* fnl,
                                const double* chi, const StepParams& sp, double* mom_part);
size_t plane_lds_for(int n, int prec, int M, int variant);
// the plane kernel's pruned first / last passes need the sphere inside
// |k| < n/4 on every axis
inline bool plane_prunable(int K, int n) { return 4 * K < n; }
// one launch of kind 0 (expand), 1 (plane; variant bit 0 the tile, bit 1
// the pruned passes) or 2 (contract; variant 1: with the fused expand) over
// the fragments whose
// argument blocks are args[0..nfrag) (device), all of grid size n
int launch_batch(int kind, int n, int prec, const void* args, const int* prefix, int nfrag,
                 int total, size_t lds, int variant, hipStream_t s);
int chunk_cols_for(int n);
int round_fft(int n_in);

// ---- GEMM (sdft_gemm.hip) -----------------------------------------------
// X / fnl in the form's type (double for prec 0, float for prec 1)
int launch_nl_project(Ctx& c, int prec, const void* X, int nstate_pad);
int launch_nl_force(Ctx& c, int prec, int nstate_pad, void* fnl);
int kb_project_split(const Ctx& c, int ns, int* per);
int launch_kb_batch(int kind, int prec, const void* args, const int* prefix, int nfrag,
                    int total, hipStream_t s);
int nl_force_sums(Ctx& c, const double* X, int ns, int nstate, const double* gint, double* Ad,
                  double* Zd, double* part, double* S);

// ---- real-space KB, per-atom form (sdft_rs.hip) --------------------------
// rs_set: validate the CSR projector lists and build the device data (the
// groups of projectors sharing one point list, the point -> projector
// inverse lists); keeps a host copy for rs_table_args.  rs_free releases it.
int rs_set(Ctx& c, int nproj, const long long* off, const int* idx, const double* val,
           const double* energy, double vol_elem);
void rs_free(Ctx& c);
bool rs_active(const Ctx& c);
// the host CSR (for building the equivalent G-space table)
void rs_host_lists(const Ctx& c, int* nproj, const long long** off, const int** idx,
                   const double** val, const double** energy, double* vol_elem);
// the four stages of the per-atom form on npairs pairs (W in: the a-expansion;
// W out: what the plane kernel would have produced), and the NL energy sum
// sum_p E_p dV |Z_p|^2 / N over the pairs' orbitals
int rs_to_grid(Ctx& c, int prec, void* W, int npairs);
int rs_project(Ctx& c, int prec, int npairs);
int rs_apply(Ctx& c, int prec, int npairs);
int rs_from_grid(Ctx& c, int prec, void* W, int npairs);
int rs_energy(Ctx& c, int npairs, int nstate, double* out);
// projector gradient tables (d beta_p / d r on the same point lists) and the
// force sums S_p,j = sum_s Z_ps dZ_ps/d r_j over the grids of npairs FP64 pairs
int rs_set_gradients(Ctx& c, int nproj, const double* dx, const double* dy, const double* dz);
bool rs_has_gradients(const Ctx& c);
int rs_force_sums(Ctx& c, int npairs, double* out);

// ---- Chebyshev scalars (sdft_cheby.hip) ---------------------------------
int cheby_coeffs_device(Ctx& c, int L, double emin, double emax, double mu, double beta,
                        int fun_flag, double* coeffs_host);
int chem_pot_device(Ctx& c, int L, double emin, double emax, double beta,
                    const double* moments, double nsto, double ne_true, double mu_init,
                    double gap, double occ, double tol, double* mu_out, double* ne_out);
// the same search on the library's fixed per-device scratch array (no device
// allocation); noalloc_scratch_setup reads that array's address on `device`
// (the current device) once, at context creation
int noalloc_scratch_setup(int device);
int chem_pot_device_noalloc(Ctx& c, int L, double emin, double emax, double beta,
                            const double* moments, double nsto, double ne_true, double mu_init,
                            double gap, double occ, double tol, double* mu_out, double* ne_out);
int moments_reduce(Ctx& c, const double* mom_part, int nsteps, int nwg, double* out3);
int energy_ke_device(Ctx& c, const double* X, int nstate, double* out);
int energy_nl_device(Ctx& c, int nstate_pad, int nstate, double* out);
int normalize0_device(Ctx& c, double* X);
int power_update_device(Ctx& c, double* X, const double* F, int mode, double emax,
                        double* energy_host);

}  // namespace sdft

// The opaque handles of the C ABI.
struct sdft_ctx {
  sdft::Ctx c;
};
struct sdft_block {
  sdft::Block b;
};
