// C ABI of libthcgw: context, workspaces and the stage pipelines of one THC GW iteration.
#include "../../include/thcgw.h"
#include "thcgw_internal.hpp"
#include "thcgw_host.hpp"
#include "lri.hpp"
#include "mlp.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <vector>

using thcgw::c128;
using thcgw::c64;
using thcgw::Cplx;
using thcgw::DT_C128;
using thcgw::DT_C64;

namespace {

thread_local std::string g_create_error;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// Staging buffer of the entry point that allocates nothing (thcgw_primary_to_aux_noalloc):
// a file-scope device array, so each GPU that loads the library holds its own copy until
// the process ends.  Host code never passes the array's name as an address: its device
// address is resolved once per device by hipGetSymbolAddress (noalloc_scratch, called by
// thcgw_create with that device current) and kept per device.
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;  // 64 MiB
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];
std::mutex g_noalloc_mu;
std::map<int, void*> g_noalloc_ptr;  // device -> address of that device's copy of scratch

void* noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  auto it = g_noalloc_ptr.find(device);
  if (it != g_noalloc_ptr.end()) return it->second;
  int cur = -1;
  THCGW_HIP(hipGetDevice(&cur));
  THCGW_REQUIRE(cur == device, "noalloc scratch: the context's device is not current");
  void* p = nullptr;
  THCGW_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  THCGW_REQUIRE(p != nullptr, "noalloc scratch: hipGetSymbolAddress returned no address");
  g_noalloc_ptr[device] = p;
  return p;
}

}  // namespace

struct thcgw_ctx {
  thcgw_dims d{};
  int nk = 0, nt_half = 0, nw_half = 0;
  long M = 0;  // nmu^2
  int device = 0;
  thcgw::host::Tables tab;  // rescaled host tables
  double* d_tau = nullptr;  // tau values (nt)
  c128* d_Tplus = nullptr;  // nw_half x nt_half (IAFT.icc:62-70)
  c128* d_Tminus = nullptr; // nt_half x nw_half (IAFT.icc:117-125)
  bool real_Tplus = false;  // the folded tables are real up to rounding: only Re T is applied
  bool real_Tminus = false;
  c128* d_Twt_ff = nullptr; // nw_f x nt
  c128* d_Ttw_ff = nullptr; // nt x nw_f
  c128* d_Tbeta = nullptr;  // 1 x nt
  long* d_wn_f = nullptr;   // nw_f fermionic Matsubara indices
  int* d_kmq = nullptr;     // k - q table [k][q] and its transpose [q][k] (k-space convolutions)
  int trev = 0;             // time-reversal symmetric input: Dyson on the irreducible q only
  int nq_ir = 0;            // irreducible q (q <= -q) under time reversal
  int* d_trq = nullptr;     // irreducible q list (nq_ir) then the -q map (nk)
  int dy_nq = 0;            // V members of the Dyson chunk functions (0 = nk)
  // q-point symmetry map (thcgw_set_qsym): W(q) = D_q P_q W^(src q) P_q^T D_q^*, W^ = W or conj W
  int sym = 0;
  int sym_nir = 0;          // irreducible q (src q == q)
  int* d_sym = nullptr;     // src (nk), conj (nk), irreducible list (sym_nir), identity (nk), perm (nk * nmu)
  c128* d_symph = nullptr;  // phases (nk * nmu)
  // Brillouin-zone symmetry (thcgw_set_bz): the wedge path's device tables
  int bz_on = 0, bz_nk_ibz = 0, bz_nq_ibz = 0, bz_nk_nt = 0, bz_nsym = 0, bz_nterms = 0;
  std::vector<double> bz_qw;  // multiplicity of each irreducible q
  int* d_bz = nullptr;        // int tables at the offsets bz_o[BZ_*]
  long bz_o[10] = {};
  c128* d_bzD = nullptr;      // orbital rotations (nsym x nk_ibz x no x no)
  int vq() const { return dy_nq > 0 ? dy_nq : nk; }
  std::map<std::string, DevBuf> ws;
  std::string err;
  bool force_fallback = false;
  int pair_chunk = 0;   // tau pairs per projection batch (0 = automatic)
  int dyson_chunk = 0;  // matrices per Dyson batch (0 = automatic)
  int mem_share = 1;    // scratch pools side by side (batch lanes): automatic chunks take 1/mem_share
  int dyson_lanes = 1;  // flag 11: resolvent Dyson chunks pipelined over two streams (2) or not (1, default)
  hipStream_t dy_s[2] = {nullptr, nullptr};
  hipEvent_t dy_ev[3] = {nullptr, nullptr, nullptr};
  unsigned long* scratch_gen = nullptr;  // bumped on every scratch (re)allocation (a batch's graph)
  int herm_g = 0;       // flag 10: G(tau, k) Hermitian, its projections formed on the lower triangle
  int dyson_algo = 0;   // 0 = LU with partial pivoting (reference algorithm), 1 = Hermitian (Cholesky),
                        // 2 = resolvent (V^-1 - Pi)^-1
  int precision = 0;    // 0 = FP64 throughout, 1 = mixed (FP32 products, FP64 factorisations and IR),
                        // 2 = mixed with split-BF16 products in place of FP32 MFMA
  int tile32 = -1;      // FP32 tile shape chosen for this context (-1: process default)
  bool deferred = false;  // enqueue without host synchronisation (stream capture): checks run later
  int nf_arrays = 0;      // arrays whose non-finite counts a deferred enqueue left in "nf_cnt"
  std::map<std::string, DevBuf>* pool = nullptr;  // scratch shared with the other systems of a batch
  // cgemm flags of this context: precision and FP32 tile shape
  int pf() const {
    return (precision == 2 ? thcgw::GEMM_SPLIT_BF16 : 0) | (tile32 >= 0 ? (tile32 + 1) << thcgw::GEMM_TILE32_SHIFT : 0);
  }
  // Arrays read back after the stream completes stay private to the context.
  static bool private_buf(const std::string& name) {
    return name == "dy_info" || name == "dy_vinfo" || name == "ec_tr" || name == "ec_ld" || name == "nf_cnt";
  }

  template <class T>
  T* buf(const std::string& name, size_t count) {
    DevBuf& b = (pool && !private_buf(name)) ? (*pool)[name] : ws[name];
    const size_t need = count * sizeof(T);
    if (b.bytes < need) {
      if (b.p) THCGW_HIP(hipFree(b.p));
      b.p = nullptr;
      b.bytes = 0;
      hipError_t e = hipMalloc(&b.p, need);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        throw thcgw::Error(THCGW_ERR_ALLOC, "hipMalloc of " + std::to_string(need) + " bytes for workspace '" +
                                                        name + "' failed: " + hipGetErrorString(e));
      }
      b.bytes = need;
      if (scratch_gen) ++*scratch_gen;
    }
    return static_cast<T*>(b.p);
  }
  // Bytes currently held by workspace `name` (0 if not allocated).
  size_t held_bytes(const std::string& name) const {
    const std::map<std::string, DevBuf>& m = (pool && !private_buf(name)) ? *pool : ws;
    const auto it = m.find(name);
    return it == m.end() ? 0 : it->second.bytes;
  }
  void release() {
    for (auto& kv : ws)
      if (kv.second.p) (void)hipFree(kv.second.p);
    ws.clear();
    if (d_tau) (void)hipFree(d_tau);
    if (d_Tplus) (void)hipFree(d_Tplus);
    if (d_Tminus) (void)hipFree(d_Tminus);
    if (d_Twt_ff) (void)hipFree(d_Twt_ff);
    if (d_Ttw_ff) (void)hipFree(d_Ttw_ff);
    if (d_Tbeta) (void)hipFree(d_Tbeta);
    if (d_wn_f) (void)hipFree(d_wn_f);
    if (d_kmq) (void)hipFree(d_kmq);
    d_kmq = nullptr;
    if (d_trq) (void)hipFree(d_trq);
    d_trq = nullptr;
    if (d_sym) (void)hipFree(d_sym);
    d_sym = nullptr;
    if (d_symph) (void)hipFree(d_symph);
    d_symph = nullptr;
    if (d_bz) (void)hipFree(d_bz);
    d_bz = nullptr;
    for (hipStream_t& x : dy_s) {
      if (x) {
        (void)hipStreamSynchronize(x);
        (void)hipStreamDestroy(x);
      }
      x = nullptr;
    }
    for (hipEvent_t& e : dy_ev) {
      if (e) (void)hipEventDestroy(e);
      e = nullptr;
    }
    if (d_bzD) (void)hipFree(d_bzD);
    d_bzD = nullptr;
    d_tau = nullptr;
    d_Tplus = d_Tminus = d_Twt_ff = d_Ttw_ff = d_Tbeta = nullptr;
    d_wn_f = nullptr;
  }
};

// A batch of independent systems (uneven sizes), each with its own context.  The systems
// are spread over `lanes` streams (one by default); the systems of a lane run one after
// another and share that lane's scratch pool.  After a first eager iteration (which
// allocates the scratch, tunes each mixed-precision system's FP32 tile shape and keeps the
// chunk sizes) the whole batch iteration is captured into one HIP graph, the lanes as
// parallel branches between a fork and a join, and replayed.
struct thcgw_batch {
  std::vector<thcgw_ctx*> ctxs;
  std::vector<std::map<std::string, DevBuf>> pools{1};  // scratch per lane
  std::vector<int> lane_of;                              // lane of each system
  std::vector<hipStream_t> lane_s;                       // lane streams (lanes > 1)
  std::vector<hipEvent_t> lane_ev;                       // join events (lanes > 1)
  hipEvent_t ev_fork = nullptr;
  unsigned long gen = 0;  // scratch allocations of the contexts: a change invalidates the graph
  bool warmed = false;
  bool use_graph = true;
  bool tune = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<const void*> sig;  // pointers the graph was captured with
  int herm = -1;
  int device = 0;
  hipStream_t cap = nullptr;  // the batch's own stream: capture and replay (the legacy stream cannot be captured)
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  std::string err;
  int lanes() const { return static_cast<int>(pools.size()); }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    sig.clear();
  }
  void free_lanes() {
    for (hipStream_t x : lane_s) (void)hipStreamDestroy(x);
    for (hipEvent_t x : lane_ev) (void)hipEventDestroy(x);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    lane_s.clear();
    lane_ev.clear();
    ev_fork = nullptr;
  }
};

namespace {

template <class F>
int guarded(thcgw_ctx* ctx, F&& f) {
  try {
    if (ctx) THCGW_HIP(hipSetDevice(ctx->device));
    f();
    return THCGW_OK;
  } catch (const thcgw::Error& e) {
    if (ctx) ctx->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    if (ctx) ctx->err = "host allocation failed";
    return THCGW_ERR_ALLOC;
  } catch (const std::exception& e) {
    if (ctx) ctx->err = e.what();
    return THCGW_ERR_ARG;
  }
}

inline hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }
inline const c128* C(const double* p) { return reinterpret_cast<const c128*>(p); }
inline c128* Cm(double* p) { return reinterpret_cast<c128*>(p); }

const Cplx kOne{1.0, 0.0};
const Cplx kZero{0.0, 0.0};
const Cplx kMinusOne{-1.0, 0.0};

// Free device memory the pipelines may use for batching, leaving headroom.
size_t free_device_bytes() {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    (void)hipGetLastError();
    return size_t(8) << 30;
  }
  return fr;
}

// O_out[b,k] = X_k O_in[b,k] X_k^H for nbk = nb*nk matrices (thc_solver_comm.hpp:422,427).
// herm: O_in is Hermitian, so is O_out; its lower tiles are formed and mirrored.
void project_p2a(thcgw_ctx* ctx, hipStream_t s, const c128* X, const c128* Oin, c128* Oout, long nbk, c128* Abuf,
                 bool herm = false, int xmod = 0) {
  const int nk = xmod > 0 ? xmod : ctx->nk, no = ctx->d.no, nmu = ctx->d.nmu;
  const long sX = static_cast<long>(nmu) * no;
  THCGW_REQUIRE(nbk < (1L << 30), "primary_to_aux: batch too large");
  thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, nmu, no, no, kOne, thcgw::bp(X, sX, nk), no,
               thcgw::bp(Oin, static_cast<long>(no) * no), no, kZero, Abuf, sX, no, static_cast<int>(nbk));
  thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_C, nmu, nmu, no, kOne, thcgw::bp(Abuf, sX), no, thcgw::bp(X, sX, nk), no,
               kZero, Oout, ctx->M, nmu, static_cast<int>(nbk), kZero,
               herm ? (thcgw::GEMM_TRI_LOWER | thcgw::GEMM_HERM_SUB) : 0);
}

// O_out[b,k] = X_k^H O_in[b,k] X_k (thc_solver_comm.hpp:560,563).
void project_a2p(thcgw_ctx* ctx, hipStream_t s, const c128* X, const c128* Oin, c128* Oout, long nbk, c128* Bbuf) {
  const int nk = ctx->nk, no = ctx->d.no, nmu = ctx->d.nmu;
  const long sX = static_cast<long>(nmu) * no;
  THCGW_REQUIRE(nbk < (1L << 30), "aux_to_primary: batch too large");
  thcgw::zgemm(s, thcgw::OP_C, thcgw::OP_N, no, nmu, nmu, kOne, thcgw::bp(X, sX, nk), no, thcgw::bp(Oin, ctx->M), nmu,
               kZero, Bbuf, sX, nmu, static_cast<int>(nbk));
  thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, no, no, nmu, kOne, thcgw::bp(Bbuf, sX), nmu, thcgw::bp(X, sX, nk), no,
               kZero, Oout, static_cast<long>(no) * no, no, static_cast<int>(nbk));
}

// The same for a Hermitian O = L + L^H given by its lower half L (the lower triangle with the
// diagonal halved, zeros above it within the diagonal 128 x 128 blocks): X^H O X = Y + Y^H for
// Y = X^H L X, whose first product has a lower triangular operand (half the work of the
// full one).
void project_a2p_herm(thcgw_ctx* ctx, hipStream_t s, const c128* X, const c128* Lin, c128* Oout, long nbk,
                      c128* Bbuf) {
  const int nk = ctx->nk, no = ctx->d.no, nmu = ctx->d.nmu;
  const long sX = static_cast<long>(nmu) * no;
  THCGW_REQUIRE(nbk < (1L << 30), "aux_to_primary: batch too large");
  const c128* Oin = Lin;
  thcgw::zgemm(s, thcgw::OP_C, thcgw::OP_N, no, nmu, nmu, kOne, thcgw::bp(X, sX, nk), no, thcgw::bp(Oin, ctx->M), nmu,
               kZero, Bbuf, sX, nmu, static_cast<int>(nbk), kZero, thcgw::GEMM_TRI_KFROM_COL);
  thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, no, no, nmu, kOne, thcgw::bp(Bbuf, sX), nmu, thcgw::bp(X, sX, nk), no,
               kZero, Oout, static_cast<long>(no) * no, no, static_cast<int>(nbk));
  thcgw::hermitize(s, Oout, no, nbk, 2.0);
}

// Mixed precision: X rounded to complex<float> once per call, FP32 products.
const c64* x_c64(thcgw_ctx* ctx, const c128* X, hipStream_t s) {
  const long n = static_cast<long>(ctx->nk) * ctx->d.nmu * ctx->d.no;
  c64* X64 = ctx->buf<c64>("mx_x64", static_cast<size_t>(n));
  thcgw::to_c64(s, X, X64, n);
  return X64;
}

// O_out[b,k] = X_k O_in[b,k] X_k^H with O_in complex<double>, result complex<float>; herm as
// for project_p2a.
void project_p2a_f32(thcgw_ctx* ctx, hipStream_t s, const c64* X64, const c128* Oin, c64* Oout, long nbk, c64* Abuf,
                     bool herm = false, bool g3 = false) {
  const int f = ctx->pf() | (g3 ? thcgw::GEMM_G3 : 0);
  const int nk = ctx->nk, no = ctx->d.no, nmu = ctx->d.nmu;
  const long sX = static_cast<long>(nmu) * no;
  THCGW_REQUIRE(nbk < (1L << 30), "primary_to_aux: batch too large");
  thcgw::cgemm(s, thcgw::OP_N, thcgw::OP_N, nmu, no, no, kOne, thcgw::bp(X64, sX, nk), DT_C64, no,
               thcgw::bp(Oin, static_cast<long>(no) * no), DT_C128, no, kZero, Abuf, DT_C64, sX, no,
               static_cast<int>(nbk), kZero, f);
  thcgw::cgemm(s, thcgw::OP_N, thcgw::OP_C, nmu, nmu, no, kOne, thcgw::bp(Abuf, sX), DT_C64, no,
               thcgw::bp(X64, sX, nk), DT_C64, no, kZero, Oout, DT_C64, ctx->M, nmu, static_cast<int>(nbk), kZero,
               f | (herm ? (thcgw::GEMM_TRI_LOWER | thcgw::GEMM_HERM_SUB) : 0));
}

// O_out[b,k] = X_k^H O_in[b,k] X_k with O_in complex<float>, result complex<double>; herm: O_in
// holds the lower half of a Hermitian matrix, as for project_a2p_herm.
void project_a2p_f32(thcgw_ctx* ctx, hipStream_t s, const c64* X64, const c64* Oin, c128* Oout, long nbk, c64* Bbuf,
                     bool herm = false) {
  const int nk = ctx->nk, no = ctx->d.no, nmu = ctx->d.nmu;
  const long sX = static_cast<long>(nmu) * no;
  THCGW_REQUIRE(nbk < (1L << 30), "aux_to_primary: batch too large");
  // Both products sum over N_mu: each K chunk is folded into FP64 accumulators, so the
  // rounding grows with the chunk length instead of with N_mu.
  thcgw::cgemm(s, thcgw::OP_C, thcgw::OP_N, no, nmu, nmu, kOne, thcgw::bp(X64, sX, nk), DT_C64, no,
               thcgw::bp(Oin, ctx->M), DT_C64, nmu, kZero, Bbuf, DT_C64, sX, nmu,
               static_cast<int>(nbk), kZero,
               ctx->pf() | thcgw::GEMM_FOLD64 | (herm ? thcgw::GEMM_TRI_KFROM_COL : 0));
  thcgw::cgemm(s, thcgw::OP_N, thcgw::OP_N, no, no, nmu, kOne, thcgw::bp(Bbuf, sX), DT_C64, nmu,
               thcgw::bp(X64, sX, nk), DT_C64, no, kZero, Oout, DT_C128, static_cast<long>(no) * no, no,
               static_cast<int>(nbk), kZero, ctx->pf() | thcgw::GEMM_FOLD64);
  if (herm) thcgw::hermitize(s, Oout, no, nbk, 2.0);
}

int choose_pair_chunk(thcgw_ctx* ctx, int per_pair_arrays) {
  if (ctx->pair_chunk > 0) return std::min(ctx->pair_chunk, ctx->nt_half);
  const double per = double(per_pair_arrays) * ctx->nk * double(ctx->M) * 16.0;
  const double budget = std::min(0.25 * double(free_device_bytes()), 24.0 * double(1L << 30)) / ctx->mem_share;
  int c = static_cast<int>(budget / per);
  return std::max(1, std::min(c, ctx->nt_half));
}

std::string sci(double v);

// Flag 10 refuses G(tau, k) that is not Hermitian: max |G - G^H| over all (tau, k) against
// 1e-12 of max |G| (G built from a Hermitian H(k) is Hermitian to rounding).
void check_herm_g(thcgw_ctx* ctx, const c128* G, hipStream_t s, int nk_g = 0) {
  if (!ctx->herm_g || ctx->deferred) return;
  const int nbx = 256;
  double* d = ctx->buf<double>("hg_chk", static_cast<size_t>(2) * nbx);
  thcgw::herm_diff(s, G, ctx->d.no, static_cast<long>(ctx->d.nt) * (nk_g > 0 ? nk_g : ctx->nk), nbx, d);
  std::vector<double> h(static_cast<size_t>(2) * nbx);
  THCGW_HIP(hipMemcpyAsync(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
  THCGW_HIP(hipStreamSynchronize(s));
  double diff = 0.0, mag = 0.0;
  for (size_t i = 0; i < h.size(); i += 2) {
    diff = std::max(diff, h[i]);
    mag = std::max(mag, h[i + 1]);
  }
  if (!(diff <= 1e-12 * mag))
    throw thcgw::Error(THCGW_ERR_ARG, "flag 10 (Hermitian G): G is not Hermitian, max |G - G^H| = " + sci(diff) +
                                          " against max |G| = " + sci(mag));
}

void run_pi_f32(thcgw_ctx* ctx, const c128* G, const c128* X, c128* Pi, hipStream_t s) {
  const int nk = ctx->nk, no = ctx->d.no, nmu = ctx->d.nmu, nt = ctx->d.nt, nth = ctx->nt_half;
  const long M = ctx->M;
  const long gs = static_cast<long>(nk) * no * no;
  const int P = choose_pair_chunk(ctx, 1);
  const c64* X64 = x_c64(ctx, X, s);
  c64* GP = ctx->buf<c64>("mx_gp", static_cast<size_t>(P) * nk * M);
  c64* GN = ctx->buf<c64>("mx_gn", static_cast<size_t>(P) * nk * M);
  c64* Ab = ctx->buf<c64>("mx_ab", static_cast<size_t>(P) * nk * nmu * no);
  const double scale = -2.0 / (double(nk) * nk);
  for (int i0 = 0; i0 < nth; i0 += P) {
    const int c = std::min(P, nth - i0);
    project_p2a_f32(ctx, s, X64, G + i0 * gs, GP, static_cast<long>(c) * nk, Ab, ctx->herm_g);
    project_p2a_f32(ctx, s, X64, G + (nt - i0 - c) * gs, GN, static_cast<long>(c) * nk, Ab, ctx->herm_g);
    for (int j = 0; j < c; ++j)
      thcgw::pi_kfft_f32(s, ctx->d.kgrid, ctx->force_fallback, GP + static_cast<long>(j) * nk * M,
                         GN + static_cast<long>(c - 1 - j) * nk * M, Pi + static_cast<long>(i0 + j) * nk * M, M, scale);
  }
}

void run_pi(thcgw_ctx* ctx, const c128* G, const c128* X, c128* Pi, hipStream_t s) {
  check_herm_g(ctx, G, s);
  if (ctx->precision != 0) return run_pi_f32(ctx, G, X, Pi, s);
  const int nk = ctx->nk, no = ctx->d.no, nmu = ctx->d.nmu, nt = ctx->d.nt, nth = ctx->nt_half;
  const long M = ctx->M;
  const long gs = static_cast<long>(nk) * no * no;  // one tau slice of G
  const int P = choose_pair_chunk(ctx, 2);
  c128* GP = ctx->buf<c128>("pi_gp", static_cast<size_t>(P) * nk * M);
  c128* GN = ctx->buf<c128>("pi_gn", static_cast<size_t>(P) * nk * M);
  c128* Ab = ctx->buf<c128>("proj_ab", static_cast<size_t>(P) * nk * nmu * no);
  const double scale = -2.0 / (double(nk) * nk);  // sp = -2 for ns = 1 (rpa_pi.icc:46)
  for (int i0 = 0; i0 < nth; i0 += P) {
    const int c = std::min(P, nth - i0);
    project_p2a(ctx, s, X, G + i0 * gs, GP, static_cast<long>(c) * nk, Ab, ctx->herm_g);
    // G(beta - tau_i) = G[nt-1-i] for i in [i0, i0+c): tau slots [nt-i0-c, nt-i0), reversed
    project_p2a(ctx, s, X, G + (nt - i0 - c) * gs, GN, static_cast<long>(c) * nk, Ab, ctx->herm_g);
    for (int j = 0; j < c; ++j)
      thcgw::pi_kfft(s, ctx->d.kgrid, ctx->force_fallback, GP + static_cast<long>(j) * nk * M,
                     GN + static_cast<long>(c - 1 - j) * nk * M, Pi + static_cast<long>(i0 + j) * nk * M, M, scale);
  }
}

// Pi(iW_n, q, PQ) = sum_i T+[n][i] Pi(tau_i, q, PQ): one GEMM (nw_half x nt_half) x (nt_half x N_q N_mu^2).
void run_tau_to_w(thcgw_ctx* ctx, const c128* Pit, c128* Piw, hipStream_t s, int nq = 0) {
  const long cols = ctx->M * (nq > 0 ? nq : ctx->nk);
  if (thcgw::ir_transform_fits(ctx->nw_half, ctx->nt_half)) {
    thcgw::ir_transform(s, ctx->d_Tplus, ctx->nw_half, ctx->nt_half, Pit, cols, 0, Piw, cols, 0, cols, 1,
                        ctx->real_Tplus);
    return;
  }
  THCGW_REQUIRE(cols < (1L << 31), "tau_to_w: N_q N_mu^2 too large for one GEMM");
  thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, ctx->nw_half, static_cast<int>(cols), ctx->nt_half, kOne,
               thcgw::bp(ctx->d_Tplus, 0), ctx->nt_half, thcgw::bp(Pit, 0), cols, kZero, Piw, 0, cols, 1);
}

// W(q, tau_i, PQ) = sum_n T-[i][n] W(iW_n, q, PQ): per q a GEMM (nt_half x nw_half) x (nw_half x N_mu^2)
// whose output rows land at stride N_mu^2 inside the q block of W_qtPQ.
void run_w_to_tau(thcgw_ctx* ctx, const c128* Ww, c128* Wqt, hipStream_t s, int nq = 0) {
  const long M = ctx->M, nk = nq > 0 ? nq : ctx->nk;
  if (thcgw::ir_transform_fits(ctx->nt_half, ctx->nw_half)) {
    thcgw::ir_transform(s, ctx->d_Tminus, ctx->nt_half, ctx->nw_half, Ww, nk * M, M, Wqt, M, ctx->nt_half * M, M,
                        static_cast<int>(nk), ctx->real_Tminus);
    return;
  }
  THCGW_REQUIRE(M < (1L << 31), "w_to_tau: N_mu^2 too large for one GEMM");
  thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, ctx->nt_half, static_cast<int>(M), ctx->nw_half, kOne,
               thcgw::bp(ctx->d_Tminus, 0), ctx->nw_half, thcgw::bp(Ww, M), nk * M, kZero, Wqt, ctx->nt_half * M, M,
               static_cast<int>(nk));
}

int choose_dyson_chunk(thcgw_ctx* ctx, int total) {
  if (ctx->dyson_chunk > 0) return std::min(ctx->dyson_chunk, total);
  if (total <= 1) return 1;
  const double per = 2.0 * double(ctx->M) * 16.0;
  // The chunk buffers already held are reused, so they count as free: otherwise the chunk
  // shrinks after the first call wherever memory is short.  Larger chunks run faster per
  // matrix (the recursion's small levels take one launch per chunk): at N_k = 64,
  // N_mu = 1600 the mixed resolvent iteration takes 1.225 s in chunks sized to 45% of the free
  // memory, 1.208 s with 437 matrices and 1.204 s with 525, while 656 (54 GB of chunk
  // buffers) left the runtime without memory for its own queues.  60% keeps 40% of the
  // free memory for the runtime and later workspaces.
  double held = 0.0;
  for (const char* name : {"dy_lu", "dy_z", "dy_z64", "dy_m64", "dy_lu1", "dy_z1"})
    held += double(ctx->held_bytes(name));
  const double budget =
      std::min(0.6 * (double(free_device_bytes()) + held), 64.0 * double(1L << 30)) / ctx->mem_share;
  const int c = std::max(1, std::min(static_cast<int>(budget / per), total));
  // equal chunks rather than full ones and a short last one
  const int nchunks = (total + c - 1) / c;
  return (total + nchunks - 1) / nchunks;
}

// One chunk of matrices b0 .. b0+cb of the LU form: A = I - V Pi (or I - Pi V) -> getrf ->
// ln|Re det|; with solve, (I - V Pi) Y = V and W - V = Y - V replace getri + GEMM
// (scr_coulomb_t.cpp:180-196, thc_rpa.icc:141-163).
void dyson_chunk_lu(thcgw_ctx* ctx, c128* PiW, const c128* V, int b0, int cb, c128* d_trace, double* d_ldet,
                    bool solve, bool pi_first, c128* LU, c128* Z, int* ipiv, int* info, hipStream_t s) {
  const int n = ctx->d.nmu, nq = ctx->vq();
  const long M = ctx->M;
  if (!pi_first)
    thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, n, n, n, kMinusOne, thcgw::bp(V, M, nq, 1, b0), n,
                 thcgw::bp(PiW + static_cast<long>(b0) * M, M), n, kZero, LU, M, n, cb, kOne);
  else
    thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, n, n, n, kMinusOne, thcgw::bp(PiW + static_cast<long>(b0) * M, M), n,
                 thcgw::bp(V, M, nq, 1, b0), n, kZero, LU, M, n, cb, kOne);
  if (d_trace) thcgw::one_minus_trace(s, LU, M, n, d_trace + b0, cb);
  thcgw::getrf_batched(s, LU, M, n, n, ipiv, info, cb);
  if (d_ldet) thcgw::log_re_det(s, LU, M, n, ipiv, d_ldet + b0, cb);
  if (solve) {
    thcgw::copy_batch(s, V, M, nq, Z, M, M, b0, cb);
    thcgw::trsm_left_lower(s, LU, M, n, Z, M, n, n, n, cb);
    thcgw::trsm_left_upper_unit(s, LU, M, n, Z, M, n, n, n, cb);
    thcgw::dyson_finish(s, Z, M, V, M, nq, ipiv, n, PiW + static_cast<long>(b0) * M, M, b0, cb);
  }
}

bool any_nonzero(const int* d_info, int count, hipStream_t s) {
  std::vector<int> h(count);
  THCGW_HIP(hipMemcpyAsync(h.data(), d_info, sizeof(int) * count, hipMemcpyDeviceToHost, s));
  THCGW_HIP(hipStreamSynchronize(s));
  for (int x : h)
    if (x != 0) return true;
  return false;
}

// Lower Cholesky factors L_q of V_q (strict upper triangles zeroed); false when some V_q
// is not positive definite.
bool factor_v(thcgw_ctx* ctx, const c128* V, c128* Lv, hipStream_t s) {
  const int n = ctx->d.nmu, nq = ctx->vq();
  const long M = ctx->M;
  int* info = ctx->buf<int>("dy_vinfo", static_cast<size_t>(nq));
  thcgw::copy_batch(s, V, M, nq, Lv, M, M, 0, nq);
  thcgw::potrf_batched(s, Lv, M, n, n, info, nq);
  thcgw::zero_upper(s, Lv, M, n, nq);
  if (ctx->deferred) return true;  // checked after the stream completes
  return !any_nonzero(info, nq, s);
}

// One chunk of the Hermitian form (see chol.hip): M = I - L^H Pi L, M = L_M L_M^H,
// ln det = 2 sum ln L_M[j][j], W = K K^H - V with K = L L_M^-H.  Returns false, leaving
// the chunk's Pi slots untouched, when some M is not positive definite.
bool dyson_chunk_hpd(thcgw_ctx* ctx, c128* PiW, const c128* V, const c128* Lv, int b0, int cb, c128* d_trace,
                     double* d_ldet, bool solve, c128* Mb, c128* Z, int* info, hipStream_t s) {
  const int n = ctx->d.nmu, nq = ctx->vq();
  const long M = ctx->M;
  c128* P = PiW + static_cast<long>(b0) * M;
  thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, n, n, n, kOne, thcgw::bp(P, M), n, thcgw::bp(Lv, M, nq, 1, b0), n, kZero,
               Z, M, n, cb, kZero, thcgw::GEMM_TRI_KFROM_COL);
  thcgw::zgemm(s, thcgw::OP_C, thcgw::OP_N, n, n, n, kMinusOne, thcgw::bp(Lv, M, nq, 1, b0), n, thcgw::bp(Z, M), n,
               kZero, Mb, M, n, cb, kOne, thcgw::GEMM_TRI_LOWER | thcgw::GEMM_TRI_KFROM_ROW);
  if (d_trace) thcgw::one_minus_trace(s, Mb, M, n, d_trace + b0, cb);
  thcgw::potrf_batched(s, Mb, M, n, n, info, cb);
  if (!ctx->deferred && any_nonzero(info, cb, s)) return false;
  if (d_ldet) thcgw::chol_logdet(s, Mb, M, n, d_ldet + b0, cb);
  if (solve) {
    thcgw::copy_batch(s, Lv, M, nq, Z, M, M, b0, cb);
    thcgw::trsm_right_lower_conj(s, Mb, M, n, Z, M, n, n, n, cb);
    // W = K K^H - V, both triangles written from the lower tiles
    thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_C, n, n, n, kOne, thcgw::bp(Z, M), n, thcgw::bp(Z, M), n, kZero, P, M, n,
                 cb, kZero, thcgw::GEMM_TRI_LOWER | thcgw::GEMM_HERM_SUB, thcgw::bp(V, M, nq, 1, b0));
  }
  return true;
}

// The Hermitian form with the products in FP32 (L, C and K stored as complex<float>) and
// M, its Cholesky factor and ln det in FP64.
bool dyson_chunk_hpd_f32(thcgw_ctx* ctx, c128* PiW, const c128* V, const c64* Lv64, int b0, int cb, c128* d_trace,
                         double* d_ldet, bool solve, c128* Mb, c64* Z, c64* Mb64, int* info, hipStream_t s) {
  const int n = ctx->d.nmu, nq = ctx->vq();
  const long M = ctx->M;
  c128* P = PiW + static_cast<long>(b0) * M;
  // Pi rounded once to complex<float> (the FP32 product rounds it on load anyway), which
  // halves the bytes of its repeated tile reads; M feeds ln det, so both products keep
  // FP64 sums across K chunks.
  thcgw::to_c64(s, P, Mb64, static_cast<long>(cb) * M);
  thcgw::cgemm(s, thcgw::OP_N, thcgw::OP_N, n, n, n, kOne, thcgw::bp(static_cast<const c64*>(Mb64), M), DT_C64, n,
               thcgw::bp(Lv64, M, nq, 1, b0), DT_C64, n, kZero, Z, DT_C64, M, n, cb, kZero,
               thcgw::GEMM_TRI_KFROM_COL | thcgw::GEMM_FOLD64 | ctx->pf());
  thcgw::cgemm(s, thcgw::OP_C, thcgw::OP_N, n, n, n, kMinusOne, thcgw::bp(Lv64, M, nq, 1, b0), DT_C64, n,
               thcgw::bp(Z, M), DT_C64, n, kZero, Mb, DT_C128, M, n, cb, kOne,
               thcgw::GEMM_TRI_LOWER | thcgw::GEMM_TRI_KFROM_ROW | thcgw::GEMM_FOLD64 | ctx->pf());
  if (d_trace) thcgw::one_minus_trace(s, Mb, M, n, d_trace + b0, cb);
  thcgw::potrf_batched(s, Mb, M, n, n, info, cb);
  if (!ctx->deferred && any_nonzero(info, cb, s)) return false;
  if (d_ldet) thcgw::chol_logdet(s, Mb, M, n, d_ldet + b0, cb);
  if (solve) {
    thcgw::copy_batch(s, Lv64, M, nq, Z, M, M, b0, cb);
    thcgw::to_c64(s, Mb, Mb64, static_cast<long>(cb) * M);  // L_M rounded once for the FP32 solve
    thcgw::trsm_right_lower_conj_f32(s, static_cast<const c64*>(Mb64), M, n, Z, M, n, n, n, cb, ctx->pf());
    if (ctx->precision == 2) {  // the split-BF16 kernel has no Hermitian epilogue
      thcgw::cgemm(s, thcgw::OP_N, thcgw::OP_C, n, n, n, kOne, thcgw::bp(Z, M), DT_C64, n, thcgw::bp(Z, M), DT_C64,
                   n, kZero, P, DT_C128, M, n, cb, kZero, thcgw::GEMM_TRI_LOWER | ctx->pf());
      thcgw::herm_finish(s, P, M, V, M, nq, b0, n, cb);
    } else {  // W = K K^H - V, both triangles written from the lower tiles
      thcgw::cgemm(s, thcgw::OP_N, thcgw::OP_C, n, n, n, kOne, thcgw::bp(Z, M), DT_C64, n, thcgw::bp(Z, M), DT_C64,
                   n, kZero, P, DT_C128, M, n, cb, kZero, thcgw::GEMM_TRI_LOWER | thcgw::GEMM_HERM_SUB | ctx->pf(),
                   thcgw::bp(V, M, nq, 1, b0));
    }
  }
  return true;
}

// One chunk of the resolvent form: W + V = (I - V Pi)^-1 V = (V^-1 - Pi)^-1 = A^-1, with
// A = V^-1 - Pi = L_A L_A^H (Hermitian positive definite for Pi <= 0), A^-1 = Z Z^H with
// Z = L_A^-H, and ln det(I - V Pi) = ln det V + ln det A.  V^-1 and ln det V are formed once
// per q (factor_dyson_v).  Per matrix this is n^3/3 complex multiply-adds for the Cholesky
// factorisation together with the inverse of its factor (cholinv_batched) and n^3/6 for
// the triangular product, all in FP64.  Returns false, leaving the chunk's Pi slots
// untouched, when some A is not positive definite.
// The chunk's front half: A and its Cholesky factor with Z (reads the chunk's Pi slots,
// writes only Ab, Z, the trace terms and the flags); `trp` names the trace-parts buffer.
void resolvent_front(thcgw_ctx* ctx, const c128* PiW, const c128* V, const c128* Vi, int b0, int cb, c128* d_trace,
                     c128* Ab, c128* Z, int* info, hipStream_t s, const char* trp) {
  const int n = ctx->d.nmu, nq = ctx->vq();
  const long M = ctx->M;
  const c128* P = PiW + static_cast<long>(b0) * M;
  c128* part = d_trace ? ctx->buf<c128>(trp, static_cast<size_t>(cb) * thcgw::resolvent_prep_parts(n)) : nullptr;
  thcgw::resolvent_prep(s, V, Vi, M, nq, b0, P, Ab, M, n, d_trace ? d_trace + b0 : nullptr, part, cb);
  thcgw::cholinv_batched(s, Ab, M, Z, M, n, info, cb);
}

// The back half, once the factorisations are known to have succeeded: ln det and
// W = Z Z^H - V into the chunk's Pi slots.
void resolvent_back(thcgw_ctx* ctx, c128* PiW, const c128* V, const double* ldv, int b0, int cb, double* d_ldet,
                    bool solve, const c128* Ab, const c128* Z, hipStream_t s) {
  const int n = ctx->d.nmu, nq = ctx->vq();
  const long M = ctx->M;
  c128* P = PiW + static_cast<long>(b0) * M;
  if (d_ldet) thcgw::chol_logdet(s, Ab, M, n, d_ldet + b0, cb, ldv, nq, b0);
  if (solve) {
    // W = Z Z^H - V from the lower tiles; Z is upper triangular, so the sum starts at the row
    thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_C, n, n, n, kOne, thcgw::bp(Z, M), n, thcgw::bp(Z, M), n, kZero, P, M, n,
                 cb, kZero, thcgw::GEMM_TRI_LOWER | thcgw::GEMM_TRI_KFROM_ROW | thcgw::GEMM_HERM_SUB,
                 thcgw::bp(V, M, nq, 1, b0));
  }
}

bool dyson_chunk_inv(thcgw_ctx* ctx, c128* PiW, const c128* V, const c128* Vi, const double* ldv, int b0, int cb,
                     c128* d_trace, double* d_ldet, bool solve, c128* Ab, c128* Z, int* info, hipStream_t s) {
  resolvent_front(ctx, PiW, V, Vi, b0, cb, d_trace, Ab, Z, info, s, "dy_trp");
  if (!ctx->deferred && any_nonzero(info, cb, s)) return false;
  resolvent_back(ctx, PiW, V, ldv, b0, cb, d_ldet, solve, Ab, Z, s);
  return true;
}

// Factors of V for the Hermitian and resolvent forms (V members: vq()).  hpd is false when
// the LU form is selected or some V_q is not positive definite; inv selects the resolvent
// form, for which Vi holds the lower triangles of V_q^-1 and ldv ln det V_q.
struct VFactors {
  bool hpd = false;
  bool inv = false;
  c128* Lv = nullptr;
  c64* Lv64 = nullptr;
  c128* Vi = nullptr;
  double* ldv = nullptr;
};

VFactors factor_dyson_v(thcgw_ctx* ctx, const c128* V, hipStream_t s) {
  VFactors f;
  const bool mixed = ctx->precision != 0;
  if (!(mixed || ctx->dyson_algo != 0)) return f;
  const int nq = ctx->vq(), n = ctx->d.nmu;
  const long M = ctx->M;
  f.Lv = ctx->buf<c128>("dy_lv", static_cast<size_t>(nq) * M);
  f.hpd = factor_v(ctx, V, f.Lv, s);
  if (f.hpd && ctx->dyson_algo == 2) {
    // V^-1 = T T^H with T = L^-H (upper triangular): lower tiles, K from the row
    f.inv = true;
    f.ldv = ctx->buf<double>("dy_ldv", static_cast<size_t>(nq));
    thcgw::chol_logdet(s, f.Lv, M, n, f.ldv, nq);
    c128* T = ctx->buf<c128>("dy_vt", static_cast<size_t>(nq) * M);
    thcgw::tri_inverse_conj(s, f.Lv, M, n, T, M, n, n, nq);
    f.Vi = f.Lv;  // L is no longer needed
    f.Lv = nullptr;
    thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_C, n, n, n, kOne, thcgw::bp(T, M), n, thcgw::bp(T, M), n, kZero, f.Vi, M, n,
                 nq, kZero, thcgw::GEMM_TRI_LOWER | thcgw::GEMM_TRI_KFROM_ROW);
  } else if (f.hpd && mixed) {
    f.Lv64 = ctx->buf<c64>("dy_lv64", static_cast<size_t>(nq) * M);
    thcgw::to_c64(s, f.Lv, f.Lv64, static_cast<long>(nq) * M);
  }
  return f;
}

// Flag 11: the resolvent chunks pipelined over two streams, each with its own buffers (the
// chunk size halved, so that both hold the memory of one).  Chunk c's front half runs on
// lane c % 2; while the host reads its factorisation flags, chunk c + 1's front half runs on
// the other lane; then chunk c's back half (or, if a factorisation failed, the LU form on
// its untouched Pi slots) follows on its own lane.  The small products of one chunk's
// recursion so overlap the other chunk's large ones.
void dyson_resolvent_lanes(thcgw_ctx* ctx, c128* PiW, const c128* V, const VFactors& f, int total, int CH,
                           c128* d_trace, double* d_ldet, bool solve, bool pi_first, int* info_all, hipStream_t s) {
  const int n = ctx->d.nmu;
  const long M = ctx->M;
  for (int l = 0; l < 2; ++l)
    if (!ctx->dy_s[l]) THCGW_HIP(hipStreamCreateWithFlags(&ctx->dy_s[l], hipStreamNonBlocking));
  for (hipEvent_t& e : ctx->dy_ev)
    if (!e) THCGW_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  c128* LU[2] = {ctx->buf<c128>("dy_lu", static_cast<size_t>(CH) * M), ctx->buf<c128>("dy_lu1", static_cast<size_t>(CH) * M)};
  c128* Z[2] = {ctx->buf<c128>("dy_z", static_cast<size_t>(CH) * M), ctx->buf<c128>("dy_z1", static_cast<size_t>(CH) * M)};
  // the lanes start after the caller's earlier work on s (Pi, the factors of V)
  THCGW_HIP(hipEventRecord(ctx->dy_ev[2], s));
  for (hipStream_t x : ctx->dy_s) THCGW_HIP(hipStreamWaitEvent(x, ctx->dy_ev[2], 0));
  const int nch = (total + CH - 1) / CH;
  auto b0 = [&](int c) { return c * CH; };
  auto cbn = [&](int c) { return std::min(CH, total - c * CH); };
  auto front = [&](int c) {
    const int l = c & 1;
    resolvent_front(ctx, PiW, V, f.Vi, b0(c), cbn(c), d_trace, LU[l], Z[l], info_all + b0(c), ctx->dy_s[l],
                    l ? "dy_trp1" : "dy_trp");
  };
  front(0);
  for (int c = 0; c < nch; ++c) {
    if (c + 1 < nch) front(c + 1);
    const int l = c & 1;
    hipStream_t ls = ctx->dy_s[l];
    if (!any_nonzero(info_all + b0(c), cbn(c), ls)) {
      resolvent_back(ctx, PiW, V, f.ldv, b0(c), cbn(c), d_ldet, solve, LU[l], Z[l], ls);
    } else {
      int* ipiv = ctx->buf<int>(l ? "dy_ipiv1" : "dy_ipiv", static_cast<size_t>(CH) * n);
      dyson_chunk_lu(ctx, PiW, V, b0(c), cbn(c), d_trace, d_ldet, solve, pi_first, LU[l], Z[l], ipiv, info_all + b0(c),
                     ls);
    }
  }
  for (int l = 0; l < 2; ++l) {
    THCGW_HIP(hipEventRecord(ctx->dy_ev[l], ctx->dy_s[l]));
    THCGW_HIP(hipStreamWaitEvent(s, ctx->dy_ev[l], 0));
  }
}

// In place Pi -> W - V for `total` matrices at PiW (matrix b uses V member b % vq()), in
// chunks; Tr, ln det and the factorisation flags at d_trace + b, d_ldet + b, info + b.
void dyson_matrices(thcgw_ctx* ctx, c128* PiW, const c128* V, const VFactors& f, int total, c128* d_trace,
                    double* d_ldet, bool solve, bool pi_first, int* info_all, hipStream_t s) {
  const int n = ctx->d.nmu;
  const long M = ctx->M;
  const int CH = choose_dyson_chunk(ctx, total);
  const bool mixed = ctx->precision != 0;
  if (f.inv && ctx->dyson_lanes > 1 && !ctx->deferred && total > 1)
    return dyson_resolvent_lanes(ctx, PiW, V, f, total, std::max(1, (CH + 1) / 2), d_trace, d_ldet, solve, pi_first,
                                 info_all, s);
  c128* LU = ctx->buf<c128>("dy_lu", static_cast<size_t>(CH) * M);
  c128* Z = nullptr;
  c64* Z64 = nullptr;
  c64* Mb64 = nullptr;
  int* ipiv = nullptr;
  for (int b0 = 0; b0 < total; b0 += CH) {
    const int cb = std::min(CH, total - b0);
    if (f.inv) {
      if (!Z) Z = ctx->buf<c128>("dy_z", static_cast<size_t>(CH) * M);
      if (dyson_chunk_inv(ctx, PiW, V, f.Vi, f.ldv, b0, cb, d_trace, d_ldet, solve, LU, Z, info_all + b0, s)) continue;
    } else if (f.hpd && mixed) {
      if (!Z64) Z64 = ctx->buf<c64>("dy_z64", static_cast<size_t>(CH) * M);
      if (!Mb64) Mb64 = ctx->buf<c64>("dy_m64", static_cast<size_t>(CH) * M);
      if (dyson_chunk_hpd_f32(ctx, PiW, V, f.Lv64, b0, cb, d_trace, d_ldet, solve, LU, Z64, Mb64, info_all + b0, s))
        continue;
    } else if (f.hpd) {
      if (!Z) Z = ctx->buf<c128>("dy_z", static_cast<size_t>(CH) * M);
      if (dyson_chunk_hpd(ctx, PiW, V, f.Lv, b0, cb, d_trace, d_ldet, solve, LU, Z, info_all + b0, s)) continue;
    }
    if (!ipiv) ipiv = ctx->buf<int>("dy_ipiv", static_cast<size_t>(CH) * n);
    if (solve && !Z) Z = ctx->buf<c128>("dy_z", static_cast<size_t>(CH) * M);
    dyson_chunk_lu(ctx, PiW, V, b0, cb, d_trace, d_ldet, solve, pi_first, LU, Z, ipiv, info_all + b0, s);
  }
}

// Time-reversal tables: the irreducible q (q <= -q, C order) and the -q map.
const int* trev_tables(thcgw_ctx* ctx) {
  if (ctx->d_trq) return ctx->d_trq;
  const int* g = ctx->d.kgrid;
  const int nq = ctx->nk;
  std::vector<int> mq(nq), ir;
  for (int i = 0; i < g[0]; ++i)
    for (int j = 0; j < g[1]; ++j)
      for (int l = 0; l < g[2]; ++l)
        mq[(i * g[1] + j) * g[2] + l] = (((g[0] - i) % g[0]) * g[1] + (g[1] - j) % g[1]) * g[2] + (g[2] - l) % g[2];
  for (int q = 0; q < nq; ++q)
    if (q <= mq[q]) ir.push_back(q);
  std::vector<int> h(ir);
  h.insert(h.end(), mq.begin(), mq.end());
  THCGW_HIP(hipMalloc(&ctx->d_trq, sizeof(int) * h.size()));
  THCGW_HIP(hipMemcpy(ctx->d_trq, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
  ctx->nq_ir = static_cast<int>(ir.size());
  return ctx->d_trq;
}

std::string sci(double v) {
  char b[32];
  std::snprintf(b, sizeof b, "%.3e", v);
  return b;
}

// Tolerance of the symmetry checks of the input: 1e-9 of max |A| in FP64; the mixed forms
// compute Pi from FP32 products, whose rounding breaks the relations at ~1e-7.
double sym_tol(const thcgw_ctx* ctx) { return ctx->precision ? 1e-5 : 1e-9; }

// Refuse input that is not time-reversal symmetric: max |A(-q) - conj A(q)| over the nq
// matrices at A against sym_tol of max |A|.
void check_trev(thcgw_ctx* ctx, const c128* A, const char* what, hipStream_t s) {
  const int* mq = trev_tables(ctx) + ctx->nq_ir;
  const int nbx = 64, nq = ctx->nk;
  double* d = ctx->buf<double>("tr_chk", static_cast<size_t>(2) * nq * nbx);
  thcgw::trev_diff(s, A, ctx->M, mq, nq, nbx, d);
  std::vector<double> h(static_cast<size_t>(2) * nq * nbx);
  THCGW_HIP(hipMemcpyAsync(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
  THCGW_HIP(hipStreamSynchronize(s));
  double diff = 0.0, mag = 0.0;
  for (size_t i = 0; i < h.size(); i += 2) {
    diff = std::max(diff, h[i]);
    mag = std::max(mag, h[i + 1]);
  }
  if (!(diff <= sym_tol(ctx) * mag))
    throw thcgw::Error(THCGW_ERR_ARG, std::string("time reversal (flag 9): ") + what +
                                          " is not symmetric, max |A(-q) - conj A(q)| = " + sci(diff) +
                                          " against max |A| = " + sci(mag));
}

// Time-reversal symmetric input: Pi(-q) = conj Pi(q) and V(-q) = conj V(q), hence
// W(-q) = conj W(q), Tr(V Pi)(-q) = conj Tr(V Pi)(q) and equal determinants.  Only the
// irreducible q are solved: the matrices of a block of frequencies are gathered into a
// staging array, solved with the irreducible V members, and scattered back to q and,
// conjugated, to -q, together with the Ec terms and flags.
void run_dyson_trev(thcgw_ctx* ctx, c128* PiW, const c128* V, c128* d_trace, double* d_ldet, bool solve,
                    bool pi_first, int* info_all, hipStream_t s) {
  const int* tq = trev_tables(ctx);
  const int nqi = ctx->nq_ir, nq = ctx->nk, nw = ctx->nw_half;
  const int *d_qir = tq, *d_mq = tq + nqi;
  const long M = ctx->M;
  if (!ctx->deferred) {
    check_trev(ctx, V, "V", s);
    check_trev(ctx, PiW, "Pi(iW_0)", s);
  }
  c128* Vir = ctx->buf<c128>("tr_v", static_cast<size_t>(nqi) * M);
  thcgw::gather_ir(s, V, M, 0, 0, d_qir, nqi, nqi, Vir);
  struct Restore {
    thcgw_ctx* c;
    ~Restore() { c->dy_nq = 0; }
  } restore{ctx};
  ctx->dy_nq = nqi;
  const VFactors f = factor_dyson_v(ctx, Vir, s);
  const int FB = std::max(1, choose_dyson_chunk(ctx, nw * nqi) / nqi);  // frequencies per block
  const int cap = std::min(FB, nw) * nqi;
  c128* S = ctx->buf<c128>("tr_s", static_cast<size_t>(cap) * M);
  c128* trs = d_trace ? ctx->buf<c128>("tr_tr", static_cast<size_t>(cap)) : nullptr;
  double* lds = d_ldet ? ctx->buf<double>("tr_ld", static_cast<size_t>(cap)) : nullptr;
  int* infos = ctx->buf<int>("tr_info", static_cast<size_t>(cap));
  for (int n0 = 0; n0 < nw; n0 += FB) {
    const int cnt = std::min(FB, nw - n0) * nqi;
    thcgw::gather_ir(s, PiW, M, n0, nq, d_qir, nqi, cnt, S);
    dyson_matrices(ctx, S, Vir, f, cnt, trs, lds, solve, pi_first, infos, s);
    if (solve) thcgw::scatter_ir(s, S, M, n0, nq, d_qir, d_mq, nqi, cnt, PiW);
    thcgw::scatter_ir_terms(s, trs, lds, infos, n0, nq, d_qir, d_mq, nqi, cnt, d_trace, d_ldet, info_all);
  }
}

// q-point symmetry (thcgw_set_qsym): the Dyson step runs on the irreducible q only, staged
// like the time-reversal path; every other q of an orbit is then filled from its source,
// W(q) = D P W^(src) P^T D^* (a unitary similarity, so Tr(V Pi) and the determinant carry
// over, conjugated where W^ = conj W).
void check_qsym(thcgw_ctx* ctx, const c128* A, const char* what, hipStream_t s) {
  const int nq = ctx->nk, n = ctx->d.nmu, nbx = 64;
  const int* src = ctx->d_sym;
  const int* cj = src + nq;
  const int* perm = cj + nq + ctx->sym_nir + nq;
  double* d = ctx->buf<double>("sy_chk", static_cast<size_t>(2) * nq * nbx);
  thcgw::qsym_diff(s, A, n, nq, src, cj, perm, ctx->d_symph, nbx, d);
  std::vector<double> h(static_cast<size_t>(2) * nq * nbx);
  THCGW_HIP(hipMemcpyAsync(h.data(), d, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
  THCGW_HIP(hipStreamSynchronize(s));
  double diff = 0.0, mag = 0.0;
  for (size_t i = 0; i < h.size(); i += 2) {
    diff = std::max(diff, h[i]);
    mag = std::max(mag, h[i + 1]);
  }
  if (!(diff <= sym_tol(ctx) * mag))
    throw thcgw::Error(THCGW_ERR_ARG, std::string("q-point symmetry: ") + what + " does not follow the map, max deviation " +
                                          sci(diff) + " against max |A| = " + sci(mag));
}

void run_dyson_sym(thcgw_ctx* ctx, c128* PiW, const c128* V, c128* d_trace, double* d_ldet, bool solve,
                   bool pi_first, int* info_all, hipStream_t s) {
  const int nqi = ctx->sym_nir, nq = ctx->nk, nw = ctx->nw_half, n = ctx->d.nmu;
  const int* src = ctx->d_sym;
  const int* cj = src + nq;
  const int* d_qir = cj + nq;
  const int* ident = d_qir + nqi;
  const int* perm = ident + nq;
  const long M = ctx->M;
  if (!ctx->deferred) {
    check_qsym(ctx, V, "V", s);
    check_qsym(ctx, PiW, "Pi(iW_0)", s);
  }
  c128* Vir = ctx->buf<c128>("tr_v", static_cast<size_t>(nqi) * M);
  thcgw::gather_ir(s, V, M, 0, 0, d_qir, nqi, nqi, Vir);
  struct Restore {
    thcgw_ctx* c;
    ~Restore() { c->dy_nq = 0; }
  } restore{ctx};
  ctx->dy_nq = nqi;
  const VFactors f = factor_dyson_v(ctx, Vir, s);
  const int FB = std::max(1, choose_dyson_chunk(ctx, nw * nqi) / nqi);
  const int cap = std::min(FB, nw) * nqi;
  c128* S = ctx->buf<c128>("tr_s", static_cast<size_t>(cap) * M);
  c128* trs = d_trace ? ctx->buf<c128>("tr_tr", static_cast<size_t>(cap)) : nullptr;
  double* lds = d_ldet ? ctx->buf<double>("tr_ld", static_cast<size_t>(cap)) : nullptr;
  int* infos = ctx->buf<int>("tr_info", static_cast<size_t>(cap));
  for (int n0 = 0; n0 < nw; n0 += FB) {
    const int cnt = std::min(FB, nw - n0) * nqi;
    thcgw::gather_ir(s, PiW, M, n0, nq, d_qir, nqi, cnt, S);
    dyson_matrices(ctx, S, Vir, f, cnt, trs, lds, solve, pi_first, infos, s);
    if (solve) thcgw::scatter_ir(s, S, M, n0, nq, d_qir, ident, nqi, cnt, PiW);
    thcgw::scatter_ir_terms(s, trs, lds, infos, n0, nq, d_qir, ident, nqi, cnt, d_trace, d_ldet, info_all);
  }
  if (solve) thcgw::qsym_fill(s, PiW, n, nq, nw, src, cj, perm, ctx->d_symph);
  thcgw::qsym_fill_terms(s, nq, nw, src, cj, d_trace, d_ldet, info_all);
}

// Input guard of the entry points that factorise: refuses arrays holding a NaN or an
// infinity (THCGW_ERR_NONFINITE) before any factorisation runs, so that no pivot search
// or Cholesky step ever sees a non-finite column.  One HBM-bound pass per array.  In
// deferred mode (graph capture) the counts stay on the device and check_deferred()
// reads them after the stream completes.
struct FiniteArg {
  const void* p;
  long ncplx;  // complex elements
  const char* name;
};
constexpr int kNonfiniteBlocks = 512;

void check_finite(thcgw_ctx* ctx, const char* entry, std::initializer_list<FiniteArg> args, hipStream_t s) {
  std::vector<FiniteArg> a;
  for (const FiniteArg& x : args)
    if (x.p && x.ncplx > 0) a.push_back(x);
  if (a.empty()) return;
  int* cnt = ctx->buf<int>("nf_cnt", a.size() * kNonfiniteBlocks);
  for (size_t i = 0; i < a.size(); ++i)
    thcgw::nonfinite_count(s, static_cast<const double*>(a[i].p), 2 * a[i].ncplx, kNonfiniteBlocks,
                           cnt + i * kNonfiniteBlocks);
  if (ctx->deferred) {
    ctx->nf_arrays = static_cast<int>(a.size());
    return;
  }
  std::vector<int> h(a.size() * kNonfiniteBlocks);
  THCGW_HIP(hipMemcpyAsync(h.data(), cnt, sizeof(int) * h.size(), hipMemcpyDeviceToHost, s));
  THCGW_HIP(hipStreamSynchronize(s));
  for (size_t i = 0; i < a.size(); ++i) {
    long bad = 0;
    for (int b = 0; b < kNonfiniteBlocks; ++b) bad += h[i * kNonfiniteBlocks + b];
    if (bad > 0)
      throw thcgw::Error(THCGW_ERR_NONFINITE, std::string(entry) + ": " + a[i].name + " holds " + std::to_string(bad) +
                                                  " NaN or infinite values");
  }
}

// In place Pi(iW) -> W(iW) - V; optional Tr(V Pi) and ln|Re det(I - V Pi)| per matrix
// (device arrays of length nw_half*nk).  pi_first selects A = I - Pi V (thc_rpa.icc:141)
// instead of I - V Pi, and solve = false stops after the determinant.  The Hermitian
// form is used when selected and V and every M are positive definite; chunks where it
// does not apply run the LU form.  With time reversal (flag 9) only the irreducible q are
// solved.
void run_dyson(thcgw_ctx* ctx, c128* PiW, const c128* V, c128* d_trace, double* d_ldet, bool solve, bool pi_first,
               hipStream_t s) {
  const int nq = ctx->nk;
  const int total = ctx->nw_half * nq;
  int* info_all = ctx->buf<int>("dy_info", static_cast<size_t>(total));
  if (ctx->sym) {
    run_dyson_sym(ctx, PiW, V, d_trace, d_ldet, solve, pi_first, info_all, s);
  } else if (ctx->trev) {
    run_dyson_trev(ctx, PiW, V, d_trace, d_ldet, solve, pi_first, info_all, s);
  } else {
    const VFactors f = factor_dyson_v(ctx, V, s);
    dyson_matrices(ctx, PiW, V, f, total, d_trace, d_ldet, solve, pi_first, info_all, s);
  }
  if (ctx->deferred) return;  // check_deferred() reads the flags after the stream completes
  std::vector<int> info(total);
  THCGW_HIP(hipMemcpyAsync(info.data(), info_all, sizeof(int) * total, hipMemcpyDeviceToHost, s));
  THCGW_HIP(hipStreamSynchronize(s));
  for (int b = 0; b < total; ++b)
    if (info[b] != 0)
      throw thcgw::Error(THCGW_ERR_SINGULAR,
                         "Dyson: (I - V Pi) is singular at (n, q) = (" + std::to_string(b / nq) + ", " +
                             std::to_string(b % nq) + "), zero pivot " + std::to_string(info[b] - 1));
}

// Ec from per-(n, q) Tr and ln|Re det| (thc_rpa.icc:130-173, q_weight = 1).
double assemble_ec(thcgw_ctx* ctx, const std::vector<std::complex<double>>& tr, const std::vector<double>& ld) {
  const int nq = ctx->nk, nwb = ctx->d.nw_b;
  std::vector<std::complex<double>> ew(nwb, 0.0);
  for (int n = 0; n < ctx->nw_half; ++n)
    for (int q = 0; q < nq; ++q) {
      const int g = n * nq + q;
      const int wn = nwb / 2 + n, mwn = nwb / 2 - n;
      const std::complex<double> v = tr[g] + ld[g];
      ew[wn] += v;
      if (mwn != wn) ew[mwn] += v;
    }
  return thcgw::host::ec_from_ew(ctx->tab, ew, nq);
}

void fetch_ec_terms(thcgw_ctx* ctx, const c128* d_tr, const double* d_ld, std::vector<std::complex<double>>& tr,
                    std::vector<double>& ld, hipStream_t s) {
  const int total = ctx->nw_half * ctx->nk;
  tr.resize(total);
  ld.resize(total);
  THCGW_HIP(hipMemcpyAsync(tr.data(), d_tr, sizeof(c128) * total, hipMemcpyDeviceToHost, s));
  THCGW_HIP(hipMemcpyAsync(ld.data(), d_ld, sizeof(double) * total, hipMemcpyDeviceToHost, s));
  THCGW_HIP(hipStreamSynchronize(s));
}

void run_sigma_f32(thcgw_ctx* ctx, const c128* G, const c128* Wqt, const c128* X, c128* Sig, bool herm,
                   hipStream_t s) {
  const int nk = ctx->nk, no = ctx->d.no, nmu = ctx->d.nmu, nt = ctx->d.nt, nth = ctx->nt_half;
  const long M = ctx->M;
  const long gs = static_cast<long>(nk) * no * no;
  const int P = choose_pair_chunk(ctx, 2);
  const c64* X64 = x_c64(ctx, X, s);
  c64* GP = ctx->buf<c64>("mx_gp", static_cast<size_t>(P) * nk * M);
  c64* GN = ctx->buf<c64>("mx_gn", static_cast<size_t>(P) * nk * M);
  c64* SP = ctx->buf<c64>("mx_sp", static_cast<size_t>(P) * nk * M);
  c64* SN = ctx->buf<c64>("mx_sn", static_cast<size_t>(P) * nk * M);
  c64* Ab = ctx->buf<c64>("mx_ab", static_cast<size_t>(P) * nk * nmu * no);
  const double scale = -1.0 / (double(nk) * nk);
  // flag 12 bit 3: these projections (Sigma's only, not chi0's) by three real products
  const bool g3 = (thcgw::gemm_complex3() & 8) != 0;
  for (int i0 = 0; i0 < nth; i0 += P) {
    const int c = std::min(P, nth - i0);
    const int c2 = std::max(0, std::min(c, nt / 2 - i0));
    project_p2a_f32(ctx, s, X64, G + i0 * gs, GP, static_cast<long>(c) * nk, Ab, ctx->herm_g, g3);
    if (c2 > 0)
      project_p2a_f32(ctx, s, X64, G + (nt - i0 - c2) * gs, GN, static_cast<long>(c2) * nk, Ab, ctx->herm_g, g3);
    for (int j = 0; j < c; ++j) {
      const int i = i0 + j;
      const bool second = j < c2;
      thcgw::sigma_kfft_f32(s, ctx->d.kgrid, ctx->force_fallback, GP + static_cast<long>(j) * nk * M,
                            second ? GN + static_cast<long>(c2 - 1 - j) * nk * M : nullptr,
                            Wqt + static_cast<long>(i) * M, static_cast<long>(nth) * M,
                            SP + static_cast<long>(j) * nk * M,
                            second ? SN + static_cast<long>(c2 - 1 - j) * nk * M : nullptr, M, scale,
                            ctx->herm_g ? nmu : 0);
    }
    project_a2p_f32(ctx, s, X64, SP, Sig + i0 * gs, static_cast<long>(c) * nk, Ab, ctx->herm_g);
    if (c2 > 0)
      project_a2p_f32(ctx, s, X64, SN, Sig + (nt - i0 - c2) * gs, static_cast<long>(c2) * nk, Ab, ctx->herm_g);
  }
  // with flag 10 Sigma_PQ is Hermitian and so is every back-projection (Y + Y^H)
  if (herm && !ctx->herm_g) thcgw::hermitize(s, Sig, no, static_cast<long>(nt) * nk);
}

void run_sigma(thcgw_ctx* ctx, const c128* G, const c128* Wqt, const c128* X, c128* Sig, bool herm, hipStream_t s) {
  check_herm_g(ctx, G, s);
  if (ctx->precision != 0) return run_sigma_f32(ctx, G, Wqt, X, Sig, herm, s);
  const int nk = ctx->nk, no = ctx->d.no, nmu = ctx->d.nmu, nt = ctx->d.nt, nth = ctx->nt_half;
  const long M = ctx->M;
  const long gs = static_cast<long>(nk) * no * no;
  const int P = choose_pair_chunk(ctx, 4);
  c128* GP = ctx->buf<c128>("sg_gp", static_cast<size_t>(P) * nk * M);
  c128* GN = ctx->buf<c128>("sg_gn", static_cast<size_t>(P) * nk * M);
  c128* SP = ctx->buf<c128>("sg_sp", static_cast<size_t>(P) * nk * M);
  c128* SN = ctx->buf<c128>("sg_sn", static_cast<size_t>(P) * nk * M);
  c128* Ab = ctx->buf<c128>("proj_ab", static_cast<size_t>(P) * nk * nmu * no);
  const double scale = -1.0 / (double(nk) * nk);
  for (int i0 = 0; i0 < nth; i0 += P) {
    const int c = std::min(P, nth - i0);
    // second pass only for i < nt/2 (thc_gw.icc:53-62 writes Sigma(beta - tau_i) at nt-1-i)
    const int c2 = std::max(0, std::min(c, nt / 2 - i0));
    project_p2a(ctx, s, X, G + i0 * gs, GP, static_cast<long>(c) * nk, Ab, ctx->herm_g);
    if (c2 > 0) project_p2a(ctx, s, X, G + (nt - i0 - c2) * gs, GN, static_cast<long>(c2) * nk, Ab, ctx->herm_g);
    for (int j = 0; j < c; ++j) {
      const int i = i0 + j;
      const bool second = j < c2;
      thcgw::sigma_kfft(s, ctx->d.kgrid, ctx->force_fallback, GP + static_cast<long>(j) * nk * M,
                        second ? GN + static_cast<long>(c2 - 1 - j) * nk * M : nullptr, Wqt + static_cast<long>(i) * M,
                        static_cast<long>(nth) * M, SP + static_cast<long>(j) * nk * M,
                        second ? SN + static_cast<long>(c2 - 1 - j) * nk * M : nullptr, M, scale,
                        ctx->herm_g ? nmu : 0);
    }
    if (ctx->herm_g) {
      project_a2p_herm(ctx, s, X, SP, Sig + i0 * gs, static_cast<long>(c) * nk, Ab);
      if (c2 > 0) project_a2p_herm(ctx, s, X, SN, Sig + (nt - i0 - c2) * gs, static_cast<long>(c2) * nk, Ab);
    } else {
      project_a2p(ctx, s, X, SP, Sig + i0 * gs, static_cast<long>(c) * nk, Ab);
      if (c2 > 0) project_a2p(ctx, s, X, SN, Sig + (nt - i0 - c2) * gs, static_cast<long>(c2) * nk, Ab);
    }
  }
  // with flag 10 Sigma_PQ is Hermitian and so is every back-projection (Y + Y^H)
  if (herm && !ctx->herm_g) thcgw::hermitize(s, Sig, no, static_cast<long>(nt) * nk);
}

// ---- k-space (convolution) variants of chi0 and Sigma ----------------------------------

const int* kmq_tables(thcgw_ctx* ctx) {
  if (!ctx->d_kmq) {
    const int nk = ctx->nk;
    std::vector<int> t = thcgw::kminus_table(ctx->d.kgrid), both(static_cast<size_t>(2) * nk * nk);
    for (int a = 0; a < nk; ++a)
      for (int b = 0; b < nk; ++b) {
        both[static_cast<size_t>(a) * nk + b] = t[static_cast<size_t>(a) * nk + b];
        both[static_cast<size_t>(nk) * nk + static_cast<size_t>(a) * nk + b] = t[static_cast<size_t>(b) * nk + a];
      }
    THCGW_HIP(hipMalloc(&ctx->d_kmq, sizeof(int) * both.size()));
    THCGW_HIP(hipMemcpy(ctx->d_kmq, both.data(), sizeof(int) * both.size(), hipMemcpyHostToDevice));
  }
  return ctx->d_kmq;
}

// Pi^k_PQ(tau_i, q) = (sp / N_k) sum_k conj(Gp(k - q)) Gn(k)   (rpa_pi.icc:243, 328-333)
void run_pi_kspace(thcgw_ctx* ctx, const c128* G, const c128* X, c128* Pi, hipStream_t s) {
  const int nk = ctx->nk, no = ctx->d.no, nmu = ctx->d.nmu, nt = ctx->d.nt, nth = ctx->nt_half;
  const long M = ctx->M;
  const long gs = static_cast<long>(nk) * no * no;
  const int* kmq = kmq_tables(ctx);
  c128* GP = ctx->buf<c128>("pi_gp", static_cast<size_t>(nk) * M);
  c128* GN = ctx->buf<c128>("pi_gn", static_cast<size_t>(nk) * M);
  c128* Ab = ctx->buf<c128>("proj_ab", static_cast<size_t>(nk) * nmu * no);
  for (int i = 0; i < nth; ++i) {
    project_p2a(ctx, s, X, G + i * gs, GP, nk, Ab);
    project_p2a(ctx, s, X, G + (nt - 1 - i) * gs, GN, nk, Ab);
    thcgw::kconv_pi(s, GP, GN, kmq + static_cast<long>(nk) * nk, nk, Pi + static_cast<long>(i) * nk * M, M,
                    -2.0 / nk);
  }
}

// Sigma_PQ(tau, k) = -(1/N_k) sum_q G(k - q) W(q), both particle-hole passes   (thc_gw.icc:342-443)
void run_sigma_kspace(thcgw_ctx* ctx, const c128* G, const c128* Wqt, const c128* X, c128* Sig, bool herm,
                      hipStream_t s) {
  const int nk = ctx->nk, no = ctx->d.no, nmu = ctx->d.nmu, nt = ctx->d.nt, nth = ctx->nt_half;
  const long M = ctx->M;
  const long gs = static_cast<long>(nk) * no * no;
  const int* kmq = kmq_tables(ctx);
  c128* GP = ctx->buf<c128>("sg_gp", static_cast<size_t>(nk) * M);
  c128* SP = ctx->buf<c128>("sg_sp", static_cast<size_t>(nk) * M);
  c128* Ab = ctx->buf<c128>("proj_ab", static_cast<size_t>(nk) * nmu * no);
  for (int pass = 0; pass < 2; ++pass) {
    const int ntp = pass == 0 ? nth : nt / 2;
    for (int it = 0; it < ntp; ++it) {
      const int itg = pass == 0 ? it : nt - it - 1;
      project_p2a(ctx, s, X, G + itg * gs, GP, nk, Ab);
      thcgw::kconv_sigma(s, GP, Wqt + static_cast<long>(it) * M, static_cast<long>(nth) * M, kmq, nk, SP, M,
                         -1.0 / nk);
      project_a2p(ctx, s, X, SP, Sig + itg * gs, nk, Ab);
    }
  }
  if (herm) thcgw::hermitize(s, Sig, no, static_cast<long>(nt) * nk);
}

// ---- Brillouin-zone symmetry: chi0 and Sigma on the irreducible wedge -------------------

enum BzTab {
  BZ_G_SRC, BZ_G_TRANS,                       // Sigma: G of the listed points without time reversal
  BZ_PG_SRC, BZ_PG_TRANS, BZ_PG_CONJ, BZ_PG_X, // chi0: G and X of every grid point (partners conjugated)
  BZ_Q_GRID, BZ_X_SRC, BZ_T_OFF, BZ_T_TERMS
};

const int* bz_tab(const thcgw_ctx* ctx, BzTab t) { return ctx->d_bz + ctx->bz_o[t]; }

void require_bz(const thcgw_ctx* ctx, const std::string& entry) {
  THCGW_REQUIRE(ctx->bz_on, entry + ": no Brillouin-zone symmetry set (thcgw_set_bz)");
  THCGW_REQUIRE(ctx->precision == 0, entry + ": the symmetry-reduced path runs in FP64 (flag 7 = 0)");
  THCGW_REQUIRE(!ctx->trev && !ctx->sym, entry + ": flag 9 and the q map do not combine with thcgw_set_bz");
}

// Pi(tau_i, q) on the irreducible q by the R-space algorithm (rpa_pi.icc:22-180 with
// nqpts_ibz < nqpts).  The reference projects G on the points without time reversal and
// takes each time-reversal point's G_PQ as the conjugate of its partner's
// (rpa_pi.icc:106-113).  Here every grid point is projected directly in C-order grid
// position, a time-reversal point from the conjugated X and G of its partner:
// conj(X) conj(G) conj(X)^H is that conjugate exactly.  Then, per tau pair, the fused
// k-grid kernel of the full grid and the irreducible q gathered from its result
// (rpa_pi.icc:160-169).  Projections are batched over tau pairs.
void run_pi_ibz(thcgw_ctx* ctx, const c128* G, const c128* X, c128* Pi, hipStream_t s) {
  check_herm_g(ctx, G, s, ctx->bz_nk_ibz);
  const int nk = ctx->nk, no = ctx->d.no, nmu = ctx->d.nmu, nt = ctx->d.nt, nth = ctx->nt_half;
  const int nqi = ctx->bz_nq_ibz;
  const long M = ctx->M, gs = static_cast<long>(ctx->bz_nk_ibz) * no * no, sX = static_cast<long>(nmu) * no;
  const long per = static_cast<long>(nk) * no * no;
  const int P = choose_pair_chunk(ctx, 2);
  c128* Xg = ctx->buf<c128>("bz_xgrid", static_cast<size_t>(nk) * sX);
  c128* Gg = ctx->buf<c128>("bz_ggrid", static_cast<size_t>(P) * per);
  c128* Ab = ctx->buf<c128>("proj_ab", static_cast<size_t>(P) * nk * nmu * no);
  c128* GP = ctx->buf<c128>("pi_gp", static_cast<size_t>(P) * nk * M);
  c128* GN = ctx->buf<c128>("pi_gn", static_cast<size_t>(P) * nk * M);
  c128* PF = ctx->buf<c128>("bz_pf", static_cast<size_t>(nk) * M);
  thcgw::gather_rows(s, X, sX, bz_tab(ctx, BZ_PG_X), bz_tab(ctx, BZ_PG_CONJ), nk, sX, Xg, sX, nullptr);
  const double scale = -2.0 / (double(nk) * nk);  // sp = -2 for ns = 1 (rpa_pi.icc:46)
  auto gather = [&](int it, c128* dst) {
    thcgw::gather_blocks(s, G + it * gs, bz_tab(ctx, BZ_PG_SRC), bz_tab(ctx, BZ_PG_TRANS), bz_tab(ctx, BZ_PG_CONJ),
                         no, nk, dst);
  };
  for (int i0 = 0; i0 < nth; i0 += P) {
    const int c = std::min(P, nth - i0);
    for (int j = 0; j < c; ++j) gather(i0 + j, Gg + j * per);
    project_p2a(ctx, s, Xg, Gg, GP, static_cast<long>(c) * nk, Ab, ctx->herm_g);
    for (int j = 0; j < c; ++j) gather(nt - 1 - (i0 + j), Gg + j * per);
    project_p2a(ctx, s, Xg, Gg, GN, static_cast<long>(c) * nk, Ab, ctx->herm_g);
    for (int j = 0; j < c; ++j) {
      thcgw::pi_kfft(s, ctx->d.kgrid, ctx->force_fallback, GP + j * nk * M, GN + j * nk * M, PF, M, scale);
      thcgw::gather_rows(s, PF, M, bz_tab(ctx, BZ_Q_GRID), nullptr, nqi, M, Pi + static_cast<long>(i0 + j) * nqi * M,
                         M, nullptr);
    }
  }
}

// X_{k S} D(S, k) for every operation S and irreducible k: the back-projection operands of
// Sigma (no rotation for the identity, thc_gw.icc:316-326).  (nsym nk_ibz) x N_mu x N_orb.
const c128* bz_xd(thcgw_ctx* ctx, const c128* X, hipStream_t s) {
  const int no = ctx->d.no, nmu = ctx->d.nmu, nki = ctx->bz_nk_ibz, nout = ctx->bz_nsym * nki;
  const long sX = static_cast<long>(nmu) * no, sD = static_cast<long>(no) * no;
  c128* XD = ctx->buf<c128>("bz_xd", static_cast<size_t>(nout) * sX);
  thcgw::gather_rows(s, X, sX, bz_tab(ctx, BZ_X_SRC), nullptr, nki, sX, XD, sX, nullptr);
  if (nout > nki) {
    c128* Xg = ctx->buf<c128>("bz_xg", static_cast<size_t>(nout - nki) * sX);
    thcgw::gather_rows(s, X, sX, bz_tab(ctx, BZ_X_SRC) + nki, nullptr, nout - nki, sX, Xg, sX, nullptr);
    thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, nmu, no, no, kOne, thcgw::bp(Xg, sX), no,
                 thcgw::bp(ctx->d_bzD + nki * sD, sD), no, kZero, XD + nki * sX, sX, no, nout - nki);
  }
  return XD;
}

// Sigma(tau, k) on the irreducible k by the k-space algorithm (thc_gw.icc:152-178, 216-443):
// per tau of both particle-hole passes, G_PQ on the points without time reversal, the
// Hadamard sums of every operation S over the q of Qs[S] at k S (time-reversal conjugates
// in the term flags, thc_gw.icc:366-395), Sigma_ab(k S) = X_{kS}^H Sigma_PQ X_{kS} rotated
// by D (one product with X_{kS} D), and the sum over S.  Projections and back-projections
// are batched over tau.
void run_sigma_ibz(thcgw_ctx* ctx, const c128* G, const c128* Wqt, const c128* X, c128* Sig, bool herm,
                   hipStream_t s) {
  check_herm_g(ctx, G, s, ctx->bz_nk_ibz);
  const int nk = ctx->nk, no = ctx->d.no, nmu = ctx->d.nmu, nt = ctx->d.nt, nth = ctx->nt_half;
  const int nki = ctx->bz_nk_ibz, nqi = ctx->bz_nq_ibz, nkn = ctx->bz_nk_nt, nsym = ctx->bz_nsym;
  const int nout = nsym * nki;
  const long M = ctx->M, gs = static_cast<long>(nki) * no * no, sX = static_cast<long>(nmu) * no;
  const long per = static_cast<long>(nkn) * no * no, so = static_cast<long>(no) * no;
  const int P = choose_pair_chunk(ctx, std::max(1, (nkn + nout + nk - 1) / nk));
  const c128* XD = bz_xd(ctx, X, s);
  c128* Gg = ctx->buf<c128>("bz_gg", static_cast<size_t>(P) * per);
  c128* B = ctx->buf<c128>("bz_buf", static_cast<size_t>(P) * nkn * M);
  c128* SP = ctx->buf<c128>("bz_sp", static_cast<size_t>(P) * nout * M);
  c128* SA = ctx->buf<c128>("bz_sa", static_cast<size_t>(P) * nout * so);
  c128* Ab = ctx->buf<c128>("proj_ab", static_cast<size_t>(P) * std::max(nkn, nout) * nmu * no);
  for (int pass = 0; pass < 2; ++pass) {
    const int ntp = pass == 0 ? nth : nt / 2;  // the middle tau of an odd grid is the first pass's
    for (int i0 = 0; i0 < ntp; i0 += P) {
      const int c = std::min(P, ntp - i0);
      auto itg = [&](int j) { return pass == 0 ? i0 + j : nt - 1 - (i0 + j); };
      for (int j = 0; j < c; ++j)
        thcgw::gather_blocks(s, G + itg(j) * gs, bz_tab(ctx, BZ_G_SRC), bz_tab(ctx, BZ_G_TRANS), nullptr, no, nkn,
                             Gg + j * per);
      project_p2a(ctx, s, X, Gg, B, static_cast<long>(c) * nkn, Ab, ctx->herm_g, nkn);
      // with flag 10 every Sigma_PQ of an operation is Hermitian (Hadamard sums of Hermitian G
      // and W): its lower half L goes back as X^H L X + (X^H L X)^H, a triangular first product
      const int hn = ctx->herm_g ? nmu : 0;
      for (int j = 0; j < c; ++j)
        thcgw::kconv_terms(s, B + j * nkn * M, nkn, Wqt + static_cast<long>(i0 + j) * M, static_cast<long>(nth) * M,
                           nqi, bz_tab(ctx, BZ_T_OFF), bz_tab(ctx, BZ_T_TERMS), nout, SP + j * nout * M, M, -1.0 / nk,
                           hn);
      const int nb = c * nout;
      thcgw::zgemm(s, thcgw::OP_C, thcgw::OP_N, no, nmu, nmu, kOne, thcgw::bp(XD, sX, nout), no, thcgw::bp(SP, M), nmu,
                   kZero, Ab, sX, nmu, nb, kZero, hn ? thcgw::GEMM_TRI_KFROM_COL : 0);
      thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, no, no, nmu, kOne, thcgw::bp(Ab, sX), nmu, thcgw::bp(XD, sX, nout), no,
                   kZero, SA, so, no, nb);
      if (hn) thcgw::hermitize(s, SA, no, nb, 2.0);
      for (int j = 0; j < c; ++j) thcgw::sum_groups(s, SA + j * nout * so, nsym, gs, Sig + itg(j) * gs);
    }
  }
  if (herm) thcgw::hermitize(s, Sig, no, static_cast<long>(nt) * nki);
}

// Pi -> W on the wedge and Ec with each irreducible q weighted by its multiplicity
// (thc_rpa.icc:130-160); Sigma on the wedge.
void run_iteration_ibz(thcgw_ctx* ctx, const c128* G, const c128* X, const c128* V, c128* Sig, c128* Wqt, double* Ec,
                       bool herm, hipStream_t s) {
  const int nqi = ctx->bz_nq_ibz, nwh = ctx->nw_half;
  const long M = ctx->M;
  const int total = nwh * nqi;
  c128* Piw = ctx->buf<c128>("it_piw", static_cast<size_t>(total) * M);
  c128* dtr = Ec ? ctx->buf<c128>("ec_tr", total) : nullptr;
  double* dld = Ec ? ctx->buf<double>("ec_ld", total) : nullptr;
  run_pi_ibz(ctx, G, X, Wqt, s);  // W_qtPQ doubles as Pi_tqPQ scratch (same size)
  run_tau_to_w(ctx, Wqt, Piw, s, nqi);
  {
    struct Restore {
      thcgw_ctx* c;
      ~Restore() { c->dy_nq = 0; }
    } restore{ctx};
    ctx->dy_nq = nqi;
    int* info = ctx->buf<int>("dy_info", static_cast<size_t>(total));
    const VFactors f = factor_dyson_v(ctx, V, s);
    dyson_matrices(ctx, Piw, V, f, total, dtr, dld, true, false, info, s);
    std::vector<int> h(total);
    THCGW_HIP(hipMemcpyAsync(h.data(), info, sizeof(int) * total, hipMemcpyDeviceToHost, s));
    THCGW_HIP(hipStreamSynchronize(s));
    for (int b = 0; b < total; ++b)
      if (h[b] != 0)
        throw thcgw::Error(THCGW_ERR_SINGULAR, "Dyson: (I - V Pi) is singular at (n, q_ibz) = (" +
                                                   std::to_string(b / nqi) + ", " + std::to_string(b % nqi) + ")");
  }
  run_w_to_tau(ctx, Piw, Wqt, s, nqi);
  run_sigma_ibz(ctx, G, Wqt, X, Sig, herm, s);
  if (Ec) {
    std::vector<std::complex<double>> tr(total);
    std::vector<double> ld(total);
    THCGW_HIP(hipMemcpyAsync(tr.data(), dtr, sizeof(c128) * total, hipMemcpyDeviceToHost, s));
    THCGW_HIP(hipMemcpyAsync(ld.data(), dld, sizeof(double) * total, hipMemcpyDeviceToHost, s));
    THCGW_HIP(hipStreamSynchronize(s));
    const int nwb = ctx->d.nw_b;
    std::vector<std::complex<double>> ew(nwb, 0.0);
    for (int n = 0; n < nwh; ++n)
      for (int q = 0; q < nqi; ++q) {
        const std::complex<double> v = ctx->bz_qw[q] * (tr[n * nqi + q] + ld[n * nqi + q]);
        const int wn = nwb / 2 + n, mwn = nwb / 2 - n;
        ew[wn] += v;
        if (mwn != wn) ew[mwn] += v;
      }
    *Ec = thcgw::host::ec_from_ew(ctx->tab, ew, ctx->nk);
  }
}

// ---- one iteration as enqueue + finish (the batch captures the enqueue part) -----------

void enqueue_iteration(thcgw_ctx* ctx, const c128* G, const c128* X, const c128* V, c128* Sig, c128* Wqt, bool want_ec,
                       bool herm, hipStream_t s) {
  const long M = ctx->M;
  const int total = ctx->nw_half * ctx->nk;
  c128* Pit = Wqt;  // W_qtPQ doubles as Pi_tqPQ scratch (same size)
  c128* Piw = ctx->buf<c128>("it_piw", static_cast<size_t>(total) * M);
  c128* dtr = want_ec ? ctx->buf<c128>("ec_tr", total) : nullptr;
  double* dld = want_ec ? ctx->buf<double>("ec_ld", total) : nullptr;
  const long nk = ctx->nk, no = ctx->d.no;
  check_finite(ctx, "gw_iteration",
               {{G, ctx->d.nt * nk * no * no, "G_tskij"}, {X, nk * ctx->d.nmu * no, "X_skPa"}, {V, ctx->vq() * M, "V_qPQ"}},
               s);
  run_pi(ctx, G, X, Pit, s);
  run_tau_to_w(ctx, Pit, Piw, s);
  run_dyson(ctx, Piw, V, dtr, dld, true, false, s);
  run_w_to_tau(ctx, Piw, Wqt, s);
  run_sigma(ctx, G, Wqt, X, Sig, herm, s);
}

// After a deferred enqueue has completed: Cholesky / LU flags of the Dyson step.
void check_deferred(thcgw_ctx* ctx, hipStream_t s) {
  const int total = ctx->nw_half * ctx->nk, nq = ctx->nk;
  if (ctx->nf_arrays > 0) {
    std::vector<int> h(static_cast<size_t>(ctx->nf_arrays) * kNonfiniteBlocks);
    THCGW_HIP(hipMemcpyAsync(h.data(), ctx->buf<int>("nf_cnt", h.size()), sizeof(int) * h.size(),
                             hipMemcpyDeviceToHost, s));
    THCGW_HIP(hipStreamSynchronize(s));
    long bad = 0;
    for (int x : h) bad += x;
    if (bad > 0)
      throw thcgw::Error(THCGW_ERR_NONFINITE,
                         "batch: the inputs G, X or V hold " + std::to_string(bad) + " NaN or infinite values");
  }
  if (ctx->precision != 0 || ctx->dyson_algo != 0) {
    std::vector<int> vi(nq);
    THCGW_HIP(hipMemcpyAsync(vi.data(), ctx->buf<int>("dy_vinfo", static_cast<size_t>(nq)), sizeof(int) * nq,
                             hipMemcpyDeviceToHost, s));
    THCGW_HIP(hipStreamSynchronize(s));
    for (int q = 0; q < nq; ++q)
      if (vi[q] != 0)
        throw thcgw::Error(THCGW_ERR_SINGULAR, "batch: V_q is not positive definite at q = " + std::to_string(q) +
                                                   " (the captured Hermitian form has no LU fallback)");
  }
  std::vector<int> info(total);
  THCGW_HIP(hipMemcpyAsync(info.data(), ctx->buf<int>("dy_info", static_cast<size_t>(total)), sizeof(int) * total,
                           hipMemcpyDeviceToHost, s));
  THCGW_HIP(hipStreamSynchronize(s));
  for (int b = 0; b < total; ++b)
    if (info[b] != 0)
      throw thcgw::Error(THCGW_ERR_SINGULAR, "batch: Dyson factorisation failed at (n, q) = (" + std::to_string(b / nq) +
                                                 ", " + std::to_string(b % nq) + ")");
}

double finish_ec(thcgw_ctx* ctx, hipStream_t s) {
  const int total = ctx->nw_half * ctx->nk;
  std::vector<std::complex<double>> tr;
  std::vector<double> ld;
  fetch_ec_terms(ctx, ctx->buf<c128>("ec_tr", total), ctx->buf<double>("ec_ld", total), tr, ld, s);
  return assemble_ec(ctx, tr, ld);
}

// ---- Dyson equation for G on the Matsubara axis ------------------------------------

// The Dyson equation for G refuses non-finite H0, F, S or Sigma (see check_finite).
void check_dyson_g_inputs(thcgw_ctx* ctx, const char* entry, const double* H0, const double* F, const double* Sov,
                          const double* Sigma, hipStream_t s) {
  const long per = static_cast<long>(ctx->nk) * ctx->d.no * ctx->d.no;
  check_finite(ctx, entry, {{H0, per, "H0_skij"}, {F, per, "F_skij"}, {Sov, per, "S_skij"},
                            {Sigma, ctx->d.nt * per, "Sigma_tskij"}}, s);
}

// Sw(iw_n) = tau_to_w(Sigma(tau), fermi) over all (k, i, j)   (simple_dyson.cpp:39-40)
c128* sigma_to_w(thcgw_ctx* ctx, const c128* Sigma, hipStream_t s) {
  const int nt = ctx->d.nt, nw = ctx->d.nw_f;
  const long dim = static_cast<long>(ctx->nk) * ctx->d.no * ctx->d.no;
  THCGW_REQUIRE(dim < (1L << 31), "Dyson for G: N_k N_o^2 too large for one GEMM");
  c128* Sw = ctx->buf<c128>("dg_sw", static_cast<size_t>(nw) * dim);
  thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, nw, static_cast<int>(dim), nt, kOne, thcgw::bp(ctx->d_Twt_ff, 0), nt,
               thcgw::bp(Sigma, 0), dim, kZero, Sw, 0, dim, 1);
  return Sw;
}

// Out[n,k] = [(iw_n + mu) S - H0 - F - Sw]^-1 Rhs, Rhs = S_k (with_s) or I, for all (n, k) by
// batched LU (A = Lt Ut Pt), two triangular solves and the row gather that undoes the swaps.
void matsubara_solve(thcgw_ctx* ctx, const c128* H0, const c128* F, const c128* S, const c128* Sw, double mu,
                     bool rhs_s, c128* Out, hipStream_t s) {
  const int nw = ctx->d.nw_f, nk = ctx->nk, no = ctx->d.no;
  const long m = static_cast<long>(no) * no;
  const long count = static_cast<long>(nw) * nk;
  THCGW_REQUIRE(count < (1L << 30), "Dyson for G: too many matrices");
  c128* X = ctx->buf<c128>("dg_x", static_cast<size_t>(count) * m);
  c128* Y = ctx->buf<c128>("dg_y", static_cast<size_t>(count) * m);
  int* ipiv = ctx->buf<int>("dg_ipiv", static_cast<size_t>(count) * no);
  int* info = ctx->buf<int>("dg_info", static_cast<size_t>(count));
  thcgw::matsubara_dyson(s, H0, F, S, Sw, ctx->d_wn_f, ctx->d.beta, mu, nw, nk, no, X);
  thcgw::getrf_batched(s, X, m, no, no, ipiv, info, static_cast<int>(count));
  thcgw::rhs_batch(s, rhs_s ? S : nullptr, nk, no, count, Y);
  thcgw::trsm_left_lower(s, X, m, no, Y, m, no, no, no, static_cast<int>(count));
  thcgw::trsm_left_upper_unit(s, X, m, no, Y, m, no, no, no, static_cast<int>(count));
  thcgw::dyson_finish(s, Y, m, nullptr, 0, 1, ipiv, no, Out, m, 0, static_cast<int>(count));
  std::vector<int> h(count);
  THCGW_HIP(hipMemcpyAsync(h.data(), info, sizeof(int) * count, hipMemcpyDeviceToHost, s));
  THCGW_HIP(hipStreamSynchronize(s));
  for (long b = 0; b < count; ++b)
    if (h[b] != 0)
      throw thcgw::Error(THCGW_ERR_SINGULAR, "Dyson for G: (iw_n + mu) S - H0 - F - Sigma is singular at (n, k) = (" +
                                                 std::to_string(b / nk) + ", " + std::to_string(b % nk) + ")");
}

// Electron number at mu (compute_Nelec, scf_common.cpp:10-42) with the spectral sum of
// compute_eigenspectra (simple_dyson.cpp:130-172) taken as the resolvent trace
// sum_b 1/(iw_n + mu - lambda_b) = Tr[((iw_n + mu) S - H0 - F - Sigma)^-1 S].
double nelec_at(thcgw_ctx* ctx, const c128* H0, const c128* F, const c128* S, const c128* Sw, double mu,
                hipStream_t s) {
  const int nw = ctx->d.nw_f, nk = ctx->nk, no = ctx->d.no, nt = ctx->d.nt;
  const long m = static_cast<long>(no) * no;
  const long count = static_cast<long>(nw) * nk;
  c128* R = ctx->buf<c128>("dg_r", static_cast<size_t>(count) * m);
  c128* tr = ctx->buf<c128>("dg_tr", static_cast<size_t>(count));
  matsubara_solve(ctx, H0, F, S, Sw, mu, true, R, s);
  thcgw::batched_trace(s, R, m, no, tr, static_cast<int>(count));
  std::vector<std::complex<double>> h(count);
  THCGW_HIP(hipMemcpyAsync(h.data(), tr, sizeof(c128) * count, hipMemcpyDeviceToHost, s));
  THCGW_HIP(hipStreamSynchronize(s));
  std::vector<std::complex<double>> Xw(nw, 0.0);
  for (int n = 0; n < nw; ++n)
    for (int k = 0; k < nk; ++k) Xw[n] += h[static_cast<size_t>(n) * nk + k] / double(nk);  // k weights 1/N_k
  std::complex<double> nel = 0.0;
  for (int t = 0; t < nt; ++t) {
    std::complex<double> xt = 0.0;
    for (int n = 0; n < nw; ++n) xt += ctx->tab.Ttw_ff[static_cast<size_t>(t) * nw + n] * Xw[n];
    nel += ctx->tab.T_beta[t] * xt;
  }
  return -2.0 * nel.real();  // spin factor 2 for ns = npol = 1
}

}  // namespace

extern "C" {

int thcgw_create(const thcgw_dims* dims, const thcgw_ir_tables* tables, thcgw_ctx** out) {
  g_create_error.clear();
  if (!dims || !tables || !out) {
    g_create_error = "thcgw_create: null argument";
    return THCGW_ERR_ARG;
  }
  *out = nullptr;
  thcgw_ctx* ctx = new (std::nothrow) thcgw_ctx();
  if (!ctx) {
    g_create_error = "thcgw_create: out of host memory";
    return THCGW_ERR_ALLOC;
  }
  int rc = THCGW_OK;
  try {
    const thcgw_dims& d = *dims;
    THCGW_REQUIRE(d.kgrid[0] >= 1 && d.kgrid[1] >= 1 && d.kgrid[2] >= 1, "kgrid entries must be >= 1");
    THCGW_REQUIRE(d.ns == 1, "only ns = 1 is supported (no spin polarisation)");
    THCGW_REQUIRE(d.no >= 1 && d.nmu >= 1 && d.nt >= 1 && d.nw_b >= 1 && d.nw_f >= 2, "sizes must be positive");
    THCGW_REQUIRE(d.beta > 0.0, "beta must be positive");
    THCGW_REQUIRE(d.nmu <= 2048, "nmu above 2048 is not supported by the batched LU");
    ctx->d = d;
    ctx->nk = d.kgrid[0] * d.kgrid[1] * d.kgrid[2];
    THCGW_REQUIRE(thcgw::kfft_supported(d.kgrid), "k-grid not supported (cubic n<=4, or nk<=16)");
    ctx->nt_half = (d.nt % 2 == 0) ? d.nt / 2 : d.nt / 2 + 1;
    ctx->nw_half = (d.nw_b % 2 == 0) ? d.nw_b / 2 : d.nw_b / 2 + 1;
    ctx->M = static_cast<long>(d.nmu) * d.nmu;
    thcgw::host::load_tables(d, *tables, ctx->tab);  // checks PH symmetry of the tau mesh
    THCGW_HIP(hipGetDevice(&ctx->device));
    noalloc_scratch(ctx->device);  // once per device: address of the no-allocation staging buffer
    std::vector<std::complex<double>> Tp, Tm;
    thcgw::host::fold_ph(ctx->tab, d, Tp, Tm);
    // Folding G(tau) and G(beta - tau) pairs e^{i W tau} with e^{-i W tau}: the tables are real
    // up to the rounding of the IR tables (|Im| ~ 1e-14 |Re|).  Below 1e-12 the transforms
    // apply the real part only, with half the MFMA work.
    auto is_real = [](const std::vector<std::complex<double>>& T) {
      double re = 0.0, im = 0.0;
      for (const auto& z : T) {
        re = std::max(re, std::fabs(z.real()));
        im = std::max(im, std::fabs(z.imag()));
      }
      return im <= 1e-12 * re;
    };
    ctx->real_Tplus = is_real(Tp);
    ctx->real_Tminus = is_real(Tm);
    THCGW_HIP(hipMalloc(&ctx->d_tau, sizeof(double) * d.nt));
    THCGW_HIP(hipMalloc(&ctx->d_Tplus, sizeof(c128) * Tp.size()));
    THCGW_HIP(hipMalloc(&ctx->d_Tminus, sizeof(c128) * Tm.size()));
    THCGW_HIP(hipMemcpy(ctx->d_tau, ctx->tab.tau.data(), sizeof(double) * d.nt, hipMemcpyHostToDevice));
    THCGW_HIP(hipMemcpy(ctx->d_Tplus, Tp.data(), sizeof(c128) * Tp.size(), hipMemcpyHostToDevice));
    THCGW_HIP(hipMemcpy(ctx->d_Tminus, Tm.data(), sizeof(c128) * Tm.size(), hipMemcpyHostToDevice));
    THCGW_HIP(hipMalloc(&ctx->d_Twt_ff, sizeof(c128) * ctx->tab.Twt_ff.size()));
    THCGW_HIP(hipMemcpy(ctx->d_Twt_ff, ctx->tab.Twt_ff.data(), sizeof(c128) * ctx->tab.Twt_ff.size(),
                        hipMemcpyHostToDevice));
    THCGW_HIP(hipMalloc(&ctx->d_Ttw_ff, sizeof(c128) * ctx->tab.Ttw_ff.size()));
    THCGW_HIP(hipMemcpy(ctx->d_Ttw_ff, ctx->tab.Ttw_ff.data(), sizeof(c128) * ctx->tab.Ttw_ff.size(),
                        hipMemcpyHostToDevice));
    THCGW_HIP(hipMalloc(&ctx->d_Tbeta, sizeof(c128) * ctx->tab.T_beta.size()));
    THCGW_HIP(hipMemcpy(ctx->d_Tbeta, ctx->tab.T_beta.data(), sizeof(c128) * ctx->tab.T_beta.size(),
                        hipMemcpyHostToDevice));
    THCGW_HIP(hipMalloc(&ctx->d_wn_f, sizeof(long) * ctx->tab.wn_f.size()));
    THCGW_HIP(hipMemcpy(ctx->d_wn_f, ctx->tab.wn_f.data(), sizeof(long) * ctx->tab.wn_f.size(),
                        hipMemcpyHostToDevice));
  } catch (const thcgw::Error& e) {
    ctx->err = e.what();
    rc = e.code;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    rc = THCGW_ERR_ARG;
  }
  if (rc != THCGW_OK) {
    g_create_error = ctx->err.empty() ? "thcgw_create failed" : ctx->err;
    ctx->release();
    delete ctx;
    return rc;
  }
  *out = ctx;
  return THCGW_OK;
}

void thcgw_destroy(thcgw_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  ctx->release();
  delete ctx;
}

const char* thcgw_last_error(const thcgw_ctx* ctx) {
  if (!ctx) return g_create_error.c_str();
  return ctx->err.c_str();
}

long thcgw_workspace_bytes(const thcgw_ctx* ctx) {
  if (!ctx) return 0;
  long b = 0;
  for (auto& kv : ctx->ws) b += static_cast<long>(kv.second.bytes);
  return b;
}

int thcgw_set_flag(thcgw_ctx* ctx, int flag, long value) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    if (flag == 1) ctx->force_fallback = value != 0;
    else if (flag == 2) ctx->pair_chunk = static_cast<int>(value);
    else if (flag == 3) ctx->dyson_chunk = static_cast<int>(value);
    else if (flag == 4) thcgw::gemm_profile_enable(value != 0);
    else if (flag == 5) {
      THCGW_REQUIRE(value >= 0 && value <= 3, "flag 5 (GEMM tile) takes 0 (automatic) to 3");
      thcgw::gemm_set_tile_mode(static_cast<int>(value));
    }
    else if (flag == 8) {
      THCGW_REQUIRE(value >= 0 && value <= 2, "flag 8 (FP32 tile) takes 0, 1 or 2");
      thcgw::gemm_set_tile_mode32(static_cast<int>(value));
    } else if (flag == 12) {
      THCGW_REQUIRE(value >= 0 && value <= 15, "flag 12 (complex products by three real products) takes a mask 0..15");
      thcgw::gemm_set_complex3(static_cast<int>(value));
    } else if (flag == 6) {
      THCGW_REQUIRE(value >= 0 && value <= 2, "flag 6 (Dyson algorithm) takes 0 (LU), 1 (Hermitian) or 2 (resolvent)");
      ctx->dyson_algo = static_cast<int>(value);
    } else if (flag == 7) {
      THCGW_REQUIRE(value >= 0 && value <= 2, "flag 7 (precision) takes 0 (FP64), 1 (mixed) or 2 (mixed, split BF16)");
      ctx->precision = static_cast<int>(value);
    } else if (flag == 11) {
      THCGW_REQUIRE(value >= 0 && value <= 2, "flag 11 (resolvent Dyson lanes) takes 0, 1 or 2");
      ctx->dyson_lanes = value == 2 ? 2 : 1;
    } else if (flag == 10) {
      THCGW_REQUIRE(value == 0 || value == 1, "flag 10 (Hermitian G) takes 0 or 1");
      ctx->herm_g = static_cast<int>(value);
    } else if (flag == 9) {
      THCGW_REQUIRE(value == 0 || value == 1, "flag 9 (time reversal) takes 0 or 1");
      ctx->trev = static_cast<int>(value);
      if (ctx->trev) trev_tables(ctx);
    }
    else THCGW_REQUIRE(false, "unknown flag");
  });
}

int thcgw_set_qsym(thcgw_ctx* ctx, const int* src, const int* perm, const double* phase, const int* conj_) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    const int nq = ctx->nk, n = ctx->d.nmu;
    if (ctx->d_sym) THCGW_HIP(hipFree(ctx->d_sym));
    if (ctx->d_symph) THCGW_HIP(hipFree(ctx->d_symph));
    ctx->d_sym = nullptr;
    ctx->d_symph = nullptr;
    ctx->sym = 0;
    ctx->sym_nir = 0;
    if (!src) return;  // clears the map
    THCGW_REQUIRE(perm, "thcgw_set_qsym: perm is required with src");
    std::vector<int> ir, cj(nq, 0), ident(nq);
    for (int q = 0; q < nq; ++q) {
      THCGW_REQUIRE(src[q] >= 0 && src[q] < nq && src[src[q]] == src[q],
                    "thcgw_set_qsym: src[q] must be an irreducible q (src[src[q]] == src[q])");
      if (conj_) cj[q] = conj_[q] != 0;
      ident[q] = q;
      std::vector<char> seen(n, 0);
      for (int P = 0; P < n; ++P) {
        const int t = perm[static_cast<size_t>(q) * n + P];
        THCGW_REQUIRE(t >= 0 && t < n && !seen[t], "thcgw_set_qsym: perm[q] must be a permutation of 0..nmu-1");
        seen[t] = 1;
        if (src[q] == q) THCGW_REQUIRE(t == P, "thcgw_set_qsym: an irreducible q must map to itself");
      }
      if (src[q] == q) {
        THCGW_REQUIRE(cj[q] == 0, "thcgw_set_qsym: an irreducible q cannot be conjugated");
        if (phase)
          for (int P = 0; P < n; ++P)
            THCGW_REQUIRE(phase[2 * (static_cast<size_t>(q) * n + P)] == 1.0 && phase[2 * (static_cast<size_t>(q) * n + P) + 1] == 0.0,
                          "thcgw_set_qsym: an irreducible q has unit phases");
        ir.push_back(q);
      } else if (phase) {
        for (int P = 0; P < n; ++P) {
          const double re = phase[2 * (static_cast<size_t>(q) * n + P)], im = phase[2 * (static_cast<size_t>(q) * n + P) + 1];
          THCGW_REQUIRE(std::fabs(re * re + im * im - 1.0) < 1e-12, "thcgw_set_qsym: phases must have modulus 1");
        }
      }
    }
    std::vector<int> h(src, src + nq);
    h.insert(h.end(), cj.begin(), cj.end());
    h.insert(h.end(), ir.begin(), ir.end());
    h.insert(h.end(), ident.begin(), ident.end());
    h.insert(h.end(), perm, perm + static_cast<size_t>(nq) * n);
    std::vector<std::complex<double>> ph(static_cast<size_t>(nq) * n, 1.0);
    if (phase) std::memcpy(ph.data(), phase, sizeof(c128) * ph.size());
    THCGW_HIP(hipMalloc(&ctx->d_sym, sizeof(int) * h.size()));
    THCGW_HIP(hipMemcpy(ctx->d_sym, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    THCGW_HIP(hipMalloc(&ctx->d_symph, sizeof(c128) * ph.size()));
    THCGW_HIP(hipMemcpy(ctx->d_symph, ph.data(), sizeof(c128) * ph.size(), hipMemcpyHostToDevice));
    ctx->sym_nir = static_cast<int>(ir.size());
    ctx->sym = 1;
  });
}

// Brillouin-zone symmetry tables (bz_symmetry.hpp:497-576): every index is checked on the
// host, then the wedge path's gather lists and Sigma term lists go to the device.
int thcgw_set_bz(thcgw_ctx* ctx, const thcgw_bz* bz) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    if (ctx->d_bz) THCGW_HIP(hipFree(ctx->d_bz));
    if (ctx->d_bzD) THCGW_HIP(hipFree(ctx->d_bzD));
    ctx->d_bz = nullptr;
    ctx->d_bzD = nullptr;
    ctx->bz_on = 0;
    if (!bz) return;
    const int nk = ctx->nk, no = ctx->d.no;
    const int nki = bz->nk_ibz, nqi = bz->nq_ibz, ntr = bz->nk_trev_pairs, nsym = bz->nsym;
    THCGW_REQUIRE(bz->k_grid_index && bz->kp_to_ibz && bz->kp_trev && bz->kp_trev_pair && bz->qk_to_k2 &&
                      bz->qminus && bz->qp_to_ibz && bz->qp_trev && bz->nq_per_s && bz->Qs && bz->ks_to_k && bz->D,
                  "thcgw_set_bz: null table");
    THCGW_REQUIRE(nk <= 64, "thcgw_set_bz: the symmetry-reduced path takes at most 64 k-points");
    THCGW_REQUIRE(ntr >= 0 && ntr < nk && nki >= 1 && nki <= nk - ntr && nqi >= 1 && nqi <= nk,
                  "thcgw_set_bz: nk_ibz, nq_ibz or nk_trev_pairs out of range");
    THCGW_REQUIRE(nsym >= 1 && nsym <= 1024, "thcgw_set_bz: nsym out of range");
    const int nkn = nk - ntr;
    auto in = [](int v, int lo, int hi) { return v >= lo && v < hi; };
    std::vector<char> seen(nk, 0);
    for (int k = 0; k < nk; ++k) {
      const int g = bz->k_grid_index[k];
      THCGW_REQUIRE(in(g, 0, nk) && !seen[g], "thcgw_set_bz: k_grid_index is not a permutation of the grid");
      seen[g] = 1;
      THCGW_REQUIRE(bz->kp_trev[k] == 0 || bz->kp_trev[k] == 1, "thcgw_set_bz: kp_trev is 0 or 1");
      THCGW_REQUIRE((bz->kp_trev[k] != 0) == (k >= nkn),
                    "thcgw_set_bz: the nk_trev_pairs time-reversal points must be listed last");
      THCGW_REQUIRE(in(bz->kp_to_ibz[k], 0, nki), "thcgw_set_bz: kp_to_ibz out of range");
      THCGW_REQUIRE(k >= nki || bz->kp_to_ibz[k] == k, "thcgw_set_bz: the irreducible k-points must be listed first");
      THCGW_REQUIRE(!bz->kp_trev[k] || in(bz->kp_trev_pair[k], 0, nkn),
                    "thcgw_set_bz: a time-reversal point needs a partner without time reversal");
      THCGW_REQUIRE(in(bz->qminus[k], 0, nk) && in(bz->qp_to_ibz[k], 0, nqi) &&
                        (bz->qp_trev[k] == 0 || bz->qp_trev[k] == 1),
                    "thcgw_set_bz: qminus, qp_to_ibz or qp_trev out of range");
      THCGW_REQUIRE(k >= nqi || (bz->qp_to_ibz[k] == k && !bz->qp_trev[k]),
                    "thcgw_set_bz: the irreducible q-points must be listed first");
    }
    for (long e = 0; e < static_cast<long>(nk) * nk; ++e)
      THCGW_REQUIRE(in(bz->qk_to_k2[e], 0, nk), "thcgw_set_bz: qk_to_k2 out of range");
    std::vector<char> used(nk, 0);
    long tot = 0;
    for (int is = 0; is < nsym; ++is) {
      const int n = bz->nq_per_s[is];
      THCGW_REQUIRE(n >= 0 && n <= nk, "thcgw_set_bz: nq_per_s out of range");
      tot += n;
      for (int j = 0; j < n; ++j) {
        const int q = bz->Qs[static_cast<long>(is) * nk + j];
        THCGW_REQUIRE(in(q, 0, nk) && !used[q], "thcgw_set_bz: every q must appear in exactly one row of Qs");
        used[q] = 1;
      }
      for (int ik = 0; ik < nki; ++ik)
        THCGW_REQUIRE(in(bz->ks_to_k[static_cast<long>(is) * nki + ik], 0, nkn),
                      "thcgw_set_bz: k S must be a listed point without time reversal");
    }
    THCGW_REQUIRE(tot == nk, "thcgw_set_bz: nq_per_s must sum to N_k");
    std::vector<double> qw(nqi, 0.0);
    for (int q = 0; q < nk; ++q) qw[bz->qp_to_ibz[q]] += 1.0;
    // device int tables: G gather (nkn, nkn), placement (nk x 3), irreducible q (nqi),
    // X_{kS} (nsym nki), Sigma term offsets (nsym nki + 1) and terms (nki nk)
    const int nout = nsym * nki;
    std::vector<int> h;
    long o[10];
    o[BZ_G_SRC] = 0;
    for (int k = 0; k < nkn; ++k) h.push_back(bz->kp_to_ibz[k]);
    o[BZ_G_TRANS] = static_cast<long>(h.size());
    for (int k = 0; k < nkn; ++k) h.push_back(bz->kp_trev[k]);
    // chi0 in grid order: grid point g holds listed point k; a time-reversal point is
    // formed from its partner p (rpa_pi.icc:106-113)
    std::vector<int> listed(nk);
    for (int k = 0; k < nk; ++k) listed[bz->k_grid_index[k]] = k;
    auto src_pt = [&](int g) { const int k = listed[g]; return bz->kp_trev[k] ? bz->kp_trev_pair[k] : k; };
    o[BZ_PG_SRC] = static_cast<long>(h.size());
    for (int g = 0; g < nk; ++g) h.push_back(bz->kp_to_ibz[src_pt(g)]);
    o[BZ_PG_TRANS] = static_cast<long>(h.size());
    for (int g = 0; g < nk; ++g) h.push_back(bz->kp_trev[src_pt(g)]);
    o[BZ_PG_CONJ] = static_cast<long>(h.size());
    for (int g = 0; g < nk; ++g) h.push_back(bz->kp_trev[listed[g]]);
    o[BZ_PG_X] = static_cast<long>(h.size());
    for (int g = 0; g < nk; ++g) h.push_back(src_pt(g));
    o[BZ_Q_GRID] = static_cast<long>(h.size());
    for (int q = 0; q < nqi; ++q) h.push_back(bz->k_grid_index[q]);
    o[BZ_X_SRC] = static_cast<long>(h.size());
    for (int e = 0; e < nout; ++e) h.push_back(bz->ks_to_k[e]);
    // Sigma terms of thc_gw.icc:366-395: A = G_PQ rows (points without time reversal),
    // B = W rows (irreducible q); conj A where k - q needs time reversal (its partner's G),
    // conj B where q needs it (W(q) = conj W(q_ibz))
    std::vector<int> off(nout + 1, 0), terms;
    for (int is = 0; is < nsym; ++is)
      for (int ik = 0; ik < nki; ++ik) {
        const int ks = bz->ks_to_k[static_cast<long>(is) * nki + ik];
        for (int j = 0; j < bz->nq_per_s[is]; ++j) {
          const int qp = bz->Qs[static_cast<long>(is) * nk + j];
          const int qs = bz->qp_to_ibz[qp];
          const bool cb = bz->qp_trev[qp] != 0;
          const int k2 = cb ? bz->qk_to_k2[static_cast<long>(bz->qminus[qs]) * nk + ks]   // ks + qs
                            : bz->qk_to_k2[static_cast<long>(qs) * nk + ks];              // ks - qs
          const bool ca = bz->kp_trev[k2] != 0;
          const int ia = ca ? bz->kp_trev_pair[k2] : k2;
          THCGW_REQUIRE(in(ia, 0, nkn) && in(qs, 0, nqi), "thcgw_set_bz: inconsistent tables");
          terms.push_back(ia | (qs << 8) | (int(ca) << 16) | (int(cb) << 17));
        }
        off[is * nki + ik + 1] = static_cast<int>(terms.size());
      }
    o[BZ_T_OFF] = static_cast<long>(h.size());
    h.insert(h.end(), off.begin(), off.end());
    o[BZ_T_TERMS] = static_cast<long>(h.size());
    h.insert(h.end(), terms.begin(), terms.end());
    int* d = nullptr;
    THCGW_HIP(hipMalloc(&d, sizeof(int) * h.size()));
    ctx->d_bz = d;
    THCGW_HIP(hipMemcpy(d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    const size_t nD = static_cast<size_t>(nout) * no * no;
    THCGW_HIP(hipMalloc(&ctx->d_bzD, sizeof(c128) * nD));
    THCGW_HIP(hipMemcpy(ctx->d_bzD, bz->D, sizeof(c128) * nD, hipMemcpyHostToDevice));
    std::copy(o, o + 10, ctx->bz_o);
    ctx->bz_nk_ibz = nki;
    ctx->bz_nq_ibz = nqi;
    ctx->bz_nk_nt = nkn;
    ctx->bz_nsym = nsym;
    ctx->bz_nterms = static_cast<int>(terms.size());
    ctx->bz_qw = qw;
    ctx->bz_on = 1;
  });
}

int thcgw_profile_read(thcgw_ctx* ctx, double* ms, double* flops, double* bytes, long* launches) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(ms && flops && launches, "thcgw_profile_read: null pointer");
    thcgw::gemm_profile_read(ms, flops, bytes, launches);
  });
}

int thcgw_update_G(thcgw_ctx* ctx, const double* Cmo, const double* E, double mu, double* G, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(Cmo && E && G, "thcgw_update_G: null pointer");
    const int nt = ctx->d.nt, nk = ctx->nk, no = ctx->d.no;
    const long sm = static_cast<long>(no) * no;
    c128* CG = ctx->buf<c128>("ug_cg", static_cast<size_t>(nt) * nk * sm);
    thcgw::scale_mo(S(stream), C(Cmo), E, ctx->d_tau, mu, ctx->d.beta, nt, nk, no, CG);
    // G = (C g) C^H   (qp_scf_common.cpp:105-115)
    thcgw::zgemm(S(stream), thcgw::OP_N, thcgw::OP_C, no, no, no, kOne, thcgw::bp(CG, sm), no,
                 thcgw::bp(C(Cmo), sm, nk), no, kZero, Cm(G), sm, no, nt * nk);
  });
}

int thcgw_primary_to_aux(thcgw_ctx* ctx, long nb, const double* X, const double* Oin, double* Oout, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(X && Oin && Oout && nb >= 0, "thcgw_primary_to_aux: bad argument");
    const long nbk = nb * ctx->nk;
    c128* Ab = ctx->buf<c128>("proj_ab", static_cast<size_t>(nbk) * ctx->d.nmu * ctx->d.no);
    project_p2a(ctx, S(stream), C(X), C(Oin), Cm(Oout), nbk, Ab);
  });
}

// thcgw_primary_to_aux without device allocation: the intermediate X_k O_{b,k} of each (b, k)
// (N_mu x N_orb) is staged in the static scratch; the batch runs in chunks of whole (b, k)
// entries that fit it (whole b blocks when a block of nk fits, else k ranges within a block,
// so that the X_k of a chunk never wraps).  The GEMM overwrites the staged values (beta = 0),
// so nothing needs clearing between calls; calls that overlap on one GPU race on scratch.
int thcgw_primary_to_aux_noalloc(thcgw_ctx* ctx, long nb, const double* X, const double* Oin, double* Oout,
                                 void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(X && Oin && Oout && nb >= 0, "thcgw_primary_to_aux_noalloc: bad argument");
    const int nk = ctx->nk, no = ctx->d.no;
    const long per = static_cast<long>(ctx->d.nmu) * no;  // c128 elements staged per (b, k)
    const long cap = static_cast<long>(kNoallocScratchBytes / (sizeof(c128) * per));
    THCGW_REQUIRE(cap >= 1, "thcgw_primary_to_aux_noalloc: one staged N_mu x N_orb block (" +
                                std::to_string(per * 16) + " bytes) exceeds the " +
                                std::to_string(kNoallocScratchBytes >> 20) + " MiB scratch");
    c128* Ab = static_cast<c128*>(noalloc_scratch(ctx->device));
    const long so = static_cast<long>(no) * no;
    hipStream_t s = S(stream);
    if (cap >= nk) {
      const long bchunk = cap / nk;
      for (long b0 = 0; b0 < nb; b0 += bchunk) {
        const long cb = std::min(bchunk, nb - b0);
        project_p2a(ctx, s, C(X), C(Oin) + b0 * nk * so, Cm(Oout) + b0 * nk * ctx->M, cb * nk, Ab);
      }
    } else {
      const long sX = per;
      for (long b = 0; b < nb; ++b)
        for (long k0 = 0; k0 < nk; k0 += cap) {
          const long ck = std::min(cap, nk - k0);
          const long e0 = b * nk + k0;
          project_p2a(ctx, s, C(X) + k0 * sX, C(Oin) + e0 * so, Cm(Oout) + e0 * ctx->M, ck, Ab);
        }
    }
  });
}

int thcgw_aux_to_primary(thcgw_ctx* ctx, long nb, const double* X, const double* Oin, double* Oout, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(X && Oin && Oout && nb >= 0, "thcgw_aux_to_primary: bad argument");
    const long nbk = nb * ctx->nk;
    c128* Bb = ctx->buf<c128>("proj_ab", static_cast<size_t>(nbk) * ctx->d.nmu * ctx->d.no);
    project_a2p(ctx, S(stream), C(X), C(Oin), Cm(Oout), nbk, Bb);
  });
}

int thcgw_pi(thcgw_ctx* ctx, const double* G, const double* X, double* Pi, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(G && X && Pi, "thcgw_pi: null pointer");
    run_pi(ctx, C(G), C(X), Cm(Pi), S(stream));
  });
}

int thcgw_pi_kspace(thcgw_ctx* ctx, const double* G, const double* X, double* Pi, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(G && X && Pi, "thcgw_pi_kspace: null pointer");
    run_pi_kspace(ctx, C(G), C(X), Cm(Pi), S(stream));
  });
}

int thcgw_sigma_kspace(thcgw_ctx* ctx, const double* G, const double* Wqt, const double* X, double* Sig, int herm,
                       void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(G && Wqt && X && Sig, "thcgw_sigma_kspace: null pointer");
    run_sigma_kspace(ctx, C(G), C(Wqt), C(X), Cm(Sig), herm != 0, S(stream));
  });
}

int thcgw_pi_ibz(thcgw_ctx* ctx, const double* G, const double* X, double* Pi, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    require_bz(ctx, "thcgw_pi_ibz");
    THCGW_REQUIRE(G && X && Pi, "thcgw_pi_ibz: null pointer");
    run_pi_ibz(ctx, C(G), C(X), Cm(Pi), S(stream));
  });
}

int thcgw_sigma_ibz(thcgw_ctx* ctx, const double* G, const double* Wqt, const double* X, double* Sig, int herm,
                    void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    require_bz(ctx, "thcgw_sigma_ibz");
    THCGW_REQUIRE(G && Wqt && X && Sig, "thcgw_sigma_ibz: null pointer");
    run_sigma_ibz(ctx, C(G), C(Wqt), C(X), Cm(Sig), herm != 0, S(stream));
  });
}

int thcgw_gw_iteration_ibz(thcgw_ctx* ctx, const double* G, const double* X, const double* V, double* Sig,
                           double* Wqt, double* Ec, int herm, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    require_bz(ctx, "thcgw_gw_iteration_ibz");
    THCGW_REQUIRE(G && X && V && Sig && Wqt, "thcgw_gw_iteration_ibz: null pointer");
    hipStream_t s = S(stream);
    const long no = ctx->d.no;
    check_finite(ctx, "gw_iteration_ibz",
                 {{G, ctx->d.nt * ctx->bz_nk_ibz * no * no, "G_tskij"},
                  {X, static_cast<long>(ctx->nk) * ctx->d.nmu * no, "X_skPa"},
                  {V, ctx->bz_nq_ibz * ctx->M, "V_qPQ"}},
                 s);
    run_iteration_ibz(ctx, C(G), C(X), C(V), Cm(Sig), Cm(Wqt), Ec, herm != 0, s);
    THCGW_HIP(hipStreamSynchronize(s));
  });
}

int thcgw_tau_to_w(thcgw_ctx* ctx, const double* Pit, double* Piw, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(Pit && Piw && Pit != Piw, "thcgw_tau_to_w: bad pointers");
    run_tau_to_w(ctx, C(Pit), Cm(Piw), S(stream));
  });
}

int thcgw_w_to_tau(thcgw_ctx* ctx, const double* Ww, double* Wqt, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(Ww && Wqt && Ww != Wqt, "thcgw_w_to_tau: bad pointers");
    run_w_to_tau(ctx, C(Ww), Cm(Wqt), S(stream));
  });
}

int thcgw_dyson_w(thcgw_ctx* ctx, double* PiW, const double* V, double* trace_out, double* logdet_out, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(PiW && V, "thcgw_dyson_w: null pointer");
    const int total = ctx->nw_half * ctx->nk;
    check_finite(ctx, "thcgw_dyson_w", {{PiW, total * ctx->M, "Pi_wqPQ"}, {V, ctx->vq() * ctx->M, "V_qPQ"}}, S(stream));
    const bool terms = trace_out || logdet_out;
    c128* dtr = terms ? ctx->buf<c128>("ec_tr", total) : nullptr;
    double* dld = terms ? ctx->buf<double>("ec_ld", total) : nullptr;
    run_dyson(ctx, Cm(PiW), C(V), dtr, dld, true, false, S(stream));
    if (terms) {
      std::vector<std::complex<double>> tr;
      std::vector<double> ld;
      fetch_ec_terms(ctx, dtr, dld, tr, ld, S(stream));
      if (trace_out) std::memcpy(trace_out, tr.data(), sizeof(double) * 2 * total);
      if (logdet_out) std::memcpy(logdet_out, ld.data(), sizeof(double) * total);
    }
  });
}

int thcgw_sigma(thcgw_ctx* ctx, const double* G, const double* Wqt, const double* X, double* Sig, int herm,
                void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(G && Wqt && X && Sig, "thcgw_sigma: null pointer");
    run_sigma(ctx, C(G), C(Wqt), C(X), Cm(Sig), herm != 0, S(stream));
  });
}

int thcgw_hermitize(thcgw_ctx* ctx, long nb, int n, double* A, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(A && nb >= 0 && n >= 0, "thcgw_hermitize: bad argument");
    thcgw::hermitize(S(stream), Cm(A), n, nb);
  });
}

int thcgw_rpa_ec(thcgw_ctx* ctx, const double* Piw, const double* V, double* Ec, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(Piw && V && Ec, "thcgw_rpa_ec: null pointer");
    const int total = ctx->nw_half * ctx->nk;
    check_finite(ctx, "thcgw_rpa_ec", {{Piw, total * ctx->M, "Pi_wqPQ"}, {V, ctx->vq() * ctx->M, "V_qPQ"}}, S(stream));
    c128* dtr = ctx->buf<c128>("ec_tr", total);
    double* dld = ctx->buf<double>("ec_ld", total);
    // Pi_wqPQ is read only: factor copies of (I - Pi V) in the Dyson workspace.
    run_dyson(ctx, const_cast<c128*>(C(Piw)), C(V), dtr, dld, false, true, S(stream));
    std::vector<std::complex<double>> tr;
    std::vector<double> ld;
    fetch_ec_terms(ctx, dtr, dld, tr, ld, S(stream));
    *Ec = assemble_ec(ctx, tr, ld);
  });
}

int thcgw_gw_iteration(thcgw_ctx* ctx, const double* G, const double* X, const double* V, double* Sig, double* Wqt,
                       double* Ec, int herm, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(G && X && V && Sig && Wqt, "thcgw_gw_iteration: null pointer");
    hipStream_t s = S(stream);
    enqueue_iteration(ctx, C(G), C(X), C(V), Cm(Sig), Cm(Wqt), Ec != nullptr, herm != 0, s);
    if (Ec) *Ec = finish_ec(ctx, s);
    THCGW_HIP(hipStreamSynchronize(s));
  });
}

int thcgw_solve_qp(thcgw_ctx* ctx, const double* Sigma, const double* Cmo, const double* Vhf, const double* E0,
                   double mu, int nfit, double eta, double tol, double* E_out, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(Sigma && Cmo && Vhf && E0 && E_out, "thcgw_solve_qp: null pointer");
    hipStream_t s = S(stream);
    const int nt = ctx->d.nt, nk = ctx->nk, no = ctx->d.no;
    const long sm = static_cast<long>(no) * no;
    c128* T1 = ctx->buf<c128>("qp_t1", static_cast<size_t>(nt) * nk * sm);
    c128* dg = ctx->buf<c128>("qp_diag", static_cast<size_t>(nt + 1) * nk * no);
    // diag(C^H Sigma C) per (tau, k) and diag(C^H Vhf C) per k  (qp_scf_common.cpp:190-205, 213-231)
    thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, no, no, no, kOne, thcgw::bp(C(Sigma), sm), no, thcgw::bp(C(Cmo), sm, nk),
                 no, kZero, T1, sm, no, nt * nk);
    thcgw::mo_diag(s, T1, C(Cmo), dg, nk, no, static_cast<long>(nt) * nk);
    thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, no, no, no, kOne, thcgw::bp(C(Vhf), sm), no, thcgw::bp(C(Cmo), sm, nk),
                 no, kZero, T1, sm, no, nk);
    thcgw::mo_diag(s, T1, C(Cmo), dg + static_cast<long>(nt) * nk * no, nk, no, nk);
    // Pade fit frequencies (pade_driver.hpp:44-62) and the QP equation per state, on the device
    const std::vector<int> idx = thcgw::host::pade_fit_indices(ctx->d.nw_f, nfit);
    THCGW_REQUIRE(nfit <= thcgw::QP_MAX_FIT, "thcgw_solve_qp: nfit above 64 is not supported");
    thcgw::QpFit fit{};
    fit.nfit = nfit;
    for (int i = 0; i < nfit; ++i) {
      fit.w[i] = idx[i];
      fit.zim[i] = double(ctx->tab.wn_f[idx[i]]) * M_PI / ctx->d.beta;  // IAFT.hpp:55-57
    }
    const long dim = static_cast<long>(nk) * no;
    double* dE = ctx->buf<double>("qp_e", static_cast<size_t>(2 * dim));
    THCGW_HIP(hipMemcpyAsync(dE, E0, sizeof(double) * dim, hipMemcpyHostToDevice, s));
    thcgw::qp_solve(s, dg, dg + static_cast<long>(nt) * dim, dE, ctx->d_Twt_ff, fit, nt, dim, mu, eta, tol, dE + dim);
    THCGW_HIP(hipMemcpyAsync(E_out, dE + dim, sizeof(double) * dim, hipMemcpyDeviceToHost, s));
    THCGW_HIP(hipStreamSynchronize(s));
  });
}

int thcgw_hf(thcgw_ctx* ctx, const double* Dm, const double* X, const double* V, int hartree, int exchange, double* F,
             void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(Dm && X && V && F, "thcgw_hf: null pointer");
    hipStream_t s = S(stream);
    const int nk = ctx->nk, nmu = ctx->d.nmu, no = ctx->d.no;
    const long M = ctx->M;
    c128* DP = ctx->buf<c128>("hf_dp", static_cast<size_t>(nk) * M);
    c128* FP = ctx->buf<c128>("hf_fp", static_cast<size_t>(nk) * M);
    c128* Ab = ctx->buf<c128>("hf_ab", static_cast<size_t>(nk) * nmu * no);
    project_p2a(ctx, s, C(X), C(Dm), DP, nk, Ab);  // thc_hf.icc:162-163
    // exchange: F^P(k) = sum_R e^{ikR} [-Dm(R) o U(R)] = -(1/N_k) sum_q Dm^P(k-q) o V(q)  (thc_hf.icc:211-232)
    if (exchange)
      thcgw::sigma_kfft(s, ctx->d.kgrid, ctx->force_fallback, DP, nullptr, C(V), M, FP, nullptr, M,
                        -1.0 / (double(nk) * nk));
    else
      thcgw::zero_c128(s, FP, static_cast<long>(nk) * M);
    if (hartree) {  // thc_hf.icc:182-205
      c128* d = ctx->buf<c128>("hf_d", static_cast<size_t>(nmu));
      thcgw::hf_hartree(s, DP, C(V), d, FP, nk, nmu);
    }
    project_a2p(ctx, s, C(X), FP, Cm(F), nk, Ab);  // thc_hf.icc:236-237
  });
}

int thcgw_dyson_g(thcgw_ctx* ctx, const double* H0, const double* F, const double* Sov, const double* Sigma, double mu,
                  double* G, double* Dm, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(F && Sigma && G, "thcgw_dyson_g: null pointer");
    hipStream_t s = S(stream);
    check_dyson_g_inputs(ctx, "thcgw_dyson_g", H0, F, Sov, Sigma, s);
    const int nt = ctx->d.nt, nw = ctx->d.nw_f;
    const long dim = static_cast<long>(ctx->nk) * ctx->d.no * ctx->d.no;
    c128* Sw = sigma_to_w(ctx, C(Sigma), s);
    c128* Gw = ctx->buf<c128>("dg_r", static_cast<size_t>(nw) * dim);
    matsubara_solve(ctx, C(H0), C(F), C(Sov), Sw, mu, false, Gw, s);  // simple_dyson.cpp:65-78
    // G(tau) = w_to_tau(G(iw), fermi)  (simple_dyson.cpp:97-98)
    thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, nt, static_cast<int>(dim), nw, kOne, thcgw::bp(ctx->d_Ttw_ff, 0), nw,
                 thcgw::bp(Gw, 0), dim, kZero, Cm(G), 0, dim, 1);
    if (Dm)  // Dm = -tau_to_beta(G)  (simple_dyson.cpp:121-126)
      thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, 1, static_cast<int>(dim), nt, kMinusOne, thcgw::bp(ctx->d_Tbeta, 0),
                   nt, thcgw::bp(C(G), 0), dim, kZero, Cm(Dm), 0, dim, 1);
  });
}

int thcgw_nelec(thcgw_ctx* ctx, const double* H0, const double* F, const double* Sov, const double* Sigma, double mu,
                double* nel, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(F && Sigma && nel, "thcgw_nelec: null pointer");
    hipStream_t s = S(stream);
    check_dyson_g_inputs(ctx, "thcgw_nelec", H0, F, Sov, Sigma, s);
    c128* Sw = sigma_to_w(ctx, C(Sigma), s);
    *nel = nelec_at(ctx, C(H0), C(F), C(Sov), Sw, mu, s);
  });
}

int thcgw_update_mu(thcgw_ctx* ctx, const double* H0, const double* F, const double* Sov, const double* Sigma,
                    double nel_target, double old_mu, double mu_tol, double* mu_out, double* nel_out, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(F && Sigma && mu_out, "thcgw_update_mu: null pointer");
    THCGW_REQUIRE(mu_tol > 0.0, "thcgw_update_mu: mu_tol must be positive");
    hipStream_t s = S(stream);
    check_dyson_g_inputs(ctx, "thcgw_update_mu", H0, F, Sov, Sigma, s);
    c128* Sw = sigma_to_w(ctx, C(Sigma), s);
    auto N = [&](double mu) { return nelec_at(ctx, C(H0), C(F), C(Sov), Sw, mu, s); };
    // scf_common.cpp:147-202: bracket in steps of 0.2, then bisect to |nel - target| < mu_tol
    const double delta = 0.2;
    double nel = N(old_mu), mu = old_mu;
    if (std::abs(nel - nel_target) >= mu_tol) {
      double mu1, mu2;
      if (nel >= nel_target) {
        mu2 = old_mu;
        mu1 = old_mu - delta;
        int it = 0;
        for (; it < 1000 && N(mu1) > nel_target; ++it) mu1 -= delta;
        THCGW_REQUIRE(it < 1000, "thcgw_update_mu: no lower bracket within 200 Ha");
      } else {
        mu1 = old_mu;
        mu2 = old_mu + delta;
        int it = 0;
        for (; it < 1000 && N(mu2) < nel_target; ++it) mu2 += delta;
        THCGW_REQUIRE(it < 1000, "thcgw_update_mu: no upper bracket within 200 Ha");
      }
      mu = 0.5 * (mu1 + mu2);
      nel = N(mu);
      for (int it = 0; it < 200 && std::abs(nel - nel_target) >= mu_tol; ++it) {
        if (nel >= nel_target) mu2 = mu;
        else mu1 = mu;
        mu = 0.5 * (mu1 + mu2);
        nel = N(mu);
      }
    }
    *mu_out = mu;
    if (nel_out) *nel_out = nel;
  });
}

int thcgw_batch_create(int nsys, thcgw_ctx* const* ctxs, thcgw_batch** out) {
  g_create_error.clear();
  if (nsys < 1 || !ctxs || !out) {
    g_create_error = "thcgw_batch_create: bad argument";
    return THCGW_ERR_ARG;
  }
  *out = nullptr;
  thcgw_batch* b = new (std::nothrow) thcgw_batch();
  if (!b) return THCGW_ERR_ALLOC;
  for (int i = 0; i < nsys; ++i) {
    if (!ctxs[i] || (i > 0 && ctxs[i]->device != ctxs[0]->device)) {
      g_create_error = "thcgw_batch_create: null context or contexts on different devices";
      delete b;
      return THCGW_ERR_ARG;
    }
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) {
        g_create_error = "thcgw_batch_create: a context appears twice";
        delete b;
        return THCGW_ERR_ARG;
      }
  }
  b->ctxs.assign(ctxs, ctxs + nsys);
  b->device = ctxs[0]->device;
  if (hipSetDevice(b->device) != hipSuccess || hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_out, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    g_create_error = "thcgw_batch_create: stream or event creation failed";
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->cap) (void)hipStreamDestroy(b->cap);
    delete b;
    return THCGW_ERR_HIP;
  }
  for (thcgw_ctx* c : b->ctxs) {  // scratch moves to the shared pool
    for (auto& kv : c->ws)
      if (!thcgw_ctx::private_buf(kv.first) && kv.second.p) (void)hipFree(kv.second.p);
    for (auto it = c->ws.begin(); it != c->ws.end();)
      it = thcgw_ctx::private_buf(it->first) ? std::next(it) : c->ws.erase(it);
    c->pool = &b->pools[0];
    c->scratch_gen = &b->gen;
  }
  b->lane_of.assign(nsys, 0);
  *out = b;
  return THCGW_OK;
}

void thcgw_batch_destroy(thcgw_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  b->drop_graph();
  if (b->cap) {
    (void)hipStreamSynchronize(b->cap);
    (void)hipStreamDestroy(b->cap);
  }
  for (hipStream_t x : b->lane_s) (void)hipStreamSynchronize(x);
  if (b->ev_in) (void)hipEventDestroy(b->ev_in);
  if (b->ev_out) (void)hipEventDestroy(b->ev_out);
  b->free_lanes();
  for (thcgw_ctx* c : b->ctxs) {
    c->pool = nullptr;
    c->scratch_gen = nullptr;
    c->mem_share = 1;
  }
  for (auto& pool : b->pools)
    for (auto& kv : pool)
      if (kv.second.p) (void)hipFree(kv.second.p);
  delete b;
}

const char* thcgw_batch_last_error(const thcgw_batch* b) { return b ? b->err.c_str() : g_create_error.c_str(); }

int thcgw_batch_set(thcgw_batch* b, int use_graph, int tune) {
  if (!b) return THCGW_ERR_ARG;
  b->use_graph = use_graph != 0;
  b->tune = tune != 0;
  if (!b->use_graph) b->drop_graph();
  return THCGW_OK;
}

// Spread the systems over `lanes` streams, largest work first onto the least loaded lane
// (work ~ N_k nw_half N_mu^3, the Dyson step).  Only before the first iteration.
int thcgw_batch_set_lanes(thcgw_batch* b, int lanes) {
  if (!b) return THCGW_ERR_ARG;
  try {
    THCGW_REQUIRE(lanes >= 1 && lanes <= 8, "thcgw_batch_set_lanes: lanes must be in 1..8");
    THCGW_REQUIRE(!b->warmed, "thcgw_batch_set_lanes: the batch has already run; set the lanes before the first iteration");
    THCGW_HIP(hipSetDevice(b->device));
    const int n = static_cast<int>(b->ctxs.size());
    lanes = std::min(lanes, n);
    b->drop_graph();
    b->free_lanes();
    THCGW_HIP(hipDeviceSynchronize());  // contexts run alone may still use the scratch
    for (auto& pool : b->pools)
      for (auto& kv : pool)
        if (kv.second.p) THCGW_HIP(hipFree(kv.second.p));
    b->pools.clear();
    b->pools.resize(lanes);
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    auto work = [&](int i) {
      const thcgw_ctx* c = b->ctxs[i];
      return double(c->nk) * c->nw_half * double(c->d.nmu) * c->d.nmu * c->d.nmu;
    };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work(x) > work(y); });
    std::vector<double> load(lanes, 0.0);
    for (int i : order) {
      const int l = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
      b->lane_of[i] = l;
      load[l] += work(i);
      b->ctxs[i]->pool = &b->pools[l];
      b->ctxs[i]->mem_share = lanes;
    }
    if (lanes > 1) {
      b->lane_s.resize(lanes, nullptr);
      b->lane_ev.resize(lanes, nullptr);
      for (int l = 0; l < lanes; ++l) {
        THCGW_HIP(hipStreamCreateWithFlags(&b->lane_s[l], hipStreamNonBlocking));
        THCGW_HIP(hipEventCreateWithFlags(&b->lane_ev[l], hipEventDisableTiming));
      }
      THCGW_HIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
    }
    return THCGW_OK;
  } catch (const thcgw::Error& e) {
    b->err = e.what();
    return e.code;
  }
}

int thcgw_batch_gw_iteration(thcgw_batch* b, const double* const* G, const double* const* X, const double* const* V,
                             double* const* Sig, double* const* Wqt, double* Ec, int herm, void* stream) {
  if (!b) return THCGW_ERR_ARG;
  const int n = static_cast<int>(b->ctxs.size());
  try {
    THCGW_HIP(hipSetDevice(b->device));
    THCGW_REQUIRE(G && X && V && Sig && Wqt, "thcgw_batch_gw_iteration: null pointer array");
    for (int i = 0; i < n; ++i)
      THCGW_REQUIRE(G[i] && X[i] && V[i] && Sig[i] && Wqt[i], "thcgw_batch_gw_iteration: null pointer");
    hipStream_t s = S(stream);
    const bool want_ec = Ec != nullptr;
    if (!b->warmed) {
      // eager pass: allocates the shared scratch at its largest size, fixes the chunk sizes
      // (with several lanes the pools live side by side, and the automatic chunk sizes of
      // each context take 1/lanes of the free memory).
      // Calls of the contexts alone before this may have grown every lane's pool to a chunk
      // sized from the free memory of that moment; start the lanes from empty pools so that
      // the batch's own chunk sizes are what they hold.
      THCGW_HIP(hipDeviceSynchronize());
      for (auto& pool : b->pools) {
        for (auto& kv : pool)
          if (kv.second.p) THCGW_HIP(hipFree(kv.second.p));
        pool.clear();
      }
      for (int i = 0; i < n; ++i) {
        thcgw_ctx* c = b->ctxs[i];
        c->pair_chunk = c->pair_chunk > 0 ? c->pair_chunk : choose_pair_chunk(c, 4);
        c->dyson_chunk = c->dyson_chunk > 0 ? c->dyson_chunk : choose_dyson_chunk(c, c->nw_half * c->nk);
        enqueue_iteration(c, C(G[i]), C(X[i]), C(V[i]), Cm(Sig[i]), Cm(Wqt[i]), true, herm != 0, s);
        finish_ec(c, s);
      }
      // tile tuning: the FP32 tile shape per mixed-precision system, timed on a full iteration
      if (b->tune) {
        hipEvent_t e0, e1;
        THCGW_HIP(hipEventCreate(&e0));
        THCGW_HIP(hipEventCreate(&e1));
        for (int i = 0; i < n; ++i) {
          thcgw_ctx* c = b->ctxs[i];
          if (c->precision == 0) continue;
          float best = 1e30f;
          int pick = 0;
          for (int t = 0; t < 3; ++t) {
            c->tile32 = t;
            THCGW_HIP(hipEventRecord(e0, s));
            enqueue_iteration(c, C(G[i]), C(X[i]), C(V[i]), Cm(Sig[i]), Cm(Wqt[i]), true, herm != 0, s);
            THCGW_HIP(hipEventRecord(e1, s));
            THCGW_HIP(hipEventSynchronize(e1));
            float ms = 0.0f;
            THCGW_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
              best = ms;
              pick = t;
            }
          }
          c->tile32 = pick;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
      }
      b->warmed = true;
    }
    std::vector<const void*> sig;
    for (int i = 0; i < n; ++i) {
      sig.push_back(G[i]);
      sig.push_back(X[i]);
      sig.push_back(V[i]);
      sig.push_back(Sig[i]);
      sig.push_back(Wqt[i]);
    }
    sig.push_back(stream);
    sig.push_back(reinterpret_cast<const void*>(static_cast<long>(want_ec)));
    // the process-wide GEMM settings (flags 5, 8, 12) pick the kernels a graph holds
    sig.push_back(reinterpret_cast<const void*>(static_cast<long>(thcgw::gemm_mode_key())));
    // A scratch buffer grown since the capture (a context run alone with other settings)
    // leaves the graph pointing at freed memory: the allocation count is part of the key.
    sig.push_back(reinterpret_cast<const void*>(b->gen));
    if (b->use_graph) {
      // the caller's earlier work first, then capture / replay on the batch's own stream
      THCGW_HIP(hipEventRecord(b->ev_in, s));
      THCGW_HIP(hipStreamWaitEvent(b->cap, b->ev_in, 0));
      hipStream_t cs = b->cap;
      // A buffer that grows while the graph is captured frees memory an earlier node of the
      // same graph uses: such a capture is dropped and taken again over the grown buffers.
      for (int attempt = 0; !b->exec || sig != b->sig || b->herm != herm; ++attempt) {
        THCGW_REQUIRE(attempt < 2, "thcgw_batch_gw_iteration: scratch kept growing during graph capture");
        b->drop_graph();
        const unsigned long gen0 = b->gen;
        for (thcgw_ctx* c : b->ctxs) c->deferred = true;
        hipError_t ce = hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed);
        if (ce != hipSuccess) {
          for (thcgw_ctx* c : b->ctxs) c->deferred = false;
          THCGW_HIP(ce);
        }
        try {
          if (b->lanes() == 1) {
            for (int i = 0; i < n; ++i)
              enqueue_iteration(b->ctxs[i], C(G[i]), C(X[i]), C(V[i]), Cm(Sig[i]), Cm(Wqt[i]), want_ec, herm != 0,
                                cs);
          } else {
            // fork: every lane stream joins the capture through the fork event; join: the
            // capture stream waits for each lane's last work
            THCGW_HIP(hipEventRecord(b->ev_fork, cs));
            for (int l = 0; l < b->lanes(); ++l) THCGW_HIP(hipStreamWaitEvent(b->lane_s[l], b->ev_fork, 0));
            for (int i = 0; i < n; ++i)
              enqueue_iteration(b->ctxs[i], C(G[i]), C(X[i]), C(V[i]), Cm(Sig[i]), Cm(Wqt[i]), want_ec, herm != 0,
                                b->lane_s[b->lane_of[i]]);
            for (int l = 0; l < b->lanes(); ++l) {
              THCGW_HIP(hipEventRecord(b->lane_ev[l], b->lane_s[l]));
              THCGW_HIP(hipStreamWaitEvent(cs, b->lane_ev[l], 0));
            }
          }
        } catch (...) {
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(cs, &g);
          if (g) (void)hipGraphDestroy(g);
          (void)hipGetLastError();
          for (thcgw_ctx* c : b->ctxs) c->deferred = false;
          throw;
        }
        for (thcgw_ctx* c : b->ctxs) c->deferred = false;
        THCGW_HIP(hipStreamEndCapture(cs, &b->graph));
        if (b->gen != gen0) {
          sig.back() = reinterpret_cast<const void*>(b->gen);
          b->drop_graph();
          continue;
        }
        THCGW_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
        b->sig = sig;
        b->herm = herm;
      }
      THCGW_HIP(hipGraphLaunch(b->exec, cs));
      THCGW_HIP(hipEventRecord(b->ev_out, cs));
      THCGW_HIP(hipStreamWaitEvent(s, b->ev_out, 0));
      THCGW_HIP(hipStreamSynchronize(cs));
      for (int i = 0; i < n; ++i) {
        check_deferred(b->ctxs[i], cs);
        if (want_ec) Ec[i] = finish_ec(b->ctxs[i], cs);
      }
    } else {
      for (int i = 0; i < n; ++i) {
        enqueue_iteration(b->ctxs[i], C(G[i]), C(X[i]), C(V[i]), Cm(Sig[i]), Cm(Wqt[i]), want_ec, herm != 0, s);
        if (want_ec) Ec[i] = finish_ec(b->ctxs[i], s);
      }
      THCGW_HIP(hipStreamSynchronize(s));
    }
    return THCGW_OK;
  } catch (const thcgw::Error& e) {
    b->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    b->err = e.what();
    return THCGW_ERR_ARG;
  }
}

int thcgw_eps_inv_head(thcgw_ctx* ctx, const double* W_qtPQ, const double* chi_bar_qP, const double* qpts_cart,
                       double volume, int div_treatment, double* eps_inv_tq, double* eps_inv_t, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(W_qtPQ && chi_bar_qP && qpts_cart && eps_inv_t, "thcgw_eps_inv_head: null pointer");
    THCGW_REQUIRE(div_treatment >= 0 && div_treatment <= 3, "thcgw_eps_inv_head: div_treatment in 0..3");
    hipStream_t s = S(stream);
    const int nq = ctx->nk, nth = ctx->nt_half, n = ctx->d.nmu;
    c128* d = ctx->buf<c128>("gy_head", static_cast<size_t>(nth) * nq);
    thcgw::prod_head(s, C(W_qtPQ), C(chi_bar_qP), n, nth, nq, d);
    std::vector<std::complex<double>> h(static_cast<size_t>(nth) * nq);
    THCGW_HIP(hipMemcpyAsync(h.data(), d, sizeof(c128) * h.size(), hipMemcpyDeviceToHost, s));
    THCGW_HIP(hipStreamSynchronize(s));
    // eps^-1(t, q) = (|q|^2 / 4 pi) Vol chi_bar^T W conj(chi_bar)   (g0_div_utils.hpp:197-209)
    const double fpi = 4.0 * M_PI;
    for (int q = 0; q < nq; ++q) {
      const double* v = qpts_cart + 3 * q;
      const double f = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / fpi * volume;
      for (int t = 0; t < nth; ++t) h[size_t(t) * nq + q] *= f;
    }
    if (eps_inv_tq) std::memcpy(eps_inv_tq, h.data(), sizeof(c128) * h.size());
    // a single q-point cannot be extrapolated (g0_div_utils.hpp:66-70)
    const int mode = (nq == 1) ? 0 : div_treatment;
    const std::vector<std::complex<double>> h0 = thcgw::host::eps_inv_q0(ctx->tab, ctx->d, h, qpts_cart, nq, mode, 0.8, 2);
    std::memcpy(eps_inv_t, h0.data(), sizeof(c128) * h0.size());
  });
}

int thcgw_sigma_div_correction(thcgw_ctx* ctx, double* Sigma, const double* G, const double* X_skPa,
                               const double* chi_head_P, double madelung, const double* eps_inv_t, int div_treatment,
                               void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(div_treatment >= 0 && div_treatment <= 3, "thcgw_sigma_div_correction: div_treatment in 0..3");
    if (div_treatment == 0 || ctx->nk == 1) return;  // ignore_g0 (thc_gw.icc:454-461)
    THCGW_REQUIRE(Sigma && G && X_skPa && chi_head_P && eps_inv_t, "thcgw_sigma_div_correction: null pointer");
    hipStream_t s = S(stream);
    const int nt = ctx->d.nt, nk = ctx->nk, no = ctx->d.no, nmu = ctx->d.nmu, nth = ctx->nt_half;
    const long O2 = static_cast<long>(no) * no, XS = static_cast<long>(nmu) * no;
    // T^k = X^H diag(conj chi_head) X   (thc_gw.icc:476-491)
    c128* Y = ctx->buf<c128>("gy_y", static_cast<size_t>(nk) * XS);
    c128* T = ctx->buf<c128>("gy_t", static_cast<size_t>(nk) * O2);
    thcgw::scale_rows_conj(s, C(X_skPa), C(chi_head_P), nk, nmu, no, Y);
    thcgw::zgemm(s, thcgw::OP_C, thcgw::OP_N, no, no, nmu, kOne, thcgw::bp(C(X_skPa), XS), no, thcgw::bp(Y, XS), no,
                 kZero, T, O2, no, nk);
    // Delta(t, k) = -madelung Re eps^-1(t') T G(t, k) T^H, t' = t or nt - 1 - t   (thc_gw.icc:496-519)
    std::vector<double> f(nt);
    for (int t = 0; t < nt; ++t) {
      const int tp = t < nth ? t : nt - t - 1;
      f[t] = -madelung * eps_inv_t[2 * tp];
    }
    double* d_f = ctx->buf<double>("gy_f", static_cast<size_t>(nt));
    THCGW_HIP(hipMemcpyAsync(d_f, f.data(), sizeof(double) * nt, hipMemcpyHostToDevice, s));
    const long per_t = static_cast<long>(nk) * O2;
    c128* Gs = ctx->buf<c128>("gy_g", static_cast<size_t>(nt) * per_t);
    c128* B1 = ctx->buf<c128>("gy_b", static_cast<size_t>(nt) * per_t);
    THCGW_HIP(hipMemcpyAsync(Gs, G, sizeof(c128) * nt * per_t, hipMemcpyDeviceToDevice, s));
    thcgw::scale_by_t(s, Gs, d_f, per_t, nt);
    thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, no, no, no, kOne, thcgw::bp(T, O2, nk), no, thcgw::bp(Gs, O2), no, kZero,
                 B1, O2, no, nt * nk);
    thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_C, no, no, no, kOne, thcgw::bp(B1, O2), no, thcgw::bp(T, O2, nk), no, kOne,
                 Cm(Sigma), O2, no, nt * nk);
    THCGW_HIP(hipStreamSynchronize(s));  // f lives on the host stack
  });
}

int thcgw_gm_energy(thcgw_ctx* ctx, const double* G, const double* Sigma, double* e_corr, void* stream) {
  if (!ctx) return THCGW_ERR_ARG;
  return guarded(ctx, [&] {
    THCGW_REQUIRE(G && Sigma && e_corr, "thcgw_gm_energy: null pointer");
    hipStream_t s = S(stream);
    const int nt = ctx->d.nt, nw = ctx->d.nw_f, nk = ctx->nk, no = ctx->d.no;
    const long dim = static_cast<long>(nk) * no * no;
    c128* Gw = ctx->buf<c128>("gm_gw", static_cast<size_t>(nw) * dim);
    c128* Sw = ctx->buf<c128>("gm_sw", static_cast<size_t>(nw) * dim);
    c128* sg = ctx->buf<c128>("gm_sg", static_cast<size_t>(nw));
    // tau -> iw_n on the fermionic mesh (IAFT.icc:9-26), then sum_k w_k Tr[Sigma G] per frequency
    thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, nw, static_cast<int>(dim), nt, kOne, thcgw::bp(ctx->d_Twt_ff, 0), nt,
                 thcgw::bp(C(G), 0), dim, kZero, Gw, 0, dim, 1);
    thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, nw, static_cast<int>(dim), nt, kOne, thcgw::bp(ctx->d_Twt_ff, 0), nt,
                 thcgw::bp(C(Sigma), 0), dim, kZero, Sw, 0, dim, 1);
    thcgw::trace_prod(s, Sw, Gw, nw, nk, no, 1.0 / nk, sg);
    std::vector<std::complex<double>> h(nw);
    THCGW_HIP(hipMemcpyAsync(h.data(), sg, sizeof(c128) * nw, hipMemcpyDeviceToHost, s));
    THCGW_HIP(hipStreamSynchronize(s));
    *e_corr = thcgw::host::gm_from_sigmag(ctx->tab, h);
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// Learned low-rank approximate inverse for the Dyson solve (Krylov form).
//
// For each (n, q), with V = L L^H and A = -L^H Pi(iW_n) L (Hermitian, positive
// semidefinite), the Hermitian Dyson form needs M^-1 L^H, M = I + A:
// W - V = L M^-1 L^H - V.  The Krylov form solves M X = [L^H | Q] column by column with
// preconditioned conjugate gradients in FP64, where Q (n x r) is an orthonormal basis of
// the sketch A Omega (Omega a fixed Gaussian n x r) and the preconditioner is the
// rank-r approximate inverse
//     P^-1 = I - Q C Q^H,   C Hermitian r x r (the core),
// also used for the warm start X0 = P^-1 B.  The core that makes P^-1 agree with M^-1 on
// span(Q) in the Frobenius sense is C* = I - Q^H M^-1 Q; the direct solve computes it
// exactly, the Krylov solve from its last r columns.  The previous call's exact core is
// carried into the current basis as C_tr = O^H C*_prev O, O = Q_prev^H Q.  The core comes
// from
//   mode 0: no preconditioner (C = 0, X0 = 0);
//   mode 1: Woodbury on span(Q), C_W = T (I + T)^-1, T = Q^H A Q (exact when span(Q) is
//           invariant under A);
//   mode 2: C_tr (mode 1 on the first call);
//   mode 3: C_W plus the correction predicted by the network from (T, C_W, C_tr).
// Every Krylov solve is accepted only when each column's true FP64 residual
// ||B - M X|| / ||B|| is at most tol; otherwise its chunk is solved again directly.
// ---------------------------------------------------------------------------------
struct thcgw_lri {
  thcgw_ctx* ctx = nullptr;
  int r = 0, width = 0;
  std::unique_ptr<thcgw::Mlp> net;
  int mode = 3;
  int max_iter = 200;
  double tol = 1e-10;     // accepted relative residual (true FP64 residual, every column)
  double cg_tol = 1e-11;  // CG stops when every recursive residual is below this
  double t_scale = 0.1;   // feature scale of T blocks
  double y_scale = 1.0;   // target scale (set from the recorded targets at the first training)
  bool y_scaled = false;
  c128* Om = nullptr;     // n x r sketch
  c128* Qprev = nullptr;  // per (n, q): n x r basis of the previous call
  c128* Cprev = nullptr;  // per (n, q): r x r exact core of the previous call
  bool has_prev = false;
  float* feat = nullptr;  // recorded training pairs
  float* targ = nullptr;
  long n_pairs = 0, cap_pairs = 0;
  float* ybuf = nullptr;  // network output / training scratch
  long ybuf_count = 0;
  std::string err;
  int nin() const { return 3 * r * r + 1; }
  int nout() const { return r * r; }
  ~thcgw_lri() {
    for (void* p : {static_cast<void*>(Om), static_cast<void*>(Qprev), static_cast<void*>(Cprev),
                    static_cast<void*>(feat), static_cast<void*>(targ), static_cast<void*>(ybuf)})
      if (p) (void)hipFree(p);
  }
  float* scratch(long count) {
    if (count > ybuf_count) {
      if (ybuf) THCGW_HIP(hipFree(ybuf));
      ybuf = nullptr;
      ybuf_count = 0;
      THCGW_HIP(hipMalloc(&ybuf, sizeof(float) * count));
      ybuf_count = count;
    }
    return ybuf;
  }
};

struct thcgw_mlp {
  std::unique_ptr<thcgw::Mlp> net;
  int device = 0;
  float* buf = nullptr;  // Y and dY of train_mse
  long buf_count = 0;
  std::string err;
  ~thcgw_mlp() {
    if (buf) (void)hipFree(buf);
  }
};

namespace {

template <class F>
int guard_err(std::string& err, int device, F&& f) {
  try {
    THCGW_HIP(hipSetDevice(device));
    f();
    return THCGW_OK;
  } catch (const thcgw::Error& e) {
    err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    err = "host allocation failed";
    return THCGW_ERR_ALLOC;
  } catch (const std::exception& e) {
    err = e.what();
    return THCGW_ERR_ARG;
  }
}

// Woodbury core C = T (I + T)^-1 = I - (I + T)^-1 of r x r Hermitian T (batch), with
// (I + T)^-1 = L^-H L^-1 from the Cholesky factor of I + T.
void woodbury_core(thcgw_ctx* ctx, hipStream_t s, const c128* T, int r, int cb, c128* C, c128* W1, c128* W2,
                   int* info) {
  const long R2 = static_cast<long>(r) * r;
  thcgw::copy_batch(s, T, R2, 1 << 30, W1, R2, R2, 0, cb);
  thcgw::add_diag(s, W1, R2, r, r, 1.0, cb);
  thcgw::potrf_batched(s, W1, R2, r, r, info, cb);
  thcgw::set_identity(s, W2, R2, r, cb);
  thcgw::trsm_right_lower_conj(s, W1, R2, r, W2, R2, r, r, r, cb);  // W2 = L^-H
  thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_C, r, r, r, kMinusOne, thcgw::bp(W2, R2), r, thcgw::bp(W2, R2), r, kZero, C,
               R2, r, cb, kOne);
  (void)ctx;
}

// Z = P^-1 Rin = Rin - Q (C (Q^H Rin)) over n x ncol blocks (ld ncol).
void apply_pinv(hipStream_t s, const c128* Q, const c128* Cc, const c128* Rin, c128* Zout, c128* U, c128* U2, int n,
                int r, int ncol, int cb) {
  const long MB = static_cast<long>(n) * ncol, QS = static_cast<long>(n) * r, R2 = static_cast<long>(r) * r,
             US = static_cast<long>(r) * ncol;
  thcgw::zgemm(s, thcgw::OP_C, thcgw::OP_N, r, ncol, n, kOne, thcgw::bp(Q, QS), r, thcgw::bp(Rin, MB), ncol, kZero, U,
               US, ncol, cb);
  thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, r, ncol, r, kOne, thcgw::bp(Cc, R2), r, thcgw::bp(U, US), ncol, kZero, U2,
               US, ncol, cb);
  if (Zout != Rin) thcgw::copy2d(s, false, Rin, MB, ncol, 1 << 30, 0, Zout, MB, ncol, n, ncol, cb);
  thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, n, ncol, r, kMinusOne, thcgw::bp(Q, QS), r, thcgw::bp(U2, US), ncol, kOne,
               Zout, MB, ncol, cb);
}

double host_max(const double* d, int count, hipStream_t s) {
  std::vector<double> h(count);
  THCGW_HIP(hipMemcpyAsync(h.data(), d, sizeof(double) * count, hipMemcpyDeviceToHost, s));
  THCGW_HIP(hipStreamSynchronize(s));
  double m = 0.0;
  for (double v : h) m = std::isnan(v) ? INFINITY : std::max(m, v);
  return m;
}

std::vector<double> host_vec(const double* d, int count, hipStream_t s) {
  std::vector<double> h(count);
  THCGW_HIP(hipMemcpyAsync(h.data(), d, sizeof(double) * count, hipMemcpyDeviceToHost, s));
  THCGW_HIP(hipStreamSynchronize(s));
  return h;
}

// ||X - Y||_F^2 and ||Y||_F^2 summed over a batch of r x r matrices (host reduction).
void core_err_sums(const c128* X, const c128* Y, long count, hipStream_t s, double& num, double& den) {
  std::vector<std::complex<double>> hx(count), hy(count);
  THCGW_HIP(hipMemcpyAsync(hx.data(), X, sizeof(c128) * count, hipMemcpyDeviceToHost, s));
  THCGW_HIP(hipMemcpyAsync(hy.data(), Y, sizeof(c128) * count, hipMemcpyDeviceToHost, s));
  THCGW_HIP(hipStreamSynchronize(s));
  for (long i = 0; i < count; ++i) {
    num += std::norm(hx[i] - hy[i]);
    den += std::norm(hy[i]);
  }
}

// features: [T t_scale | C_W | C_tr | has_prev]; target: (C* - C_W) y_scale
void lri_features(thcgw_lri* L, hipStream_t s, const c128* T, const c128* Cw, const c128* Ctr, bool prev, float* F,
                  int cb) {
  const int r = L->r, nin = L->nin();
  const long R2 = static_cast<long>(r) * r;
  thcgw::pack_herm(s, T, nullptr, R2, r, F, nin, 0, L->t_scale, cb);
  thcgw::pack_herm(s, Cw, nullptr, R2, r, F, nin, R2, 1.0, cb);
  thcgw::pack_herm(s, Ctr, nullptr, R2, r, F, nin, 2 * R2, 1.0, cb);
  thcgw::set_col(s, F, nin, 3 * R2, prev ? 1.0f : 0.0f, cb);
}

void lri_record(thcgw_lri* L, hipStream_t s, const c128* T, const c128* Cw, const c128* Ctr, bool prev,
                const c128* Cst, int cb) {
  const int r = L->r, nin = L->nin(), nout = L->nout();
  const long R2 = static_cast<long>(r) * r;
  if (L->n_pairs + cb > L->cap_pairs) {
    const long cap = std::max(L->n_pairs + cb, 2 * L->cap_pairs + 256);
    float *f = nullptr, *t = nullptr;
    THCGW_HIP(hipMalloc(&f, sizeof(float) * cap * nin));
    THCGW_HIP(hipMalloc(&t, sizeof(float) * cap * nout));
    if (L->n_pairs) {
      THCGW_HIP(hipMemcpyAsync(f, L->feat, sizeof(float) * L->n_pairs * nin, hipMemcpyDeviceToDevice, s));
      THCGW_HIP(hipMemcpyAsync(t, L->targ, sizeof(float) * L->n_pairs * nout, hipMemcpyDeviceToDevice, s));
      THCGW_HIP(hipStreamSynchronize(s));
    }
    if (L->feat) THCGW_HIP(hipFree(L->feat));
    if (L->targ) THCGW_HIP(hipFree(L->targ));
    L->feat = f;
    L->targ = t;
    L->cap_pairs = cap;
  }
  float* F = L->feat + L->n_pairs * nin;
  float* Y = L->targ + L->n_pairs * nout;
  lri_features(L, s, T, Cw, Ctr, prev, F, cb);
  thcgw::pack_herm(s, Cst, Cw, R2, r, Y, nout, 0, L->y_scale, cb);
  L->n_pairs += cb;
}

}  // namespace

extern "C" {

int thcgw_lri_create(thcgw_ctx* ctx, int rank, int width, unsigned long long seed, thcgw_lri** out) {
  if (!ctx || !out) return THCGW_ERR_ARG;
  *out = nullptr;
  auto L = std::make_unique<thcgw_lri>();
  const int rc = guard_err(ctx->err, ctx->device, [&] {
    THCGW_REQUIRE(rank >= 1 && rank <= 128 && rank <= ctx->d.nmu, "thcgw_lri_create: rank must be in [1, min(128, nmu)]");
    THCGW_REQUIRE(width >= 1 && width <= 4096, "thcgw_lri_create: width must be in [1, 4096]");
    L->ctx = ctx;
    L->r = rank;
    L->width = width;
    L->net = std::make_unique<thcgw::Mlp>(L->nin(), width, L->nout(), seed);
    const int n = ctx->d.nmu, total = ctx->nw_half * ctx->nk;
    const long R2 = static_cast<long>(rank) * rank;
    // Omega: complex Gaussian with unit variance per entry (Box-Muller on a seeded generator)
    std::vector<std::complex<double>> h(static_cast<size_t>(n) * rank);
    std::mt19937_64 rng(seed ^ 0x5bd1e995ULL);
    std::normal_distribution<double> g(0.0, std::sqrt(0.5));
    for (auto& z : h) z = std::complex<double>(g(rng), g(rng));
    THCGW_HIP(hipMalloc(&L->Om, sizeof(c128) * h.size()));
    THCGW_HIP(hipMemcpy(L->Om, h.data(), sizeof(c128) * h.size(), hipMemcpyHostToDevice));
    THCGW_HIP(hipMalloc(&L->Qprev, sizeof(c128) * static_cast<long>(n) * rank * total));
    THCGW_HIP(hipMalloc(&L->Cprev, sizeof(c128) * R2 * total));
  });
  if (rc == THCGW_OK) *out = L.release();
  return rc;
}

void thcgw_lri_destroy(thcgw_lri* L) { delete L; }

const char* thcgw_lri_last_error(const thcgw_lri* L) { return L ? L->err.c_str() : "null handle"; }

int thcgw_lri_set(thcgw_lri* L, int mode, int max_iter, double tol) {
  if (!L) return THCGW_ERR_ARG;
  if (mode < 0 || mode > 3 || max_iter < 0 || !(tol > 0.0)) {
    L->err = "thcgw_lri_set: mode in 0..3, max_iter >= 0, tol > 0";
    return THCGW_ERR_ARG;
  }
  L->mode = mode;
  L->max_iter = max_iter;
  L->tol = tol;
  L->cg_tol = 0.1 * tol;
  return THCGW_OK;
}

long thcgw_lri_num_pairs(const thcgw_lri* L) { return L ? L->n_pairs : 0; }

int thcgw_lri_clear_pairs(thcgw_lri* L) {
  if (!L) return THCGW_ERR_ARG;
  L->n_pairs = 0;
  return THCGW_OK;
}

int thcgw_lri_reset_state(thcgw_lri* L) {
  if (!L) return THCGW_ERR_ARG;
  L->has_prev = false;
  return THCGW_OK;
}

int thcgw_lri_dyson(thcgw_lri* L, double* PiW, const double* V, int krylov, int record, thcgw_lri_stats* st,
                    void* stream) {
  if (!L) return THCGW_ERR_ARG;
  thcgw_ctx* ctx = L->ctx;
  return guard_err(L->err, ctx->device, [&] {
    THCGW_REQUIRE(PiW && V, "thcgw_lri_dyson: null pointer");
    hipStream_t s = S(stream);
    check_finite(ctx, "thcgw_lri_dyson",
                 {{PiW, static_cast<long>(ctx->nw_half) * ctx->nk * ctx->M, "Pi_wqPQ"}, {V, ctx->vq() * ctx->M, "V_qPQ"}},
                 s);
    const int n = ctx->d.nmu, nq = ctx->nk, r = L->r, ncol = n + r;
    const long M = ctx->M, R2 = static_cast<long>(r) * r, QS = static_cast<long>(n) * r,
               MB = static_cast<long>(n) * ncol, US = static_cast<long>(r) * ncol;
    const int total = ctx->nw_half * nq;
    thcgw_lri_stats acc{};
    double core_num = 0.0, core_den = 0.0, iter_sum = 0.0;
    int chunks = 0;
    // chunk: A, Zb (n x n); B, X, R, Zp, P, AP (n x ncol) per matrix
    const double per = 16.0 * (2.0 * M + (krylov ? 6.0 * MB : 0.0) + 6.0 * QS + 2.0 * US + 8.0 * R2);
    const double budget = std::min(0.3 * double(free_device_bytes()), 24.0 * double(1L << 30));
    const int CH = std::max(1, std::min(total, static_cast<int>(budget / per)));
    c128* Lv = ctx->buf<c128>("dy_lv", static_cast<size_t>(nq) * M);
    if (!factor_v(ctx, C(V), Lv, s)) throw thcgw::Error(THCGW_ERR_SINGULAR, "thcgw_lri_dyson: some V_q is not positive definite");
    c128* A = ctx->buf<c128>("lr_a", static_cast<size_t>(CH) * M);
    c128* Zb = ctx->buf<c128>("lr_z", static_cast<size_t>(CH) * M);
    c128* Sk = ctx->buf<c128>("lr_s", static_cast<size_t>(CH) * QS);
    c128* Q = ctx->buf<c128>("lr_q", static_cast<size_t>(CH) * QS);
    c128* AQ = ctx->buf<c128>("lr_aq", static_cast<size_t>(CH) * QS);
    c128* Gs = ctx->buf<c128>("lr_g", static_cast<size_t>(CH) * R2);
    c128* T = ctx->buf<c128>("lr_t", static_cast<size_t>(CH) * R2);
    c128* Cc = ctx->buf<c128>("lr_c", static_cast<size_t>(CH) * R2);
    c128* Cst = ctx->buf<c128>("lr_cst", static_cast<size_t>(CH) * R2);
    c128* Cw = ctx->buf<c128>("lr_cw", static_cast<size_t>(CH) * R2);
    c128* Ctr = ctx->buf<c128>("lr_ctr", static_cast<size_t>(CH) * R2);
    c128* W1 = ctx->buf<c128>("lr_w1", static_cast<size_t>(CH) * R2);
    c128* W2 = ctx->buf<c128>("lr_w2", static_cast<size_t>(CH) * R2);
    int* info = ctx->buf<int>("lr_info", static_cast<size_t>(CH));
    c128 *B = nullptr, *X = nullptr, *R = nullptr, *Zp = nullptr, *P = nullptr, *AP = nullptr, *U = nullptr,
         *U2 = nullptr;
    double *bn = nullptr, *rel = nullptr, *rz = nullptr, *rmax = nullptr;
    if (krylov) {
      B = ctx->buf<c128>("lr_b", static_cast<size_t>(CH) * MB);
      X = ctx->buf<c128>("lr_x", static_cast<size_t>(CH) * MB);
      R = ctx->buf<c128>("lr_r", static_cast<size_t>(CH) * MB);
      Zp = ctx->buf<c128>("lr_zp", static_cast<size_t>(CH) * MB);
      P = ctx->buf<c128>("lr_p", static_cast<size_t>(CH) * MB);
      AP = ctx->buf<c128>("lr_ap", static_cast<size_t>(CH) * MB);
      U = ctx->buf<c128>("lr_u", static_cast<size_t>(CH) * US);
      U2 = ctx->buf<c128>("lr_u2", static_cast<size_t>(CH) * US);
      bn = ctx->buf<double>("lr_bn", static_cast<size_t>(CH) * ncol);
      rel = ctx->buf<double>("lr_rel", static_cast<size_t>(CH) * ncol);
      rz = ctx->buf<double>("lr_rz", static_cast<size_t>(CH) * ncol);
      rmax = ctx->buf<double>("lr_rmax", static_cast<size_t>(CH));
    }
    const bool prev = L->has_prev;
    for (int b0 = 0; b0 < total; b0 += CH) {
      const int cb = std::min(CH, total - b0);
      c128* Pw = Cm(PiW) + static_cast<long>(b0) * M;
      // A = -L^H Pi L (lower tiles, then the upper triangle by reflection)
      thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, n, n, n, kOne, thcgw::bp(Pw, M), n, thcgw::bp(Lv, M, nq, 1, b0), n,
                   kZero, Zb, M, n, cb, kZero, thcgw::GEMM_TRI_KFROM_COL);
      thcgw::zgemm(s, thcgw::OP_C, thcgw::OP_N, n, n, n, kMinusOne, thcgw::bp(Lv, M, nq, 1, b0), n, thcgw::bp(Zb, M), n,
                   kZero, A, M, n, cb, kZero, thcgw::GEMM_TRI_LOWER | thcgw::GEMM_TRI_KFROM_ROW);
      thcgw::fill_upper(s, A, M, n, cb);
      // Q = orth(A Omega) by Cholesky QR; T = Q^H A Q
      thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, n, r, n, kOne, thcgw::bp(A, M), n, thcgw::bp(L->Om, 0), r, kZero, Sk,
                   QS, r, cb);
      thcgw::zgemm(s, thcgw::OP_C, thcgw::OP_N, r, r, n, kOne, thcgw::bp(Sk, QS), r, thcgw::bp(Sk, QS), r, kZero, Gs,
                   R2, r, cb);
      thcgw::potrf_batched(s, Gs, R2, r, r, info, cb);
      if (any_nonzero(info, cb, s))
        throw thcgw::Error(THCGW_ERR_SINGULAR, "thcgw_lri_dyson: the sketch A Omega is rank deficient");
      thcgw::copy_batch(s, Sk, QS, 1 << 30, Q, QS, QS, 0, cb);
      thcgw::trsm_right_lower_conj(s, Gs, R2, r, Q, QS, r, n, r, cb);
      thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, n, r, n, kOne, thcgw::bp(A, M), n, thcgw::bp(Q, QS), r, kZero, AQ, QS,
                   r, cb);
      thcgw::zgemm(s, thcgw::OP_C, thcgw::OP_N, r, r, n, kOne, thcgw::bp(Q, QS), r, thcgw::bp(AQ, QS), r, kZero, T, R2,
                   r, cb);
      thcgw::hermitize(s, T, r, cb);
      c128* Qp = L->Qprev + static_cast<long>(b0) * QS;
      c128* Cp = L->Cprev + static_cast<long>(b0) * R2;
      // Woodbury core, and the previous exact core carried into this basis
      woodbury_core(ctx, s, T, r, cb, Cw, W1, W2, info);
      if (prev) {
        thcgw::zgemm(s, thcgw::OP_C, thcgw::OP_N, r, r, n, kOne, thcgw::bp(Qp, QS), r, thcgw::bp(Q, QS), r, kZero, W1,
                     R2, r, cb);  // O = Q_prev^H Q
        thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, r, r, r, kOne, thcgw::bp(Cp, R2), r, thcgw::bp(W1, R2), r, kZero, W2,
                     R2, r, cb);
        thcgw::zgemm(s, thcgw::OP_C, thcgw::OP_N, r, r, r, kOne, thcgw::bp(W1, R2), r, thcgw::bp(W2, R2), r, kZero, Ctr,
                     R2, r, cb);
        thcgw::hermitize(s, Ctr, r, cb);
      } else {
        thcgw::copy_batch(s, Cw, R2, 1 << 30, Ctr, R2, R2, 0, cb);
      }
      // the core of the selected mode
      const int mode = L->mode;
      if (mode == 1) {
        thcgw::copy_batch(s, Cw, R2, 1 << 30, Cc, R2, R2, 0, cb);
      } else if (mode == 2) {
        thcgw::copy_batch(s, Ctr, R2, 1 << 30, Cc, R2, R2, 0, cb);
      } else if (mode == 3) {
        const int nin = L->nin(), nout = L->nout();
        float* F = L->scratch(static_cast<long>(cb) * (nin + nout));
        float* Yn = F + static_cast<long>(cb) * nin;
        lri_features(L, s, T, Cw, Ctr, prev, F, cb);
        L->net->forward(s, cb, F, Yn);
        thcgw::unpack_herm(s, Yn, nout, 1.0 / L->y_scale, Cw, Cc, R2, r, cb);
      }
      bool solved = false;
      if (krylov) {
        const long nmat = static_cast<long>(cb);
        // B = [L^H | Q]
        thcgw::copy2d(s, true, Lv, M, n, nq, b0 % nq, B, MB, ncol, n, n, cb);
        thcgw::copy2d(s, false, Q, QS, r, 1 << 30, 0, B + n, MB, ncol, n, r, cb);
        thcgw::colnorm(s, B, MB, ncol, n, ncol, bn, cb);
        if (mode == 0) {
          thcgw::zero_c128(s, X, nmat * MB);
          thcgw::copy2d(s, false, B, MB, ncol, 1 << 30, 0, R, MB, ncol, n, ncol, cb);
          thcgw::zero_c128(s, AP, nmat * MB);
          thcgw::residual(s, B, X, AP, nullptr, MB, ncol, n, ncol, bn, rel, cb);
        } else {
          apply_pinv(s, Q, Cc, B, X, U, U2, n, r, ncol, cb);  // warm start X0 = P^-1 B
          thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, n, ncol, n, kOne, thcgw::bp(A, M), n, thcgw::bp(X, MB), ncol,
                       kZero, AP, MB, ncol, cb);
          thcgw::residual(s, B, X, AP, R, MB, ncol, n, ncol, bn, rel, cb);
        }
        c128* Zr = mode == 0 ? R : Zp;
        if (mode != 0) apply_pinv(s, Q, Cc, R, Zp, U, U2, n, r, ncol, cb);
        thcgw::cg_init(s, R, Zr, P, MB, ncol, n, ncol, rz, cb);
        int it = 0;
        for (; it < L->max_iter; ++it) {
          thcgw::rowmax(s, rel, ncol, rmax, cb);
          if (host_max(rmax, cb, s) <= L->cg_tol) break;
          thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, n, ncol, n, kOne, thcgw::bp(A, M), n, thcgw::bp(P, MB), ncol, kZero,
                       AP, MB, ncol, cb);
          thcgw::cg_step(s, X, R, P, AP, MB, ncol, n, ncol, rz, bn, rel, cb);
          if (mode != 0) apply_pinv(s, Q, Cc, R, Zp, U, U2, n, r, ncol, cb);
          thcgw::cg_dir(s, R, Zr, P, MB, ncol, n, ncol, rz, cb);
        }
        // acceptance: true FP64 residual of every column
        thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, n, ncol, n, kOne, thcgw::bp(A, M), n, thcgw::bp(X, MB), ncol, kZero,
                     AP, MB, ncol, cb);
        thcgw::residual(s, B, X, AP, nullptr, MB, ncol, n, ncol, bn, rel, cb);
        thcgw::rowmax(s, rel, ncol, rmax, cb);
        const std::vector<double> res = host_vec(rmax, cb, s);
        int bad = 0;
        double worst = 0.0;
        for (double v : res) {
          if (!(v <= L->tol)) ++bad;
          else worst = std::max(worst, v);
        }
        iter_sum += it;
        ++chunks;
        acc.iterations = std::max(acc.iterations, it);
        if (bad == 0) {
          // W = L X_B - V;  C* = I - Q^H X_Q
          thcgw::copy_batch(s, C(V), M, nq, Pw, M, M, b0, cb);
          thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_N, n, n, n, kOne, thcgw::bp(Lv, M, nq, 1, b0), n, thcgw::bp(X, MB),
                       ncol, kMinusOne, Pw, M, n, cb);
          thcgw::zgemm(s, thcgw::OP_C, thcgw::OP_N, r, r, n, kMinusOne, thcgw::bp(Q, QS), r, thcgw::bp(X + n, MB), ncol,
                       kZero, Cst, R2, r, cb, kOne);
          thcgw::hermitize(s, Cst, r, cb);
          acc.krylov_accepted += cb;
          acc.resid_max = std::max(acc.resid_max, worst);
          solved = true;
        } else {
          acc.fallbacks += cb;
        }
      }
      if (!solved) {
        // direct: M = I + A = L_M L_M^H; W = K K^H - V, K = L L_M^-H; C* = I - Y Y^H, Y = Q^H L_M^-H
        thcgw::add_diag(s, A, M, n, n, 1.0, cb);
        thcgw::potrf_batched(s, A, M, n, n, info, cb);
        if (any_nonzero(info, cb, s))
          throw thcgw::Error(THCGW_ERR_SINGULAR, "thcgw_lri_dyson: I - L^H Pi L is not positive definite");
        thcgw::copy_batch(s, Lv, M, nq, Zb, M, M, b0, cb);
        thcgw::trsm_right_lower_conj(s, A, M, n, Zb, M, n, n, n, cb);
        thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_C, n, n, n, kOne, thcgw::bp(Zb, M), n, thcgw::bp(Zb, M), n, kZero, Pw, M,
                     n, cb, kZero, thcgw::GEMM_TRI_LOWER | thcgw::GEMM_HERM_SUB, thcgw::bp(C(V), M, nq, 1, b0));
        thcgw::copy2d(s, true, Q, QS, r, 1 << 30, 0, Sk, QS, n, r, n, cb);  // Q^H (r x n)
        thcgw::trsm_right_lower_conj(s, A, M, n, Sk, QS, n, r, n, cb);
        thcgw::zgemm(s, thcgw::OP_N, thcgw::OP_C, r, r, n, kMinusOne, thcgw::bp(Sk, QS), n, thcgw::bp(Sk, QS), n, kZero,
                     Cst, R2, r, cb, kOne);
        thcgw::hermitize(s, Cst, r, cb);
        if (st && mode != 0) core_err_sums(Cc, Cst, static_cast<long>(cb) * R2, s, core_num, core_den);
      }
      if (record & 1) lri_record(L, s, T, Cw, Ctr, prev, Cst, cb);
      if (!(record & 2)) {
        thcgw::copy_batch(s, Q, QS, 1 << 30, Qp, QS, QS, 0, cb);
        thcgw::copy_batch(s, Cst, R2, 1 << 30, Cp, R2, R2, 0, cb);
      }
      acc.solves += cb;
    }
    if (!(record & 2)) L->has_prev = true;
    THCGW_HIP(hipStreamSynchronize(s));
    if (st) {
      acc.iter_mean = chunks ? iter_sum / chunks : 0.0;
      acc.core_err = core_den > 0.0 ? std::sqrt(core_num / core_den) : 0.0;
      *st = acc;
    }
  });
}

int thcgw_lri_train(thcgw_lri* L, int epochs, int batch, float lr, unsigned long long seed, float* loss_first,
                    float* loss_last, void* stream) {
  if (!L) return THCGW_ERR_ARG;
  return guard_err(L->err, L->ctx->device, [&] {
    THCGW_REQUIRE(L->n_pairs > 0, "thcgw_lri_train: no recorded training pairs");
    THCGW_REQUIRE(epochs >= 1 && batch >= 1 && lr > 0.0f, "thcgw_lri_train: epochs, batch and lr must be positive");
    hipStream_t s = S(stream);
    const int nin = L->nin(), nout = L->nout();
    const long np = L->n_pairs;
    if (!L->y_scaled) {  // targets scaled to unit RMS once, before the first step
      std::vector<float> h(static_cast<size_t>(np) * nout);
      THCGW_HIP(hipMemcpyAsync(h.data(), L->targ, sizeof(float) * h.size(), hipMemcpyDeviceToHost, s));
      THCGW_HIP(hipStreamSynchronize(s));
      double ss = 0.0;
      for (float v : h) ss += double(v) * v;
      const double rms = std::sqrt(ss / double(h.size()));
      const double a = rms > 0.0 ? 1.0 / rms : 1.0;
      thcgw::scale_f32(s, L->targ, static_cast<long>(h.size()), static_cast<float>(a));
      L->y_scale *= a;
      L->y_scaled = true;
    }
    const int bs = static_cast<int>(std::min<long>(batch, np));
    const int steps = static_cast<int>((np + bs - 1) / bs);
    float* buf = L->scratch(static_cast<long>(bs) * (nin + 3L * nout));
    float *Xb = buf, *Tb = Xb + static_cast<long>(bs) * nin, *Yb = Tb + static_cast<long>(bs) * nout,
          *dYb = Yb + static_cast<long>(bs) * nout;
    int* d_idx = L->ctx->buf<int>("lr_idx", static_cast<size_t>(bs));
    float* d_loss = L->ctx->buf<float>("lr_loss", static_cast<size_t>(epochs) * steps);
    std::vector<int> perm(np);
    for (long i = 0; i < np; ++i) perm[i] = static_cast<int>(i);
    std::mt19937_64 rng(seed);
    for (int e = 0; e < epochs; ++e) {
      std::shuffle(perm.begin(), perm.end(), rng);
      for (int st = 0; st < steps; ++st) {
        const long lo = static_cast<long>(st) * bs;
        const int cnt = static_cast<int>(std::min<long>(bs, np - lo));
        THCGW_HIP(hipMemcpyAsync(d_idx, perm.data() + lo, sizeof(int) * cnt, hipMemcpyHostToDevice, s));
        thcgw::gather_rows(s, L->feat, nin, d_idx, Xb, cnt);
        thcgw::gather_rows(s, L->targ, nout, d_idx, Tb, cnt);
        L->net->train_mse(s, cnt, Xb, Tb, lr, Yb, dYb);
        THCGW_HIP(hipMemcpyAsync(d_loss + static_cast<long>(e) * steps + st, L->net->loss, sizeof(float),
                                 hipMemcpyDeviceToDevice, s));
        THCGW_HIP(hipStreamSynchronize(s));  // the index array is reused by the next step
      }
    }
    std::vector<float> h(static_cast<size_t>(epochs) * steps);
    THCGW_HIP(hipMemcpyAsync(h.data(), d_loss, sizeof(float) * h.size(), hipMemcpyDeviceToHost, s));
    THCGW_HIP(hipStreamSynchronize(s));
    double f = 0.0, l = 0.0;
    for (int st = 0; st < steps; ++st) {
      f += h[st];
      l += h[static_cast<size_t>(epochs - 1) * steps + st];
    }
    if (loss_first) *loss_first = static_cast<float>(f / steps);
    if (loss_last) *loss_last = static_cast<float>(l / steps);
  });
}

// ---- the network alone (checks of its kernels) ----

int thcgw_mlp_create(int n_in, int width, int n_out, unsigned long long seed, thcgw_mlp** out) {
  if (!out) return THCGW_ERR_ARG;
  *out = nullptr;
  auto m = std::make_unique<thcgw_mlp>();
  (void)hipGetDevice(&m->device);
  const int rc = guard_err(g_create_error, m->device, [&] { m->net = std::make_unique<thcgw::Mlp>(n_in, width, n_out, seed); });
  if (rc == THCGW_OK) *out = m.release();
  return rc;
}

void thcgw_mlp_destroy(thcgw_mlp* m) { delete m; }

const char* thcgw_mlp_last_error(const thcgw_mlp* m) { return m ? m->err.c_str() : g_create_error.c_str(); }

long thcgw_mlp_num_params(const thcgw_mlp* m) { return m ? m->net->n_params : 0; }

int thcgw_mlp_get_params(thcgw_mlp* m, float* host) {
  if (!m || !host) return THCGW_ERR_ARG;
  return guard_err(m->err, m->device, [&] { m->net->get_params(host, nullptr); });
}

int thcgw_mlp_set_params(thcgw_mlp* m, const float* host) {
  if (!m || !host) return THCGW_ERR_ARG;
  return guard_err(m->err, m->device, [&] {
    m->net->set_params(host, nullptr);
    THCGW_HIP(hipStreamSynchronize(nullptr));
  });
}

int thcgw_mlp_forward(thcgw_mlp* m, int batch, const float* X, float* Y, void* stream) {
  if (!m || !X || !Y) return THCGW_ERR_ARG;
  return guard_err(m->err, m->device, [&] { m->net->forward(S(stream), batch, X, Y); });
}

int thcgw_mlp_backward(thcgw_mlp* m, int batch, const float* X, const float* dY, float* grad, void* stream) {
  if (!m || !X || !dY || !grad) return THCGW_ERR_ARG;
  return guard_err(m->err, m->device, [&] {
    hipStream_t s = S(stream);
    m->net->backward(s, batch, X, dY);
    THCGW_HIP(hipMemcpyAsync(grad, m->net->g, sizeof(float) * m->net->n_params, hipMemcpyDeviceToDevice, s));
  });
}

int thcgw_mlp_adam_step(thcgw_mlp* m, const float* grad, float lr, void* stream) {
  if (!m || !grad) return THCGW_ERR_ARG;
  return guard_err(m->err, m->device, [&] { m->net->adam(S(stream), grad, lr); });
}

int thcgw_mlp_train_mse(thcgw_mlp* m, int batch, const float* X, const float* T, float lr, float* loss, void* stream) {
  if (!m || !X || !T) return THCGW_ERR_ARG;
  return guard_err(m->err, m->device, [&] {
    hipStream_t s = S(stream);
    const long need = 2L * batch * m->net->n_out;
    if (need > m->buf_count) {
      if (m->buf) THCGW_HIP(hipFree(m->buf));
      m->buf = nullptr;
      m->buf_count = 0;
      THCGW_HIP(hipMalloc(&m->buf, sizeof(float) * need));
      m->buf_count = need;
    }
    m->net->train_mse(s, batch, X, T, lr, m->buf, m->buf + static_cast<long>(batch) * m->net->n_out);
    if (loss) {
      THCGW_HIP(hipMemcpyAsync(loss, m->net->loss, sizeof(float), hipMemcpyDeviceToHost, s));
      THCGW_HIP(hipStreamSynchronize(s));
    }
  });
}

}  // extern "C"
