// Internal declarations shared by the HIP translation units of libthcgw.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/thcgw.h"

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace thcgw {

// Device complex<double>: (re, im), 16-byte aligned, same layout as std::complex<double>.
using c128 = double2;
// Device complex<float>: (re, im), 8-byte aligned, same layout as std::complex<float>.
using c64 = float2;

// Storage type of a complex operand.
enum Dt : int { DT_C128 = 0, DT_C64 = 1 };

struct Cplx {
  double re, im;
};

enum Op : int { OP_N = 0, OP_T = 1, OP_C = 2 };

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};


// A failed HIP call is reported with its message and the runtime's last-error state is
// cleared, so that the caller's next HIP call does not see it again.
#define THCGW_HIP(call)                                                                           \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) {                                                                       \
      (void)hipGetLastError();                                                                    \
      throw ::thcgw::Error(THCGW_ERR_HIP, std::string(#call) + " failed: " + hipGetErrorString(e_)); \
    }                                                                                             \
  } while (0)

#define THCGW_REQUIRE(cond, msg)                                                                  \
  do {                                                                                            \
    if (!(cond)) throw ::thcgw::Error(THCGW_ERR_ARG, (msg));                             \
  } while (0)

// Pointer of batch member b: p + (((b + off) / div) % mod) * stride, in elements of the
// operand's storage type.
struct BatchPtr {
  const void* p;
  long stride;
  int mod;
  int div;
  int off;
};

inline BatchPtr bp(const c128* p, long stride, int mod = 1 << 30, int div = 1, int off = 0) {
  return BatchPtr{p, stride, mod, div, off};
}
inline BatchPtr bp(const c64* p, long stride, int mod = 1 << 30, int div = 1, int off = 0) {
  return BatchPtr{p, stride, mod, div, off};
}

// Triangular structure flags of zgemm (tile granular, 64 x 64 tiles):
//   GEMM_TRI_LOWER     only tiles holding some element on or below the diagonal are computed;
//                      elements above the diagonal inside those tiles are unspecified.
//   GEMM_TRI_KFROM_ROW the K sum of a tile starts at its first row (op(A) is upper triangular,
//                      e.g. A = L^H with L lower); the skipped operand entries must be zero.
//   GEMM_TRI_KFROM_COL the K sum of a tile starts at its first column (op(B) lower triangular).
//   GEMM_TRI_KTO_ROW   the K sum of a tile ends after its last row (op(A) lower triangular).
//   GEMM_TRI_KTO_COL   the K sum of a tile ends after its last column (op(B) upper triangular).
constexpr int GEMM_TRI_LOWER = 1;
constexpr int GEMM_TRI_KFROM_ROW = 2;
constexpr int GEMM_TRI_KFROM_COL = 4;
constexpr int GEMM_TRI_KTO_ROW = 512;
constexpr int GEMM_TRI_KTO_COL = 1024;
// cgemm only: accumulate the FP32 MFMA partial sums of every K chunk in FP64.
constexpr int GEMM_FOLD64 = 8;
// cgemm only: split-BF16 products (three bf16 parts per value, six part products, FP32
// accumulation) instead of FP32 MFMA.
constexpr int GEMM_SPLIT_BF16 = 16;
// cgemm only: bits 6-7 hold (FP32 tile shape + 1) for products without FP64 folding
// (0 = the process default, see gemm_set_tile_mode32).
constexpr int GEMM_TILE32_SHIFT = 6;
// With GEMM_TRI_LOWER on a square product X = alpha op(A) op(B) that is Hermitian: write
// C = X - H on both triangles from the lower tiles (H complex<double>, same layout as C, or
// none for C = X; no beta, gamma or split BF16).
constexpr int GEMM_HERM_SUB = 256;
// cgemm only: three real MFMA products per complex product for this product, whatever the
// process-wide setting (gemm_set_complex3) says for FP32 products.
constexpr int GEMM_G3 = 2048;

// C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b] + gamma * I   (row-major, M x N)
void zgemm(hipStream_t s, int opA, int opB, int M, int N, int K, Cplx alpha, BatchPtr A, long lda,
           BatchPtr B, long ldb, Cplx beta, c128* C, long strideC, long ldc, int batch,
           Cplx gamma = Cplx{0.0, 0.0}, int tri = 0, BatchPtr hsub = BatchPtr{nullptr, 0, 1, 1, 0});

// The same product in FP32 arithmetic (complex<float> MFMA, FP32 accumulation) with each
// operand and the result stored as complex<double> or complex<float> (ta, tb, tc); alpha,
// beta and gamma are applied in double.  Ops N/T/C for A and B in {N, C} x {N, C}.
void cgemm(hipStream_t s, int opA, int opB, int M, int N, int K, Cplx alpha, BatchPtr A, Dt ta, long lda,
           BatchPtr B, Dt tb, long ldb, Cplx beta, void* C, Dt tc, long strideC, long ldc, int batch,
           Cplx gamma = Cplx{0.0, 0.0}, int tri = 0, BatchPtr hsub = BatchPtr{nullptr, 0, 1, 1, 0});

// Number of FP64 GEMM launches and their modelled flops since the last reset (diagnostics).
struct GemmStats {
  long launches;
  double flops;
};
GemmStats gemm_stats();
void gemm_stats_reset();
void gemm_profile_enable(bool on);
void gemm_set_tile_mode(int mode);    // 0 automatic, 1 = 64x64, 2 = 128x64, 3 = 128x128
void gemm_set_tile_mode32(int mode);  // FP32 products without FP64 folding: 0 = 64x64, 1 = 128x64, 2 = 128x128
void gemm_set_complex3(int on);       // 64x64 products by three real MFMA products: mask, see flag 12
int gemm_complex3();                  // the mask set last
int gemm_mode_key();                  // the process-wide settings above, packed (a captured graph's key)
constexpr int GEMM_PROFILE_CLASSES = 17;
void gemm_profile_read(double* ms, double* flops, double* bytes, long* launches);

// kfft.hip
bool kfft_supported(const int* kgrid);
void pi_kfft(hipStream_t s, const int* kgrid, bool force_fallback, const c128* Gp, const c128* Gn, c128* Pi,
             long M, double scale);
// herm_n > 0: S Hermitian of order herm_n, only its lower half L (S = L + L^H) is written
void sigma_kfft(hipStream_t s, const int* kgrid, bool force_fallback, const c128* G1, const c128* G2,
                const c128* W, long w_kstride, c128* S1, c128* S2, long M, double scale, int herm_n = 0);
// FP32 arithmetic (the generic-grid fallback computes in FP64), complex<float> G and Sigma.
void pi_kfft_f32(hipStream_t s, const int* kgrid, bool force_fallback, const c64* Gp, const c64* Gn, c128* Pi,
                 long M, double scale);
void sigma_kfft_f32(hipStream_t s, const int* kgrid, bool force_fallback, const c64* G1, const c64* G2,
                    const c128* W, long w_kstride, c64* S1, c64* S2, long M, double scale, int herm_n = 0);

// kconv.hip: k-space convolutions (chi0 uses the [q][k] table, Sigma the [k][q] one)
std::vector<int> kminus_table(const int* kgrid);  // t[a * nk + b] = index of k_a - k_b
void kconv_pi(hipStream_t s, const c128* Gp, const c128* Gn, const int* kmq_qk, int nk, c128* Pi, long M,
              double scale);
void kconv_sigma(hipStream_t s, const c128* G, const c128* W, long w_kstride, const int* kmq_kq, int nk, c128* S,
                 long M, double scale);

// symk.hip: Brillouin-zone symmetry (chi0 and Sigma on the irreducible wedge)
void gather_blocks(hipStream_t s, const c128* src, const int* sidx, const int* trans, const int* cj, int n, int count,
                   c128* dst);
void gather_rows(hipStream_t s, const c128* src, long sstride, const int* sidx, const int* cj, int count, long len,
                 c128* dst, long dstride, const int* dmap);
void kconv_terms(hipStream_t s, const c128* A, int nA, const c128* B, long b_stride, int nB, const int* off,
                 const int* terms, int nout, c128* out, long M, double scale, int herm_n = 0);
void sum_groups(hipStream_t s, const c128* src, int ng, long per, c128* dst);

// lu.hip
void getrf_batched(hipStream_t s, c128* A, long sA, int lda, int n, int* ipiv, int* info, int batch);
void trsm_right_upper_unit(hipStream_t s, const c128* U, long sU, int ldu, c128* T, long sT, int ldt, int m, int nb,
                           int batch);
void trsm_left_lower(hipStream_t s, const c128* L, long sL, int ldl, c128* B, long sB, int ldb, int n, int ncol,
                     int batch);
void trsm_left_upper_unit(hipStream_t s, const c128* U, long sU, int ldu, c128* B, long sB, int ldb, int n, int ncol,
                          int batch);

// chol.hip (lower Cholesky; only lower triangles of factors are defined)
void potrf_batched(hipStream_t s, c128* A, long sA, int lda, int n, int* info, int batch);
void trsm_right_lower_conj(hipStream_t s, const c128* L, long sL, int ldl, c128* T, long sT, int ldt, int m, int nb,
                           int batch);
void trsm_right_lower_conj_f32(hipStream_t s, const c128* L, long sL, int ldl, c64* T, long sT, int ldt, int m,
                               int nb, int batch, int gemm_flags = 0);
void trsm_right_lower_conj_f32(hipStream_t s, const c64* L, long sL, int ldl, c64* T, long sT, int ldt, int m,
                               int nb, int batch, int gemm_flags = 0);
void zero_upper(hipStream_t s, c128* A, long sA, int n, int batch);
void chol_logdet(hipStream_t s, const c128* L, long sL, int n, double* ldet, int batch, const double* add = nullptr,
                 int amod = 1, int aoff = 0);
// A = L L^H (lower) and Z = L^-H (upper, valid for products with triangular-K flags: zeros
// only in the band of 128 below the diagonal) by GEMM-only recursion
void cholinv_batched(hipStream_t s, c128* A, long sA, c128* Z, long sZ, int n, int* info, int batch);
// Z = L^-H for lower non-unit L (Z packed, ldz = n)
void tri_inverse_conj(hipStream_t s, const c128* L, long sL, int ldl, c128* Z, long sZ, int ldz, int n, int batch);
// lower triangle of A[b] = Vi[m] - P[b] and, with tr, tr[b] = Tr(V[m] P[b]) for Hermitian V
// (m = (b + voff) % vmod) in one pass; part holds resolvent_prep_parts(n) values per matrix
int resolvent_prep_parts(int n);
void resolvent_prep(hipStream_t s, const c128* V, const c128* Vi, long sV, int vmod, int voff, const c128* P, c128* A,
                    long sA, int n, c128* tr, c128* part, int batch);
void herm_finish(hipStream_t s, c128* W, long sW, const c128* V, long sV, int vmod, int voff, int n, int batch);

// irt.hip: out[b][r][c] = sum_i T[r][i] in[b][i][c] (T R x I row-major, R <= 48, I <= 44; rows of in
// and out ld_in / ld_out apart, matrices bs_in / bs_out apart); real_t: only Re T is used
bool ir_transform_fits(int R, int I);
void ir_transform(hipStream_t s, const c128* T, int R, int I, const c128* in, long ld_in, long bs_in, c128* out,
                  long ld_out, long bs_out, long cols, int batch, bool real_t = false);

// misc.hip
void scale_mo(hipStream_t s, const c128* C, const double* E, const double* tau, double mu, double beta, int nt, int nk,
              int no, c128* CG);
void hermitize(hipStream_t s, c128* A, int n, long nb, double scale = 1.0);
// L with A = L + L^H in place (strict upper triangle zeroed, diagonal halved)
void lower_half(hipStream_t s, c128* A, int n, long nb);
void lower_half(hipStream_t s, c64* A, int n, long nb);
void one_minus_trace(hipStream_t s, const c128* A, long sA, int n, c128* trace, int batch);
void log_re_det(hipStream_t s, const c128* A, long sA, int n, const int* ipiv, double* ldet, int batch);
// W = P Z - V: rows of Z in the order that undoes the LU column swaps, minus V (nullable).
void dyson_finish(hipStream_t s, const c128* Z, long sZ, const c128* V, long sV, int vmod, const int* ipiv, int n,
                  c128* W, long sW, int b0, int batch);
void copy_batch(hipStream_t s, const c128* src, long sS, int smod, c128* dst, long sD, long per, int b0, int batch);
void copy_batch(hipStream_t s, const c64* src, long sS, int smod, c64* dst, long sD, long per, int b0, int batch);
void to_c64(hipStream_t s, const c128* src, c64* dst, long n);
void zero_c128(hipStream_t s, c128* p, long n);
// X = (iw_n + mu) S - H0 - F - Sw for all (n, k) (S, H0 nullable); Y = S_k or I per (n, k).
void matsubara_dyson(hipStream_t s, const c128* H0, const c128* F, const c128* S, const c128* Sw, const long* wn,
                     double beta, double mu, int nw, int nk, int no, c128* X);
void rhs_batch(hipStream_t s, const c128* S, int nk, int no, long count, c128* Y);
void batched_trace(hipStream_t s, const c128* A, long sA, int n, c128* tr, int batch);
// THC Hartree term: d = twice the R = 0 diagonal of DP (nk, nmu, nmu); FP_kPP += (V0 d)_P.
void hf_hartree(hipStream_t s, const c128* DP, const c128* V0, c128* d, c128* FP, int nk, int nmu);
void mo_diag(hipStream_t s, const c128* T1, const c128* C, c128* out, int nk, int no, long nb);
void trace_prod(hipStream_t s, const c128* A, const c128* B, int nw, int nk, int no, double scale, c128* out);
// Time reversal: matrix t of a staging block is (n, q) = (n0 + t / nqi, qir[t % nqi]) of an
// (n, nq) array of M-element matrices.  gather copies it in; scatter writes it back to q and,
// conjugated, to mq[q]; scatter_ir_terms does the same for Tr (conjugated), ln det and flags
// (tr, ld nullable).  trev_diff: per (q, block) max |A[mq q] - conj A[q]| and max |A[q]|.
void gather_ir(hipStream_t s, const c128* src, long M, int n0, int nq, const int* qir, int nqi, int cnt, c128* dst);
void scatter_ir(hipStream_t s, const c128* S, long M, int n0, int nq, const int* qir, const int* mq, int nqi, int cnt,
                c128* dst);
void scatter_ir_terms(hipStream_t s, const c128* trs, const double* lds, const int* infos, int n0, int nq,
                      const int* qir, const int* mq, int nqi, int cnt, c128* tr, double* ld, int* info);
void trev_diff(hipStream_t s, const c128* A, long M, const int* mq, int nq, int nbx, double* out);
// out[b] (b < nbx) = number of non-finite doubles among x[0, n) in block b's grid-stride share
void nonfinite_count(hipStream_t s, const double* x, long n, int nbx, int* out);

// Quasiparticle equation per (k, n) state on the device (qp.hip): Pade fit frequencies as
// rows of Twt_ff (w) with z_i = i * zim[i].
constexpr int QP_MAX_FIT = 64;
struct QpFit {
  int nfit;
  int w[QP_MAX_FIT];
  double zim[QP_MAX_FIT];
};
// E_out[I] for I < dim from sig_t (nt, dim) diagonal Sigma_nn(tau), vhf (dim) diagonal Vhf_nn,
// E0 (dim) starting energies (all device arrays); Twt (nw_f, nt) the fermionic table.
void qp_solve(hipStream_t s, const c128* sig_t, const c128* vhf, const double* E0, const c128* Twt, const QpFit& fit,
              int nt, long dim, double mu, double eta, double tol, double* E_out);
// per block b' < nbx: max |A_ij - conj(A_ji)| and max |A_ij| over nb matrices of order n
void herm_diff(hipStream_t s, const c128* A, int n, long nb, int nbx, double* out);
// q-point symmetry map: W[w][q] = ph_q conj(ph_q)^T o W^[w][src q](perm_q, perm_q) for the q with
// src q != q (W^ = conj W where cj[q]); the Ec terms and flags likewise; and the check
// max |A[q] - map(A[src q])| per (q, block) with max |A[q]|.
void qsym_fill(hipStream_t s, c128* W, int n, int nq, int nw, const int* src, const int* cj, const int* perm,
               const c128* ph);
void qsym_fill_terms(hipStream_t s, int nq, int nw, const int* src, const int* cj, c128* tr, double* ld, int* info);
void qsym_diff(hipStream_t s, const c128* A, int n, int nq, const int* src, const int* cj, const int* perm,
               const c128* ph, int nbx, double* out);
// Gygi correction: out[t nq + q] = sum_PQ cbar[q][P] W[q][t][P][Q] conj(cbar[q][Q]);
// Y[k][P][a] = conj(h[P]) X[k][P][a]; A[t][...] *= f[t] (per_t elements per t).
void prod_head(hipStream_t s, const c128* W, const c128* cbar, int n, int nt_half, int nq, c128* out);
void scale_rows_conj(hipStream_t s, const c128* X, const c128* h, int nk, int nmu, int no, c128* Y);
void scale_by_t(hipStream_t s, c128* A, const double* f, long per_t, int nt);

}  // namespace thcgw
