// Internal declarations shared by the HIP translation units of libthcgw.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/thcgw.h"

#include <cstdint>
#include <stdexcept>
#include <string>

namespace thcgw {

// Device complex<double>: (re, im), 16-byte aligned, same layout as std::complex<double>.
using c128 = double2;

struct Cplx {
  double re, im;
};

enum Op : int { OP_N = 0, OP_T = 1, OP_C = 2 };

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};


#define THCGW_HIP(call)                                                                           \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      throw ::thcgw::Error(THCGW_ERR_HIP,                                                \
                           std::string(#call) + " failed: " + hipGetErrorString(e_));             \
  } while (0)

#define THCGW_REQUIRE(cond, msg)                                                                  \
  do {                                                                                            \
    if (!(cond)) throw ::thcgw::Error(THCGW_ERR_ARG, (msg));                             \
  } while (0)

// Pointer of batch member b: p + (((b + off) / div) % mod) * stride.
struct BatchPtr {
  const c128* p;
  long stride;
  int mod;
  int div;
  int off;
};

inline BatchPtr bp(const c128* p, long stride, int mod = 1 << 30, int div = 1, int off = 0) {
  return BatchPtr{p, stride, mod, div, off};
}

// C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b] + gamma * I   (row-major, M x N)
void zgemm(hipStream_t s, int opA, int opB, int M, int N, int K, Cplx alpha, BatchPtr A, long lda,
           BatchPtr B, long ldb, Cplx beta, c128* C, long strideC, long ldc, int batch,
           Cplx gamma = Cplx{0.0, 0.0});

// Number of FP64 GEMM launches and their modelled flops since the last reset (diagnostics).
struct GemmStats {
  long launches;
  double flops;
};
GemmStats gemm_stats();
void gemm_stats_reset();
void gemm_profile_enable(bool on);
void gemm_set_tile_mode(int mode);  // 0 automatic, 1 = 64x64, 2 = 128x64, 3 = 128x128
void gemm_profile_read(double* ms, double* flops, long* launches);

// kfft.hip
bool kfft_supported(const int* kgrid);
void pi_kfft(hipStream_t s, const int* kgrid, bool force_fallback, const c128* Gp, const c128* Gn, c128* Pi,
             long M, double scale);
void sigma_kfft(hipStream_t s, const int* kgrid, bool force_fallback, const c128* G1, const c128* G2,
                const c128* W, long w_kstride, c128* S1, c128* S2, long M, double scale);

// lu.hip
void getrf_batched(hipStream_t s, c128* A, long sA, int lda, int n, int* ipiv, int* info, int batch);
void trsm_right_upper_unit(hipStream_t s, const c128* U, long sU, int ldu, c128* T, long sT, int ldt, int m, int nb,
                           int batch);
void trsm_left_lower(hipStream_t s, const c128* L, long sL, int ldl, c128* B, long sB, int ldb, int n, int ncol,
                     int batch);
void trsm_left_upper_unit(hipStream_t s, const c128* U, long sU, int ldu, c128* B, long sB, int ldb, int n, int ncol,
                          int batch);

// misc.hip
void scale_mo(hipStream_t s, const c128* C, const double* E, const double* tau, double mu, double beta, int nt, int nk,
              int no, c128* CG);
void hermitize(hipStream_t s, c128* A, int n, long nb);
void one_minus_trace(hipStream_t s, const c128* A, long sA, int n, c128* trace, int batch);
void log_re_det(hipStream_t s, const c128* A, long sA, int n, const int* ipiv, double* ldet, int batch);
void dyson_finish(hipStream_t s, const c128* Z, long sZ, const c128* V, long sV, int vmod, const int* ipiv, int n,
                  c128* W, long sW, int b0, int batch);
void copy_batch(hipStream_t s, const c128* src, long sS, int smod, c128* dst, long sD, long per, int b0, int batch);
void mo_diag(hipStream_t s, const c128* T1, const c128* C, c128* out, int nk, int no, long nb);
void trace_prod(hipStream_t s, const c128* A, const c128* B, int nw, int nk, int no, double scale, c128* out);

}  // namespace thcgw
