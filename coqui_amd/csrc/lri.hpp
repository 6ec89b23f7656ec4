// Kernels of the Krylov Dyson solve with a low-rank approximate inverse (lri.hip).
#pragma once

#include "thcgw_internal.hpp"

namespace thcgw {

// A(j, i) = conj(A(i, j)) for j > i, Im A(i, i) = 0 (n x n, ld n)
void fill_upper(hipStream_t s, c128* A, long sA, int n, int batch);
void add_diag(hipStream_t s, c128* A, long sA, int lda, int n, double v, int batch);
void set_identity(hipStream_t s, c128* A, long sA, int n, int batch);
// dst[b] (rows x cols, ldd) = src[(b + soff) % smod] or its conjugate transpose
void copy2d(hipStream_t s, bool conj_transpose, const c128* src, long sS, int lds, int smod, int soff, c128* dst,
            long sD, int ldd, int rows, int cols, int batch);
// Hermitian r x r (H - Hsub) packed into r*r reals at out[b * ldo + off]: Re below and on
// the diagonal, Im of the lower element above it.
void pack_herm(hipStream_t s, const c128* H, const c128* Hsub, long sH, int r, float* out, long ldo, long off,
               double scale, int batch);
void unpack_herm(hipStream_t s, const float* Y, long ldy, double scale, const c128* Cbase, c128* C, long sC, int r,
                 int batch);
void set_col(hipStream_t s, float* out, long ldo, long col, float v, int batch);
// column-wise CG over n x ncol blocks (ld) of matrices sM apart
void colnorm(hipStream_t s, const c128* X, long sX, int ld, int n, int ncol, double* nrm, int batch);
void residual(hipStream_t s, const c128* B, const c128* X, const c128* AX, c128* R, long sM, int ld, int n, int ncol,
              const double* bn, double* rel, int batch);
void cg_init(hipStream_t s, const c128* R, const c128* Z, c128* P, long sM, int ld, int n, int ncol, double* rz,
             int batch);
void cg_step(hipStream_t s, c128* X, c128* R, const c128* P, const c128* AP, long sM, int ld, int n, int ncol,
             const double* rz, const double* bn, double* rel, int batch);
void cg_dir(hipStream_t s, const c128* R, const c128* Z, c128* P, long sM, int ld, int n, int ncol, double* rz,
            int batch);
void rowmax(hipStream_t s, const double* v, int ncol, double* out, int batch);
void gather_rows(hipStream_t s, const float* src, long cols, const int* idx, float* dst, int rows);
void scale_f32(hipStream_t s, float* x, long n, float a);

}  // namespace thcgw
