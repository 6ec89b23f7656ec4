// Kernels of the Krylov Dyson solve with a low-rank approximate inverse (api.hip runs the
// solve): Hermitian fills and shifts, strided (conjugate-)transposed copies, packing of
// r x r Hermitian cores into network features, and the column-wise preconditioned
// conjugate-gradient steps.
//
// CG acts on the columns of an n x ncol block (row-major, leading dimension ld) of each
// matrix of a batch; every column is an independent system with its own alpha and beta.
// A workgroup owns 64 columns of one matrix: lane x = column, and its four waves split the
// rows and combine their partial sums in LDS, so each column's reduction stays in one
// workgroup (no atomics, fixed summation order).
#include "thcgw_internal.hpp"
#include "lri.hpp"

#include <algorithm>

namespace thcgw {
namespace {

constexpr int CX = 64;  // columns per workgroup
constexpr int CY = 4;   // row slices (waves)

__device__ __forceinline__ c128 cadd(c128 a, c128 b) { return c128{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ double re_conj_mul(c128 a, c128 b) { return a.x * b.x + a.y * b.y; }  // Re(conj(a) b)

// Reduce v over the four row slices; every thread of the column gets the sum.
__device__ __forceinline__ double col_reduce(double v, double (*red)[CX]) {
  red[threadIdx.y][threadIdx.x] = v;
  __syncthreads();
  const double s = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
  __syncthreads();
  return s;
}

__global__ void fill_upper_kernel(c128* A, long sA, int n) {
  c128* a = A + static_cast<long>(blockIdx.y) * sA;
  const long e = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<long>(n) * n) return;
  const int i = static_cast<int>(e / n), j = static_cast<int>(e % n);
  if (j > i) {
    const c128 v = a[static_cast<long>(j) * n + i];
    a[e] = c128{v.x, -v.y};
  } else if (j == i) {
    a[e].y = 0.0;
  }
}

__global__ void add_diag_kernel(c128* A, long sA, int lda, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[static_cast<long>(blockIdx.y) * sA + static_cast<long>(i) * lda + i].x += v;
}

__global__ void set_identity_kernel(c128* A, long sA, int n) {
  const long e = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<long>(n) * n) return;
  const int i = static_cast<int>(e / n), j = static_cast<int>(e % n);
  A[static_cast<long>(blockIdx.y) * sA + e] = c128{i == j ? 1.0 : 0.0, 0.0};
}

// dst(i, j) = op(src)(i, j) for a rows x cols destination; src member ((b + off) % mod).
template <bool CT>
__global__ void copy2d_kernel(const c128* __restrict__ src, long sS, int lds, int smod, int soff, c128* __restrict__ dst,
                              long sD, int ldd, int rows, int cols) {
  const int b = blockIdx.y;
  const long e = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= static_cast<long>(rows) * cols) return;
  const int i = static_cast<int>(e / cols), j = static_cast<int>(e % cols);
  const c128* s = src + static_cast<long>((b + soff) % smod) * sS;
  c128 v;
  if constexpr (CT) {
    v = s[static_cast<long>(j) * lds + i];
    v.y = -v.y;
  } else {
    v = s[static_cast<long>(i) * lds + j];
  }
  dst[static_cast<long>(b) * sD + static_cast<long>(i) * ldd + j] = v;
}

// out[b][off + i r + j] = scale * (j <= i ? Re D(i, j) : Im D(j, i)), D = H - Hsub (Hsub nullable)
__global__ void pack_herm_kernel(const c128* __restrict__ H, const c128* __restrict__ Hsub, long sH, int r,
                                 float* __restrict__ out, long ldo, long off, double scale) {
  const int b = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= r * r) return;
  const int i = e / r, j = e % r;
  const long src = static_cast<long>(b) * sH + (j <= i ? static_cast<long>(i) * r + j : static_cast<long>(j) * r + i);
  c128 v = H[src];
  if (Hsub) {
    v.x -= Hsub[src].x;
    v.y -= Hsub[src].y;
  }
  out[static_cast<long>(b) * ldo + off + e] = static_cast<float>(scale * (j <= i ? v.x : v.y));
}

// C = Cbase + scale * unpack(Y) (Hermitian; Cbase nullable)
__global__ void unpack_herm_kernel(const float* __restrict__ Y, long ldy, double scale, const c128* __restrict__ Cbase,
                                   c128* __restrict__ C, long sC, int r) {
  const int b = blockIdx.y;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= r * r) return;
  const int i = e / r, j = e % r;
  const float* y = Y + static_cast<long>(b) * ldy;
  const int lo_i = i >= j ? i : j, lo_j = i >= j ? j : i;  // the lower-triangle position
  double re = y[lo_i * r + lo_j], im = lo_i == lo_j ? 0.0 : y[lo_j * r + lo_i];
  if (i < j) im = -im;
  c128 v{scale * re, scale * im};
  if (Cbase) v = cadd(v, Cbase[static_cast<long>(b) * sC + e]);
  C[static_cast<long>(b) * sC + e] = v;
}

__global__ void set_col_kernel(float* out, long ldo, long col, float v) { out[static_cast<long>(blockIdx.x) * ldo + col] = v; }

// nrm[b][j] = ||X(:, j)||
__global__ __launch_bounds__(CX* CY) void colnorm_kernel(const c128* __restrict__ X, long sX, int ld, int n, int ncol,
                                                          double* __restrict__ nrm) {
  __shared__ double red[CY][CX];
  const int b = blockIdx.y, j = blockIdx.x * CX + threadIdx.x;
  const c128* x = X + static_cast<long>(b) * sX;
  double s = 0.0;
  if (j < ncol)
    for (int i = threadIdx.y; i < n; i += CY) {
      const c128 v = x[static_cast<long>(i) * ld + j];
      s += v.x * v.x + v.y * v.y;
    }
  s = col_reduce(s, red);
  if (j < ncol && threadIdx.y == 0) nrm[static_cast<long>(b) * ncol + j] = sqrt(s);
}

// R = B - X - AX; rel[b][j] = ||R(:, j)|| / bn[b][j]
__global__ __launch_bounds__(CX* CY) void residual_kernel(const c128* __restrict__ B, const c128* __restrict__ X,
                                                           const c128* __restrict__ AX, c128* __restrict__ R, long sM,
                                                           int ld, int n, int ncol, const double* __restrict__ bn,
                                                           double* __restrict__ rel) {
  __shared__ double red[CY][CX];
  const int b = blockIdx.y, j = blockIdx.x * CX + threadIdx.x;
  const long base = static_cast<long>(b) * sM;
  double s = 0.0;
  if (j < ncol)
    for (int i = threadIdx.y; i < n; i += CY) {
      const long o = base + static_cast<long>(i) * ld + j;
      const c128 v{B[o].x - X[o].x - AX[o].x, B[o].y - X[o].y - AX[o].y};
      if (R) R[o] = v;
      s += v.x * v.x + v.y * v.y;
    }
  s = col_reduce(s, red);
  if (j < ncol && threadIdx.y == 0) {
    const double d = bn[static_cast<long>(b) * ncol + j];
    rel[static_cast<long>(b) * ncol + j] = d > 0.0 ? sqrt(s) / d : sqrt(s);
  }
}

// P = Z, rz[j] = Re <R_j, Z_j>
__global__ __launch_bounds__(CX* CY) void cg_init_kernel(const c128* __restrict__ R, const c128* __restrict__ Z,
                                                          c128* __restrict__ P, long sM, int ld, int n, int ncol,
                                                          double* __restrict__ rz) {
  __shared__ double red[CY][CX];
  const int b = blockIdx.y, j = blockIdx.x * CX + threadIdx.x;
  const long base = static_cast<long>(b) * sM;
  double s = 0.0;
  if (j < ncol)
    for (int i = threadIdx.y; i < n; i += CY) {
      const long o = base + static_cast<long>(i) * ld + j;
      const c128 z = Z[o];
      P[o] = z;
      s += re_conj_mul(R[o], z);
    }
  s = col_reduce(s, red);
  if (j < ncol && threadIdx.y == 0) rz[static_cast<long>(b) * ncol + j] = s;
}

// With MP = P + AP: alpha = rz / Re <P, MP>; X += alpha P; R -= alpha MP; rel = ||R|| / bn.
// Columns with a non-positive curvature or zero rz stay where they are.
__global__ __launch_bounds__(CX* CY) void cg_step_kernel(c128* __restrict__ X, c128* __restrict__ R,
                                                          const c128* __restrict__ P, const c128* __restrict__ AP,
                                                          long sM, int ld, int n, int ncol,
                                                          const double* __restrict__ rz, const double* __restrict__ bn,
                                                          double* __restrict__ rel) {
  __shared__ double red[CY][CX];
  const int b = blockIdx.y, j = blockIdx.x * CX + threadIdx.x;
  const long base = static_cast<long>(b) * sM;
  double s = 0.0;
  if (j < ncol)
    for (int i = threadIdx.y; i < n; i += CY) {
      const long o = base + static_cast<long>(i) * ld + j;
      s += re_conj_mul(P[o], cadd(P[o], AP[o]));
    }
  const double pmp = col_reduce(s, red);
  double alpha = 0.0;
  if (j < ncol) {
    const double r0 = rz[static_cast<long>(b) * ncol + j];
    alpha = (pmp > 0.0 && r0 > 0.0) ? r0 / pmp : 0.0;
  }
  s = 0.0;
  if (j < ncol)
    for (int i = threadIdx.y; i < n; i += CY) {
      const long o = base + static_cast<long>(i) * ld + j;
      const c128 p = P[o], mp = cadd(p, AP[o]);
      const c128 x = X[o], r = R[o];
      X[o] = c128{x.x + alpha * p.x, x.y + alpha * p.y};
      const c128 rn{r.x - alpha * mp.x, r.y - alpha * mp.y};
      R[o] = rn;
      s += rn.x * rn.x + rn.y * rn.y;
    }
  s = col_reduce(s, red);
  if (j < ncol && threadIdx.y == 0) {
    const double d = bn[static_cast<long>(b) * ncol + j];
    rel[static_cast<long>(b) * ncol + j] = d > 0.0 ? sqrt(s) / d : sqrt(s);
  }
}

// rz' = Re <R, Z>; beta = rz' / rz; P = Z + beta P; rz = rz'.
__global__ __launch_bounds__(CX* CY) void cg_dir_kernel(const c128* __restrict__ R, const c128* __restrict__ Z,
                                                         c128* __restrict__ P, long sM, int ld, int n, int ncol,
                                                         double* __restrict__ rz) {
  __shared__ double red[CY][CX];
  const int b = blockIdx.y, j = blockIdx.x * CX + threadIdx.x;
  const long base = static_cast<long>(b) * sM;
  double s = 0.0;
  if (j < ncol)
    for (int i = threadIdx.y; i < n; i += CY) {
      const long o = base + static_cast<long>(i) * ld + j;
      s += re_conj_mul(R[o], Z[o]);
    }
  const double rz1 = col_reduce(s, red);
  double beta = 0.0;
  if (j < ncol) {
    const double r0 = rz[static_cast<long>(b) * ncol + j];
    beta = r0 > 0.0 ? rz1 / r0 : 0.0;
  }
  if (j < ncol)
    for (int i = threadIdx.y; i < n; i += CY) {
      const long o = base + static_cast<long>(i) * ld + j;
      const c128 z = Z[o], p = P[o];
      P[o] = c128{z.x + beta * p.x, z.y + beta * p.y};
    }
  if (j < ncol && threadIdx.y == 0) rz[static_cast<long>(b) * ncol + j] = rz1;
}

// out[b] = max_j v[b][j]
__global__ __launch_bounds__(256) void rowmax_kernel(const double* __restrict__ v, int ncol, double* __restrict__ out) {
  __shared__ double red[256];
  const double* x = v + static_cast<long>(blockIdx.x) * ncol;
  double m = 0.0;
  for (int j = threadIdx.x; j < ncol; j += 256) m = fmax(m, x[j]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// dst[i][:] = src[idx[i]][:] (cols floats per row)
__global__ void gather_rows_kernel(const float* __restrict__ src, long cols, const int* __restrict__ idx,
                                   float* __restrict__ dst) {
  const long row = blockIdx.y;
  const float* s = src + static_cast<long>(idx[row]) * cols;
  for (long c = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; c < cols; c += static_cast<long>(gridDim.x) * blockDim.x)
    dst[row * cols + c] = s[c];
}

__global__ void scale_f32_kernel(float* x, long n, float a) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) x[i] *= a;
}

unsigned nblk(long n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

}  // namespace

void gather_rows(hipStream_t s, const float* src, long cols, const int* idx, float* dst, int rows) {
  if (rows <= 0) return;
  hipLaunchKernelGGL(gather_rows_kernel, dim3(std::min<unsigned>(nblk(cols, 256), 64u), rows), dim3(256), 0, s, src,
                     cols, idx, dst);
  THCGW_HIP(hipGetLastError());
}

void scale_f32(hipStream_t s, float* x, long n, float a) {
  if (n <= 0) return;
  hipLaunchKernelGGL(scale_f32_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, x, n, a);
  THCGW_HIP(hipGetLastError());
}

void fill_upper(hipStream_t s, c128* A, long sA, int n, int batch) {
  hipLaunchKernelGGL(fill_upper_kernel, dim3(nblk(static_cast<long>(n) * n, 256), batch), dim3(256), 0, s, A, sA, n);
  THCGW_HIP(hipGetLastError());
}

void add_diag(hipStream_t s, c128* A, long sA, int lda, int n, double v, int batch) {
  hipLaunchKernelGGL(add_diag_kernel, dim3(nblk(n, 256), batch), dim3(256), 0, s, A, sA, lda, n, v);
  THCGW_HIP(hipGetLastError());
}

void set_identity(hipStream_t s, c128* A, long sA, int n, int batch) {
  hipLaunchKernelGGL(set_identity_kernel, dim3(nblk(static_cast<long>(n) * n, 256), batch), dim3(256), 0, s, A, sA, n);
  THCGW_HIP(hipGetLastError());
}

void copy2d(hipStream_t s, bool conj_transpose, const c128* src, long sS, int lds, int smod, int soff, c128* dst,
            long sD, int ldd, int rows, int cols, int batch) {
  const dim3 g(nblk(static_cast<long>(rows) * cols, 256), batch);
  if (conj_transpose)
    hipLaunchKernelGGL(copy2d_kernel<true>, g, dim3(256), 0, s, src, sS, lds, smod, soff, dst, sD, ldd, rows, cols);
  else
    hipLaunchKernelGGL(copy2d_kernel<false>, g, dim3(256), 0, s, src, sS, lds, smod, soff, dst, sD, ldd, rows, cols);
  THCGW_HIP(hipGetLastError());
}

void pack_herm(hipStream_t s, const c128* H, const c128* Hsub, long sH, int r, float* out, long ldo, long off,
               double scale, int batch) {
  hipLaunchKernelGGL(pack_herm_kernel, dim3(nblk(static_cast<long>(r) * r, 256), batch), dim3(256), 0, s, H, Hsub, sH,
                     r, out, ldo, off, scale);
  THCGW_HIP(hipGetLastError());
}

void unpack_herm(hipStream_t s, const float* Y, long ldy, double scale, const c128* Cbase, c128* C, long sC, int r,
                 int batch) {
  hipLaunchKernelGGL(unpack_herm_kernel, dim3(nblk(static_cast<long>(r) * r, 256), batch), dim3(256), 0, s, Y, ldy,
                     scale, Cbase, C, sC, r);
  THCGW_HIP(hipGetLastError());
}

void set_col(hipStream_t s, float* out, long ldo, long col, float v, int batch) {
  hipLaunchKernelGGL(set_col_kernel, dim3(batch), dim3(1), 0, s, out, ldo, col, v);
  THCGW_HIP(hipGetLastError());
}

void colnorm(hipStream_t s, const c128* X, long sX, int ld, int n, int ncol, double* nrm, int batch) {
  hipLaunchKernelGGL(colnorm_kernel, dim3(nblk(ncol, CX), batch), dim3(CX, CY), 0, s, X, sX, ld, n, ncol, nrm);
  THCGW_HIP(hipGetLastError());
}

void residual(hipStream_t s, const c128* B, const c128* X, const c128* AX, c128* R, long sM, int ld, int n, int ncol,
              const double* bn, double* rel, int batch) {
  hipLaunchKernelGGL(residual_kernel, dim3(nblk(ncol, CX), batch), dim3(CX, CY), 0, s, B, X, AX, R, sM, ld, n, ncol,
                     bn, rel);
  THCGW_HIP(hipGetLastError());
}

void cg_init(hipStream_t s, const c128* R, const c128* Z, c128* P, long sM, int ld, int n, int ncol, double* rz,
             int batch) {
  hipLaunchKernelGGL(cg_init_kernel, dim3(nblk(ncol, CX), batch), dim3(CX, CY), 0, s, R, Z, P, sM, ld, n, ncol, rz);
  THCGW_HIP(hipGetLastError());
}

void cg_step(hipStream_t s, c128* X, c128* R, const c128* P, const c128* AP, long sM, int ld, int n, int ncol,
             const double* rz, const double* bn, double* rel, int batch) {
  hipLaunchKernelGGL(cg_step_kernel, dim3(nblk(ncol, CX), batch), dim3(CX, CY), 0, s, X, R, P, AP, sM, ld, n, ncol,
                     rz, bn, rel);
  THCGW_HIP(hipGetLastError());
}

void cg_dir(hipStream_t s, const c128* R, const c128* Z, c128* P, long sM, int ld, int n, int ncol, double* rz,
            int batch) {
  hipLaunchKernelGGL(cg_dir_kernel, dim3(nblk(ncol, CX), batch), dim3(CX, CY), 0, s, R, Z, P, sM, ld, n, ncol, rz);
  THCGW_HIP(hipGetLastError());
}

void rowmax(hipStream_t s, const double* v, int ncol, double* out, int batch) {
  hipLaunchKernelGGL(rowmax_kernel, dim3(batch), dim3(256), 0, s, v, ncol, out);
  THCGW_HIP(hipGetLastError());
}

}  // namespace thcgw
