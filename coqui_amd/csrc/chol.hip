// Batched Cholesky factorisation and the triangular kernels of the Hermitian Dyson solve,
// complex<double>, row-major.
//
// For Hermitian positive definite V_q = L L^H and Hermitian negative semidefinite Pi(iW_n)
// (the case for any G built from a Hermitian H(k) and Fermi occupations), the Dyson
// solve of scr_coulomb_t.cpp:180-196 and the determinant of thc_rpa.icc:141-163 are
//
//     M = I - L^H Pi L          (Hermitian, eigenvalues >= 1),
//     det(I - V Pi) = det(I - Pi V) = det M = prod_j L_M[j][j]^2,   M = L_M L_M^H,
//     Tr(V Pi) = Tr(L^H Pi L) = N - Tr M,
//     W = (I - V Pi)^-1 V - V = L M^-1 L^H - V = K K^H - V,   K = L L_M^-H,
//
// which replaces LU with partial pivoting by two Cholesky factorisations that need no
// pivoting, and turns the dense products into triangular ones (see the GEMM_TRI flags).
//
// The factorisation is recursive on the lower triangle: factor A11, A21 <- A21 L11^-H,
// A22 <- A22 - A21 A21^H (lower tiles only), factor A22.  Leaves of order <= 32 are
// factored in LDS by one workgroup per matrix.  Only the lower triangle of a factored
// matrix is defined.
#include "thcgw_internal.hpp"

#include <cmath>

namespace thcgw {
namespace {

__device__ __forceinline__ c128 cmul(c128 a, c128 b) { return c128{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// a * conj(b)
__device__ __forceinline__ c128 cmulc(c128 a, c128 b) { return c128{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y}; }

constexpr int CB = 32;  // leaf order

// In-place Cholesky of the nb x nb (nb <= 32) diagonal block at A[j0][j0] of each matrix.
// info[b] receives j0 + j + 1 for the first non-positive (or NaN) pivot j; the factor then
// continues with a unit pivot so that the launch stays finite.
__global__ __launch_bounds__(CB* CB) void potrf_leaf_kernel(c128* A, long sA, int lda, int j0, int nb, int* info) {
  __shared__ c128 S[CB][CB + 1];
  __shared__ double s_inv;
  const int b = blockIdx.x;
  c128* Ab = A + static_cast<long>(b) * sA + static_cast<long>(j0) * lda + j0;
  const int t = threadIdx.x;
  const int i = t / CB, k = t % CB;
  S[i][k] = (i < nb && k <= i) ? Ab[static_cast<long>(i) * lda + k] : c128{0.0, 0.0};
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    if (t == 0) {
      double d = S[j][j].x;
      if (!(d > 0.0)) {
        if (info[b] == 0) info[b] = j0 + j + 1;
        d = 1.0;
      }
      const double l = sqrt(d);
      S[j][j] = c128{l, 0.0};
      s_inv = 1.0 / l;
    }
    __syncthreads();
    if (k == j && i > j && i < nb) {
      const double r = s_inv;
      S[i][j] = c128{S[i][j].x * r, S[i][j].y * r};
    }
    __syncthreads();
    if (i > j && k > j && k <= i && i < nb) {
      const c128 u = cmulc(S[i][j], S[k][j]);
      S[i][k] = c128{S[i][k].x - u.x, S[i][k].y - u.y};
    }
    __syncthreads();
  }
  if (i < nb && k <= i) Ab[static_cast<long>(i) * lda + k] = S[i][k];
}

// X L^H = T for the lower non-unit nb x nb block L (nb <= CBL); one thread per row of T.
// RC is the complex arithmetic type; T and L may be stored in another precision.
template <typename RC, typename TT, typename TL, int CBL = CB>
__global__ __launch_bounds__(256) void trsm_rlc_leaf_kernel(const TL* L, long sL, int ldl, TT* T, long sT, int ldt,
                                                            int m, int nb) {
  using S = decltype(RC{}.x);
  constexpr int CB = CBL;
  __shared__ RC Ls[CB][CB];
  const int b = blockIdx.y;
  const TL* Lb = L + static_cast<long>(b) * sL;
  for (int e = threadIdx.x; e < CB * CB; e += 256) {
    const int r = e / CB, c = e % CB;
    if (r < nb && c <= r) {
      const TL v = Lb[static_cast<long>(r) * ldl + c];
      Ls[r][c] = RC{static_cast<S>(v.x), static_cast<S>(v.y)};
    } else {
      Ls[r][c] = RC{0, 0};
    }
  }
  __syncthreads();
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= m) return;
  TT* Tr = T + static_cast<long>(b) * sT + static_cast<long>(row) * ldt;
  RC x[CB];
#pragma unroll
  for (int c = 0; c < CB; ++c) {
    if (c < nb) {
      const TT v = Tr[c];
      x[c] = RC{static_cast<S>(v.x), static_cast<S>(v.y)};
    } else {
      x[c] = RC{0, 0};
    }
  }
#pragma unroll
  for (int c = 0; c < CB; ++c) {
    if (c < nb) {
      RC s = x[c];
#pragma unroll
      for (int cp = 0; cp < c; ++cp) {  // s -= x[cp] * conj(L[c][cp])
        const RC a = x[cp], l = Ls[c][cp];
        s = RC{s.x - (a.x * l.x + a.y * l.y), s.y - (a.y * l.x - a.x * l.y)};
      }
      // divide by conj(L[c][c])
      const RC d = Ls[c][c];
      const S inv = S(1) / (d.x * d.x + d.y * d.y);
      x[c] = RC{(s.x * d.x - s.y * d.y) * inv, (s.y * d.x + s.x * d.y) * inv};
    }
  }
  using ST = decltype(TT{}.x);
#pragma unroll
  for (int c = 0; c < CB; ++c)
    if (c < nb) Tr[c] = TT{static_cast<ST>(x[c].x), static_cast<ST>(x[c].y)};
}

__global__ void fill_int0_kernel(int* p, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Zero the strict upper triangle of each n x n matrix.
__global__ __launch_bounds__(256) void zero_upper_kernel(c128* A, long sA, int n) {
  const long e = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  const long nn = static_cast<long>(n) * n;
  if (e >= nn) return;
  const int r = static_cast<int>(e / n), c = static_cast<int>(e % n);
  if (c > r) A[static_cast<long>(blockIdx.y) * sA + e] = c128{0.0, 0.0};
}

// ldet[b] = 2 sum_j ln Re L[j][j]
__global__ __launch_bounds__(256) void chol_logdet_kernel(const c128* L, long sL, int n, double* ldet) {
  __shared__ double red[256];
  const c128* Lb = L + static_cast<long>(blockIdx.x) * sL;
  double acc = 0.0;
  for (int j = threadIdx.x; j < n; j += 256) acc += log(Lb[static_cast<long>(j) * n + j].x);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) ldet[blockIdx.x] = 2.0 * red[0];
}

// W <- H - V where H is the Hermitian matrix whose lower triangle W holds on entry.
// One workgroup per 32 x 32 tile pair (ti >= tj); the upper tile is written through LDS
// so that both reads and writes are row-contiguous.
__global__ __launch_bounds__(256) void herm_finish_kernel(c128* W, long sW, const c128* V, long sV, int vmod, int voff,
                                                          int n) {
  const int ti = blockIdx.x, tj = blockIdx.y;
  if (tj > ti) return;
  const int b = blockIdx.z;
  c128* Wb = W + static_cast<long>(b) * sW;
  const c128* Vb = V + static_cast<long>((b + voff) % vmod) * sV;
  __shared__ c128 T[32][33];
  const int r0 = ti * 32, c0 = tj * 32;
  for (int e = threadIdx.x; e < 32 * 32; e += 256) {
    const int r = e / 32, c = e % 32;
    const int gr = r0 + r, gc = c0 + c;
    if (gr < n && gc < n) {
      c128 w;
      if (gr >= gc) {
        w = Wb[static_cast<long>(gr) * n + gc];
      } else {  // only in a diagonal tile: mirror of the lower element
        const c128 m = Wb[static_cast<long>(gc) * n + gr];
        w = c128{m.x, -m.y};
      }
      T[r][c] = w;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 32 * 32; e += 256) {
    const int r = e / 32, c = e % 32;
    const int gr = r0 + r, gc = c0 + c;
    if (gr < n && gc < n) {
      const long o = static_cast<long>(gr) * n + gc;
      const c128 v = Vb[o];
      Wb[o] = c128{T[r][c].x - v.x, T[r][c].y - v.y};
    }
    if (ti != tj) {  // upper tile (tj, ti): element (c0 + r, r0 + c) = conj(T[c][r])
      const int ur = c0 + r, uc = r0 + c;
      if (ur < n && uc < n) {
        const long o = static_cast<long>(ur) * n + uc;
        const c128 v = Vb[o];
        Wb[o] = c128{T[c][r].x - v.x, -T[c][r].y - v.y};
      }
    }
  }
}

const Cplx kOne{1.0, 0.0};
const Cplx kMinusOne{-1.0, 0.0};

int split_point(int n, int unit) {
  int h = (n / 2 / unit) * unit;
  return h < unit ? unit : h;
}

void potrf_rec(hipStream_t s, c128* A, long sA, int lda, int j0, int n, int* info, int batch) {
  if (n <= CB) {
    hipLaunchKernelGGL(potrf_leaf_kernel, dim3(batch), dim3(CB * CB), 0, s, A, sA, lda, j0, n, info);
    THCGW_HIP(hipGetLastError());
    return;
  }
  const int n1 = split_point(n, CB), n2 = n - n1;
  potrf_rec(s, A, sA, lda, j0, n1, info, batch);
  c128* A11 = A + static_cast<long>(j0) * lda + j0;
  c128* A21 = A11 + static_cast<long>(n1) * lda;
  c128* A22 = A21 + n1;
  trsm_right_lower_conj(s, A11, sA, lda, A21, sA, lda, n2, n1, batch);
  zgemm(s, OP_N, OP_C, n2, n2, n1, kMinusOne, bp(A21, sA), lda, bp(A21, sA), lda, kOne, A22, sA, lda, batch,
        Cplx{0.0, 0.0}, GEMM_TRI_LOWER);
  potrf_rec(s, A, sA, lda, j0 + n1, n2, info, batch);
}

}  // namespace

// X L^H = T, L lower non-unit nb x nb, T m x nb; T is overwritten with X.
void trsm_right_lower_conj(hipStream_t s, const c128* L, long sL, int ldl, c128* T, long sT, int ldt, int m, int nb,
                           int batch) {
  if (m <= 0 || nb <= 0 || batch <= 0) return;
  if (nb <= CB) {
    hipLaunchKernelGGL((trsm_rlc_leaf_kernel<c128, c128, c128>), dim3((m + 255) / 256, batch), dim3(256), 0, s, L, sL,
                       ldl, T, sT, ldt, m, nb);
    THCGW_HIP(hipGetLastError());
    return;
  }
  const int n1 = split_point(nb, CB), n2 = nb - n1;
  trsm_right_lower_conj(s, L, sL, ldl, T, sT, ldt, m, n1, batch);
  // T2 -= X1 (L21)^H
  zgemm(s, OP_N, OP_C, m, n2, n1, kMinusOne, bp(T, sT), ldt, bp(L + static_cast<long>(n1) * ldl, sL), ldl, kOne,
        T + n1, sT, ldt, batch);
  trsm_right_lower_conj(s, L + static_cast<long>(n1) * ldl + n1, sL, ldl, T + n1, sT, ldt, m, n2, batch);
}

// The same in FP32 arithmetic with T stored as complex<float> and L as complex<double> or
// complex<float>; gemm_flags selects the product of the updates (0: FP32 MFMA,
// GEMM_SPLIT_BF16).  The leaves are 64 wide (64 complex<float> per thread), so that no
// update product is narrower than the 64 x 64 GEMM tile.
constexpr int CB32 = 64;

template <typename TL>
void trsm_rlc_f32(hipStream_t s, const TL* L, long sL, int ldl, c64* T, long sT, int ldt, int m, int nb, int batch,
                  int gemm_flags) {
  if (m <= 0 || nb <= 0 || batch <= 0) return;
  if (nb <= CB32) {
    hipLaunchKernelGGL((trsm_rlc_leaf_kernel<c64, c64, TL, CB32>), dim3((m + 255) / 256, batch), dim3(256), 0, s, L,
                       sL, ldl, T, sT, ldt, m, nb);
    THCGW_HIP(hipGetLastError());
    return;
  }
  const int n1 = split_point(nb, CB32), n2 = nb - n1;
  trsm_rlc_f32(s, L, sL, ldl, T, sT, ldt, m, n1, batch, gemm_flags);
  cgemm(s, OP_N, OP_C, m, n2, n1, kMinusOne, bp(T, sT), DT_C64, ldt, bp(L + static_cast<long>(n1) * ldl, sL),
        sizeof(TL) == sizeof(c128) ? DT_C128 : DT_C64, ldl, kOne, T + n1, DT_C64, sT, ldt, batch, Cplx{0.0, 0.0},
        gemm_flags);
  trsm_rlc_f32(s, L + static_cast<long>(n1) * ldl + n1, sL, ldl, T + n1, sT, ldt, m, n2, batch, gemm_flags);
}

void trsm_right_lower_conj_f32(hipStream_t s, const c128* L, long sL, int ldl, c64* T, long sT, int ldt, int m,
                               int nb, int batch, int gemm_flags) {
  trsm_rlc_f32(s, L, sL, ldl, T, sT, ldt, m, nb, batch, gemm_flags);
}

void trsm_right_lower_conj_f32(hipStream_t s, const c64* L, long sL, int ldl, c64* T, long sT, int ldt, int m,
                               int nb, int batch, int gemm_flags) {
  trsm_rlc_f32(s, L, sL, ldl, T, sT, ldt, m, nb, batch, gemm_flags);
}

// In-place lower Cholesky A = L L^H of `batch` Hermitian n x n matrices (lower triangle read).
// info (batch ints) is cleared, then set to the first failing pivot + 1 where A is not
// positive definite.
void potrf_batched(hipStream_t s, c128* A, long sA, int lda, int n, int* info, int batch) {
  THCGW_REQUIRE(n >= 1, "potrf: order must be positive");
  if (batch <= 0) return;
  hipLaunchKernelGGL(fill_int0_kernel, dim3((batch + 255) / 256), dim3(256), 0, s, info, batch);
  THCGW_HIP(hipGetLastError());
  potrf_rec(s, A, sA, lda, 0, n, info, batch);
}

void zero_upper(hipStream_t s, c128* A, long sA, int n, int batch) {
  if (batch <= 0 || n <= 0) return;
  const long nn = static_cast<long>(n) * n;
  hipLaunchKernelGGL(zero_upper_kernel, dim3(static_cast<unsigned>((nn + 255) / 256), batch), dim3(256), 0, s, A, sA,
                     n);
  THCGW_HIP(hipGetLastError());
}

void chol_logdet(hipStream_t s, const c128* L, long sL, int n, double* ldet, int batch) {
  if (batch <= 0) return;
  hipLaunchKernelGGL(chol_logdet_kernel, dim3(batch), dim3(256), 0, s, L, sL, n, ldet);
  THCGW_HIP(hipGetLastError());
}

void herm_finish(hipStream_t s, c128* W, long sW, const c128* V, long sV, int vmod, int voff, int n, int batch) {
  if (batch <= 0 || n <= 0) return;
  const int nt = (n + 31) / 32;
  hipLaunchKernelGGL(herm_finish_kernel, dim3(nt, nt, batch), dim3(256), 0, s, W, sW, V, sV, vmod, voff, n);
  THCGW_HIP(hipGetLastError());
}

}  // namespace thcgw
