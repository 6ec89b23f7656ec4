// Batched dense LU with partial pivoting and triangular solves, complex<double>.
//
// The reference inverts (I - V Pi) with getrf + getri per (q, iW_n)
// (scr_coulomb_t.cpp:188 -> slate_ops.hpp:324-333) and takes det(I - Pi V) from the
// LU factors (thc_rpa.icc:151, slate_ops.hpp:397-403).  nda hands its C-order
// matrices to column-major LAPACK, i.e. LAPACK factors the transpose.  The same is
// done here: a row-major n x n matrix A is viewed as the column-major M = A^T and
// factored as M = P L U, which in row-major terms is
//
//     A = Lt Ut Pt,   Lt = U^T (lower, non-unit),  Ut = L^T (upper, unit diagonal),
//
// with Pt a sequence of column swaps.  Pivot rows of M are contiguous rows of A, so
// the pivot search and the rank-1 updates of the panel read contiguous memory.
//
// The factorisation is recursive (split the column panel in two, factor the left
// half, update the right half with a triangular solve and a GEMM, factor it, swap
// back) so that most flops run in the MFMA GEMM kernel with large K.  The leaves
// are 16-wide panels factored in registers by one 1024-thread workgroup per matrix.
#include "thcgw_internal.hpp"

#include <climits>

namespace thcgw {
namespace {

__device__ __forceinline__ c128 cmul(c128 a, c128 b) {
  return c128{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ c128 csub(c128 a, c128 b) { return c128{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ c128 crecip(c128 d) {
  const double s = 1.0 / (d.x * d.x + d.y * d.y);
  return c128{d.x * s, -d.y * s};
}

constexpr int PT = 1024;  // panel kernel threads

// Factor the M-column panel [j0, j0+w) over M-rows [j0, n), i.e. the A-rows [j0, j0+w)
// restricted to A-columns [j0, n).  Thread t owns A-columns j0 + t + s*PT, s < R, and
// holds their W panel entries in registers.  Column swaps are applied inside the
// panel rows only; the caller applies them to the other rows.
template <int R, int W>
__global__ __launch_bounds__(PT) void getf2_panel_kernel(c128* A, long sA, int lda, int n, int j0, int w,
                                                         int* ipiv, int* info) {
  const int b = blockIdx.x;
  c128* Ab = A + static_cast<long>(b) * sA;
  int* piv = ipiv + static_cast<long>(b) * n;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;

  // Per-wave pivot candidates, double buffered by column parity.  A column's candidates are
  // written before its first barrier and read between its two barriers; the next write to
  // the same buffer comes two columns later, after a barrier that every reader has passed.
  __shared__ double red_val[2][PT / 64];
  __shared__ int red_idx[2][PT / 64];
  __shared__ c128 pcol[W], jcol[W];

  c128 v[R][W];
#pragma unroll
  for (int s = 0; s < R; ++s) {
    const int r = j0 + t + s * PT;
#pragma unroll
    for (int c = 0; c < W; ++c)
      v[s][c] = (r < n && c < w) ? Ab[static_cast<long>(j0 + c) * lda + r] : c128{0.0, 0.0};
  }

  int first_zero = 0;
#pragma unroll
  for (int c = 0; c < W; ++c) {
    if (c < w) {
      const int jc = j0 + c;
      // 1. pivot: largest |Re|+|Im| in M-column jc over M-rows >= jc, first index on ties.
      //    A thread's candidate index starts at its first row >= jc, so a column whose
      //    entries are all NaN (no comparison succeeds) still pivots on row jc, as LAPACK's
      //    izamax does; threads with no row >= jc keep INT_MAX and lose every reduction.
      double best = -1.0;
      int bidx = INT_MAX;
#pragma unroll
      for (int s = 0; s < R; ++s) {
        const int r = j0 + t + s * PT;
        if (r >= jc && r < n) {
          const double val = fabs(v[s][c].x) + fabs(v[s][c].y);
          if (bidx == INT_MAX) bidx = r;
          if (val > best) {
            best = val;
            bidx = r;
          }
        }
      }
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) {
        const double ov = __shfl_xor(best, off);
        const int oi = __shfl_xor(bidx, off);
        if (ov > best || (ov == best && oi < bidx)) {
          best = ov;
          bidx = oi;
        }
      }
      const int buf = c & 1;
      if (lane == 0) {
        red_val[buf][wave] = best;
        red_idx[buf][wave] = bidx;
      }
      __syncthreads();
      // every thread reduces the per-wave candidates itself (broadcast LDS reads, the same
      // order and tie rule everywhere), so no second barrier hands the pivot around
      double bv = red_val[buf][0];
      int p = red_idx[buf][0];
      for (int q = 1; q < PT / 64; ++q) {
        const double qv = red_val[buf][q];
        const int qi = red_idx[buf][q];
        if (qv > bv || (qv == bv && qi < p)) {
          bv = qv;
          p = qi;
        }
      }
      if (t == 0) piv[jc] = p;
      // 2. exchange the panel entries of columns p and jc
      const int op = (p - j0) % PT, sp_ = (p - j0) / PT;
      const int oj = (jc - j0) % PT, sj = (jc - j0) / PT;
#pragma unroll
      for (int s = 0; s < R; ++s) {
        if (t == op && s == sp_)
#pragma unroll
          for (int cc = 0; cc < W; ++cc) pcol[cc] = v[s][cc];
        if (t == oj && s == sj)
#pragma unroll
          for (int cc = 0; cc < W; ++cc) jcol[cc] = v[s][cc];
      }
      __syncthreads();
      if (p != jc) {
#pragma unroll
        for (int s = 0; s < R; ++s) {
          if (t == oj && s == sj)
#pragma unroll
            for (int cc = 0; cc < W; ++cc) v[s][cc] = pcol[cc];
          if (t == op && s == sp_)
#pragma unroll
            for (int cc = 0; cc < W; ++cc) v[s][cc] = jcol[cc];
        }
      }
      // 3. scale below the pivot and update the rest of the panel.  A zero or non-finite
      //    pivot is reported in info (first such column + 1), as a singular matrix.
      const c128 d = pcol[c];
      if (!(isfinite(d.x) && isfinite(d.y)) && first_zero == 0) first_zero = jc + 1;
      if (d.x != 0.0 || d.y != 0.0) {
        const c128 inv = crecip(d);
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const int r = j0 + t + s * PT;
          if (r > jc && r < n) {
            const c128 l = cmul(v[s][c], inv);
            v[s][c] = l;
#pragma unroll
            for (int cc = c + 1; cc < W; ++cc)
              if (cc < w) v[s][cc] = csub(v[s][cc], cmul(l, pcol[cc]));
          }
        }
      } else if (first_zero == 0) {
        first_zero = jc + 1;
      }
      // No barrier closes the column: the next one writes pcol/jcol only after its first
      // barrier, which every thread reaches after finishing this column's reads.
    }
  }

#pragma unroll
  for (int s = 0; s < R; ++s) {
    const int r = j0 + t + s * PT;
    if (r < n)
#pragma unroll
      for (int c = 0; c < W; ++c)
        if (c < w) Ab[static_cast<long>(j0 + c) * lda + r] = v[s][c];
  }
  if (t == 0 && first_zero != 0 && info[b] == 0) info[b] = first_zero;
}

__global__ void fill_int_kernel(int* p, int n, int v) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// Apply the column swaps ipiv[k0..k1) (in order) to A-rows [r0, r1).
__global__ __launch_bounds__(256) void laswp_kernel(c128* A, long sA, int lda, int n, int r0, int r1, int k0, int k1,
                                                    const int* ipiv) {
  const int b = blockIdx.y;
  const int row = r0 + blockIdx.x * 256 + threadIdx.x;
  if (row >= r1) return;
  c128* Ar = A + static_cast<long>(b) * sA + static_cast<long>(row) * lda;
  const int* piv = ipiv + static_cast<long>(b) * n;
  for (int k = k0; k < k1; ++k) {
    const int p = piv[k];
    if (p != k && p >= 0 && p < n) {  // the panel writes pivots in [k, n); the range check guards the read
      const c128 x = Ar[k];
      Ar[k] = Ar[p];
      Ar[p] = x;
    }
  }
}

constexpr int LB = 32;  // triangular-solve leaf size

// X U = T for the unit upper nb x nb block U; thread per row of T (m rows).
__global__ __launch_bounds__(256) void trsm_ruu_leaf(const c128* U, long sU, int ldu, c128* T, long sT, int ldt,
                                                     int m, int nb) {
  __shared__ c128 Us[LB][LB];
  const int b = blockIdx.y;
  const c128* Ub = U + static_cast<long>(b) * sU;
  for (int e = threadIdx.x; e < LB * LB; e += 256) {
    const int i = e / LB, j = e % LB;
    Us[i][j] = (i < nb && j < nb) ? Ub[static_cast<long>(i) * ldu + j] : c128{0.0, 0.0};
  }
  __syncthreads();
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= m) return;
  c128* Tr = T + static_cast<long>(b) * sT + static_cast<long>(row) * ldt;
  c128 x[LB];
#pragma unroll
  for (int c = 0; c < LB; ++c) x[c] = (c < nb) ? Tr[c] : c128{0.0, 0.0};
#pragma unroll
  for (int c = 1; c < LB; ++c) {
    if (c < nb) {
      c128 s = x[c];
#pragma unroll
      for (int cp = 0; cp < c; ++cp) s = csub(s, cmul(x[cp], Us[cp][c]));
      x[c] = s;
    }
  }
#pragma unroll
  for (int c = 0; c < LB; ++c)
    if (c < nb) Tr[c] = x[c];
}

// L X = B, L lower non-unit (LOWER = true) or U X = B, U upper unit (LOWER = false);
// nb x nb triangle, thread per right-hand-side column (ncol columns).
template <bool LOWER>
__global__ __launch_bounds__(256) void trsm_left_leaf(const c128* Tm, long sTm, int ldm, c128* B, long sB, int ldb,
                                                      int nb, int ncol) {
  __shared__ c128 Ts[LB][LB];
  const int b = blockIdx.y;
  const c128* Tb = Tm + static_cast<long>(b) * sTm;
  for (int e = threadIdx.x; e < LB * LB; e += 256) {
    const int i = e / LB, j = e % LB;
    Ts[i][j] = (i < nb && j < nb) ? Tb[static_cast<long>(i) * ldm + j] : c128{0.0, 0.0};
  }
  __syncthreads();
  const int col = blockIdx.x * 256 + threadIdx.x;
  if (col >= ncol) return;
  c128* Bc = B + static_cast<long>(b) * sB + col;
  c128 x[LB];
#pragma unroll
  for (int r = 0; r < LB; ++r) x[r] = (r < nb) ? Bc[static_cast<long>(r) * ldb] : c128{0.0, 0.0};
  if (LOWER) {
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      if (r < nb) {
        c128 s = x[r];
#pragma unroll
        for (int c = 0; c < r; ++c) s = csub(s, cmul(Ts[r][c], x[c]));
        x[r] = cmul(s, crecip(Ts[r][r]));
      }
    }
  } else {
#pragma unroll
    for (int r = LB - 1; r >= 0; --r) {
      if (r < nb) {
        c128 s = x[r];
#pragma unroll
        for (int c = r + 1; c < LB; ++c)
          if (c < nb) s = csub(s, cmul(Ts[r][c], x[c]));
        x[r] = s;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < LB; ++r)
    if (r < nb) Bc[static_cast<long>(r) * ldb] = x[r];
}

const Cplx kOne{1.0, 0.0};
const Cplx kMinusOne{-1.0, 0.0};

int split_point(int n, int unit) {
  int h = (n / 2 / unit) * unit;
  return h < unit ? unit : h;
}

// Above order 128 the recursions split at multiples of the 64-wide GEMM tile, so that no
// update product has a half-filled tile row or column at the split.
int tile_split(int n, int unit) { return split_point(n, n > 128 ? 64 : unit); }

}  // namespace

// ---- triangular solves (batched, strides between matrices) ----------------------

// X U = T, U unit upper (nb x nb), T m x nb; T is overwritten with X.
void trsm_right_upper_unit(hipStream_t s, const c128* U, long sU, int ldu, c128* T, long sT, int ldt, int m, int nb,
                           int batch) {
  if (m <= 0 || nb <= 0 || batch <= 0) return;
  if (nb <= LB) {
    hipLaunchKernelGGL(trsm_ruu_leaf, dim3((m + 255) / 256, batch), dim3(256), 0, s, U, sU, ldu, T, sT, ldt, m, nb);
    THCGW_HIP(hipGetLastError());
    return;
  }
  const int n1 = tile_split(nb, LB), n2 = nb - n1;
  trsm_right_upper_unit(s, U, sU, ldu, T, sT, ldt, m, n1, batch);
  // T2 -= X1 U12
  zgemm(s, OP_N, OP_N, m, n2, n1, kMinusOne, bp(T, sT), ldt, bp(U + n1, sU), ldu, kOne, T + n1, sT, ldt, batch);
  trsm_right_upper_unit(s, U + static_cast<long>(n1) * ldu + n1, sU, ldu, T + n1, sT, ldt, m, n2, batch);
}

// L X = B, L lower non-unit n x n, B n x ncol.
void trsm_left_lower(hipStream_t s, const c128* L, long sL, int ldl, c128* B, long sB, int ldb, int n, int ncol,
                     int batch) {
  if (n <= 0 || ncol <= 0 || batch <= 0) return;
  if (n <= LB) {
    hipLaunchKernelGGL((trsm_left_leaf<true>), dim3((ncol + 255) / 256, batch), dim3(256), 0, s, L, sL, ldl, B, sB,
                       ldb, n, ncol);
    THCGW_HIP(hipGetLastError());
    return;
  }
  const int n1 = tile_split(n, LB), n2 = n - n1;
  trsm_left_lower(s, L, sL, ldl, B, sB, ldb, n1, ncol, batch);
  zgemm(s, OP_N, OP_N, n2, ncol, n1, kMinusOne, bp(L + static_cast<long>(n1) * ldl, sL), ldl, bp(B, sB), ldb, kOne,
        B + static_cast<long>(n1) * ldb, sB, ldb, batch);
  trsm_left_lower(s, L + static_cast<long>(n1) * ldl + n1, sL, ldl, B + static_cast<long>(n1) * ldb, sB, ldb, n2,
                  ncol, batch);
}

// U X = B, U upper unit n x n, B n x ncol.
void trsm_left_upper_unit(hipStream_t s, const c128* U, long sU, int ldu, c128* B, long sB, int ldb, int n, int ncol,
                          int batch) {
  if (n <= 0 || ncol <= 0 || batch <= 0) return;
  if (n <= LB) {
    hipLaunchKernelGGL((trsm_left_leaf<false>), dim3((ncol + 255) / 256, batch), dim3(256), 0, s, U, sU, ldu, B, sB,
                       ldb, n, ncol);
    THCGW_HIP(hipGetLastError());
    return;
  }
  const int n1 = tile_split(n, LB), n2 = n - n1;
  trsm_left_upper_unit(s, U + static_cast<long>(n1) * ldu + n1, sU, ldu, B + static_cast<long>(n1) * ldb, sB, ldb, n2,
                       ncol, batch);
  zgemm(s, OP_N, OP_N, n1, ncol, n2, kMinusOne, bp(U + n1, sU), ldu, bp(B + static_cast<long>(n1) * ldb, sB), ldb,
        kOne, B, sB, ldb, batch);
  trsm_left_upper_unit(s, U, sU, ldu, B, sB, ldb, n1, ncol, batch);
}

// ---- LU ------------------------------------------------------------------------

namespace {
void laswp(hipStream_t s, c128* A, long sA, int lda, int n, int r0, int r1, int k0, int k1, const int* ipiv,
           int batch) {
  if (r1 <= r0 || k1 <= k0) return;
  hipLaunchKernelGGL(laswp_kernel, dim3((r1 - r0 + 255) / 256, batch), dim3(256), 0, s, A, sA, lda, n, r0, r1, k0,
                     k1, ipiv);
  THCGW_HIP(hipGetLastError());
}

void panel(hipStream_t s, c128* A, long sA, int lda, int n, int j0, int w, int* ipiv, int* info, int batch) {
  const int cols = n - j0;
  if (cols <= PT) {
    hipLaunchKernelGGL((getf2_panel_kernel<1, 16>), dim3(batch), dim3(PT), 0, s, A, sA, lda, n, j0, w, ipiv, info);
  } else {
    THCGW_REQUIRE(cols <= 2 * PT, "getrf: matrix order above 2048 is not supported");
    hipLaunchKernelGGL((getf2_panel_kernel<2, 8>), dim3(batch), dim3(PT), 0, s, A, sA, lda, n, j0, w, ipiv, info);
  }
  THCGW_HIP(hipGetLastError());
}

void getrf_rec(hipStream_t s, c128* A, long sA, int lda, int n, int j0, int w, int* ipiv, int* info, int batch) {
  const int pw = (n - j0 <= PT) ? 16 : 8;
  if (w <= pw) {
    panel(s, A, sA, lda, n, j0, w, ipiv, info, batch);
    return;
  }
  const int w1 = tile_split(w, pw), w2 = w - w1;
  getrf_rec(s, A, sA, lda, n, j0, w1, ipiv, info, batch);
  laswp(s, A, sA, lda, n, j0 + w1, j0 + w, j0, j0 + w1, ipiv, batch);
  c128* diag = A + static_cast<long>(j0) * lda + j0;
  c128* T = A + static_cast<long>(j0 + w1) * lda + j0;
  trsm_right_upper_unit(s, diag, sA, lda, T, sA, lda, w2, w1, batch);
  zgemm(s, OP_N, OP_N, w2, n - j0 - w1, w1, kMinusOne, bp(T, sA), lda, bp(diag + w1, sA), lda, kOne, T + w1, sA, lda,
        batch);
  getrf_rec(s, A, sA, lda, n, j0 + w1, w2, ipiv, info, batch);
  laswp(s, A, sA, lda, n, j0, j0 + w1, j0 + w1, j0 + w, ipiv, batch);
}
}  // namespace

// In-place LU of `batch` row-major n x n matrices (stride sA): A = Lt Ut Pt as described
// above.  ipiv (batch x n) receives the column swaps, info (batch) 0 or first zero pivot + 1.
void getrf_batched(hipStream_t s, c128* A, long sA, int lda, int n, int* ipiv, int* info, int batch) {
  THCGW_REQUIRE(n >= 1 && n <= 2 * PT, "getrf: order must be in [1, 2048]");
  hipLaunchKernelGGL(fill_int_kernel, dim3((batch + 255) / 256), dim3(256), 0, s, info, batch, 0);
  THCGW_HIP(hipGetLastError());
  getrf_rec(s, A, sA, lda, n, 0, n, ipiv, info, batch);
}

}  // namespace thcgw
