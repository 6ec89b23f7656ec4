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
#include "lri.hpp"
#include "thcgw_internal.hpp"

#include <algorithm>
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

// The leaf of the Cholesky-inverse recursion, order nb <= NB, one workgroup of 1024
// threads per matrix block with NB * NB / 1024 elements each: the diagonal block at
// [j0][j0] is factored in place (A = L L^H, lower) and the block of Z at [j0][j0] receives
// L^-H (upper triangular, zeros below the diagonal).  One LDS array holds the block: in and
// below the diagonal the factorisation's working values S, above it T = the unscaled rows
// of L^-1, transposed (T[r][c] at [c][r]; T[j][j] = 1).  Column j of S and row j of T are
// used unscaled, divided by the pivot d_j = S[j][j] on the fly, so that each column costs
// one barrier (no thread waits for another's square root or scaling):
//   i >= k > j:  S[i][k] -= S[i][j] conj(S[k][j]) / d_j        (right-looking Cholesky)
//   r > j >= c:  T[r][c] -= S[r][j] T[j][c] / d_j               (right-looking L^-1)
// Neither reads an element written in the same step.  At the end L[i][k] = S[i][k] / sqrt(d_k)
// (L[k][k] = sqrt(d_k)) and L^-1[r][c] = T[r][c] / sqrt(d_r).  info as for potrf_leaf_kernel
// (a non-positive or NaN pivot is recorded and replaced by 1).  Order 32: 20.7 us per call on
// 200 matrices, against 46.1 us for a leaf that scales column j and forms L^-1 by one
// thread's forward substitution per column (three barriers per column).  Order-64 leaves
// took 89 us (this scheme, four elements per thread, 66.5 KB of LDS) and 78 us (the block
// in registers, one published column per step): VALU issue, not LDS, bounds them, and
// either only matched the order-32 leaf together with the products of the level it
// replaces.
template <int NB>
__global__ __launch_bounds__(1024) void cholinv_leaf_kernel(c128* A, long sA, int lda, c128* Z, long sZ, int ldz,
                                                            int j0, int nb, int* info) {
  constexpr int EPT = NB * NB / 1024;
  static_assert(EPT >= 1 && NB * NB == EPT * 1024, "NB * NB must be a multiple of 1024");
  __shared__ c128 S[NB][NB + 1];
  const int b = blockIdx.x, t = threadIdx.x;
  c128* Ab = A + static_cast<long>(b) * sA + static_cast<long>(j0) * lda + j0;
  c128* Zb = Z + static_cast<long>(b) * sZ + static_cast<long>(j0) * ldz + j0;
  int ei[EPT], ek[EPT];
#pragma unroll
  for (int m = 0; m < EPT; ++m) {
    const int e = t + 1024 * m;
    ei[m] = e / NB;
    ek[m] = e % NB;
    const int i = ei[m], k = ek[m];
    S[i][k] = (i < nb && k <= i) ? Ab[static_cast<long>(i) * lda + k] : c128{0.0, 0.0};
  }
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    double d = S[j][j].x;
    if (!(d > 0.0)) {
      if (t == 0 && info[b] == 0) info[b] = j0 + j + 1;
      d = 1.0;
    }
    const double inv = 1.0 / d;
    // every read of this step before any write (no element read here is written here), so
    // that the LDS reads of the thread's elements are in flight together
    c128 p[EPT], q[EPT], o[EPT];
    bool upd[EPT];
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
      const int i = ei[m], k = ek[m];
      const bool in = i < nb && k < nb;
      const bool lo = in && k <= i && k > j;            // S[i][k], i >= k > j
      const bool hi = in && k > i && k > j && i <= j;   // T[r][c], r = k > j >= c = i
      upd[m] = lo || hi;
      p[m] = q[m] = o[m] = c128{0.0, 0.0};
      if (lo) {
        p[m] = S[i][j];
        const c128 v = S[k][j];
        q[m] = c128{v.x, -v.y};
      } else if (hi) {
        p[m] = S[k][j];
        q[m] = (i == j) ? c128{1.0, 0.0} : S[i][j];
      }
      if (upd[m]) o[m] = S[i][k];
    }
#pragma unroll
    for (int m = 0; m < EPT; ++m)
      if (upd[m]) {
        const c128 u = cmul(p[m], q[m]);
        S[ei[m]][ek[m]] = c128{o[m].x - u.x * inv, o[m].y - u.y * inv};
      }
    __syncthreads();
  }
#pragma unroll
  for (int m = 0; m < EPT; ++m) {
    const int i = ei[m], k = ek[m];
    if (i >= nb || k >= nb) continue;
    // the pivots as the factorisation used them
    const double di = S[i][i].x > 0.0 ? S[i][i].x : 1.0;
    const double dk = S[k][k].x > 0.0 ? S[k][k].x : 1.0;
    if (k <= i) {
      const double sk = sqrt(dk);
      Ab[static_cast<long>(i) * lda + k] = (k == i) ? c128{sk, 0.0} : c128{S[i][k].x / sk, S[i][k].y / sk};
    }
    // Z[i][k] = conj(L^-1[k][i]) on and above the diagonal, 0 below
    c128 z{0.0, 0.0};
    if (k == i) z = c128{1.0 / sqrt(di), 0.0};
    else if (k > i) {
      const double r = 1.0 / sqrt(dk);
      z = c128{S[i][k].x * r, -S[i][k].y * r};
    }
    Zb[static_cast<long>(i) * ldz + k] = z;
  }
}

// Zero a rows x cols block (leading dimension ld) of each matrix.
__global__ __launch_bounds__(256) void zero_block_kernel(c128* A, long sA, int ld, int rows, int cols) {
  const long e = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= static_cast<long>(rows) * cols) return;
  const int r = static_cast<int>(e / cols), c = static_cast<int>(e % cols);
  A[static_cast<long>(blockIdx.y) * sA + static_cast<long>(r) * ld + c] = c128{0.0, 0.0};
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

// ldet[b] = 2 sum_j ln Re L[j][j] (+ add[(b + aoff) % amod] when add is given)
__global__ __launch_bounds__(256) void chol_logdet_kernel(const c128* L, long sL, int n, double* ldet,
                                                          const double* add, int amod, int aoff) {
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
  if (threadIdx.x == 0)
    ldet[blockIdx.x] = 2.0 * red[0] + (add ? add[(static_cast<int>(blockIdx.x) + aoff) % amod] : 0.0);
}

// One pass over the lower triangles of P[b], V[m] and Vi[m] (m = (b + voff) % vmod): the
// lower triangle of A[b] = Vi[m] - P[b], and, when part is given, per workgroup the partial
// sum of Tr(V P) for Hermitian V and P, sum_i V_ii P_ii + 2 sum_{i>j} Re(conj(V_ij) P_ij)
// (real), at part[b * gridDim.x + blockIdx.x].  Each workgroup covers RP_ELEMS consecutive
// elements of the matrices b = j, j + vmod, j + 2 vmod, ... (j = blockIdx.y) that share
// V[m] and Vi[m], RP_GROUP of them per pass: the V and Vi elements are read once per pass
// instead of once per matrix.  The form reads only these lower triangles.
constexpr int RP_ELEMS = 4096;
constexpr int RP_GROUP = 8;
__global__ __launch_bounds__(256) void resolvent_prep_kernel(const c128* V, const c128* Vi, long sV, int vmod,
                                                             int voff, const c128* P, c128* A, long sA, int n,
                                                             c128* part, int batch) {
  __shared__ double red[RP_GROUP][256];
  const int j = blockIdx.y;
  const long nn = static_cast<long>(n) * n;
  const long m = static_cast<long>((j + voff) % vmod) * sV;
  const long e0 = static_cast<long>(blockIdx.x) * RP_ELEMS;
  const long e1 = e0 + RP_ELEMS < nn ? e0 + RP_ELEMS : nn;
  for (int bg = j; bg < batch; bg += RP_GROUP * vmod) {
    int nb = 0;  // matrices in this pass: bg + k vmod for k < nb
    while (nb < RP_GROUP && bg + nb * vmod < batch) ++nb;
    double acc[RP_GROUP];
#pragma unroll
    for (int k = 0; k < RP_GROUP; ++k) acc[k] = 0.0;
    const long first = e0 + threadIdx.x;
    int r = static_cast<int>(first / n), c = static_cast<int>(first % n);
    for (long e = first; e < e1; e += 256) {
      if (c <= r) {
        const c128 w = Vi[m + e];
        const c128 v = part ? V[m + e] : c128{0.0, 0.0};
        const double f = (c == r) ? 1.0 : 2.0;
#pragma unroll
        for (int k = 0; k < RP_GROUP; ++k) {
          if (k < nb) {
            const long off = static_cast<long>(bg + k * vmod) * sA + e;
            const c128 pv = P[off];
            A[off] = c128{w.x - pv.x, w.y - pv.y};
            acc[k] += f * (v.x * pv.x + v.y * pv.y);  // Re(conj(v) p)
          }
        }
      }
      c += 256;  // next element of this thread: 256 further in row-major order
      while (c >= n) {
        c -= n;
        ++r;
      }
    }
    if (!part) continue;
#pragma unroll
    for (int k = 0; k < RP_GROUP; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (threadIdx.x < w) {
#pragma unroll
        for (int k = 0; k < RP_GROUP; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
      }
      __syncthreads();
    }
    if (threadIdx.x < nb)
      part[static_cast<long>(bg + threadIdx.x * vmod) * gridDim.x + blockIdx.x] = c128{red[threadIdx.x][0], 0.0};
    __syncthreads();  // red is reused by the next pass
  }
}

// tr[b] = sum of the np partial sums part[b * np ..]; one workgroup per matrix.
__global__ __launch_bounds__(256) void sum_parts_kernel(const c128* part, int np, c128* tr) {
  __shared__ double red[2][256];
  const int b = blockIdx.x;
  double ar = 0.0, ai = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) {
    const c128 v = part[static_cast<long>(b) * np + i];
    ar += v.x;
    ai += v.y;
  }
  red[0][threadIdx.x] = ar;
  red[1][threadIdx.x] = ai;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) tr[b] = c128{red[0][0], red[1][0]};
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

// Above order 128 the recursions split at multiples of the 64-wide GEMM tile, so that no
// product of the recursion has a half-filled tile row or column at the split.
int tile_split(int n, int unit) { return split_point(n, n > 128 ? 64 : unit); }

void potrf_rec(hipStream_t s, c128* A, long sA, int lda, int j0, int n, int* info, int batch) {
  if (n <= CB) {
    hipLaunchKernelGGL(potrf_leaf_kernel, dim3(batch), dim3(CB * CB), 0, s, A, sA, lda, j0, n, info);
    THCGW_HIP(hipGetLastError());
    return;
  }
  const int n1 = tile_split(n, CB), n2 = n - n1;
  potrf_rec(s, A, sA, lda, j0, n1, info, batch);
  c128* A11 = A + static_cast<long>(j0) * lda + j0;
  c128* A21 = A11 + static_cast<long>(n1) * lda;
  c128* A22 = A21 + n1;
  trsm_right_lower_conj(s, A11, sA, lda, A21, sA, lda, n2, n1, batch);
  zgemm(s, OP_N, OP_C, n2, n2, n1, kMinusOne, bp(A21, sA), lda, bp(A21, sA), lda, kOne, A22, sA, lda, batch,
        Cplx{0.0, 0.0}, GEMM_TRI_LOWER);
  potrf_rec(s, A, sA, lda, j0 + n1, n2, info, batch);
}

// Cholesky factorisation with the inverse of the factor, every off-diagonal step a GEMM:
// for the diagonal block [j0, j0 + n) split as n1 + n2,
//   (L11, Z11) of A11;  L21 = A21 L11^-H = A21 Z11 -> Z21 (scratch in the zero lower block of Z);
//   A22 -= L21 L21^H (lower tiles);  (L22, Z22) of A22;
//   T = L21^H Z22 -> A12 (the unused upper block of A);  Z12 = -Z11 T;  Z21 <- 0,
// which is Z = L^-H = [[Z11, -Z11 L21^H Z22], [0, Z22]].  n^3/3 complex multiply-adds, as
// for a factorisation plus a triangular inverse, without the triangular solves' narrow
// updates; the four products are N,N / N,C / C,N / N,N with triangular-K flags.  On exit A
// holds L on and below the diagonal except L21 (the diagonal blocks are exact; only their
// diagonal is read afterwards) and Z holds L^-H with zeros in the band of 128 below the
// diagonal (the only part of its lower triangle a K-bounded product with tiles up to 128
// wide reads, whatever the tiles' origin); further below it holds scratch.
void cholinv_rec(hipStream_t s, c128* A, long sA, int lda, c128* Z, long sZ, int ldz, int j0, int n, int* info,
                 int batch) {
  if (n <= CB) {
    hipLaunchKernelGGL(cholinv_leaf_kernel<CB>, dim3(batch), dim3(1024), 0, s, A, sA, lda, Z, sZ, ldz, j0, n, info);
    THCGW_HIP(hipGetLastError());
    return;
  }
  const int n1 = tile_split(n, CB), n2 = n - n1;
  cholinv_rec(s, A, sA, lda, Z, sZ, ldz, j0, n1, info, batch);
  c128* A11 = A + static_cast<long>(j0) * lda + j0;
  c128* A12 = A11 + n1;
  c128* A21 = A11 + static_cast<long>(n1) * lda;
  c128* A22 = A21 + n1;
  c128* Z11 = Z + static_cast<long>(j0) * ldz + j0;
  c128* Z12 = Z11 + n1;
  c128* Z21 = Z11 + static_cast<long>(n1) * ldz;
  c128* Z22 = Z21 + n1;
  const Cplx zero{0.0, 0.0};
  zgemm(s, OP_N, OP_N, n2, n1, n1, kOne, bp(A21, sA), lda, bp(Z11, sZ), ldz, zero, Z21, sZ, ldz, batch, zero,
        GEMM_TRI_KTO_COL);
  zgemm(s, OP_N, OP_C, n2, n2, n1, kMinusOne, bp(Z21, sZ), ldz, bp(Z21, sZ), ldz, kOne, A22, sA, lda, batch, zero,
        GEMM_TRI_LOWER);
  cholinv_rec(s, A, sA, lda, Z, sZ, ldz, j0 + n1, n2, info, batch);
  zgemm(s, OP_C, OP_N, n1, n2, n2, kOne, bp(Z21, sZ), ldz, bp(Z22, sZ), ldz, zero, A12, sA, lda, batch, zero,
        GEMM_TRI_KTO_COL);
  zgemm(s, OP_N, OP_N, n1, n2, n1, kMinusOne, bp(Z11, sZ), ldz, bp(A12, sA), lda, zero, Z12, sZ, ldz, batch, zero,
        GEMM_TRI_KFROM_ROW);
  // Z21 held a temporary and lies below the diagonal of Z.  The products that read Z later
  // (the K-bounded ones of this recursion's ancestors, each tiled from its own sub-block
  // origin, and Z Z^H with K from the row) read its lower part only inside a diagonal tile
  // of theirs, at most 128 wide: element (r, c) is read only if r - c < 128.  So the zeros
  // they need are the corner of Z21 within 128 rows and columns of the split point.
  const int rows = std::min(128, n2), cols = std::min(128, n1), c0 = n1 - cols;
  const long nb = static_cast<long>(rows) * cols;
  hipLaunchKernelGGL(zero_block_kernel, dim3(static_cast<unsigned>((nb + 255) / 256), batch), dim3(256), 0, s, Z21 + c0,
                     sZ, ldz, rows, cols);
  THCGW_HIP(hipGetLastError());
}

}  // namespace

// A = L L^H in place (lower; the strict upper triangle is overwritten with scratch) and
// Z = L^-H (upper; zeros in the band of 128 below the diagonal, scratch further below,
// which no K-bounded product reads) for `batch` Hermitian n x n matrices,
// both packed (leading dimension n).  info as for potrf_batched.
void cholinv_batched(hipStream_t s, c128* A, long sA, c128* Z, long sZ, int n, int* info, int batch) {
  THCGW_REQUIRE(n >= 1, "cholinv: order must be positive");
  if (batch <= 0) return;
  hipLaunchKernelGGL(fill_int0_kernel, dim3((batch + 255) / 256), dim3(256), 0, s, info, batch);
  THCGW_HIP(hipGetLastError());
  cholinv_rec(s, A, sA, n, Z, sZ, n, 0, n, info, batch);
}

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
  const int n1 = tile_split(nb, CB), n2 = nb - n1;
  trsm_right_lower_conj(s, L, sL, ldl, T, sT, ldt, m, n1, batch);
  // T2 -= X1 (L21)^H
  zgemm(s, OP_N, OP_C, m, n2, n1, kMinusOne, bp(T, sT), ldt, bp(L + static_cast<long>(n1) * ldl, sL), ldl, kOne,
        T + n1, sT, ldt, batch);
  trsm_right_lower_conj(s, L + static_cast<long>(n1) * ldl + n1, sL, ldl, T + n1, sT, ldt, m, n2, batch);
}

namespace {
// Column block c0 .. c0+nb-1 of X L^H = T for upper triangular T (T[r][c] = 0 for r > c):
// X is upper triangular as well, so only rows 0 .. c0+nb-1 of the block can be nonzero, and
// only those rows are solved and updated.  T and L point at column c0 and at L[c0][c0].
void trsm_rlc_utri(hipStream_t s, const c128* L, long sL, int ldl, c128* T, long sT, int ldt, int c0, int nb,
                   int batch) {
  const int m = c0 + nb;
  if (nb <= CB) {
    hipLaunchKernelGGL((trsm_rlc_leaf_kernel<c128, c128, c128>), dim3((m + 255) / 256, batch), dim3(256), 0, s, L, sL,
                       ldl, T, sT, ldt, m, nb);
    THCGW_HIP(hipGetLastError());
    return;
  }
  const int n1 = tile_split(nb, CB), n2 = nb - n1;
  trsm_rlc_utri(s, L, sL, ldl, T, sT, ldt, c0, n1, batch);
  // T2 -= X1 (L21)^H over the rows where X1 can be nonzero
  zgemm(s, OP_N, OP_C, c0 + n1, n2, n1, kMinusOne, bp(T, sT), ldt, bp(L + static_cast<long>(n1) * ldl, sL), ldl, kOne,
        T + n1, sT, ldt, batch);
  trsm_rlc_utri(s, L + static_cast<long>(n1) * ldl + n1, sL, ldl, T + n1, sT, ldt, c0 + n1, n2, batch);
}
}  // namespace

// Z = L^-H (upper triangular) for lower non-unit n x n L: the triangular solve of
// Z L^H = I restricted to the nonzero rows of each column block (n^3/6 complex
// multiply-adds instead of n^3/2 for a full right-hand side).  Z is overwritten.
void tri_inverse_conj(hipStream_t s, const c128* L, long sL, int ldl, c128* Z, long sZ, int ldz, int n, int batch) {
  if (n <= 0 || batch <= 0) return;
  THCGW_REQUIRE(ldz == n, "tri_inverse_conj: Z must be packed (ldz = n)");
  set_identity(s, Z, sZ, n, batch);
  trsm_rlc_utri(s, L, sL, ldl, Z, sZ, ldz, 0, n, batch);
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

void chol_logdet(hipStream_t s, const c128* L, long sL, int n, double* ldet, int batch, const double* add, int amod,
                 int aoff) {
  if (batch <= 0) return;
  hipLaunchKernelGGL(chol_logdet_kernel, dim3(batch), dim3(256), 0, s, L, sL, n, ldet, add, amod, aoff);
  THCGW_HIP(hipGetLastError());
}

int resolvent_prep_parts(int n) {
  return static_cast<int>((static_cast<long>(n) * n + RP_ELEMS - 1) / RP_ELEMS);
}

void resolvent_prep(hipStream_t s, const c128* V, const c128* Vi, long sV, int vmod, int voff, const c128* P, c128* A,
                    long sA, int n, c128* tr, c128* part, int batch) {
  if (batch <= 0 || n <= 0) return;
  THCGW_REQUIRE(!tr || part, "resolvent_prep: the trace needs the partial-sum buffer");
  THCGW_REQUIRE(batch <= 65535, "resolvent_prep: at most 65535 matrices per launch");
  const int np = resolvent_prep_parts(n);
  THCGW_REQUIRE(vmod >= 1 && voff >= 0, "resolvent_prep: bad V indexing");
  const int slots = batch < vmod ? batch : vmod;  // one workgroup row per V member in use
  hipLaunchKernelGGL(resolvent_prep_kernel, dim3(np, slots), dim3(256), 0, s, V, Vi, sV, vmod, voff, P, A, sA, n,
                     tr ? part : nullptr, batch);
  THCGW_HIP(hipGetLastError());
  if (tr) {
    hipLaunchKernelGGL(sum_parts_kernel, dim3(batch), dim3(256), 0, s, part, np, tr);
    THCGW_HIP(hipGetLastError());
  }
}

void herm_finish(hipStream_t s, c128* W, long sW, const c128* V, long sV, int vmod, int voff, int n, int batch) {
  if (batch <= 0 || n <= 0) return;
  const int nt = (n + 31) / 32;
  hipLaunchKernelGGL(herm_finish_kernel, dim3(nt, nt, batch), dim3(256), 0, s, W, sW, V, sV, vmod, voff, n);
  THCGW_HIP(hipGetLastError());
}

}  // namespace thcgw
