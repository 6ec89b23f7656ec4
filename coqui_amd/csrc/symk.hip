// Brillouin-zone symmetry: the kernels of chi0 and Sigma on the irreducible wedge.
//
//   gather_blocks: G(k) = op(G_ibz[kp_to_ibz(k)]), op = transpose where the k-point needs
//                  time reversal (thc_solver_comm.hpp:418-422), conjugated where cj[k]
//                  (a time-reversal point formed from its partner)
//   gather_rows:   row i of dst (at row dmap[i]) = src row sidx[i], conjugated where cj[i]:
//                  X in grid order with the partners' conjugates, the irreducible q of the
//                  R -> q transform (rpa_pi.icc:160-169), the X_{kS} of the back-projection
//   kconv_terms:   out[o] = scale sum_{t in terms(o)} opA(A[ia_t]) opB(B[ib_t]), the Hadamard
//                  sums of Sigma over the q of one symmetry operation (thc_gw.icc:366-395);
//                  each term packs ia | ib << 8 | conjA << 16 | conjB << 17; with herm_n > 0
//                  (Hermitian outputs of order herm_n) only the lower half L of out = L + L^H
//                  is written: columns above the diagonal are not formed (zeros only inside
//                  the diagonal 128 x 128 blocks, which a product with K from the tile column
//                  reads) and the diagonal is halved
//   sum_groups:    dst[e] = sum_g src[g per + e], Sigma(k) summed over the operations
//
// Arrays are (rows, M), M = N_mu^2, (P,Q) fastest.  kconv_terms stages the 16-column panels
// of A and B (at most 64 rows each) in LDS, one HBM read per element, and every thread forms
// the outputs of one column from its term list: O(N_k_ibz N_k) per column, HBM-bound on
// reading A, B and writing out.
#include "thcgw_internal.hpp"

#include <algorithm>

namespace thcgw {
namespace {

constexpr int ST = 256;
constexpr int TC_COLS = 16;
constexpr int TC_MAXR = 64;

__global__ __launch_bounds__(ST) void gather_blocks_kernel(const c128* __restrict__ src, const int* __restrict__ sidx,
                                                           const int* __restrict__ trans, const int* __restrict__ cj,
                                                           int n, int count, c128* __restrict__ dst) {
  const long per = static_cast<long>(n) * n;
  const long e = static_cast<long>(blockIdx.x) * ST + threadIdx.x;
  if (e >= per * count) return;
  const int b = static_cast<int>(e / per);
  const int i = static_cast<int>((e % per) / n), j = static_cast<int>(e % n);
  const c128* S = src + static_cast<long>(sidx[b]) * per;
  const c128 v = (trans && trans[b]) ? S[static_cast<long>(j) * n + i] : S[static_cast<long>(i) * n + j];
  dst[e] = (cj && cj[b]) ? c128{v.x, -v.y} : v;
}

__global__ __launch_bounds__(ST) void gather_rows_kernel(const c128* __restrict__ src, long sstride,
                                                         const int* __restrict__ sidx, const int* __restrict__ cj,
                                                         long len, c128* __restrict__ dst, long dstride,
                                                         const int* __restrict__ dmap) {
  const int i = blockIdx.y;
  const c128* S = src + static_cast<long>(sidx[i]) * sstride;
  c128* D = dst + static_cast<long>(dmap ? dmap[i] : i) * dstride;
  const bool c = cj && cj[i];
  for (long e = static_cast<long>(blockIdx.x) * ST + threadIdx.x; e < len; e += static_cast<long>(gridDim.x) * ST) {
    const c128 v = S[e];
    D[e] = c ? c128{v.x, -v.y} : v;
  }
}

__device__ __forceinline__ c128 cmul_acc(c128 acc, c128 a, c128 b) {
  return c128{acc.x + a.x * b.x - a.y * b.y, acc.y + a.x * b.y + a.y * b.x};
}

__global__ __launch_bounds__(ST) void kconv_terms_kernel(const c128* __restrict__ A, int nA, const c128* __restrict__ B,
                                                         long b_stride, int nB, const int* __restrict__ off,
                                                         const int* __restrict__ terms, int nout,
                                                         c128* __restrict__ out, long M, double scale, int herm_n) {
  __shared__ c128 As[TC_MAXR][TC_COLS];
  __shared__ c128 Bs[TC_MAXR][TC_COLS];
  const long col0 = static_cast<long>(blockIdx.x) * TC_COLS;
  for (int e = threadIdx.x; e < nA * TC_COLS; e += ST) {
    const int r = e / TC_COLS, c = e % TC_COLS;
    const long col = col0 + c;
    As[r][c] = col < M ? A[static_cast<long>(r) * M + col] : c128{0.0, 0.0};
  }
  for (int e = threadIdx.x; e < nB * TC_COLS; e += ST) {
    const int r = e / TC_COLS, c = e % TC_COLS;
    const long col = col0 + c;
    Bs[r][c] = col < M ? B[static_cast<long>(r) * b_stride + col] : c128{0.0, 0.0};
  }
  __syncthreads();
  const int c = threadIdx.x % TC_COLS, g = threadIdx.x / TC_COLS;
  const long col = col0 + c;
  if (col >= M) return;
  double sc = scale;
  if (herm_n > 0) {
    const int P = static_cast<int>(col / herm_n), Q = static_cast<int>(col % herm_n);
    if (Q > P) {
      if ((Q >> 7) == (P >> 7))
        for (int o = g; o < nout; o += ST / TC_COLS) out[static_cast<long>(o) * M + col] = c128{0.0, 0.0};
      return;
    }
    if (P == Q) sc *= 0.5;
  }
  for (int o = g; o < nout; o += ST / TC_COLS) {
    c128 acc{0.0, 0.0};
    const int t1 = off[o + 1];
    for (int t = off[o]; t < t1; ++t) {
      const int w = terms[t];
      c128 a = As[w & 0xff][c], b = Bs[(w >> 8) & 0xff][c];
      if (w & (1 << 16)) a.y = -a.y;
      if (w & (1 << 17)) b.y = -b.y;
      acc = cmul_acc(acc, a, b);
    }
    out[static_cast<long>(o) * M + col] = c128{acc.x * sc, acc.y * sc};
  }
}

__global__ __launch_bounds__(ST) void sum_groups_kernel(const c128* __restrict__ src, int ng, long per,
                                                        c128* __restrict__ dst) {
  const long e = static_cast<long>(blockIdx.x) * ST + threadIdx.x;
  if (e >= per) return;
  c128 acc{0.0, 0.0};
  for (int g = 0; g < ng; ++g) {
    const c128 v = src[static_cast<long>(g) * per + e];
    acc.x += v.x;
    acc.y += v.y;
  }
  dst[e] = acc;
}

unsigned blocks_for(long n) { return static_cast<unsigned>((n + ST - 1) / ST); }

}  // namespace

void gather_blocks(hipStream_t s, const c128* src, const int* sidx, const int* trans, const int* cj, int n, int count,
                   c128* dst) {
  if (count <= 0) return;
  hipLaunchKernelGGL(gather_blocks_kernel, dim3(blocks_for(static_cast<long>(n) * n * count)), dim3(ST), 0, s, src,
                     sidx, trans, cj, n, count, dst);
  THCGW_HIP(hipGetLastError());
}

void gather_rows(hipStream_t s, const c128* src, long sstride, const int* sidx, const int* cj, int count, long len,
                 c128* dst, long dstride, const int* dmap) {
  if (count <= 0) return;
  THCGW_REQUIRE(count < 65536, "gather_rows: too many rows");
  const unsigned gx = static_cast<unsigned>(std::min<long>((len + ST - 1) / ST, 2048));
  hipLaunchKernelGGL(gather_rows_kernel, dim3(gx, count), dim3(ST), 0, s, src, sstride, sidx, cj, len, dst, dstride,
                     dmap);
  THCGW_HIP(hipGetLastError());
}

void kconv_terms(hipStream_t s, const c128* A, int nA, const c128* B, long b_stride, int nB, const int* off,
                 const int* terms, int nout, c128* out, long M, double scale, int herm_n) {
  THCGW_REQUIRE(nA >= 1 && nA <= TC_MAXR && nB >= 1 && nB <= TC_MAXR,
                "symmetry-reduced convolution: at most 64 k-points and 64 irreducible q-points");
  if (nout <= 0) return;
  const unsigned grid = static_cast<unsigned>((M + TC_COLS - 1) / TC_COLS);
  THCGW_REQUIRE(herm_n == 0 || static_cast<long>(herm_n) * herm_n == M, "kconv_terms: herm_n^2 must equal M");
  hipLaunchKernelGGL(kconv_terms_kernel, dim3(grid), dim3(ST), 0, s, A, nA, B, b_stride, nB, off, terms, nout, out, M,
                     scale, herm_n);
  THCGW_HIP(hipGetLastError());
}

void sum_groups(hipStream_t s, const c128* src, int ng, long per, c128* dst) {
  hipLaunchKernelGGL(sum_groups_kernel, dim3(blocks_for(per)), dim3(ST), 0, s, src, ng, per, dst);
  THCGW_HIP(hipGetLastError());
}

}  // namespace thcgw
