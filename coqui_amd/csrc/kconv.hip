// k-space convolutions of the THC polarizability and self-energy, the reference's
// alternative algorithms for symmetry-reduced grids, here without symmetry (isym = 0, no
// time reversal):
//
//   chi0:  Pi_PQ(tau, q)  = (sp / N_k) sum_k conj(Gp_PQ(k - q)) Gn_PQ(k)   (rpa_pi.icc:182-336)
//   Sigma: Sigma_PQ(tau,k) = -(1 / N_k) sum_q G_PQ(k - q) W_PQ(q)           (thc_gw.icc:342-443)
//
// Arrays are (N_k, M), M = N_mu^2, (P,Q) fastest.  A 256-thread workgroup owns 16 (P,Q)
// columns: it stages the two operands' N_k x 16 panels in LDS (one HBM read each), then
// every thread forms N_k/16 outputs of one column by a direct sum over N_k terms with the
// k - q index from a table built on the host.  Work is O(N_k^2) per column, against
// O(N_k log N_k) for the fused transforms, so these are reference-algorithm checks and
// symmetry-ready building blocks rather than the fast path.
#include "thcgw_internal.hpp"

namespace thcgw {
namespace {

constexpr int KC_COLS = 16;
constexpr int KC_T = 256;
constexpr int KC_MAXK = 64;

__device__ __forceinline__ c128 cfma(c128 acc, c128 a, c128 b) {
  return c128{acc.x + a.x * b.x - a.y * b.y, acc.y + a.x * b.y + a.y * b.x};
}

// MODE 0: out[q] = scale * sum_k conj(A[kmq[q][k]]) B[k]   (chi0)
// MODE 1: out[k] = scale * sum_q A[kmq[k][q]] B[q]         (Sigma; B rows at stride b_kstride)
template <int MODE>
__global__ __launch_bounds__(KC_T) void kconv_kernel(const c128* __restrict__ A, const c128* __restrict__ B,
                                                      long b_kstride, const int* __restrict__ kmq, int nk,
                                                      c128* __restrict__ out, long M, double scale) {
  __shared__ c128 As[KC_MAXK][KC_COLS];
  __shared__ c128 Bs[KC_MAXK][KC_COLS];
  __shared__ int T[KC_MAXK * KC_MAXK];
  const long col0 = static_cast<long>(blockIdx.x) * KC_COLS;
  for (int e = threadIdx.x; e < nk * KC_COLS; e += KC_T) {
    const int k = e / KC_COLS, c = e % KC_COLS;
    const long col = col0 + c;
    As[k][c] = col < M ? A[static_cast<long>(k) * M + col] : c128{0.0, 0.0};
    Bs[k][c] = col < M ? B[static_cast<long>(k) * b_kstride + col] : c128{0.0, 0.0};
  }
  for (int e = threadIdx.x; e < nk * nk; e += KC_T) T[e] = kmq[e];
  __syncthreads();
  const int c = threadIdx.x % KC_COLS, g = threadIdx.x / KC_COLS;  // 16 output groups
  const long col = col0 + c;
  if (col >= M) return;
  for (int o = g; o < nk; o += KC_T / KC_COLS) {
    c128 acc{0.0, 0.0};
    for (int j = 0; j < nk; ++j) {
      const c128 a = As[T[o * nk + j]][c];
      acc = cfma(acc, MODE == 0 ? c128{a.x, -a.y} : a, Bs[j][c]);
    }
    out[static_cast<long>(o) * M + col] = c128{acc.x * scale, acc.y * scale};
  }
}

}  // namespace

// Table of k - q indices (bz_symmetry.hpp:517-524): kmq[a * nk + b] = index of k_a - k_b.
std::vector<int> kminus_table(const int* kg) {
  const int nk = kg[0] * kg[1] * kg[2];
  std::vector<int> t(static_cast<size_t>(nk) * nk);
  for (int a = 0; a < nk; ++a)
    for (int b = 0; b < nk; ++b) {
      const int ai = a / (kg[1] * kg[2]), aj = (a / kg[2]) % kg[1], al = a % kg[2];
      const int bi = b / (kg[1] * kg[2]), bj = (b / kg[2]) % kg[1], bl = b % kg[2];
      const int i = (ai - bi + kg[0]) % kg[0], j = (aj - bj + kg[1]) % kg[1], l = (al - bl + kg[2]) % kg[2];
      t[static_cast<size_t>(a) * nk + b] = (i * kg[1] + j) * kg[2] + l;
    }
  return t;
}

// The chi0 table is indexed [q][k] -> k - q, i.e. kmq[q * nk + k] = index(k_k - k_q): the
// caller passes the transpose of kminus_table for MODE 0.
void kconv_pi(hipStream_t s, const c128* Gp, const c128* Gn, const int* kmq_qk, int nk, c128* Pi, long M,
              double scale) {
  THCGW_REQUIRE(nk <= KC_MAXK, "k-space convolution: N_k above 64 is not supported");
  const unsigned grid = static_cast<unsigned>((M + KC_COLS - 1) / KC_COLS);
  hipLaunchKernelGGL(kconv_kernel<0>, dim3(grid), dim3(KC_T), 0, s, Gp, Gn, M, kmq_qk, nk, Pi, M, scale);
  THCGW_HIP(hipGetLastError());
}

void kconv_sigma(hipStream_t s, const c128* G, const c128* W, long w_kstride, const int* kmq_kq, int nk, c128* S,
                 long M, double scale) {
  THCGW_REQUIRE(nk <= KC_MAXK, "k-space convolution: N_k above 64 is not supported");
  const unsigned grid = static_cast<unsigned>((M + KC_COLS - 1) / KC_COLS);
  hipLaunchKernelGGL(kconv_kernel<1>, dim3(grid), dim3(KC_T), 0, s, G, W, w_kstride, kmq_kq, nk, S, M, scale);
  THCGW_HIP(hipGetLastError());
}

}  // namespace thcgw
