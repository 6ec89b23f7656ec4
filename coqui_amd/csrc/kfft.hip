// Fused k-grid transforms for the THC polarizability and self-energy.
//
// The reference applies the k<->R transforms as dense N_k x N_k GEMMs over all
// (P,Q) columns (rpa_pi.icc:115-124,139-148,165-172; thc_gw.icc:137-155,171-191)
// with f_Rk(R,k) = e^{-ik.R}/N_k and f_kR(k,R) = e^{+ik.R} on a Gamma-centred
// n1 x n2 x n3 grid (kpoint_utils.hpp:461-545,588-680).  With lattv = I these are
// separable 3-D DFTs, so each (P,Q) column is transformed here by small in-register
// DFTs, the Hadamard product is taken in R space and the result is transformed back,
// all in one pass over HBM:
//
//   chi0:  Pi(q)    = sp * F+[ F-[Gp] o conj(F-[Gn]) ]        (rpa_pi.icc:150-172)
//   Sigma: Sigma(k) = F+[ -F-[G] o F-[W] ]                    (thc_gw.icc:158-182)
//
// F-[x](R) = (1/N_k) sum_k e^{-ik.R} x(k),  F+[y](q) = sum_R e^{+iq.R} y(R).
// Arrays are (N_k, M) with the (P,Q) index fastest, M = N_mu^2.  One thread owns one
// column; for N_k = 64 a quad of lanes shares a column, each lane holding one plane
// of the first k index, and the transform along that index uses DPP quad permutes.
#include "thcgw_internal.hpp"

namespace thcgw {
namespace {

template <typename RC>
using scal_t = decltype(RC{}.x);

template <typename RC>
__device__ __forceinline__ RC cadd(RC a, RC b) { return RC{a.x + b.x, a.y + b.y}; }
template <typename RC>
__device__ __forceinline__ RC csub(RC a, RC b) { return RC{a.x - b.x, a.y - b.y}; }
template <typename RC>
__device__ __forceinline__ RC cmul(RC a, RC b) {
  return RC{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename RC>
__device__ __forceinline__ RC cmulconj(RC a, RC b) {  // a * conj(b)
  return RC{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename RC>
__device__ __forceinline__ RC cscale(RC a, scal_t<RC> s) { return RC{a.x * s, a.y * s}; }
// multiply by SIGN * i
template <int SIGN, typename RC>
__device__ __forceinline__ RC mul_si(RC a) {
  return SIGN > 0 ? RC{-a.y, a.x} : RC{a.y, -a.x};
}
template <typename RC, typename T>
__device__ __forceinline__ RC cvt(T v) {
  return RC{static_cast<scal_t<RC>>(v.x), static_cast<scal_t<RC>>(v.y)};
}

// y_m = sum_j x_j w^{jm}, w = e^{SIGN 2 pi i / N}, on elements base + j*stride of x.
template <int N, int SIGN, int L, typename RC>
__device__ __forceinline__ void dft_line(RC (&x)[L], int base, int stride) {
  using S = scal_t<RC>;
  if constexpr (N == 2) {
    RC a = x[base], b = x[base + stride];
    x[base] = cadd(a, b);
    x[base + stride] = csub(a, b);
  } else if constexpr (N == 3) {
    const S h = static_cast<S>(0.86602540378443864676);  // sqrt(3)/2
    RC x0 = x[base], x1 = x[base + stride], x2 = x[base + 2 * stride];
    RC a = cadd(x1, x2), b = csub(x1, x2);
    RC m = RC{x0.x - S(0.5) * a.x, x0.y - S(0.5) * a.y};
    RC t = mul_si<SIGN>(cscale(b, h));
    x[base] = cadd(x0, a);
    x[base + stride] = cadd(m, t);
    x[base + 2 * stride] = csub(m, t);
  } else if constexpr (N == 4) {
    RC x0 = x[base], x1 = x[base + stride], x2 = x[base + 2 * stride], x3 = x[base + 3 * stride];
    RC a0 = cadd(x0, x2), a1 = csub(x0, x2), b0 = cadd(x1, x3), b1 = mul_si<SIGN>(csub(x1, x3));
    x[base] = cadd(a0, b0);
    x[base + stride] = cadd(a1, b1);
    x[base + 2 * stride] = csub(a0, b0);
    x[base + 3 * stride] = csub(a1, b1);
  }
}

// Value of the lane selected by the quad_perm control CTRL within each group of 4 lanes.
template <int CTRL>
__device__ __forceinline__ double dpp_s(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, false);
  hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_s(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL, typename RC>
__device__ __forceinline__ RC dpp_c(RC v) {
  return RC{dpp_s<CTRL>(v.x), dpp_s<CTRL>(v.y)};
}

// DFT along the first k index when it is split over S lanes (S == N1): lane h holds
// plane i = h.  Every lane forms all N1 outputs and keeps output plane h.
template <int N1, int SIGN, int L, typename RC>
__device__ __forceinline__ void dft_split(RC (&x)[L], int h) {
  static_assert(N1 == 2 || N1 == 4, "split DFT needs N1 in {2,4}");
#pragma unroll
  for (int e = 0; e < L; ++e) {
    RC v = x[e];
    if constexpr (N1 == 2) {
      RC v0 = dpp_c<0xA0>(v);  // quad_perm [0,0,2,2]
      RC v1 = dpp_c<0xF5>(v);  // quad_perm [1,1,3,3]
      x[e] = (h == 0) ? cadd(v0, v1) : csub(v0, v1);
    } else {
      RC v0 = dpp_c<0x00>(v), v1 = dpp_c<0x55>(v), v2 = dpp_c<0xAA>(v), v3 = dpp_c<0xFF>(v);
      RC a0 = cadd(v0, v2), a1 = csub(v0, v2), b0 = cadd(v1, v3), b1 = mul_si<SIGN>(csub(v1, v3));
      RC y;
      if (h == 0) y = cadd(a0, b0);
      else if (h == 1) y = cadd(a1, b1);
      else if (h == 2) y = csub(a0, b0);
      else y = csub(a1, b1);
      x[e] = y;
    }
  }
}

// Separable 3-D DFT of the thread's share: P = N1/S planes of N2 x N3 points,
// element index e = (p * N2 + j) * N3 + l.
template <int N1, int N2, int N3, int S, int SIGN, typename RC>
__device__ __forceinline__ void dft3(RC (&x)[(N1 / S) * N2 * N3], int h) {
  constexpr int P = N1 / S;
  constexpr int L = P * N2 * N3;
  if constexpr (N3 > 1) {
#pragma unroll
    for (int pj = 0; pj < P * N2; ++pj) dft_line<N3, SIGN, L>(x, pj * N3, 1);
  }
  if constexpr (N2 > 1) {
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int l = 0; l < N3; ++l) dft_line<N2, SIGN, L>(x, p * N2 * N3 + l, N3);
  }
  if constexpr (N1 > 1) {
    if constexpr (S == 1) {
#pragma unroll
      for (int jl = 0; jl < N2 * N3; ++jl) dft_line<N1, SIGN, L>(x, jl, N2 * N3);
    } else {
      dft_split<N1, SIGN, L>(x, h);
    }
  }
}

template <int N1, int N2, int N3, int S, typename RC, typename T>
__device__ __forceinline__ void load_col(RC (&x)[(N1 / S) * N2 * N3], const T* src, long kstride, long col,
                                         int h) {
  constexpr int P = N1 / S;
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int jl = 0; jl < N2 * N3; ++jl) {
      const int i = (S == 1) ? p : h;
      x[p * N2 * N3 + jl] = cvt<RC>(src[static_cast<long>(i * N2 * N3 + jl) * kstride + col]);
    }
}

template <int N1, int N2, int N3, int S, typename RC, typename T>
__device__ __forceinline__ void store_col(const RC (&x)[(N1 / S) * N2 * N3], T* dst, long kstride, long col,
                                          int h) {
  constexpr int P = N1 / S;
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int jl = 0; jl < N2 * N3; ++jl) {
      const int i = (S == 1) ? p : h;
      dst[static_cast<long>(i * N2 * N3 + jl) * kstride + col] = cvt<T>(x[p * N2 * N3 + jl]);
    }
}

constexpr int KT = 256;

// Pi(:, col) = scale * F+[ F-'[Gp] o conj(F-'[Gn]) ], F-' the unscaled forward DFT; the
// caller passes scale = sp / N_k^2.  Arithmetic in the scalar type of RC; TG and TO are
// the storage types of G and Pi.
template <int N1, int N2, int N3, int S, typename RC, typename TG, typename TO>
__global__ __launch_bounds__(KT) void pi_kfft_kernel(const TG* __restrict__ Gp, const TG* __restrict__ Gn,
                                                     TO* __restrict__ Pi, long M, double scale) {
  constexpr int L = (N1 / S) * N2 * N3;
  const long gid = static_cast<long>(blockIdx.x) * KT + threadIdx.x;
  const long col = gid / S;
  const int h = static_cast<int>(gid % S);
  if (col >= M) return;  // a column's S lanes are adjacent, so groups exit together
  RC a[L], b[L];
  load_col<N1, N2, N3, S>(a, Gp, M, col, h);
  load_col<N1, N2, N3, S>(b, Gn, M, col, h);
  dft3<N1, N2, N3, S, -1>(a, h);
  dft3<N1, N2, N3, S, -1>(b, h);
  const scal_t<RC> sc = static_cast<scal_t<RC>>(scale);
#pragma unroll
  for (int e = 0; e < L; ++e) a[e] = cscale(cmulconj(a[e], b[e]), sc);
  dft3<N1, N2, N3, S, +1>(a, h);
  store_col<N1, N2, N3, S>(a, Pi, M, col, h);
}

// S(:, col) = scale * F+[ F-'[G] o F-'[W] ] for up to two G arrays sharing W; scale = -1/N_k^2.
// herm_n > 0 (Hermitian S, order herm_n): only the lower half L of S = L + L^H is written: the
// columns above the diagonal are not formed (zeros are written only inside the diagonal
// 128 x 128 blocks, the part that a product with K from the tile column reads with tiles up
// to 128 wide) and the diagonal is halved.
template <int N1, int N2, int N3, int S, typename RC, typename TG, typename TW, typename TO>
__global__ __launch_bounds__(KT) void sigma_kfft_kernel(const TG* __restrict__ G1, const TG* __restrict__ G2,
                                                        const TW* __restrict__ W, long w_kstride,
                                                        TO* __restrict__ S1, TO* __restrict__ S2, long M,
                                                        double scale, int herm_n) {
  constexpr int L = (N1 / S) * N2 * N3;
  const long gid = static_cast<long>(blockIdx.x) * KT + threadIdx.x;
  const long col = gid / S;
  const int h = static_cast<int>(gid % S);
  if (col >= M) return;
  bool half = false;
  if (herm_n > 0) {  // a column's S lanes share (P, Q) and take the same branch
    const int P = static_cast<int>(col / herm_n), Q = static_cast<int>(col % herm_n);
    if (Q > P) {
      if ((Q >> 7) == (P >> 7)) {
        RC z[L];
#pragma unroll
        for (int e = 0; e < L; ++e) z[e] = RC{0, 0};
        store_col<N1, N2, N3, S>(z, S1, M, col, h);
        if (G2 != nullptr) store_col<N1, N2, N3, S>(z, S2, M, col, h);
      }
      return;
    }
    half = (P == Q);
  }
  RC w[L], g[L];
  load_col<N1, N2, N3, S>(w, W, w_kstride, col, h);
  dft3<N1, N2, N3, S, -1>(w, h);
  const scal_t<RC> sc = static_cast<scal_t<RC>>(scale);
#pragma unroll
  for (int e = 0; e < L; ++e) w[e] = cscale(w[e], sc);
  for (int pass = 0; pass < 2; ++pass) {
    const TG* G = pass == 0 ? G1 : G2;
    TO* O = pass == 0 ? S1 : S2;
    if (G == nullptr) break;
    load_col<N1, N2, N3, S>(g, G, M, col, h);
    dft3<N1, N2, N3, S, -1>(g, h);
#pragma unroll
    for (int e = 0; e < L; ++e) g[e] = cmul(g[e], w[e]);
    dft3<N1, N2, N3, S, +1>(g, h);
    if (half) {
#pragma unroll
      for (int e = 0; e < L; ++e) g[e] = cscale(g[e], static_cast<scal_t<RC>>(0.5));
    }
    store_col<N1, N2, N3, S>(g, O, M, col, h);
  }
}

// ---------------------------------------------------------------------------------
// Fallback for any grid with N_k <= 16: one thread per column, the column staged in
// LDS, direct separable DFTs with twiddles from sincospi.  Used for grids without a
// specialised kernel and as an independent check of the specialised ones.
// ---------------------------------------------------------------------------------
constexpr int FB_T = 64;
constexpr int FB_NK = 16;

__device__ void dft3_lds(c128* x /*stride FB_T*/, c128* tmp, int n1, int n2, int n3, int sign) {
  const int dims[3] = {n1, n2, n3};
  const int strides[3] = {n2 * n3, n3, 1};
  const int nk = n1 * n2 * n3;
  for (int d = 0; d < 3; ++d) {
    const int n = dims[d], st = strides[d];
    if (n == 1) continue;
    for (int e = 0; e < nk; ++e) tmp[e * FB_T] = x[e * FB_T];
    for (int e = 0; e < nk; ++e) {
      const int m = (e / st) % n;
      const int base = e - m * st;
      c128 acc{0.0, 0.0};
      for (int j = 0; j < n; ++j) {
        double s, c;
        sincospi(2.0 * double(sign) * double((j * m) % n) / double(n), &s, &c);
        acc = cadd(acc, cmul(tmp[(base + j * st) * FB_T], c128{c, s}));
      }
      x[e * FB_T] = acc;
    }
  }
}

template <typename TG, typename TW, typename TO>
__global__ __launch_bounds__(FB_T) void kfft_fallback_kernel(int mode, int n1, int n2, int n3, const TG* A1,
                                                             const TG* A2, const TW* W, long w_kstride, TO* O1,
                                                             TO* O2, long M, double scale) {
  __shared__ c128 xs[FB_NK * FB_T], ys[FB_NK * FB_T], ts[FB_NK * FB_T];
  const long col = static_cast<long>(blockIdx.x) * FB_T + threadIdx.x;
  const int nk = n1 * n2 * n3;
  c128* x = xs + threadIdx.x;
  c128* y = ys + threadIdx.x;
  c128* t = ts + threadIdx.x;
  if (col >= M) return;
  if (mode == 0) {  // chi0
    for (int k = 0; k < nk; ++k) {
      x[k * FB_T] = cvt<c128>(A1[k * M + col]);
      y[k * FB_T] = cvt<c128>(A2[k * M + col]);
    }
    dft3_lds(x, t, n1, n2, n3, -1);
    dft3_lds(y, t, n1, n2, n3, -1);
    for (int k = 0; k < nk; ++k) x[k * FB_T] = cscale(cmulconj(x[k * FB_T], y[k * FB_T]), scale);
    dft3_lds(x, t, n1, n2, n3, +1);
    for (int k = 0; k < nk; ++k) O1[k * M + col] = cvt<TO>(x[k * FB_T]);
  } else {  // Sigma
    for (int k = 0; k < nk; ++k) y[k * FB_T] = cvt<c128>(W[k * w_kstride + col]);
    dft3_lds(y, t, n1, n2, n3, -1);
    for (int pass = 0; pass < 2; ++pass) {
      const TG* G = pass == 0 ? A1 : A2;
      TO* O = pass == 0 ? O1 : O2;
      if (G == nullptr) break;
      for (int k = 0; k < nk; ++k) x[k * FB_T] = cvt<c128>(G[k * M + col]);
      dft3_lds(x, t, n1, n2, n3, -1);
      for (int k = 0; k < nk; ++k) x[k * FB_T] = cscale(cmul(x[k * FB_T], y[k * FB_T]), scale);
      dft3_lds(x, t, n1, n2, n3, +1);
      for (int k = 0; k < nk; ++k) O[k * M + col] = cvt<TO>(x[k * FB_T]);
    }
  }
}

template <int N1, int N2, int N3, int S, typename RC, typename TG, typename TO>
void launch_pi(hipStream_t s, const TG* Gp, const TG* Gn, TO* Pi, long M, double scale) {
  const long threads = M * S;
  const unsigned grid = static_cast<unsigned>((threads + KT - 1) / KT);
  hipLaunchKernelGGL((pi_kfft_kernel<N1, N2, N3, S, RC, TG, TO>), dim3(grid), dim3(KT), 0, s, Gp, Gn, Pi, M, scale);
}

template <int N1, int N2, int N3, int S, typename RC, typename TG, typename TW, typename TO>
void launch_sigma(hipStream_t s, const TG* G1, const TG* G2, const TW* W, long wks, TO* S1, TO* S2, long M,
                  double scale, int herm_n) {
  const long threads = M * S;
  const unsigned grid = static_cast<unsigned>((threads + KT - 1) / KT);
  hipLaunchKernelGGL((sigma_kfft_kernel<N1, N2, N3, S, RC, TG, TW, TO>), dim3(grid), dim3(KT), 0, s, G1, G2, W, wks,
                     S1, S2, M, scale, herm_n);
}

int grid_code(const int* kg) {
  if (kg[0] == kg[1] && kg[1] == kg[2] && kg[0] >= 1 && kg[0] <= 4) return kg[0];
  return 0;
}

}  // namespace

bool kfft_supported(const int* kgrid) {
  const int nk = kgrid[0] * kgrid[1] * kgrid[2];
  return grid_code(kgrid) != 0 || (nk >= 1 && nk <= FB_NK);
}

namespace {

template <typename RC, typename TG, typename TO>
void pi_kfft_t(hipStream_t s, const int* kgrid, bool force_fallback, const TG* Gp, const TG* Gn, TO* Pi, long M,
               double scale) {
  const int code = force_fallback ? 0 : grid_code(kgrid);
  switch (code) {
    case 1: launch_pi<1, 1, 1, 1, RC>(s, Gp, Gn, Pi, M, scale); break;
    case 2: launch_pi<2, 2, 2, 1, RC>(s, Gp, Gn, Pi, M, scale); break;
    case 3: launch_pi<3, 3, 3, 1, RC>(s, Gp, Gn, Pi, M, scale); break;
    case 4: launch_pi<4, 4, 4, 4, RC>(s, Gp, Gn, Pi, M, scale); break;
    default: {
      const int nk = kgrid[0] * kgrid[1] * kgrid[2];
      THCGW_REQUIRE(nk <= FB_NK, "k-grid not supported by the fused transform kernels");
      const unsigned grid = static_cast<unsigned>((M + FB_T - 1) / FB_T);
      hipLaunchKernelGGL((kfft_fallback_kernel<TG, TG, TO>), dim3(grid), dim3(FB_T), 0, s, 0, kgrid[0], kgrid[1],
                         kgrid[2], Gp, Gn, static_cast<const TG*>(nullptr), 0L, Pi, static_cast<TO*>(nullptr), M,
                         scale);
    }
  }
  THCGW_HIP(hipGetLastError());
}

template <typename RC, typename TG, typename TW, typename TO>
void sigma_kfft_t(hipStream_t s, const int* kgrid, bool force_fallback, const TG* G1, const TG* G2, const TW* W,
                  long w_kstride, TO* S1, TO* S2, long M, double scale, int herm_n) {
  const int code = force_fallback ? 0 : grid_code(kgrid);
  switch (code) {
    case 1: launch_sigma<1, 1, 1, 1, RC>(s, G1, G2, W, w_kstride, S1, S2, M, scale, herm_n); break;
    case 2: launch_sigma<2, 2, 2, 1, RC>(s, G1, G2, W, w_kstride, S1, S2, M, scale, herm_n); break;
    case 3: launch_sigma<3, 3, 3, 1, RC>(s, G1, G2, W, w_kstride, S1, S2, M, scale, herm_n); break;
    case 4: launch_sigma<4, 4, 4, 4, RC>(s, G1, G2, W, w_kstride, S1, S2, M, scale, herm_n); break;
    default: {
      const int nk = kgrid[0] * kgrid[1] * kgrid[2];
      THCGW_REQUIRE(nk <= FB_NK, "k-grid not supported by the fused transform kernels");
      const unsigned grid = static_cast<unsigned>((M + FB_T - 1) / FB_T);
      hipLaunchKernelGGL((kfft_fallback_kernel<TG, TW, TO>), dim3(grid), dim3(FB_T), 0, s, 1, kgrid[0], kgrid[1],
                         kgrid[2], G1, G2, W, w_kstride, S1, S2, M, scale);
      THCGW_HIP(hipGetLastError());
      if (herm_n > 0) {  // the lower half of the full result
        lower_half(s, S1, herm_n, nk);
        if (G2 != nullptr) lower_half(s, S2, herm_n, nk);
      }
    }
  }
  THCGW_HIP(hipGetLastError());
}

}  // namespace

void pi_kfft(hipStream_t s, const int* kgrid, bool force_fallback, const c128* Gp, const c128* Gn, c128* Pi,
             long M, double scale) {
  pi_kfft_t<c128>(s, kgrid, force_fallback, Gp, Gn, Pi, M, scale);
}

void sigma_kfft(hipStream_t s, const int* kgrid, bool force_fallback, const c128* G1, const c128* G2,
                const c128* W, long w_kstride, c128* S1, c128* S2, long M, double scale, int herm_n) {
  sigma_kfft_t<c128>(s, kgrid, force_fallback, G1, G2, W, w_kstride, S1, S2, M, scale, herm_n);
}

void pi_kfft_f32(hipStream_t s, const int* kgrid, bool force_fallback, const c64* Gp, const c64* Gn, c128* Pi,
                 long M, double scale) {
  pi_kfft_t<c64>(s, kgrid, force_fallback, Gp, Gn, Pi, M, scale);
}

void sigma_kfft_f32(hipStream_t s, const int* kgrid, bool force_fallback, const c64* G1, const c64* G2,
                    const c128* W, long w_kstride, c64* S1, c64* S2, long M, double scale, int herm_n) {
  sigma_kfft_t<c64>(s, kgrid, force_fallback, G1, G2, W, w_kstride, S1, S2, M, scale, herm_n);
}

}  // namespace thcgw
