// Elementwise and small-reduction kernels of the THC GW/RPA path.
#include "thcgw_internal.hpp"

#include <algorithm>

namespace thcgw {
namespace {

constexpr int T256 = 256;

// G0 of qp_scf_common.cpp:86-93 for x = e - mu at imaginary time t.
__device__ __forceinline__ double g0(double x, double t, double beta) {
  if (x > 0) return -exp(-x * t) / (1.0 + exp(-x * beta));
  return -exp(x * (beta - t)) / (1.0 + exp(x * beta));
}

// CG[t][k][i][a] = C[k][i][a] * g0(E[k][a] - mu, tau[t])
__global__ __launch_bounds__(T256) void scale_mo_kernel(const c128* C, const double* E, const double* tau, double mu,
                                                        double beta, int nt, int nk, int no, c128* CG) {
  const long total = static_cast<long>(nt) * nk * no * no;
  const long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x;
  if (e >= total) return;
  const int a = static_cast<int>(e % no);
  const long ki = e / no;  // (t, k, i)
  const int k = static_cast<int>((ki / no) % nk);
  const int t = static_cast<int>(ki / (static_cast<long>(no) * nk));
  const c128 c = C[(static_cast<long>(k) * no + (ki % no)) * no + a];
  const double g = g0(E[static_cast<long>(k) * no + a] - mu, tau[t], beta);
  CG[e] = c128{c.x * g, c.y * g};
}

// A <- scale (A + A^H)/2 for each n x n matrix (scf_common.hpp:70-86 with scale = 1); thread
// per (i, j), i <= j.
__global__ __launch_bounds__(T256) void hermitize_kernel(c128* A, int n, long nb, double scale) {
  const long per = static_cast<long>(n) * n;
  const long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x;
  if (e >= per * nb) return;
  const long b = e / per;
  const int i = static_cast<int>((e % per) / n), j = static_cast<int>(e % n);
  if (i > j) return;
  c128* Ab = A + b * per;
  const c128 x = Ab[static_cast<long>(i) * n + j], y = Ab[static_cast<long>(j) * n + i];
  const c128 h{0.5 * scale * (x.x + y.x), 0.5 * scale * (x.y - y.y)};
  Ab[static_cast<long>(i) * n + j] = h;
  Ab[static_cast<long>(j) * n + i] = c128{h.x, -h.y};
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < T256 / 64; ++w) s += sh[w];
  return s;
}

// trace[b] = sum_c (1 - A[b][c][c]) = Tr(V Pi) when A = I - V Pi.
__global__ __launch_bounds__(T256) void one_minus_trace_kernel(const c128* A, long sA, int n, c128* trace) {
  __shared__ double sh[T256 / 64];
  const int b = blockIdx.x;
  const c128* Ab = A + static_cast<long>(b) * sA;
  double re = 0.0, im = 0.0;
  for (int c = threadIdx.x; c < n; c += T256) {
    const c128 d = Ab[static_cast<long>(c) * n + c];
    re += 1.0 - d.x;
    im += -d.y;
  }
  re = block_sum(re, sh);
  im = block_sum(im, sh);
  if (threadIdx.x == 0) trace[b] = c128{re, im};
}

// log|Re det| of an LU factorisation with diagonal d_c and column swaps ipiv:
// sum ln|d_c| + ln|cos(sum arg d_c + pi * #swaps)|  (finite for any n, no overflow).
__global__ __launch_bounds__(T256) void log_re_det_kernel(const c128* A, long sA, int n, const int* ipiv,
                                                          double* ldet) {
  __shared__ double sh[T256 / 64];
  const int b = blockIdx.x;
  const c128* Ab = A + static_cast<long>(b) * sA;
  const int* piv = ipiv + static_cast<long>(b) * n;
  double la = 0.0, ph = 0.0, sw = 0.0;
  for (int c = threadIdx.x; c < n; c += T256) {
    const c128 d = Ab[static_cast<long>(c) * n + c];
    la += log(hypot(d.x, d.y));
    ph += atan2(d.y, d.x);
    sw += (piv[c] != c) ? 1.0 : 0.0;
  }
  la = block_sum(la, sh);
  ph = block_sum(ph, sh);
  sw = block_sum(sw, sh);
  if (threadIdx.x == 0) {
    const double phase = ph + M_PI * fmod(sw, 2.0);
    ldet[b] = la + log(fabs(cos(phase)));
  }
}

// Row gather that undoes the column pivoting of A = Lt Ut Pt in a solve:
// Y = S_0 S_1 ... S_{n-1} Z, i.e. Y[r] = Z[idx[r]] with idx built by applying the swaps in
// reverse order to the identity; then W = Y - V (scr_coulomb_t.cpp:191-196 stores W - V).
__global__ __launch_bounds__(T256) void dyson_finish_kernel(const c128* Z, long sZ, const c128* V, long sV, int vmod,
                                                            const int* ipiv, int n, c128* W, long sW, int b0) {
  extern __shared__ int idx[];
  const int b = blockIdx.y;
  const int* piv = ipiv + static_cast<long>(b) * n;
  if (threadIdx.x == 0) {
    for (int r = 0; r < n; ++r) idx[r] = r;
    for (int c = n - 1; c >= 0; --c) {
      const int p = piv[c];
      if (p < 0 || p >= n) continue;  // the panel writes pivots in [c, n); out-of-range ones are skipped
      const int tmp = idx[c];
      idx[c] = idx[p];
      idx[p] = tmp;
    }
  }
  __syncthreads();
  const c128* Zb = Z + static_cast<long>(b) * sZ;
  const c128* Vb = V ? V + static_cast<long>((b0 + b) % vmod) * sV : nullptr;
  c128* Wb = W + static_cast<long>(b) * sW;
  const long per = static_cast<long>(n) * n;
  for (long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x; e < per; e += static_cast<long>(gridDim.x) * T256) {
    const int r = static_cast<int>(e / n), c = static_cast<int>(e % n);
    const c128 z = Zb[static_cast<long>(idx[r]) * n + c];
    const c128 v = Vb ? Vb[e] : c128{0.0, 0.0};
    Wb[e] = c128{z.x - v.x, z.y - v.y};
  }
}

template <typename T>
__global__ __launch_bounds__(T256) void copy_batch_kernel(const T* src, long sS, int smod, T* dst, long sD, long per,
                                                          int b0) {
  const int b = blockIdx.y;
  const T* s = src + static_cast<long>((b0 + b) % smod) * sS;
  T* d = dst + static_cast<long>(b) * sD;
  for (long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x; e < per; e += static_cast<long>(gridDim.x) * T256)
    d[e] = s[e];
}

// X[n][k] = (iw_n + mu) S_k - H0_k - F_k - Sw[n][k], iw_n = i wn[n] pi / beta; S = NULL is the
// identity and H0 = NULL zero.  One thread per element of the (nw, nk, no, no) array.
__global__ __launch_bounds__(T256) void matsubara_dyson_kernel(const c128* H0, const c128* F, const c128* S,
                                                               const c128* Sw, const long* wn, double beta, double mu,
                                                               int nw, int nk, int no, c128* X) {
  const long m = static_cast<long>(no) * no, dim = nk * m;
  const long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x;
  if (e >= nw * dim) return;
  const int n = static_cast<int>(e / dim);
  const long ke = e % dim;
  const int i = static_cast<int>((ke % m) / no), j = static_cast<int>(ke % no);
  const double wi = double(wn[n]) * M_PI / beta;
  const c128 sij = S ? S[ke] : c128{i == j ? 1.0 : 0.0, 0.0};
  const c128 h = H0 ? H0[ke] : c128{0.0, 0.0};
  const c128 f = F[ke], sw = Sw[e];
  // (mu + i wi) * s
  X[e] = c128{mu * sij.x - wi * sij.y - h.x - f.x - sw.x, mu * sij.y + wi * sij.x - h.y - f.y - sw.y};
}

// RHS of the inverse: S_k (or the identity) for every (n, k).
__global__ __launch_bounds__(T256) void rhs_batch_kernel(const c128* S, int nk, int no, long count, c128* Y) {
  const long m = static_cast<long>(no) * no;
  const long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x;
  if (e >= count * m) return;
  const long b = e / m, ij = e % m;
  const int i = static_cast<int>(ij / no), j = static_cast<int>(ij % no);
  Y[e] = S ? S[(b % nk) * m + ij] : c128{i == j ? 1.0 : 0.0, 0.0};
}

// tr[b] = sum_j A[b][j][j]
__global__ __launch_bounds__(T256) void batched_trace_kernel(const c128* A, long sA, int n, c128* tr) {
  __shared__ double sh[T256 / 64];
  const c128* Ab = A + static_cast<long>(blockIdx.x) * sA;
  double re = 0.0, im = 0.0;
  for (int c = threadIdx.x; c < n; c += T256) {
    const c128 d = Ab[static_cast<long>(c) * n + c];
    re += d.x;
    im += d.y;
  }
  re = block_sum(re, sh);
  im = block_sum(im, sh);
  if (threadIdx.x == 0) tr[blockIdx.x] = c128{re, im};
}

// d[Q] = (2 / nk) sum_k DP[k][Q][Q]: twice the R = 0 diagonal of the projected density.
__global__ __launch_bounds__(T256) void hf_diag_kernel(const c128* DP, int nk, int nmu, c128* d) {
  const int Q = blockIdx.x * T256 + threadIdx.x;
  if (Q >= nmu) return;
  const long M = static_cast<long>(nmu) * nmu;
  double re = 0.0, im = 0.0;
  for (int k = 0; k < nk; ++k) {
    const c128 v = DP[k * M + static_cast<long>(Q) * nmu + Q];
    re += v.x;
    im += v.y;
  }
  d[Q] = c128{2.0 * re / nk, 2.0 * im / nk};
}

// J_P = sum_Q V0[P][Q] d[Q];  FP[k][P][P] += J_P for every k.
__global__ __launch_bounds__(T256) void hf_hartree_kernel(const c128* V0, const c128* d, c128* FP, int nk, int nmu) {
  const int P = blockIdx.x * T256 + threadIdx.x;
  if (P >= nmu) return;
  const long M = static_cast<long>(nmu) * nmu;
  double re = 0.0, im = 0.0;
  const c128* row = V0 + static_cast<long>(P) * nmu;
  for (int Q = 0; Q < nmu; ++Q) {
    const c128 v = row[Q], x = d[Q];
    re += v.x * x.x - v.y * x.y;
    im += v.x * x.y + v.y * x.x;
  }
  for (int k = 0; k < nk; ++k) {
    c128* f = FP + k * M + static_cast<long>(P) * nmu + P;
    *f = c128{f->x + re, f->y + im};
  }
}

__global__ __launch_bounds__(T256) void zero_c128_kernel(c128* p, long n) {
  for (long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x; e < n; e += static_cast<long>(gridDim.x) * T256)
    p[e] = c128{0.0, 0.0};
}

// Round complex<double> to complex<float> (round to nearest).
__global__ __launch_bounds__(T256) void to_c64_kernel(const c128* src, c64* dst, long n) {
  for (long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x; e < n; e += static_cast<long>(gridDim.x) * T256)
    dst[e] = c64{static_cast<float>(src[e].x), static_cast<float>(src[e].y)};
}

// out[b][a] = sum_i conj(C[k][i][a]) T1[b][i][a], k = b % nk: the diagonal of C^H T1.
__global__ __launch_bounds__(T256) void mo_diag_kernel(const c128* T1, const c128* C, c128* out, int nk, int no,
                                                       long total) {
  const long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x;
  if (e >= total) return;
  const long b = e / no;
  const int a = static_cast<int>(e % no);
  const int k = static_cast<int>(b % nk);
  const c128* Ck = C + static_cast<long>(k) * no * no;
  const c128* Tb = T1 + b * no * no;
  double re = 0.0, im = 0.0;
  for (int i = 0; i < no; ++i) {
    const c128 c = Ck[static_cast<long>(i) * no + a], t = Tb[static_cast<long>(i) * no + a];
    re += c.x * t.x + c.y * t.y;
    im += c.x * t.y - c.y * t.x;
  }
  out[e] = c128{re, im};
}

// out[w] = scale * sum_{k,i,j} A[w][k][i][j] B[w][k][j][i]   (one block per w).
__global__ __launch_bounds__(T256) void trace_prod_kernel(const c128* A, const c128* B, int nk, int no, double scale,
                                                          c128* out) {
  __shared__ double sh[T256 / 64];
  const long per = static_cast<long>(no) * no;
  const c128* Aw = A + static_cast<long>(blockIdx.x) * nk * per;
  const c128* Bw = B + static_cast<long>(blockIdx.x) * nk * per;
  double re = 0.0, im = 0.0;
  for (long e = threadIdx.x; e < nk * per; e += T256) {
    const long k = e / per;
    const int i = static_cast<int>((e % per) / no), j = static_cast<int>(e % no);
    const c128 a = Aw[e], b = Bw[k * per + static_cast<long>(j) * no + i];
    re += a.x * b.x - a.y * b.y;
    im += a.x * b.y + a.y * b.x;
  }
  re = block_sum(re, sh);
  im = block_sum(im, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = c128{re * scale, im * scale};
}


// ---- time reversal: staged Dyson on the irreducible q ----
// Matrix t of a staging block is (n, q) = (n0 + t / nqi, qir[t % nqi]).

// dst[t] = src[n q]   (M elements per matrix; grid.y = matrix)
__global__ __launch_bounds__(T256) void gather_ir_kernel(const c128* __restrict__ src, long M, int n0, int nq,
                                                         const int* __restrict__ qir, int nqi, c128* __restrict__ dst) {
  const int t = blockIdx.y;
  const long from = static_cast<long>(n0 + t / nqi) * nq + qir[t % nqi];
  const c128* a = src + from * M;
  c128* b = dst + static_cast<long>(t) * M;
  for (long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x; e < M; e += static_cast<long>(gridDim.x) * T256)
    b[e] = a[e];
}

// dst[n q] = S[t]; dst[n mq(q)] = conj S[t] when mq(q) != q
__global__ __launch_bounds__(T256) void scatter_ir_kernel(const c128* __restrict__ S, long M, int n0, int nq,
                                                          const int* __restrict__ qir, const int* __restrict__ mq,
                                                          int nqi, c128* __restrict__ dst) {
  const int t = blockIdx.y;
  const int q = qir[t % nqi], p = mq[q];
  const long base = static_cast<long>(n0 + t / nqi) * nq;
  const c128* a = S + static_cast<long>(t) * M;
  c128* b = dst + (base + q) * M;
  c128* c = dst + (base + p) * M;
  for (long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x; e < M; e += static_cast<long>(gridDim.x) * T256) {
    const c128 v = a[e];
    b[e] = v;
    if (p != q) c[e] = c128{v.x, -v.y};
  }
}

__global__ void scatter_ir_terms_kernel(const c128* __restrict__ trs, const double* __restrict__ lds,
                                        const int* __restrict__ infos, int n0, int nq, const int* __restrict__ qir,
                                        const int* __restrict__ mq, int nqi, int cnt, c128* __restrict__ tr,
                                        double* __restrict__ ld, int* __restrict__ info) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int q = qir[t % nqi], p = mq[q];
  const long base = static_cast<long>(n0 + t / nqi) * nq;
  info[base + q] = infos[t];
  info[base + p] = infos[t];
  if (tr) {
    tr[base + q] = trs[t];
    tr[base + p] = c128{trs[t].x, -trs[t].y};
  }
  if (ld) {
    ld[base + q] = lds[t];
    ld[base + p] = lds[t];
  }
}

// out[2 (q nbx + bx)] = max |A[mq(q)] - conj A[q]|, out[... + 1] = max |A[q]| over the block's slice
// out[2 b'] = max |A_ij - conj(A_ji)|, out[2 b' + 1] = max |A_ij| over the elements that block
// b' visits (grid stride over all nb matrices of order n).
__global__ __launch_bounds__(T256) void herm_diff_kernel(const c128* __restrict__ A, int n, long nb,
                                                         double* __restrict__ out) {
  __shared__ double r0[T256], r1[T256];
  const long nn = static_cast<long>(n) * n, total = nn * nb;
  double d = 0.0, m = 0.0;
  for (long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x; e < total;
       e += static_cast<long>(gridDim.x) * T256) {
    const long b = e / nn;
    const int ij = static_cast<int>(e - b * nn), i = ij / n, j = ij % n;
    const c128 u = A[e], v = A[b * nn + static_cast<long>(j) * n + i];
    d = fmax(d, fmax(fabs(u.x - v.x), fabs(u.y + v.y)));
    m = fmax(m, fmax(fabs(u.x), fabs(u.y)));
  }
  r0[threadIdx.x] = d;
  r1[threadIdx.x] = m;
  __syncthreads();
  for (int w = T256 / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      r0[threadIdx.x] = fmax(r0[threadIdx.x], r0[threadIdx.x + w]);
      r1[threadIdx.x] = fmax(r1[threadIdx.x], r1[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2L * blockIdx.x] = r0[0];
    out[2L * blockIdx.x + 1] = r1[0];
  }
}

__global__ __launch_bounds__(T256) void trev_diff_kernel(const c128* __restrict__ A, long M, const int* __restrict__ mq,
                                                         double* __restrict__ out) {
  __shared__ double r0[T256], r1[T256];
  const int q = blockIdx.y;
  const c128* a = A + static_cast<long>(q) * M;
  const c128* b = A + static_cast<long>(mq[q]) * M;
  double d = 0.0, m = 0.0;
  for (long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x; e < M; e += static_cast<long>(gridDim.x) * T256) {
    const c128 u = a[e], v = b[e];
    d = fmax(d, fmax(fabs(v.x - u.x), fabs(v.y + u.y)));
    m = fmax(m, fmax(fabs(u.x), fabs(u.y)));
  }
  r0[threadIdx.x] = d;
  r1[threadIdx.x] = m;
  __syncthreads();
  for (int w = T256 / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      r0[threadIdx.x] = fmax(r0[threadIdx.x], r0[threadIdx.x + w]);
      r1[threadIdx.x] = fmax(r1[threadIdx.x], r1[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const long o = 2L * (static_cast<long>(q) * gridDim.x + blockIdx.x);
    out[o] = r0[0];
    out[o + 1] = r1[0];
  }
}


// Input guard: out[block] = number of NaN / infinite doubles among x[0, n) in the block's
// grid-stride share (blocks read 256 consecutive doubles per step).
__global__ __launch_bounds__(T256) void nonfinite_kernel(const double* __restrict__ x, long n, int* __restrict__ out) {
  __shared__ int r[T256];
  int c = 0;
  for (long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x; e < n; e += static_cast<long>(gridDim.x) * T256)
    c += isfinite(x[e]) ? 0 : 1;
  r[threadIdx.x] = c;
  __syncthreads();
  for (int w = T256 / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) r[threadIdx.x] += r[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = r[0];
}

// ---- long-wavelength (Gygi) correction ----
// out[t * nq + q] = sum_P c[q][P] sum_Q W[q][t][P][Q] conj(c[q][Q])   (g0_div_utils.hpp:199-209)
// One workgroup per (q, t) matrix; each wave takes rows P = wave, wave + 4, ..., its lanes
// read consecutive Q.
__global__ __launch_bounds__(T256) void prod_head_kernel(const c128* __restrict__ W, const c128* __restrict__ cbar,
                                                         int n, int nt_half, int nq, c128* __restrict__ out) {
  __shared__ double red[2][T256];
  const int t = blockIdx.x, q = blockIdx.y;
  const c128* w = W + (static_cast<long>(q) * nt_half + t) * static_cast<long>(n) * n;
  const c128* c = cbar + static_cast<long>(q) * n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double ar = 0.0, ai = 0.0;
  for (int P = wave; P < n; P += 4) {
    const c128* row = w + static_cast<long>(P) * n;
    double rr = 0.0, ri = 0.0;  // sum_Q W[P][Q] conj(c[Q]), this lane's share
    for (int Q = lane; Q < n; Q += 64) {
      const c128 a = row[Q], b = c[Q];
      rr += a.x * b.x + a.y * b.y;
      ri += a.y * b.x - a.x * b.y;
    }
    const c128 cp = c[P];  // each lane adds c[P] times its share
    ar += cp.x * rr - cp.y * ri;
    ai += cp.x * ri + cp.y * rr;
  }
  red[0][threadIdx.x] = ar;
  red[1][threadIdx.x] = ai;
  __syncthreads();
  for (int h = T256 / 2; h > 0; h >>= 1) {
    if (threadIdx.x < h) {
      red[0][threadIdx.x] += red[0][threadIdx.x + h];
      red[1][threadIdx.x] += red[1][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[static_cast<long>(t) * nq + q] = c128{red[0][0], red[1][0]};
}

// Y[k][P][a] = conj(h[P]) X[k][P][a]
__global__ __launch_bounds__(T256) void scale_rows_conj_kernel(const c128* __restrict__ X, const c128* __restrict__ h,
                                                               int nmu, int no, long total, c128* __restrict__ Y) {
  const long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x;
  if (e >= total) return;
  const int P = static_cast<int>((e / no) % nmu);
  const c128 x = X[e], g = h[P];
  Y[e] = c128{g.x * x.x + g.y * x.y, g.x * x.y - g.y * x.x};
}

// A[t][k][i][j] *= f[t]
__global__ __launch_bounds__(T256) void scale_by_t_kernel(c128* A, const double* __restrict__ f, long per_t, long total) {
  const long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x;
  if (e >= total) return;
  const double g = f[e / per_t];
  A[e] = c128{A[e].x * g, A[e].y * g};
}


// ---- q-point symmetry map: A[q][i][j] = ph[q][i] conj(ph[q][j]) A^[src q][perm[q][i]][perm[q][j]] ----
__device__ __forceinline__ c128 qsym_value(const c128* __restrict__ A, long M, int n, int q, int sq, bool cj,
                                           const int* __restrict__ perm, const c128* __restrict__ ph, long e) {
  const int i = static_cast<int>(e / n), j = static_cast<int>(e % n);
  const int* pq = perm + static_cast<long>(q) * n;
  const c128 a = A[static_cast<long>(sq) * M + static_cast<long>(pq[i]) * n + pq[j]];
  const c128 v{a.x, cj ? -a.y : a.y};
  const c128 pi = ph[static_cast<long>(q) * n + i], pj = ph[static_cast<long>(q) * n + j];
  const c128 w{pi.x * pj.x + pi.y * pj.y, pi.y * pj.x - pi.x * pj.y};  // ph_i conj(ph_j)
  return c128{w.x * v.x - w.y * v.y, w.x * v.y + w.y * v.x};
}

// grid.y = nw * nq matrices (n, q); the irreducible q (src q == q) are left as they are
__global__ __launch_bounds__(T256) void qsym_fill_kernel(c128* W, int n, int nq, const int* __restrict__ src,
                                                         const int* __restrict__ cj, const int* __restrict__ perm,
                                                         const c128* __restrict__ ph) {
  const int nqw = blockIdx.y, q = nqw % nq, w = nqw / nq;
  const int sq = src[q];
  if (sq == q) return;
  const long M = static_cast<long>(n) * n;
  c128* base = W + static_cast<long>(w) * nq * M;  // the nq matrices of frequency w
  c128* out = base + static_cast<long>(q) * M;
  for (long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x; e < M; e += static_cast<long>(gridDim.x) * T256)
    out[e] = qsym_value(base, M, n, q, sq, cj[q] != 0, perm, ph, e);
}

__global__ void qsym_fill_terms_kernel(int nq, int total, const int* __restrict__ src, const int* __restrict__ cj,
                                       c128* tr, double* ld, int* info) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= total) return;
  const int q = b % nq, sq = src[q];
  if (sq == q) return;
  const int from = b - q + sq;
  if (tr) tr[b] = c128{tr[from].x, cj[q] ? -tr[from].y : tr[from].y};
  if (ld) ld[b] = ld[from];
  info[b] = info[from];
}

// out[2 (q nbx + bx)] = max |A[q] - map(A[src q])|, out[... + 1] = max |A[q]|
__global__ __launch_bounds__(T256) void qsym_diff_kernel(const c128* __restrict__ A, int n, const int* __restrict__ src,
                                                         const int* __restrict__ cj, const int* __restrict__ perm,
                                                         const c128* __restrict__ ph, double* __restrict__ out) {
  __shared__ double r0[T256], r1[T256];
  const int q = blockIdx.y, sq = src[q];
  const long M = static_cast<long>(n) * n;
  double d = 0.0, m = 0.0;
  for (long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x; e < M; e += static_cast<long>(gridDim.x) * T256) {
    const c128 a = A[static_cast<long>(q) * M + e];
    m = fmax(m, fmax(fabs(a.x), fabs(a.y)));
    if (sq != q) {
      const c128 v = qsym_value(A, M, n, q, sq, cj[q] != 0, perm, ph, e);
      d = fmax(d, fmax(fabs(a.x - v.x), fabs(a.y - v.y)));
    }
  }
  r0[threadIdx.x] = d;
  r1[threadIdx.x] = m;
  __syncthreads();
  for (int w = T256 / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      r0[threadIdx.x] = fmax(r0[threadIdx.x], r0[threadIdx.x + w]);
      r1[threadIdx.x] = fmax(r1[threadIdx.x], r1[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const long o = 2L * (static_cast<long>(q) * gridDim.x + blockIdx.x);
    out[o] = r0[0];
    out[o + 1] = r1[0];
  }
}

}  // namespace

void mo_diag(hipStream_t s, const c128* T1, const c128* C, c128* out, int nk, int no, long nb) {
  const long total = nb * no;
  if (total == 0) return;
  hipLaunchKernelGGL(mo_diag_kernel, dim3(static_cast<unsigned>((total + T256 - 1) / T256)), dim3(T256), 0, s, T1, C,
                     out, nk, no, total);
  THCGW_HIP(hipGetLastError());
}

void trace_prod(hipStream_t s, const c128* A, const c128* B, int nw, int nk, int no, double scale, c128* out) {
  hipLaunchKernelGGL(trace_prod_kernel, dim3(nw), dim3(T256), 0, s, A, B, nk, no, scale, out);
  THCGW_HIP(hipGetLastError());
}

void scale_mo(hipStream_t s, const c128* C, const double* E, const double* tau, double mu, double beta, int nt, int nk,
              int no, c128* CG) {
  const long total = static_cast<long>(nt) * nk * no * no;
  hipLaunchKernelGGL(scale_mo_kernel, dim3(static_cast<unsigned>((total + T256 - 1) / T256)), dim3(T256), 0, s, C, E,
                     tau, mu, beta, nt, nk, no, CG);
  THCGW_HIP(hipGetLastError());
}

void hermitize(hipStream_t s, c128* A, int n, long nb, double scale) {
  const long total = static_cast<long>(n) * n * nb;
  if (total == 0) return;
  hipLaunchKernelGGL(hermitize_kernel, dim3(static_cast<unsigned>((total + T256 - 1) / T256)), dim3(T256), 0, s, A, n,
                     nb, scale);
  THCGW_HIP(hipGetLastError());
}

// Lower half of a Hermitian matrix, L with A = L + L^H: the strict upper triangle set to zero
// and the diagonal halved, for nb n x n matrices in place.
template <typename T>
__global__ __launch_bounds__(T256) void lower_half_kernel(T* A, int n, long nb) {
  const long per = static_cast<long>(n) * n;
  const long e = static_cast<long>(blockIdx.x) * T256 + threadIdx.x;
  if (e >= per * nb) return;
  const int i = static_cast<int>((e % per) / n), j = static_cast<int>(e % n);
  if (i > j) return;
  using S = decltype(T{}.x);
  const T x = A[e];
  A[e] = (i == j) ? T{static_cast<S>(0.5) * x.x, static_cast<S>(0.5) * x.y} : T{0, 0};
}

template <typename T>
void lower_half_t(hipStream_t s, T* A, int n, long nb) {
  const long total = static_cast<long>(n) * n * nb;
  if (total == 0) return;
  THCGW_REQUIRE((total + T256 - 1) / T256 < (1L << 31), "lower_half: too many elements");
  hipLaunchKernelGGL(lower_half_kernel<T>, dim3(static_cast<unsigned>((total + T256 - 1) / T256)), dim3(T256), 0, s, A,
                     n, nb);
  THCGW_HIP(hipGetLastError());
}

void lower_half(hipStream_t s, c128* A, int n, long nb) { lower_half_t(s, A, n, nb); }
void lower_half(hipStream_t s, c64* A, int n, long nb) { lower_half_t(s, A, n, nb); }

void one_minus_trace(hipStream_t s, const c128* A, long sA, int n, c128* trace, int batch) {
  hipLaunchKernelGGL(one_minus_trace_kernel, dim3(batch), dim3(T256), 0, s, A, sA, n, trace);
  THCGW_HIP(hipGetLastError());
}

void log_re_det(hipStream_t s, const c128* A, long sA, int n, const int* ipiv, double* ldet, int batch) {
  hipLaunchKernelGGL(log_re_det_kernel, dim3(batch), dim3(T256), 0, s, A, sA, n, ipiv, ldet);
  THCGW_HIP(hipGetLastError());
}

void dyson_finish(hipStream_t s, const c128* Z, long sZ, const c128* V, long sV, int vmod, const int* ipiv, int n,
                  c128* W, long sW, int b0, int batch) {
  const long per = static_cast<long>(n) * n;
  const unsigned gx = static_cast<unsigned>(std::min<long>((per + T256 - 1) / T256, 64));
  hipLaunchKernelGGL(dyson_finish_kernel, dim3(gx, batch), dim3(T256), sizeof(int) * n, s, Z, sZ, V, sV, vmod, ipiv, n,
                     W, sW, b0);
  THCGW_HIP(hipGetLastError());
}

void copy_batch(hipStream_t s, const c128* src, long sS, int smod, c128* dst, long sD, long per, int b0, int batch) {
  const unsigned gx = static_cast<unsigned>(std::min<long>((per + T256 - 1) / T256, 64));
  hipLaunchKernelGGL(copy_batch_kernel<c128>, dim3(gx, batch), dim3(T256), 0, s, src, sS, smod, dst, sD, per, b0);
  THCGW_HIP(hipGetLastError());
}

void copy_batch(hipStream_t s, const c64* src, long sS, int smod, c64* dst, long sD, long per, int b0, int batch) {
  const unsigned gx = static_cast<unsigned>(std::min<long>((per + T256 - 1) / T256, 64));
  hipLaunchKernelGGL(copy_batch_kernel<c64>, dim3(gx, batch), dim3(T256), 0, s, src, sS, smod, dst, sD, per, b0);
  THCGW_HIP(hipGetLastError());
}

void matsubara_dyson(hipStream_t s, const c128* H0, const c128* F, const c128* S, const c128* Sw, const long* wn,
                     double beta, double mu, int nw, int nk, int no, c128* X) {
  const long total = static_cast<long>(nw) * nk * no * no;
  hipLaunchKernelGGL(matsubara_dyson_kernel, dim3(static_cast<unsigned>((total + T256 - 1) / T256)), dim3(T256), 0, s,
                     H0, F, S, Sw, wn, beta, mu, nw, nk, no, X);
  THCGW_HIP(hipGetLastError());
}

void rhs_batch(hipStream_t s, const c128* S, int nk, int no, long count, c128* Y) {
  const long total = count * no * no;
  hipLaunchKernelGGL(rhs_batch_kernel, dim3(static_cast<unsigned>((total + T256 - 1) / T256)), dim3(T256), 0, s, S,
                     nk, no, count, Y);
  THCGW_HIP(hipGetLastError());
}

void batched_trace(hipStream_t s, const c128* A, long sA, int n, c128* tr, int batch) {
  hipLaunchKernelGGL(batched_trace_kernel, dim3(batch), dim3(T256), 0, s, A, sA, n, tr);
  THCGW_HIP(hipGetLastError());
}

void hf_hartree(hipStream_t s, const c128* DP, const c128* V0, c128* d, c128* FP, int nk, int nmu) {
  const unsigned g = static_cast<unsigned>((nmu + T256 - 1) / T256);
  hipLaunchKernelGGL(hf_diag_kernel, dim3(g), dim3(T256), 0, s, DP, nk, nmu, d);
  THCGW_HIP(hipGetLastError());
  hipLaunchKernelGGL(hf_hartree_kernel, dim3(g), dim3(T256), 0, s, V0, static_cast<const c128*>(d), FP, nk, nmu);
  THCGW_HIP(hipGetLastError());
}

void zero_c128(hipStream_t s, c128* p, long n) {
  if (n <= 0) return;
  const unsigned g = static_cast<unsigned>(std::min<long>((n + T256 - 1) / T256, 1L << 20));
  hipLaunchKernelGGL(zero_c128_kernel, dim3(g), dim3(T256), 0, s, p, n);
  THCGW_HIP(hipGetLastError());
}

void to_c64(hipStream_t s, const c128* src, c64* dst, long n) {
  if (n <= 0) return;
  const unsigned g = static_cast<unsigned>(std::min<long>((n + T256 - 1) / T256, 1L << 20));
  hipLaunchKernelGGL(to_c64_kernel, dim3(g), dim3(T256), 0, s, src, dst, n);
  THCGW_HIP(hipGetLastError());
}

unsigned slice_blocks(long M) { return static_cast<unsigned>(std::min<long>((M + T256 - 1) / T256, 1024)); }

void gather_ir(hipStream_t s, const c128* src, long M, int n0, int nq, const int* qir, int nqi, int cnt, c128* dst) {
  if (cnt <= 0) return;
  hipLaunchKernelGGL(gather_ir_kernel, dim3(slice_blocks(M), cnt), dim3(T256), 0, s, src, M, n0, nq, qir, nqi, dst);
  THCGW_HIP(hipGetLastError());
}

void scatter_ir(hipStream_t s, const c128* S, long M, int n0, int nq, const int* qir, const int* mq, int nqi, int cnt,
                c128* dst) {
  if (cnt <= 0) return;
  hipLaunchKernelGGL(scatter_ir_kernel, dim3(slice_blocks(M), cnt), dim3(T256), 0, s, S, M, n0, nq, qir, mq, nqi, dst);
  THCGW_HIP(hipGetLastError());
}

void scatter_ir_terms(hipStream_t s, const c128* trs, const double* lds, const int* infos, int n0, int nq,
                      const int* qir, const int* mq, int nqi, int cnt, c128* tr, double* ld, int* info) {
  if (cnt <= 0) return;
  hipLaunchKernelGGL(scatter_ir_terms_kernel, dim3((cnt + T256 - 1) / T256), dim3(T256), 0, s, trs, lds, infos, n0, nq,
                     qir, mq, nqi, cnt, tr, ld, info);
  THCGW_HIP(hipGetLastError());
}

void prod_head(hipStream_t s, const c128* W, const c128* cbar, int n, int nt_half, int nq, c128* out) {
  hipLaunchKernelGGL(prod_head_kernel, dim3(nt_half, nq), dim3(T256), 0, s, W, cbar, n, nt_half, nq, out);
  THCGW_HIP(hipGetLastError());
}

void scale_rows_conj(hipStream_t s, const c128* X, const c128* h, int nk, int nmu, int no, c128* Y) {
  const long total = static_cast<long>(nk) * nmu * no;
  hipLaunchKernelGGL(scale_rows_conj_kernel, dim3(static_cast<unsigned>((total + T256 - 1) / T256)), dim3(T256), 0, s, X,
                     h, nmu, no, total, Y);
  THCGW_HIP(hipGetLastError());
}

void scale_by_t(hipStream_t s, c128* A, const double* f, long per_t, int nt) {
  const long total = per_t * nt;
  hipLaunchKernelGGL(scale_by_t_kernel, dim3(static_cast<unsigned>((total + T256 - 1) / T256)), dim3(T256), 0, s, A, f,
                     per_t, total);
  THCGW_HIP(hipGetLastError());
}

void qsym_fill(hipStream_t s, c128* W, int n, int nq, int nw, const int* src, const int* cj, const int* perm,
               const c128* ph) {
  const long M = static_cast<long>(n) * n;
  hipLaunchKernelGGL(qsym_fill_kernel, dim3(slice_blocks(M), nw * nq), dim3(T256), 0, s, W, n, nq, src, cj, perm, ph);
  THCGW_HIP(hipGetLastError());
}

void qsym_fill_terms(hipStream_t s, int nq, int nw, const int* src, const int* cj, c128* tr, double* ld, int* info) {
  const int total = nq * nw;
  hipLaunchKernelGGL(qsym_fill_terms_kernel, dim3((total + T256 - 1) / T256), dim3(T256), 0, s, nq, total, src, cj, tr,
                     ld, info);
  THCGW_HIP(hipGetLastError());
}

void qsym_diff(hipStream_t s, const c128* A, int n, int nq, const int* src, const int* cj, const int* perm,
               const c128* ph, int nbx, double* out) {
  hipLaunchKernelGGL(qsym_diff_kernel, dim3(nbx, nq), dim3(T256), 0, s, A, n, src, cj, perm, ph, out);
  THCGW_HIP(hipGetLastError());
}

void herm_diff(hipStream_t s, const c128* A, int n, long nb, int nbx, double* out) {
  hipLaunchKernelGGL(herm_diff_kernel, dim3(nbx), dim3(T256), 0, s, A, n, nb, out);
  THCGW_HIP(hipGetLastError());
}

void nonfinite_count(hipStream_t s, const double* x, long n, int nbx, int* out) {
  hipLaunchKernelGGL(nonfinite_kernel, dim3(nbx), dim3(T256), 0, s, x, n, out);
  THCGW_HIP(hipGetLastError());
}

void trev_diff(hipStream_t s, const c128* A, long M, const int* mq, int nq, int nbx, double* out) {
  hipLaunchKernelGGL(trev_diff_kernel, dim3(nbx, nq), dim3(T256), 0, s, A, M, mq, out);
  THCGW_HIP(hipGetLastError());
}

}  // namespace thcgw
