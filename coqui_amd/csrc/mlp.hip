// Small multilayer perceptron with BF16 weights and FP32 accumulation, trained and run by
// the kernels of this file (forward, backward, MSE loss, Adam).  It is the learned part of
// the Krylov Dyson solve: one row of the batch is one (n, q) Dyson problem.
//
//   h0 = silu(x W0^T + b0)     W0: width x n_in
//   h1 = silu(h0 W1^T + b1)    W1: width x width
//   y  = h1 W2^T + b2          W2: n_out x width
//
// Parameters live in one FP32 master array (W0, b0, W1, b1, W2, b2, row-major, the layout
// of torch.nn.Linear) with a BF16 mirror that the products read.  Every product runs on
// v_mfma_f32_16x16x32_bf16: its operands are rounded to BF16 (round to nearest even) when
// they are staged in LDS, and it accumulates in FP32.
#include "thcgw_internal.hpp"
#include "mlp.hpp"

#include <cmath>
#include <random>

namespace thcgw {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MT = 256;  // threads per tile (four waves as 2 x 2)
constexpr int MB = 64;   // tile rows and columns
constexpr int MK = 32;   // K chunk
constexpr int MLD = 40;  // LDS row stride in bf16 values (16-byte fragment reads spread over the banks)

// C(m, n) = sum_k A(m, k) B(n, k), with A(m, k) at A[m sam + k sak] and B(n, k) at
// B[n sbn + k sbk]: any transposition of either operand is a choice of strides.
struct MlpGemm {
  int M, N, K;
  const void* A;
  long sam, sak;
  const void* B;
  long sbn, sbk;
  float* C;  // ldc = N
  const float* bias;  // per column n (EPI 1 and 2)
  float* post;        // silu(C) (EPI 1), ldc = N
};

template <typename T>
__device__ __forceinline__ float ldf(const T* p, long i) {
  return static_cast<float>(p[i]);
}

// One operand's 64 x 32 chunk into S[x][k] (bf16): 8 values per thread.  When k is the
// unit-stride index a thread reads 8 consecutive k of one row, otherwise 8 consecutive x
// of one k, so that neighbouring threads read neighbouring addresses either way.
template <typename T>
__device__ __forceinline__ void stage(__bf16 (*S)[MLD], const T* X, long sx, long sk, int nx, int nk, int x0, int k0) {
  const int t = threadIdx.x;
  if (sk == 1) {
    const int x = t >> 2, k = (t & 3) * 8;
    const int gx = x0 + x;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int gk = k0 + k + e;
      S[x][k + e] = static_cast<__bf16>((gx < nx && gk < nk) ? ldf(X, gx * sx + static_cast<long>(gk)) : 0.0f);
    }
  } else {
    const int k = t >> 3, x = (t & 7) * 8;
    const int gk = k0 + k;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int gx = x0 + x + e;
      S[x + e][k] = static_cast<__bf16>((gx < nx && gk < nk) ? ldf(X, gx * sx + gk * sk) : 0.0f);
    }
  }
}

__device__ __forceinline__ float silu(float x) { return x / (1.0f + expf(-x)); }

// EPI 0: C = AB^T;  1: C = AB^T + bias, post = silu(C);  2: C = AB^T + bias;  3: C += AB^T.
template <typename TA, typename TB, int EPI>
__global__ __launch_bounds__(MT) void mlp_gemm_kernel(MlpGemm p) {
  __shared__ __attribute__((aligned(16))) __bf16 As[MB][MLD];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[MB][MLD];
  const int tn = (p.N + MB - 1) / MB;
  const int m0 = (blockIdx.x / tn) * MB, n0 = (blockIdx.x % tn) * MB;
  const TA* A = static_cast<const TA*>(p.A);
  const TB* B = static_cast<const TB*>(p.B);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < p.K; k0 += MK) {
    stage(As, A, p.sam, p.sak, p.M, p.K, m0, k0);
    stage(Bs, B, p.sbn, p.sbk, p.N, p.K, n0, k0);
    __syncthreads();
    bf16x8 fa[2], fb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) fa[i] = *reinterpret_cast<const bf16x8*>(&As[wm * 32 + i * 16 + lr][lk * 8]);
#pragma unroll
    for (int j = 0; j < 2; ++j) fb[j] = *reinterpret_cast<const bf16x8*>(&Bs[wn * 32 + j * 16 + lr][lk * 8]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[i], fb[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // accumulator register v of lane l holds row 4(l/16) + v, column l%16
        const int r = m0 + wm * 32 + i * 16 + 4 * lk + v;
        const int c = n0 + wn * 32 + j * 16 + lr;
        if (r < p.M && c < p.N) {
          const long o = static_cast<long>(r) * p.N + c;
          float x = acc[i][j][v];
          if constexpr (EPI == 3) {
            p.C[o] += x;
          } else {
            if constexpr (EPI == 1 || EPI == 2) x += p.bias[c];
            p.C[o] = x;
            if constexpr (EPI == 1) p.post[o] = silu(x);
          }
        }
      }
}

// dpre = dpost * silu'(pre), silu'(x) = s (1 + x (1 - s)), s = 1 / (1 + e^-x)
__global__ void silu_back_kernel(const float* __restrict__ pre, const float* __restrict__ dpost,
                                 float* __restrict__ dpre, long n) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = pre[i];
  const float s = 1.0f / (1.0f + expf(-x));
  dpre[i] = dpost[i] * s * (1.0f + x * (1.0f - s));
}

// g[c] = sum_r D[r][c] (bias gradients); one thread per column, rows in order.
__global__ void colsum_kernel(const float* __restrict__ D, int rows, int cols, float* __restrict__ g) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.0f;
  for (int r = 0; r < rows; ++r) s += D[static_cast<long>(r) * cols + c];
  g[c] = s;
}

// dY = 2 (Y - T) / n and per-block partial sums of (Y - T)^2.
constexpr int LT = 256;
__global__ __launch_bounds__(LT) void mse_kernel(const float* __restrict__ Y, const float* __restrict__ T, long n,
                                                 float* __restrict__ dY, float* __restrict__ part) {
  __shared__ float red[LT];
  float s = 0.0f;
  const float sc = 2.0f / static_cast<float>(n);
  for (long i = static_cast<long>(blockIdx.x) * LT + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * LT) {
    const float d = Y[i] - T[i];
    dY[i] = sc * d;
    s += d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = LT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(LT) void sum_kernel(const float* __restrict__ part, int n, float scale,
                                                 float* __restrict__ out) {
  __shared__ float red[LT];
  float s = 0.0f;
  for (int i = threadIdx.x; i < n; i += LT) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = LT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] * scale;
}

// Adam (Kingma and Ba 2015) with the bias corrections folded into lr_t and eps_t by the
// host, then the BF16 mirror of the updated value.
__global__ void adam_kernel(float* __restrict__ w, __bf16* __restrict__ wb, float* __restrict__ m,
                            float* __restrict__ v, const float* __restrict__ g, long n, float b1, float b2,
                            float lr_t, float eps_t) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr_t * mi / (sqrtf(vi) + eps_t);
  w[i] = wi;
  wb[i] = static_cast<__bf16>(wi);
}

__global__ void mirror_kernel(const float* __restrict__ w, __bf16* __restrict__ wb, long n) {
  const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) wb[i] = static_cast<__bf16>(w[i]);
}

unsigned blocks(long n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

template <typename TA, typename TB>
void gemm(hipStream_t s, int epi, const MlpGemm& g) {
  if (g.M <= 0 || g.N <= 0) return;
  const unsigned grid = static_cast<unsigned>(((g.M + MB - 1) / MB) * ((g.N + MB - 1) / MB));
  switch (epi) {
    case 0: hipLaunchKernelGGL((mlp_gemm_kernel<TA, TB, 0>), dim3(grid), dim3(MT), 0, s, g); break;
    case 1: hipLaunchKernelGGL((mlp_gemm_kernel<TA, TB, 1>), dim3(grid), dim3(MT), 0, s, g); break;
    case 2: hipLaunchKernelGGL((mlp_gemm_kernel<TA, TB, 2>), dim3(grid), dim3(MT), 0, s, g); break;
    default: hipLaunchKernelGGL((mlp_gemm_kernel<TA, TB, 3>), dim3(grid), dim3(MT), 0, s, g); break;
  }
  THCGW_HIP(hipGetLastError());
}

}  // namespace

Mlp::Mlp(int n_in, int width, int n_out, unsigned long long seed) : n_in(n_in), width(width), n_out(n_out) {
  THCGW_REQUIRE(n_in > 0 && width > 0 && n_out > 0, "mlp: layer sizes must be positive");
  oW0 = 0;
  ob0 = oW0 + static_cast<long>(width) * n_in;
  oW1 = ob0 + width;
  ob1 = oW1 + static_cast<long>(width) * width;
  oW2 = ob1 + width;
  ob2 = oW2 + static_cast<long>(n_out) * width;
  n_params = ob2 + n_out;
  // U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for every layer (the torch.nn.Linear default); the
  // output layer is scaled by 1/10 so that an untrained network predicts small values.
  std::vector<float> h(static_cast<size_t>(n_params));
  std::mt19937_64 rng(seed);
  auto fill = [&](long off, long count, int fan_in, float extra) {
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    const double a = extra / std::sqrt(static_cast<double>(fan_in));
    for (long i = 0; i < count; ++i) h[static_cast<size_t>(off + i)] = static_cast<float>(a * u(rng));
  };
  fill(oW0, static_cast<long>(width) * n_in, n_in, 1.0f);
  fill(ob0, width, n_in, 1.0f);
  fill(oW1, static_cast<long>(width) * width, width, 1.0f);
  fill(ob1, width, width, 1.0f);
  fill(oW2, static_cast<long>(n_out) * width, width, 0.1f);
  fill(ob2, n_out, width, 0.1f);
  THCGW_HIP(hipMalloc(&w, sizeof(float) * n_params));
  THCGW_HIP(hipMalloc(&wb, sizeof(__bf16) * n_params));
  THCGW_HIP(hipMalloc(&m, sizeof(float) * n_params));
  THCGW_HIP(hipMalloc(&v, sizeof(float) * n_params));
  THCGW_HIP(hipMalloc(&g, sizeof(float) * n_params));
  THCGW_HIP(hipMalloc(&loss, sizeof(float) * 1024));
  set_params(h.data(), nullptr);
  THCGW_HIP(hipDeviceSynchronize());
}

Mlp::~Mlp() {
  for (void* p : {static_cast<void*>(w), static_cast<void*>(wb), static_cast<void*>(m), static_cast<void*>(v),
                  static_cast<void*>(g), static_cast<void*>(loss), static_cast<void*>(act)})
    if (p) (void)hipFree(p);
}

void Mlp::set_params(const float* host, hipStream_t s) {
  THCGW_HIP(hipMemcpyAsync(w, host, sizeof(float) * n_params, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(mirror_kernel, dim3(blocks(n_params, 256)), dim3(256), 0, s, w, wb, n_params);
  THCGW_HIP(hipGetLastError());
  THCGW_HIP(hipMemsetAsync(m, 0, sizeof(float) * n_params, s));
  THCGW_HIP(hipMemsetAsync(v, 0, sizeof(float) * n_params, s));
  step = 0;
}

void Mlp::get_params(float* host, hipStream_t s) const {
  THCGW_HIP(hipMemcpyAsync(host, w, sizeof(float) * n_params, hipMemcpyDeviceToHost, s));
  THCGW_HIP(hipStreamSynchronize(s));
}

float* Mlp::acts(int batch) {
  // h0pre, h0, h1pre, h1 (batch x width each) and two batch x max(width, n_out) gradients
  const long wmax = std::max(width, n_out);
  const long need = static_cast<long>(batch) * (4L * width + 2L * wmax);
  if (need > act_count) {
    if (act) THCGW_HIP(hipFree(act));
    act = nullptr;
    act_count = 0;
    THCGW_HIP(hipMalloc(&act, sizeof(float) * need));
    act_count = need;
  }
  return act;
}

void Mlp::forward(hipStream_t s, int batch, const float* X, float* Y) {
  THCGW_REQUIRE(batch > 0, "mlp: batch must be positive");
  float* a = acts(batch);
  const long bw = static_cast<long>(batch) * width;
  float *h0p = a, *h0 = a + bw, *h1p = a + 2 * bw, *h1 = a + 3 * bw;
  gemm<float, __bf16>(s, 1, MlpGemm{batch, width, n_in, X, n_in, 1, wb + oW0, n_in, 1, h0p, w + ob0, h0});
  gemm<float, __bf16>(s, 1, MlpGemm{batch, width, width, h0, width, 1, wb + oW1, width, 1, h1p, w + ob1, h1});
  gemm<float, __bf16>(s, 2, MlpGemm{batch, n_out, width, h1, width, 1, wb + oW2, width, 1, Y, w + ob2, nullptr});
}

// Gradients of sum(Y * dY) with respect to every parameter, into g (the activations of the
// last forward() on the same X are reused).
void Mlp::backward(hipStream_t s, int batch, const float* X, const float* dY) {
  float* a = acts(batch);
  const long bw = static_cast<long>(batch) * width;
  const long wmax = std::max(width, n_out);
  float *h0p = a, *h0 = a + bw, *h1p = a + 2 * bw, *h1 = a + 3 * bw;
  float* d1 = a + 4 * bw;                                // batch x width
  float* d2 = d1 + static_cast<long>(batch) * wmax;      // batch x width
  // layer 2: dW2 = dY^T h1, db2 = colsum dY, dh1 = dY W2
  gemm<float, float>(s, 0, MlpGemm{n_out, width, batch, dY, 1, n_out, h1, 1, width, g + oW2, nullptr, nullptr});
  hipLaunchKernelGGL(colsum_kernel, dim3(blocks(n_out, 256)), dim3(256), 0, s, dY, batch, n_out, g + ob2);
  gemm<float, __bf16>(s, 0, MlpGemm{batch, width, n_out, dY, n_out, 1, wb + oW2, 1, width, d1, nullptr, nullptr});
  hipLaunchKernelGGL(silu_back_kernel, dim3(blocks(bw, 256)), dim3(256), 0, s, h1p, d1, d1, bw);
  // layer 1
  gemm<float, float>(s, 0, MlpGemm{width, width, batch, d1, 1, width, h0, 1, width, g + oW1, nullptr, nullptr});
  hipLaunchKernelGGL(colsum_kernel, dim3(blocks(width, 256)), dim3(256), 0, s, d1, batch, width, g + ob1);
  gemm<float, __bf16>(s, 0, MlpGemm{batch, width, width, d1, width, 1, wb + oW1, 1, width, d2, nullptr, nullptr});
  hipLaunchKernelGGL(silu_back_kernel, dim3(blocks(bw, 256)), dim3(256), 0, s, h0p, d2, d2, bw);
  // layer 0
  gemm<float, float>(s, 0, MlpGemm{width, n_in, batch, d2, 1, width, X, 1, n_in, g + oW0, nullptr, nullptr});
  hipLaunchKernelGGL(colsum_kernel, dim3(blocks(width, 256)), dim3(256), 0, s, d2, batch, width, g + ob0);
  THCGW_HIP(hipGetLastError());
}

void Mlp::adam(hipStream_t s, const float* grad, float lr) {
  ++step;
  const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
  const double c1 = 1.0 - std::pow(b1, static_cast<double>(step));
  const double c2 = 1.0 - std::pow(b2, static_cast<double>(step));
  // w -= lr m/c1 / (sqrt(v/c2) + eps)  ==  w -= (lr sqrt(c2)/c1) m / (sqrt(v) + eps sqrt(c2))
  const float lr_t = static_cast<float>(lr * std::sqrt(c2) / c1);
  const float eps_t = static_cast<float>(eps * std::sqrt(c2));
  hipLaunchKernelGGL(adam_kernel, dim3(blocks(n_params, 256)), dim3(256), 0, s, w, wb, m, v, grad ? grad : g,
                     n_params, static_cast<float>(b1), static_cast<float>(b2), lr_t, eps_t);
  THCGW_HIP(hipGetLastError());
}

// One Adam step on the mean squared error of the batch; the loss (before the step) is left
// in loss[0] on the device.
void Mlp::train_mse(hipStream_t s, int batch, const float* X, const float* T, float lr, float* Ybuf, float* dYbuf) {
  forward(s, batch, X, Ybuf);
  const long n = static_cast<long>(batch) * n_out;
  const int nb = static_cast<int>(std::min<long>(1023, (n + LT - 1) / LT));
  hipLaunchKernelGGL(mse_kernel, dim3(nb), dim3(LT), 0, s, Ybuf, T, n, dYbuf, loss + 1);
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(LT), 0, s, loss + 1, nb, 1.0f / static_cast<float>(n), loss);
  THCGW_HIP(hipGetLastError());
  backward(s, batch, X, dYbuf);
  adam(s, nullptr, lr);
}

}  // namespace thcgw
