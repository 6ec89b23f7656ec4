// Three-layer perceptron (BF16 weights, FP32 accumulation) used by the learned Krylov
// Dyson solve; the kernels are in mlp.hip.
#pragma once

#include <hip/hip_runtime.h>

namespace thcgw {

struct Mlp {
  int n_in, width, n_out;
  long oW0 = 0, ob0 = 0, oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, n_params = 0;
  float* w = nullptr;    // FP32 master parameters
  __bf16* wb = nullptr;  // BF16 mirror read by the products
  float* m = nullptr;    // Adam moments
  float* v = nullptr;
  float* g = nullptr;    // gradients of the last backward()
  float* loss = nullptr; // loss[0]: last MSE; loss[1..]: partial sums
  float* act = nullptr;  // activations of the last forward() and backward scratch
  long act_count = 0;
  long step = 0;

  Mlp(int n_in, int width, int n_out, unsigned long long seed);
  ~Mlp();
  Mlp(const Mlp&) = delete;
  Mlp& operator=(const Mlp&) = delete;

  void set_params(const float* host, hipStream_t s);  // also resets the Adam state
  void get_params(float* host, hipStream_t s) const;
  void forward(hipStream_t s, int batch, const float* X, float* Y);
  void backward(hipStream_t s, int batch, const float* X, const float* dY);
  void adam(hipStream_t s, const float* grad, float lr);
  void train_mse(hipStream_t s, int batch, const float* X, const float* T, float lr, float* Ybuf, float* dYbuf);

 private:
  float* acts(int batch);
};

}  // namespace thcgw
