// mlp.hpp -- the planner's cost-model network and its device kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "engine.hpp"

namespace qsvh {

// kernels_mlp.hip
hipError_t mlp_gemm(const float* A, const float* B, const float* bias, float* C, int M, int N,
                    int K, bool ta, bool tb, bool relu, hipStream_t s);
hipError_t mlp_relu_back(float* d, const float* z, int n, hipStream_t s);
hipError_t mlp_col_sum(const float* d, float* out, int M, int N, hipStream_t s);
hipError_t mlp_mse(const float* p, const float* y, float* g, float* loss, int M, hipStream_t s);
hipError_t mlp_adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1,
                    float b2, float eps, int step, hipStream_t s);

int set_err(Ctx* ctx, int code, const std::string& msg);
int hip_err(Ctx* ctx, hipError_t e, const char* where);

// features -> H -> H -> 1 with ReLU after the two hidden layers.
struct Mlp {
  Ctx* ctx = nullptr;
  int F = 12, H = 64;
  float* w = nullptr;  // parameters
  float* g = nullptr;  // gradients
  float* m = nullptr;  // Adam first moments
  float* v = nullptr;  // Adam second moments
  int step = 0;
  float y_mean = 0.f, y_std = 1.f;  // target standardisation
  // per-row workspace
  size_t rows_cap = 0;
  float *x = nullptr, *y = nullptr, *a1 = nullptr, *a2 = nullptr, *p_out = nullptr;
  float *dp = nullptr, *d2 = nullptr, *d1 = nullptr, *loss = nullptr;

  ~Mlp();
  size_t n_params() const;
  float* w1() const { return w; }
  float* b1() const { return w + F * H; }
  float* w2() const { return b1() + H; }
  float* b2() const { return w2() + H * H; }
  float* w3() const { return b2() + H; }
  float* b3() const { return w3() + H; }
  int init(uint64_t seed);
  hipError_t reserve(size_t rows);
  void release_rows();
  hipError_t forward(size_t rows);
  hipError_t backward(size_t rows);
  int predict(const float* hx, size_t rows, float* out, bool raw);
  int gradients(const float* hx, const float* hy, size_t rows, float* grads, float* loss_out);
  int train(const float* hx, const float* hy, size_t rows, uint32_t epochs, float lr,
            float* final_loss);
};

}  // namespace qsvh

struct qsvh_mlp {
  qsvh::Mlp m;
};
