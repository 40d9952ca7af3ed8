// kernels_mlp.hip -- forward, backward and Adam kernels of the planner's cost model (a
// small fully connected network in FP32).  Shapes are tiny (rows = training samples or
// candidate groups, widths <= 64), so a 16 x 16 LDS-tiled GEMM with fused bias/ReLU
// epilogue is enough; every row-major operand may be read transposed.
#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace qsvh {
namespace {

constexpr int kT = 16;

// C[M x N] = op(A) op(B) (+ bias[col]) (ReLU if relu); op(A) is M x K: A is M x K
// row-major, or K x M when ta; op(B) is K x N: B is K x N, or N x K when tb.
__global__ __launch_bounds__(kT * kT) void k_gemm(const float* __restrict__ A,
                                                  const float* __restrict__ B,
                                                  const float* __restrict__ bias,
                                                  float* __restrict__ C, int M, int N, int K,
                                                  int ta, int tb, int relu) {
  __shared__ float as[kT][kT + 1], bs[kT][kT + 1];
  const int tx = threadIdx.x % kT, ty = threadIdx.x / kT;
  const int row = blockIdx.y * kT + ty, col = blockIdx.x * kT + tx;
  float acc = 0.f;
  for (int k0 = 0; k0 < K; k0 += kT) {
    const int ka = k0 + tx, kb = k0 + ty;
    as[ty][tx] = (row < M && ka < K) ? (ta ? A[size_t(ka) * M + row] : A[size_t(row) * K + ka]) : 0.f;
    bs[ty][tx] = (kb < K && col < N) ? (tb ? B[size_t(col) * K + kb] : B[size_t(kb) * N + col]) : 0.f;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kT; ++k) acc += as[ty][k] * bs[k][tx];
    __syncthreads();
  }
  if (row < M && col < N) {
    if (bias) acc += bias[col];
    if (relu) acc = acc > 0.f ? acc : 0.f;
    C[size_t(row) * N + col] = acc;
  }
}

// dZ = dA * (Z > 0): the ReLU backward, in place on dA.
__global__ void k_relu_back(float* __restrict__ d, const float* __restrict__ z, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !(z[i] > 0.f)) d[i] = 0.f;
}

// db[c] = sum over rows of dZ[r][c] (fixed order).
__global__ void k_col_sum(const float* __restrict__ d, float* __restrict__ out, int M, int N) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  float s = 0.f;
  for (int r = 0; r < M; ++r) s += d[size_t(r) * N + c];
  out[c] = s;
}

// Mean squared error over M predictions: g = 2 (p - y) / M; loss partial per block.
__global__ void k_mse(const float* __restrict__ p, const float* __restrict__ y,
                      float* __restrict__ g, float* __restrict__ loss, int M) {
  __shared__ float red[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float e = 0.f;
  if (i < M) {
    const float d = p[i] - y[i];
    g[i] = 2.f * d / float(M);
    e = d * d / float(M);
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba 2015) on a flat parameter vector; step is 1-based.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float b1, float b2, float eps,
                       int step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.f - powf(b1, float(step)));
  const float vh = vi / (1.f - powf(b2, float(step)));
  w[i] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace

hipError_t mlp_gemm(const float* A, const float* B, const float* bias, float* C, int M, int N,
                    int K, bool ta, bool tb, bool relu, hipStream_t s) {
  const dim3 grid((N + kT - 1) / kT, (M + kT - 1) / kT);
  k_gemm<<<grid, kT * kT, 0, s>>>(A, B, bias, C, M, N, K, ta, tb, relu);
  return hipGetLastError();
}

hipError_t mlp_relu_back(float* d, const float* z, int n, hipStream_t s) {
  k_relu_back<<<(n + 255) / 256, 256, 0, s>>>(d, z, n);
  return hipGetLastError();
}

hipError_t mlp_col_sum(const float* d, float* out, int M, int N, hipStream_t s) {
  k_col_sum<<<(N + 63) / 64, 64, 0, s>>>(d, out, M, N);
  return hipGetLastError();
}

// `loss` receives ceil(M / 256) partial sums.
hipError_t mlp_mse(const float* p, const float* y, float* g, float* loss, int M, hipStream_t s) {
  k_mse<<<(M + 255) / 256, 256, 0, s>>>(p, y, g, loss, M);
  return hipGetLastError();
}

hipError_t mlp_adam(float* w, const float* g, float* m, float* v, int n, float lr, float b1,
                    float b2, float eps, int step, hipStream_t s) {
  k_adam<<<(n + 255) / 256, 256, 0, s>>>(w, g, m, v, n, lr, b1, b2, eps, step);
  return hipGetLastError();
}

}  // namespace qsvh
