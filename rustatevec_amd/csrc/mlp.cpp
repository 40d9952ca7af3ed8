// mlp.cpp -- the planner's cost model: a fully connected network
// (features -> hidden -> hidden -> 1, ReLU) in FP32, trained on the device with the
// kernels of kernels_mlp.hip (forward, backward, Adam on mean squared error).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "../../include/qsvh.h"
#include "handles.hpp"
#include "mlp.hpp"

namespace qsvh {

namespace {

template <typename X>
hipError_t dalloc(X** p, size_t n) {
  return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(n, 1) * sizeof(X));
}

}  // namespace

Mlp::~Mlp() {
  release_rows();
  hipFree(w);
  hipFree(g);
  hipFree(m);
  hipFree(v);
}

void Mlp::release_rows() {
  for (float* p : {x, y, a1, a2, p_out, dp, d2, d1, loss}) hipFree(p);
  x = y = a1 = a2 = p_out = dp = d2 = d1 = loss = nullptr;
  rows_cap = 0;
}

hipError_t Mlp::reserve(size_t rows) {
  if (rows <= rows_cap) return hipSuccess;
  release_rows();
  hipError_t e = dalloc(&x, rows * F);
  if (e == hipSuccess) e = dalloc(&y, rows);
  if (e == hipSuccess) e = dalloc(&a1, rows * H);
  if (e == hipSuccess) e = dalloc(&a2, rows * H);
  if (e == hipSuccess) e = dalloc(&p_out, rows);
  if (e == hipSuccess) e = dalloc(&dp, rows);
  if (e == hipSuccess) e = dalloc(&d2, rows * H);
  if (e == hipSuccess) e = dalloc(&d1, rows * H);
  if (e == hipSuccess) e = dalloc(&loss, (rows + 255) / 256);
  if (e == hipSuccess) rows_cap = rows;
  return e;
}

// Parameter layout: W1[F x H], b1[H], W2[H x H], b2[H], W3[H x 1], b3[1] (row-major).
size_t Mlp::n_params() const { return size_t(F) * H + H + size_t(H) * H + H + H + 1; }

hipError_t Mlp::forward(size_t rows) {
  hipStream_t s = ctx->stream;
  const int M = int(rows);
  hipError_t e = mlp_gemm(x, w1(), b1(), a1, M, H, F, false, false, true, s);
  if (e == hipSuccess) e = mlp_gemm(a1, w2(), b2(), a2, M, H, H, false, false, true, s);
  if (e == hipSuccess) e = mlp_gemm(a2, w3(), b3(), p_out, M, 1, H, false, false, false, s);
  return e;
}

hipError_t Mlp::backward(size_t rows) {
  hipStream_t s = ctx->stream;
  const int M = int(rows);
  float* gw1 = g;
  float* gb1 = gw1 + F * H;
  float* gw2 = gb1 + H;
  float* gb2 = gw2 + H * H;
  float* gw3 = gb2 + H;
  float* gb3 = gw3 + H;
  hipError_t e = mlp_gemm(a2, dp, nullptr, gw3, H, 1, M, true, false, false, s);  // A2^T dP
  if (e == hipSuccess) e = mlp_col_sum(dp, gb3, M, 1, s);
  if (e == hipSuccess) e = mlp_gemm(dp, w3(), nullptr, d2, M, H, 1, false, true, false, s);
  if (e == hipSuccess) e = mlp_relu_back(d2, a2, M * H, s);
  if (e == hipSuccess) e = mlp_gemm(a1, d2, nullptr, gw2, H, H, M, true, false, false, s);
  if (e == hipSuccess) e = mlp_col_sum(d2, gb2, M, H, s);
  if (e == hipSuccess) e = mlp_gemm(d2, w2(), nullptr, d1, M, H, H, false, true, false, s);
  if (e == hipSuccess) e = mlp_relu_back(d1, a1, M * H, s);
  if (e == hipSuccess) e = mlp_gemm(x, d1, nullptr, gw1, F, H, M, true, false, false, s);
  if (e == hipSuccess) e = mlp_col_sum(d1, gb1, M, H, s);
  return e;
}

int Mlp::init(uint64_t seed) {
  const size_t np = n_params();
  hipError_t e = dalloc(&w, np);
  if (e == hipSuccess) e = dalloc(&g, np);
  if (e == hipSuccess) e = dalloc(&m, np);
  if (e == hipSuccess) e = dalloc(&v, np);
  if (e != hipSuccess) return hip_err(ctx, e, "mlp alloc");
  // He-uniform weights (limit sqrt(6 / fan_in)), zero biases, from SplitMix64(seed).
  std::vector<float> h(np, 0.f);
  SplitMix64 rng(seed);
  auto fill = [&](size_t off, size_t count, int fan_in) {
    const double lim = std::sqrt(6.0 / fan_in);
    for (size_t i = 0; i < count; ++i) h[off + i] = float((2.0 * rng.unit() - 1.0) * lim);
  };
  fill(0, size_t(F) * H, F);
  fill(size_t(F) * H + H, size_t(H) * H, H);
  fill(size_t(F) * H + H + size_t(H) * H + H, H, H);
  e = hipMemcpy(w, h.data(), np * sizeof(float), hipMemcpyHostToDevice);
  std::vector<float> z(np, 0.f);
  if (e == hipSuccess) e = hipMemcpy(m, z.data(), np * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(v, z.data(), np * sizeof(float), hipMemcpyHostToDevice);
  step = 0;
  return hip_err(ctx, e, "mlp init");
}

int Mlp::predict(const float* hx, size_t rows, float* out, bool raw) {
  hipError_t e = reserve(rows);
  if (e == hipSuccess) e = hipMemcpyAsync(x, hx, rows * F * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = forward(rows);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, p_out, rows * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_err(ctx, e, "mlp predict");
  if (!raw)
    for (size_t i = 0; i < rows; ++i) out[i] = out[i] * y_std + y_mean;
  return kOk;
}

int Mlp::gradients(const float* hx, const float* hy, size_t rows, float* grads, float* loss_out) {
  hipError_t e = reserve(rows);
  if (e == hipSuccess) e = hipMemcpyAsync(x, hx, rows * F * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(y, hy, rows * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = forward(rows);
  if (e == hipSuccess) e = mlp_mse(p_out, y, dp, loss, int(rows), ctx->stream);
  if (e == hipSuccess) e = backward(rows);
  const size_t nl = (rows + 255) / 256;
  std::vector<float> lp(nl);
  if (e == hipSuccess && grads)
    e = hipMemcpyAsync(grads, g, n_params() * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(lp.data(), loss, nl * sizeof(float), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_err(ctx, e, "mlp gradients");
  if (loss_out) {
    double s = 0;
    for (float f : lp) s += f;
    *loss_out = float(s);
  }
  return kOk;
}

int Mlp::train(const float* hx, const float* hy, size_t rows, uint32_t epochs, float lr,
               float* final_loss) {
  if (rows == 0) return set_err(ctx, kInvalidArg, "mlp train: no rows");
  // standardise the targets; the network learns (y - mean) / std
  double mean = 0, var = 0;
  for (size_t i = 0; i < rows; ++i) mean += hy[i];
  mean /= double(rows);
  for (size_t i = 0; i < rows; ++i) var += (hy[i] - mean) * (hy[i] - mean);
  const double sd = std::sqrt(var / double(rows)) + 1e-6;
  y_mean = float(mean);
  y_std = float(sd);
  std::vector<float> ys(rows);
  for (size_t i = 0; i < rows; ++i) ys[i] = float((hy[i] - mean) / sd);
  hipError_t e = reserve(rows);
  if (e == hipSuccess) e = hipMemcpyAsync(x, hx, rows * F * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(y, ys.data(), rows * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
  for (uint32_t ep = 0; ep < epochs && e == hipSuccess; ++ep) {
    e = forward(rows);
    if (e == hipSuccess) e = mlp_mse(p_out, y, dp, loss, int(rows), ctx->stream);
    if (e == hipSuccess) e = backward(rows);
    if (e == hipSuccess)
      e = mlp_adam(w, g, m, v, int(n_params()), lr, 0.9f, 0.999f, 1e-8f, ++step, ctx->stream);
  }
  if (e != hipSuccess) return hip_err(ctx, e, "mlp train");
  if (final_loss) {
    float l = 0;
    const int rc = gradients(hx, ys.data(), rows, nullptr, &l);
    if (rc != kOk) return rc;
    *final_loss = l;  // on the standardised targets
  }
  return kOk;
}

}  // namespace qsvh

using namespace qsvh;

extern "C" {

int qsvh_mlp_create(qsvh_ctx* h, uint32_t n_features, uint32_t hidden, uint64_t seed,
                    qsvh_mlp** out) {
  if (!h || !out || n_features == 0 || n_features > 64 || hidden == 0 || hidden > 256)
    return kInvalidArg;
  *out = nullptr;
  qsvh_mlp* p = new (std::nothrow) qsvh_mlp;
  if (!p) return kOom;
  p->m.ctx = &h->c;
  p->m.F = int(n_features);
  p->m.H = int(hidden);
  const int rc = p->m.init(seed);
  if (rc != kOk) {
    delete p;
    return rc;
  }
  *out = p;
  return kOk;
}

void qsvh_mlp_destroy(qsvh_mlp* p) {
  if (!p) return;
  hipStreamSynchronize(p->m.ctx->stream);
  delete p;
}

int qsvh_mlp_num_params(const qsvh_mlp* p, size_t* n) {
  if (!p || !n) return kInvalidArg;
  *n = p->m.n_params();
  return kOk;
}

int qsvh_mlp_get_params(qsvh_mlp* p, float* out) {
  if (!p || !out) return kInvalidArg;
  hipStreamSynchronize(p->m.ctx->stream);
  return hip_err(p->m.ctx, hipMemcpy(out, p->m.w, p->m.n_params() * sizeof(float),
                                     hipMemcpyDeviceToHost), "mlp get params");
}

int qsvh_mlp_set_params(qsvh_mlp* p, const float* in) {
  if (!p || !in) return kInvalidArg;
  hipStreamSynchronize(p->m.ctx->stream);
  return hip_err(p->m.ctx, hipMemcpy(p->m.w, in, p->m.n_params() * sizeof(float),
                                     hipMemcpyHostToDevice), "mlp set params");
}

int qsvh_mlp_forward(qsvh_mlp* p, const float* x, size_t rows, float* out) {
  if (!p || !x || !out || rows == 0) return kInvalidArg;
  return p->m.predict(x, rows, out, true);
}

int qsvh_mlp_predict(qsvh_mlp* p, const float* x, size_t rows, float* out) {
  if (!p || !x || !out || rows == 0) return kInvalidArg;
  return p->m.predict(x, rows, out, false);
}

int qsvh_mlp_gradients(qsvh_mlp* p, const float* x, const float* y, size_t rows, float* grads,
                       float* loss) {
  if (!p || !x || !y || rows == 0) return kInvalidArg;
  return p->m.gradients(x, y, rows, grads, loss);
}

int qsvh_mlp_train(qsvh_mlp* p, const float* x, const float* y, size_t rows, uint32_t epochs,
                   float lr, float* final_loss) {
  if (!p || !x || !y) return kInvalidArg;
  return p->m.train(x, y, rows, epochs, lr, final_loss);
}

int qsvh_mlp_target_scale(const qsvh_mlp* p, float* mean, float* stddev) {
  if (!p) return kInvalidArg;
  if (mean) *mean = p->m.y_mean;
  if (stddev) *stddev = p->m.y_std;
  return kOk;
}

}  // extern "C"
