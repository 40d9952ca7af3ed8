// kernels_compose.hip -- composition of fused gate groups in f64 on the device.
//
// build_fused (crates/qsv-core/src/fusion.rs:102-139) starts from the identity on the
// group's sorted qubits and left-multiplies each member gate embedded at its local
// positions: new[r][c] = sum over k_sub (ascending, zero entries skipped) of
// g[r_sub][k_sub] * mat[k][c], where k is r with the gate's local bits set to k_sub.
// Here one workgroup owns one group: the 2^k x 2^k product lives in LDS (64 KiB at
// k = 6) and each gate is applied in place to every column, one lane per (column, block
// of 2^m rows), which is exactly that update.  Same operations in the same order and no
// contraction (-ffp-contract=off), so the result is bit-identical to build_fused in f64.
#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace qsvh {
namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_compose(const ComposeGroup* __restrict__ groups,
                                                    const ComposeGate* __restrict__ gates,
                                                    const double* __restrict__ mats,
                                                    double* __restrict__ out) {
  extern __shared__ double lds[];
  const ComposeGroup grp = groups[blockIdx.x];
  const uint32_t D = 1u << grp.k;
  double* mr = lds;          // [r * D + c]
  double* mi = lds + D * D;  // imaginary parts
  for (uint32_t e = threadIdx.x; e < D * D; e += kBlock) {
    mr[e] = (e / D == e % D) ? 1.0 : 0.0;
    mi[e] = 0.0;
  }
  __syncthreads();
  for (uint32_t gi = 0; gi < grp.n_gates; ++gi) {
    const ComposeGate g = gates[grp.first_gate + gi];
    const uint32_t sub = 1u << g.m;
    const uint32_t nblk = D >> g.m;  // row blocks per column
    // sorted local positions for enumerating the block anchors
    uint32_t srt[kMaxGateQubits];
    for (uint32_t j = 0; j < g.m; ++j) srt[j] = g.local[j];
    for (uint32_t a = 1; a < g.m; ++a)
      for (uint32_t b = a; b > 0 && srt[b - 1] > srt[b]; --b) {
        const uint32_t t = srt[b];
        srt[b] = srt[b - 1];
        srt[b - 1] = t;
      }
    const double* gr = mats + 2 * g.mat_off;
    for (uint32_t w = threadIdx.x; w < nblk * D; w += kBlock) {
      const uint32_t c = w % D, blk = w / D;
      uint32_t base = blk;
      for (uint32_t j = 0; j < g.m; ++j)
        base = ((base >> srt[j]) << (srt[j] + 1)) | (base & ((1u << srt[j]) - 1));
      double xr[64], xi[64];
      for (uint32_t s = 0; s < sub; ++s) {
        uint32_t r = base;
        for (uint32_t j = 0; j < g.m; ++j) r |= ((s >> j) & 1u) << g.local[j];
        xr[s] = mr[r * D + c];
        xi[s] = mi[r * D + c];
      }
      for (uint32_t rs = 0; rs < sub; ++rs) {
        double ar = 0.0, ai = 0.0;
        for (uint32_t ks = 0; ks < sub; ++ks) {
          const double er = gr[2 * (rs * sub + ks)], ei = gr[2 * (rs * sub + ks) + 1];
          if (er == 0.0 && ei == 0.0) continue;
          const double pr = er * xr[ks] - ei * xi[ks];
          const double pi = er * xi[ks] + ei * xr[ks];
          ar = ar + pr;
          ai = ai + pi;
        }
        uint32_t r = base;
        for (uint32_t j = 0; j < g.m; ++j) r |= ((rs >> j) & 1u) << g.local[j];
        mr[r * D + c] = ar;
        mi[r * D + c] = ai;
      }
    }
    __syncthreads();
  }
  double* o = out + 2 * grp.out_off;
  for (uint32_t e = threadIdx.x; e < D * D; e += kBlock) {
    o[2 * e] = mr[e];
    o[2 * e + 1] = mi[e];
  }
}

}  // namespace

int compose_groups(Ctx* ctx, const std::vector<ComposeGroup>& groups,
                   const std::vector<ComposeGate>& gates, const std::vector<double>& mats,
                   std::vector<double>* out) {
  if (groups.empty()) return kOk;
  uint32_t kmax = 0;
  uint64_t out_len = 0;
  for (const ComposeGroup& g : groups) {
    kmax = g.k > kmax ? g.k : kmax;
    out_len = g.out_off + (uint64_t(1) << (2 * g.k));
  }
  const size_t shmem = 2 * sizeof(double) * (size_t(1) << (2 * kmax));
  ComposeGroup* d_groups = nullptr;
  ComposeGate* d_gates = nullptr;
  double* d_mats = nullptr;
  double* d_out = nullptr;
  hipStream_t s = ctx->stream;
  auto fail = [&](hipError_t e) {
    ctx->last_error = std::string("compose: ") + hipGetErrorString(e);
    hipFree(d_groups);
    hipFree(d_gates);
    hipFree(d_mats);
    hipFree(d_out);
    return e == hipErrorOutOfMemory ? kOom : kHip;
  };
  hipError_t e;
  if ((e = hipMalloc(&d_groups, groups.size() * sizeof(ComposeGroup))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&d_gates, (gates.size() + 1) * sizeof(ComposeGate))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&d_mats, (mats.size() + 1) * sizeof(double))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&d_out, 2 * out_len * sizeof(double))) != hipSuccess) return fail(e);
  if ((e = hipMemcpyAsync(d_groups, groups.data(), groups.size() * sizeof(ComposeGroup),
                          hipMemcpyHostToDevice, s)) != hipSuccess)
    return fail(e);
  if (!gates.empty() &&
      (e = hipMemcpyAsync(d_gates, gates.data(), gates.size() * sizeof(ComposeGate),
                          hipMemcpyHostToDevice, s)) != hipSuccess)
    return fail(e);
  if (!mats.empty() && (e = hipMemcpyAsync(d_mats, mats.data(), mats.size() * sizeof(double),
                                           hipMemcpyHostToDevice, s)) != hipSuccess)
    return fail(e);
  if ((e = hipFuncSetAttribute(reinterpret_cast<const void*>(k_compose),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem))) !=
      hipSuccess)
    return fail(e);
  k_compose<<<unsigned(groups.size()), kBlock, shmem, s>>>(d_groups, d_gates, d_mats, d_out);
  if ((e = hipGetLastError()) != hipSuccess) return fail(e);
  out->assign(2 * out_len, 0.0);
  if ((e = hipMemcpyAsync(out->data(), d_out, 2 * out_len * sizeof(double),
                          hipMemcpyDeviceToHost, s)) != hipSuccess)
    return fail(e);
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail(e);
  hipFree(d_groups);
  hipFree(d_gates);
  hipFree(d_mats);
  hipFree(d_out);
  return kOk;
}

}  // namespace qsvh
