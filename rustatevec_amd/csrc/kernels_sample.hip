// kernels_sample.hip -- measurement sampling on the device (Backend::sample,
// crates/qsv-core/src/backend/mod.rs:52-54), by inverse CDF as in cdf_sample
// (crates/qsv-core/src/sample.rs:140-153):
//
//  1. k_scan_local: p_i = |psi_i|^2 (formed in the state's precision, then widened to f64 as
//     StateVector::probabilities + to_f64 do), inclusive f64 scan within chunks of 2048;
//     chunk totals aside.
//  2. k_scan_offsets: one lane adds the chunk totals in order: offset[c + 1] = offset[c] +
//     total[c], so the global CDF cdf_i = offset[c(i)] + local_i is non-decreasing and its
//     last value of chunk c is exactly offset[c + 1].
//  3. k_draw: one lane per shot: u = unit_k * cdf_last with unit_k the k-th unit() of
//     SplitMix64(seed) (counter form: output k = mix(seed + (k + 1) gamma)), outcome =
//     min(first i with cdf_i >= u, dim - 1) -- partition_point(|c| c < u) -- by binary search
//     over the chunk ends, then inside the chunk.
//
// The scan adds in a different order than the reference's serial loop, so an outcome can
// differ from the CPU's only when u falls within rounding distance of a CDF boundary.
#include <hip/hip_runtime.h>

#include "amp.hpp"
#include "engine.hpp"

namespace qsvh {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr uint32_t kScanChunk = kScanThreads * kScanPer;  // 2048 amplitudes per chunk

__device__ __forceinline__ uint64_t splitmix_output(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_scan_local(const T* __restrict__ amp,
                                                             uint64_t dim,
                                                             double* __restrict__ local,
                                                             double* __restrict__ totals) {
  __shared__ double sp[kScanChunk];
  __shared__ double wsum[kScanThreads / 64];
  const uint64_t c0 = uint64_t(blockIdx.x) * kScanChunk;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    const uint32_t e = uint32_t(k) * kScanThreads + uint32_t(tid);
    const uint64_t i = c0 + e;
    double p = 0.0;
    if (i < dim) {
      const T r = amp[2 * i], m = amp[2 * i + 1];
      p = double(T(r * r + m * m));
    }
    sp[e] = p;
  }
  __syncthreads();
  double v[kScanPer];
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < kScanPer; ++j) {
    acc += sp[kScanPer * tid + j];
    v[j] = acc;
  }
  // exclusive prefix of the thread sums over the workgroup
  double x = acc;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  double ex = __shfl_up(x, 1, 64);
  if (lane == 0) ex = 0.0;
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  double woff = 0.0;
  for (int w = 0; w < wave; ++w) woff += wsum[w];
  const double base = woff + ex;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kScanPer; ++j) sp[kScanPer * tid + j] = base + v[j];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    const uint32_t e = uint32_t(k) * kScanThreads + uint32_t(tid);
    if (c0 + e < dim) local[c0 + e] = sp[e];
  }
  if (tid == kScanThreads - 1) {
    const uint64_t len = dim - c0 < kScanChunk ? dim - c0 : kScanChunk;
    totals[blockIdx.x] = sp[len - 1];
  }
}

// offset[0] = 0, offset[c + 1] = offset[c] + totals[c] (nc + 1 entries), in chunk order.
__global__ void k_scan_offsets(const double* __restrict__ totals, uint64_t nc,
                               double* __restrict__ offset) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double acc = 0.0;
  offset[0] = 0.0;
  for (uint64_t c = 0; c < nc; ++c) {
    acc = acc + totals[c];
    offset[c + 1] = acc;
  }
}

__global__ __launch_bounds__(256) void k_draw(const double* __restrict__ local,
                                              const double* __restrict__ offset, uint64_t nc,
                                              uint64_t dim, uint64_t seed, uint64_t shots,
                                              uint64_t* __restrict__ out) {
  const uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= shots) return;
  const double total = offset[nc];
  const double unit = double(splitmix_output(seed, k) >> 11) * (1.0 / 9007199254740992.0);
  const double u = unit * total;
  // first chunk whose last CDF value (offset[c + 1]) is >= u
  uint64_t lo = 0, hi = nc;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (offset[mid + 1] < u)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo >= nc) {  // u above every value: partition_point = dim, clamped
    out[k] = dim - 1;
    return;
  }
  const uint64_t c0 = lo * kScanChunk;
  const uint64_t len = dim - c0 < kScanChunk ? dim - c0 : kScanChunk;
  const double off = offset[lo];
  uint64_t a = 0, b = len;
  while (a < b) {
    const uint64_t mid = a + (b - a) / 2;
    if (off + local[c0 + mid] < u)
      a = mid + 1;
    else
      b = mid;
  }
  const uint64_t idx = c0 + a;
  out[k] = idx < dim - 1 ? idx : dim - 1;
}

}  // namespace

uint64_t sample_chunks(uint64_t dim) { return (dim + kScanChunk - 1) / kScanChunk; }

hipError_t launch_sample(const State& st, uint64_t shots, uint64_t seed, double* local,
                         double* totals, double* offset, uint64_t* out, hipStream_t s) {
  const uint64_t dim = st.dim();
  const uint64_t nc = sample_chunks(dim);
  if (st.dtype == 0)
    k_scan_local<double><<<unsigned(nc), kScanThreads, 0, s>>>(
        static_cast<const double*>(st.amp), dim, local, totals);
  else
    k_scan_local<float><<<unsigned(nc), kScanThreads, 0, s>>>(
        static_cast<const float*>(st.amp), dim, local, totals);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_scan_offsets<<<1, 64, 0, s>>>(totals, nc, offset);
  e = hipGetLastError();
  if (e != hipSuccess || shots == 0) return e;
  const uint64_t g = (shots + 255) / 256;
  k_draw<<<unsigned(g), 256, 0, s>>>(local, offset, nc, dim, seed, shots, out);
  return hipGetLastError();
}

}  // namespace qsvh
