// kernels_sample.hip -- measurement sampling on the device (Backend::sample,
// crates/qsv-core/src/backend/mod.rs:52-54), by inverse CDF as in cdf_sample
// (crates/qsv-core/src/sample.rs:140-153):
//
//  1. k_scan_local: p_i = |psi_i|^2 (formed in the state's precision, then widened to f64 as
//     StateVector::probabilities + to_f64 do), inclusive f64 scan within chunks of 2048;
//     chunk totals aside.
//  2. k_scan_totals, k_scan_blocks, k_add_blocks: offset[c + 1] = total[0] + ... + total[c]
//     by a three-level scan (2048 totals per workgroup, then the workgroup sums, then the
//     sums added back), so the global CDF cdf_i = offset[c(i)] + local_i is non-decreasing
//     within each chunk and across chunk boundaries to within an ulp (the last value of
//     chunk c, offset[c] + total[c], and offset[c + 1] are the same sum in two orders).
//  3. k_draw: one lane per shot: u = unit_k * cdf_last with unit_k the k-th unit() of
//     SplitMix64(seed) (counter form: output k = mix(seed + (k + 1) gamma)), outcome =
//     min(first i with cdf_i >= u, dim - 1) -- partition_point(|c| c < u) -- by binary search
//     over the chunk ends, then inside the chunk.
//
// The scan adds in a different order than the reference's serial loop, so an outcome can
// differ from the CPU's only when u falls within rounding distance of a CDF boundary.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

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

// Inclusive scan over a workgroup of kScanThreads threads, thread t holding PER values in
// order (elements t PER .. t PER + PER - 1): v becomes the inclusive prefixes and *start the
// sum before the thread's first element; returns the workgroup total.  wsum:
// kScanThreads / 64 doubles of LDS.
template <int PER>
__device__ __forceinline__ double block_scan(double (&v)[PER], double* wsum, double* start) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    acc += v[j];
    v[j] = acc;
  }
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
  double woff = 0.0, total = 0.0;
  for (int w = 0; w < kScanThreads / 64; ++w) {
    if (w < wave) woff += wsum[w];
    total += wsum[w];
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) v[j] += woff + ex;
  *start = woff + ex;
  __syncthreads();
  return total;
}

// Level 1: offset[c + 1] = the inclusive scan of totals within each run of kScanChunk
// chunks; bsum[b] = the run's sum.
__global__ __launch_bounds__(kScanThreads) void k_scan_totals(const double* __restrict__ totals,
                                                              uint64_t nc,
                                                              double* __restrict__ offset,
                                                              double* __restrict__ bsum) {
  __shared__ double wsum[kScanThreads / 64];
  const uint64_t c0 = uint64_t(blockIdx.x) * kScanChunk + uint64_t(threadIdx.x) * kScanPer;
  double v[kScanPer];
#pragma unroll
  for (int j = 0; j < kScanPer; ++j) v[j] = c0 + j < nc ? totals[c0 + j] : 0.0;
  double start;
  const double total = block_scan<kScanPer>(v, wsum, &start);
#pragma unroll
  for (int j = 0; j < kScanPer; ++j)
    if (c0 + j < nc) offset[c0 + j + 1] = v[j];
  if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Level 2 (one workgroup): bsum becomes its exclusive prefix (nb <= kScanThreads * 16).
__global__ __launch_bounds__(kScanThreads) void k_scan_blocks(double* __restrict__ bsum,
                                                              uint64_t nb) {
  constexpr int PER = 16;
  __shared__ double wsum[kScanThreads / 64];
  const uint64_t b0 = uint64_t(threadIdx.x) * PER;
  double v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) v[j] = b0 + j < nb ? bsum[b0 + j] : 0.0;
  double start;
  block_scan<PER>(v, wsum, &start);
#pragma unroll
  for (int j = 0; j < PER; ++j)  // exclusive: the inclusive prefix of the element before
    if (b0 + j < nb) bsum[b0 + j] = j == 0 ? start : v[j - 1];
}

// Level 3: offset[c + 1] += the prefix of the runs before chunk c's; offset[0] = 0.
__global__ __launch_bounds__(kScanThreads) void k_add_blocks(const double* __restrict__ bsum,
                                                             uint64_t nc,
                                                             double* __restrict__ offset) {
  const uint64_t c = uint64_t(blockIdx.x) * kScanThreads + threadIdx.x;
  if (c == 0) offset[0] = 0.0;
  if (c < nc) offset[c + 1] += bsum[c / kScanChunk];
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

// Scratch of the allocation-free sampler (qsvh_sample_noalloc): the in-chunk CDF, the chunk
// totals, their offsets and the run sums of a state of up to kNoAllocMaxQubits qubits, laid
// out as launch_sample takes them.  Every kernel writes what it later reads (nothing has to
// start at zero).  Each GPU holds one copy for the life of the process; host code reaches it
// only through the address hipGetSymbolAddress returns for the current device.
constexpr uint64_t kNoAllocAmps = uint64_t(1) << kNoAllocMaxQubits;
constexpr uint64_t kNoAllocChunks = kNoAllocAmps / kScanChunk;
constexpr uint64_t kNoAllocDoubles =
    kNoAllocAmps + 2 * kNoAllocChunks + 1 + (kNoAllocChunks + kScanChunk - 1) / kScanChunk;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

}  // namespace

uint64_t sample_chunks(uint64_t dim) { return (dim + kScanChunk - 1) / kScanChunk; }

uint64_t sample_noalloc_doubles() { return kNoAllocDoubles; }

hipError_t sample_noalloc_scratch(int device, double** ptr) {
  // looked up once per device (each device holds its own copy at its own address); the
  // caller has made `device` current
  static std::mutex mu;
  static std::map<int, double*> addr;
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(device);
  if (it != addr.end()) {
    *ptr = it->second;
    return hipSuccess;
  }
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  addr[device] = static_cast<double*>(p);
  *ptr = static_cast<double*>(p);
  return hipSuccess;
}

uint64_t sample_blocks(uint64_t dim) {
  return (sample_chunks(dim) + kScanChunk - 1) / kScanChunk;
}

hipError_t launch_sample(const State& st, uint64_t shots, uint64_t seed, double* local,
                         double* totals, double* offset, double* bsum, uint64_t* out,
                         hipStream_t s) {
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
  const uint64_t nb = (nc + kScanChunk - 1) / kScanChunk;
  k_scan_totals<<<unsigned(nb), kScanThreads, 0, s>>>(totals, nc, offset, bsum);
  k_scan_blocks<<<1, kScanThreads, 0, s>>>(bsum, nb);
  k_add_blocks<<<unsigned((nc + kScanThreads - 1) / kScanThreads), kScanThreads, 0, s>>>(
      bsum, nc, offset);
  e = hipGetLastError();
  if (e != hipSuccess || shots == 0) return e;
  const uint64_t g = (shots + 255) / 256;
  k_draw<<<unsigned(g), 256, 0, s>>>(local, offset, nc, dim, seed, shots, out);
  return hipGetLastError();
}

}  // namespace qsvh
