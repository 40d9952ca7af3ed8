// kernels_util.hip -- device-side state preparation and state comparison.
//
// * Random normalised start state: psi_i = a_i + i b_i, (a_i, b_i) from outputs 2i and 2i+1
//   of SplitMix64(seed) (crates/qsv-core/src/circuits.rs:12-47) by Box-Muller.  The k-th
//   output of SplitMix64 is mix(seed + (k + 1) * gamma), so every amplitude is generated
//   independently (counter-based) yet equals what one sequential stream would give.
// * Comparison of two states in f64: max_i |a_i - b_i|, <a|b>, <a|a>, <b|b>.
// * The ABI's layout: amplitudes [first, first + count) of the interleaved state to / from
//   separate re and im arrays in device memory (staging of the host transfers).
#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace qsvh {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 512;

__device__ __forceinline__ uint64_t splitmix_output(uint64_t seed, uint64_t k) {
  uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_init_random(T* __restrict__ amp, uint64_t dim,
                                                        uint64_t seed) {
  const double inv53 = 1.0 / 9007199254740992.0;  // 2^-53
  const double two_pi = 6.28318530717958647692;
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < dim; i += nthr) {
    const double u1 = double(splitmix_output(seed, 2 * i) >> 11) * inv53;
    const double u2 = double(splitmix_output(seed, 2 * i + 1) >> 11) * inv53;
    const double r = sqrt(-2.0 * log(1.0 - u1));
    double s, c;
    sincos(two_pi * u2, &s, &c);
    amp[2 * i] = T(r * c);
    amp[2 * i + 1] = T(r * s);
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale(T* __restrict__ amp, uint64_t dim,
                                                  const double* __restrict__ norm2) {
  const double inv = 1.0 / sqrt(*norm2);
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < dim; i += nthr) {
    amp[2 * i] = T(double(amp[2 * i]) * inv);
    amp[2 * i + 1] = T(double(amp[2 * i + 1]) * inv);
  }
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}
__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = fmax(x, __shfl_xor(x, o, 64));
  return x;
}

// Five f64 partials per block: max |a - b|, Re <a|b>, Im <a|b>, <a|a>, <b|b>.
template <typename A, typename B>
__global__ __launch_bounds__(kBlock) void k_compare(const A* __restrict__ a,
                                                    const B* __restrict__ b, uint64_t dim,
                                                    double* __restrict__ part) {
  __shared__ double ws[5][kBlock / 64];
  double mx = 0, sr = 0, si = 0, na = 0, nb = 0;
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < dim; i += nthr) {
    const double xr = a[2 * i], xi = a[2 * i + 1], yr = b[2 * i], yi = b[2 * i + 1];
    const double dr = xr - yr, di = xi - yi;
    mx = fmax(mx, sqrt(dr * dr + di * di));
    sr += xr * yr + xi * yi;  // conj(a) * b
    si += xr * yi - xi * yr;
    na += xr * xr + xi * xi;
    nb += yr * yr + yi * yi;
  }
  mx = wave_max(mx);
  sr = wave_sum(sr);
  si = wave_sum(si);
  na = wave_sum(na);
  nb = wave_sum(nb);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    ws[0][w] = mx;
    ws[1][w] = sr;
    ws[2][w] = si;
    ws[3][w] = na;
    ws[4][w] = nb;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    double v = ws[threadIdx.x][0];
    for (int k = 1; k < kBlock / 64; ++k)
      v = threadIdx.x == 0 ? fmax(v, ws[0][k]) : v + ws[threadIdx.x][k];
    part[5 * blockIdx.x + threadIdx.x] = v;
  }
}

__global__ void k_compare_final(const double* __restrict__ part, int nblocks,
                                double* __restrict__ out) {
  if (threadIdx.x < 5) {
    double v = part[threadIdx.x];
    for (int b = 1; b < nblocks; ++b)
      v = threadIdx.x == 0 ? fmax(v, part[5 * b]) : v + part[5 * b + threadIdx.x];
    out[threadIdx.x] = v;
  }
}

unsigned grid_of(uint64_t dim) {
  uint64_t g = (dim + kBlock * 4 - 1) / (kBlock * 4);
  if (g < 1) g = 1;
  if (g > kMaxBlocks) g = kMaxBlocks;
  return unsigned(g);
}

template <typename T>
hipError_t init_random_t(const State& st, uint64_t seed, double* part, int max_part,
                         double* d_norm, hipStream_t s) {
  T* amp = static_cast<T*>(st.amp);
  uint64_t g = (st.dim() + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  k_init_random<T><<<unsigned(g), kBlock, 0, s>>>(amp, st.dim(), seed);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = launch_reduce(st, 0, part, max_part, d_norm, s);
  if (e != hipSuccess) return e;
  k_scale<T><<<unsigned(g), kBlock, 0, s>>>(amp, st.dim(), d_norm);
  return hipGetLastError();
}

template <typename A, typename B>
hipError_t compare_t(const State& a, const State& b, double* part, double* out, hipStream_t s) {
  const unsigned g = grid_of(a.dim());
  k_compare<A, B><<<g, kBlock, 0, s>>>(static_cast<const A*>(a.amp),
                                       static_cast<const B*>(b.amp), a.dim(), part);
  k_compare_final<<<1, 64, 0, s>>>(part, int(g), out);
  return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_split(const T* __restrict__ amp, T* __restrict__ re,
                                                  T* __restrict__ im, uint64_t count) {
  using V = T __attribute__((ext_vector_type(2)));
  const V* a = reinterpret_cast<const V*>(amp);
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const V v = a[i];
    re[i] = v.x;
    im[i] = v.y;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_join(T* __restrict__ amp, const T* __restrict__ re,
                                                 const T* __restrict__ im, uint64_t count) {
  using V = T __attribute__((ext_vector_type(2)));
  V* a = reinterpret_cast<V*>(amp);
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += uint64_t(gridDim.x) * blockDim.x) {
    V v;
    v.x = re[i];
    v.y = im[i];
    a[i] = v;
  }
}

unsigned stage_grid(uint64_t count) {
  const uint64_t g = (count + kBlock - 1) / kBlock;
  return unsigned(g < 65536 ? (g ? g : 1) : 65536);
}

}  // namespace

hipError_t launch_split(const State& st, uint64_t first, uint64_t count, void* re, void* im,
                        hipStream_t s) {
  if (st.dtype == 0)
    k_split<double><<<stage_grid(count), kBlock, 0, s>>>(
        static_cast<const double*>(st.amp) + 2 * first, static_cast<double*>(re),
        static_cast<double*>(im), count);
  else
    k_split<float><<<stage_grid(count), kBlock, 0, s>>>(
        static_cast<const float*>(st.amp) + 2 * first, static_cast<float*>(re),
        static_cast<float*>(im), count);
  return hipGetLastError();
}

hipError_t launch_join(const State& st, uint64_t first, uint64_t count, const void* re,
                       const void* im, hipStream_t s) {
  if (st.dtype == 0)
    k_join<double><<<stage_grid(count), kBlock, 0, s>>>(
        static_cast<double*>(st.amp) + 2 * first, static_cast<const double*>(re),
        static_cast<const double*>(im), count);
  else
    k_join<float><<<stage_grid(count), kBlock, 0, s>>>(
        static_cast<float*>(st.amp) + 2 * first, static_cast<const float*>(re),
        static_cast<const float*>(im), count);
  return hipGetLastError();
}

hipError_t launch_init_random(const State& st, uint64_t seed, double* part, int max_part,
                              double* d_norm, hipStream_t s) {
  return st.dtype == 0 ? init_random_t<double>(st, seed, part, max_part, d_norm, s)
                       : init_random_t<float>(st, seed, part, max_part, d_norm, s);
}

// `part` needs 5 * 512 doubles, `out` 5.
hipError_t launch_compare(const State& a, const State& b, double* part, double* out,
                          hipStream_t s) {
  if (a.dtype == 0 && b.dtype == 0) return compare_t<double, double>(a, b, part, out, s);
  if (a.dtype == 0 && b.dtype == 1) return compare_t<double, float>(a, b, part, out, s);
  if (a.dtype == 1 && b.dtype == 0) return compare_t<float, double>(a, b, part, out, s);
  return compare_t<float, float>(a, b, part, out, s);
}

}  // namespace qsvh
