// amp.hpp -- device access to interleaved complex amplitudes (State::amp).
//
// Amplitude i of a state of T (double: complex128, float: complex64) is the pair
// amp[2 i], amp[2 i + 1].  Kernels move 16-byte chunks: A = 16 / (2 sizeof(T)) consecutive
// amplitudes (one complex128, two complex64), so a wave64 instruction covers 1 KiB of
// contiguous HBM.  Interleaving halves the number of address streams of an in-place
// update against separate re / im arrays; on MI355X a 30-qubit in-place pass measured
// 6.3 TB/s interleaved (with XCD-contiguous block order) against 5.5-5.7 TB/s split.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qsvh {

template <typename T>
constexpr int amps_per_chunk() {
  return 16 / (2 * int(sizeof(T)));
}

template <typename T, int N>
using VecT = T __attribute__((ext_vector_type(N)));

// Load A consecutive amplitudes starting at amplitude index i (16-byte aligned when
// 2 A sizeof(T) == 16); NT: non-temporal.
template <typename T, int A, bool NT = false>
__device__ __forceinline__ void cload(const T* __restrict__ amp, uint64_t i, T (&r)[A],
                                      T (&m)[A]) {
  using V = VecT<T, 2 * A>;
  const V* p = reinterpret_cast<const V*>(amp + 2 * i);
  const V v = NT ? __builtin_nontemporal_load(p) : *p;
#pragma unroll
  for (int a = 0; a < A; ++a) {
    r[a] = v[2 * a];
    m[a] = v[2 * a + 1];
  }
}

template <typename T, int A, bool NT = false>
__device__ __forceinline__ void cstore(T* __restrict__ amp, uint64_t i, const T (&r)[A],
                                       const T (&m)[A]) {
  using V = VecT<T, 2 * A>;
  V v;
#pragma unroll
  for (int a = 0; a < A; ++a) {
    v[2 * a] = r[a];
    v[2 * a + 1] = m[a];
  }
  V* p = reinterpret_cast<V*>(amp + 2 * i);
  if constexpr (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// One 16-byte vector (4 floats) at p; NT: non-temporal.
template <bool NT, typename V>
__device__ __forceinline__ V vload(const V* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, typename V>
__device__ __forceinline__ void vstore(V* p, const V& v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Block index in launch order remapped so that each XCD's blocks cover one contiguous
// range of the grid (the hardware deals block b to XCD b % 8); identity when the grid is
// not a multiple of 8 or `on` is 0.
__device__ __forceinline__ uint64_t xcd_block(int on) {
  const uint64_t g = gridDim.x, b = blockIdx.x;
  if (!on || (g & 7u)) return b;
  return (b & 7u) * (g >> 3) + (b >> 3);
}

}  // namespace qsvh
