// kernels_batch.hip -- many independent circuits in one persistent launch per pass.
//
// Every circuit of a batch has its own state in a shared pool and its own tile-pass
// schedule (tile_ops.hpp).  Launch j runs pass j of every circuit that has one: its work
// items are (circuit pass, tile) pairs, handed out through an atomic counter so that
// uneven circuits (2^10 .. 2^22 amplitudes) load-balance over the workgroups.  A circuit
// whose state fits one tile (n <= tile bits) is a single pass: it is built from |0..0>
// in LDS, run to the end and written once.  The last pass of every circuit also forms
// per-tile partial sums of <Z_0> = sum_i |psi_i|^2 (-1)^(i & 1) (sample.rs:177-200), which
// a final kernel adds up per circuit in a fixed order.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "amp.hpp"
#include "engine.hpp"
#include "tile_ops.hpp"

namespace qsvh {
namespace {

constexpr int kThreads = 1 << kTileLaneBits;

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// Two workgroups per CU (the 64 KiB tiles allow it) need <= 128 VGPRs per lane: without the
// cap the compiler took 140 and one workgroup per CU ran (512 HEA circuits 53 -> 95 ms).  The
// interpreter runs without its free lane runs (run_ops<..., false>), which would not fit.
template <typename T, int MAXTB>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_batch(
    const BatchPass* __restrict__ passes,
                                                    uint32_t n_passes, uint64_t total_work,
                                                    unsigned long long* __restrict__ counter,
                                                    T* __restrict__ amp_pool,
                                                    const TileOp* __restrict__ ops,
                                                    const TileDiag* __restrict__ diags,
                                                    const T* __restrict__ pool,
                                                    double* __restrict__ partials) {
  constexpr uint32_t NMAX = 1u << MAXTB;
  constexpr int A = amps_per_chunk<T>();  // batch circuits have >= 4 qubits: a chunk fits
  __shared__ T sr[NMAX], si[NMAX];
  __shared__ unsigned long long w_sh;
  __shared__ double red[kThreads / 64];

  for (;;) {
    if (threadIdx.x == 0) w_sh = atomicAdd(counter, 1ull);
    __syncthreads();
    const uint64_t w = w_sh;
    __syncthreads();
    if (w >= total_work) break;
    // the pass owning work item w (passes sorted by tile_begin)
    uint32_t lo = 0, hi = n_passes - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) / 2;
      if (passes[mid].tile_begin <= w)
        lo = mid;
      else
        hi = mid - 1;
    }
    const BatchPass* bp = passes + lo;
    // A full tile (tile bits = MAXTB, every work item of the circuits that exceed one
    // tile) takes a copy of the body with the tile size a compile-time constant, so the
    // per-amplitude bounds checks and chunk loops fold away; smaller tiles take the other.
    auto tile_body = [&](auto fullc) __attribute__((always_inline)) {
      constexpr bool FULL = decltype(fullc)::value;
      const uint32_t tb = FULL ? uint32_t(MAXTB) : bp->tb;
      const uint32_t N = FULL ? NMAX : (1u << tb);
      const uint32_t* ts = bp->ts;  // uniform: read through the scalar cache
      const uint64_t t = w - bp->tile_begin;
      const uint64_t base = tile::insert_zero_bits(t, bp->tsort, tb);
      T* amp = amp_pool + 2 * bp->state_off;
      const uint32_t flags = bp->flags;
      // chunk c of lane tid (e = (tid + c NT) A): offset = per-lane part | uniform part
      constexpr uint32_t LA = tile::ilog2c(kThreads * A);
      const uint32_t la = tb < LA ? tb : LA;
      const uint64_t off_lane = tile::offset((threadIdx.x * A) & (N - 1), ts, la);
      if constexpr (FULL) {
        // a full tile: every chunk's load issued before any is written to LDS (a chunk
        // offset's opaque bit reads keep the scheduler from grouping the loads otherwise)
        constexpr uint32_t PER = NMAX / (kThreads * A);
        if (!(flags & kBatchInit)) {
          T r[PER][A], m[PER][A];
#pragma unroll
          for (uint32_t c = 0; c < PER; ++c)
            cload<T, A>(amp, base + (off_lane | tile::chunk_offset(c, ts, LA)), r[c], m[c]);
#pragma unroll
          for (uint32_t c = 0; c < PER; ++c) {
            const uint32_t e = (threadIdx.x + kThreads * c) * A;
#pragma unroll
            for (int v = 0; v < A; ++v) {
              sr[e + v] = r[c][v];
              si[e + v] = m[c][v];
            }
          }
        }
      }
      for (uint32_t e = threadIdx.x * A, c = 0; e < N; e += kThreads * A, ++c) {
        if (FULL && !(flags & kBatchInit)) break;  // loaded above
        const uint64_t off = off_lane | tile::chunk_offset(c, ts, LA);
        if (flags & kBatchInit) {
          // |0...0>: amplitude 1 at state index 0, which is tile 0, local index 0
  #pragma unroll
          for (int v = 0; v < A; ++v) {
            sr[e + v] = (t == 0 && e + v == 0) ? T(1) : T(0);
            si[e + v] = T(0);
          }
        } else {
          T r[A], m[A];
          cload<T, A>(amp, base + off, r, m);
  #pragma unroll
          for (int v = 0; v < A; ++v) {
            sr[e + v] = r[v];
            si[e + v] = m[v];
          }
        }
      }
      __syncthreads();
      tile::run_ops<T, kThreads, int(NMAX / kThreads), false>(sr, si, N, base, ops + bp->op_begin,
                                                       bp->n_ops, diags, pool);
      if (flags & kBatchExpZ0) {
        double acc = 0.0;
        for (uint32_t e = threadIdx.x; e < N; e += kThreads) {
          const double a = double(sr[e]), b = double(si[e]);
          const double p = a * a + b * b;
          acc += ((base | tile::offset(e, ts, tb)) & 1u) ? -p : p;
        }
        acc = wave_sum(acc);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
          double s = 0.0;
          for (int k = 0; k < kThreads / 64; ++k) s += red[k];
          partials[bp->part_off + t] = s;
        }
      }
      for (uint32_t e = threadIdx.x * A, c = 0; e < N; e += kThreads * A, ++c) {
        T r[A], m[A];
  #pragma unroll
        for (int v = 0; v < A; ++v) {
          r[v] = sr[e + v];
          m[v] = si[e + v];
        }
        cstore<T, A>(amp, base + (off_lane | tile::chunk_offset(c, ts, LA)), r, m);
      }
    };
    if (bp->tb == uint32_t(MAXTB))
      tile_body(std::true_type{});
    else
      tile_body(std::false_type{});
    __syncthreads();
  }
}

__global__ void k_reset_counter(unsigned long long* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0ull;
}

// out[c] = sum of the circuit's last-pass partials, in tile order.
__global__ void k_batch_reduce(const uint32_t* __restrict__ part_off,
                               const uint32_t* __restrict__ n_part, uint32_t n_circ,
                               const double* __restrict__ partials, double* __restrict__ out) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_circ) return;
  double s = 0.0;
  for (uint32_t k = 0; k < n_part[c]; ++k) s += partials[part_off[c] + k];
  out[c] = s;
}

}  // namespace

hipError_t launch_batch_pass(int dtype, const BatchPass* passes, uint32_t n_passes,
                             uint64_t total_work, unsigned long long* counter, void* amp_pool,
                             const TileOp* ops, const TileDiag* diags,
                             const void* pool, double* partials, int blocks, hipStream_t s) {
  k_reset_counter<<<1, 64, 0, s>>>(counter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (dtype == 0) {
    k_batch<double, 12><<<blocks, kThreads, 0, s>>>(
        passes, n_passes, total_work, counter, static_cast<double*>(amp_pool), ops, diags,
        static_cast<const double*>(pool), partials);
  } else {
    k_batch<float, 13><<<blocks, kThreads, 0, s>>>(
        passes, n_passes, total_work, counter, static_cast<float*>(amp_pool), ops, diags,
        static_cast<const float*>(pool), partials);
  }
  return hipGetLastError();
}

hipError_t launch_batch_reduce(const uint32_t* part_off, const uint32_t* n_part,
                               uint32_t n_circ, const double* partials, double* out,
                               hipStream_t s) {
  k_batch_reduce<<<(n_circ + 255) / 256, 256, 0, s>>>(part_off, n_part, n_circ, partials, out);
  return hipGetLastError();
}

}  // namespace qsvh
