// kernels_tile.hip -- tile passes: many gates per HBM round trip, exact arithmetic.
//
// A pass owns a set of "tile bits" ts[0..TB) (ascending; always including the lowest bits
// so that a tile is made of contiguous runs of at least 16 amplitudes).  The state is cut
// into 2^(n-TB) tiles, each the 2^TB amplitudes sharing the values of all other bits.  A
// workgroup loads a tile into LDS, interprets the pass's op list on it (tile_ops.hpp), and
// stores it:
//
//  * kind 0: a 2x2 gate on one tile bit (pairs in LDS), optionally conditioned on bits
//    outside the tile (the controls of a controlled gate whose target is in the tile);
//  * kind 2, 3: a dense gate on 2 or 3 tile bits, optionally conditioned the same way;
//  * kind 1: a run of diagonal gates on any qubits -- each amplitude is read once,
//    multiplied by every gate's phase d[gather_bits(i, qs)] in order, and written once.
//
// Every op performs the operations of the per-gate kernels (kernels_plain.hip) in the
// same order, without contraction, so a pass is bit-identical to applying its ops one by
// one (crates/qsv-core/src/backend/cpu.rs:71-114,221-247).
#include <hip/hip_runtime.h>

#include "amp.hpp"
#include "engine.hpp"
#include "tile_ops.hpp"

namespace qsvh {
namespace {

constexpr int kThreads = 1 << kTileLaneBits;

template <typename T, int TB>
// Two workgroups per CU (LDS allows it up to 64 KiB tiles) need <= 128 VGPRs per lane.
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(4))) void k_tile(
                                                   T* __restrict__ amp, uint64_t n_tiles,
                                                   TilePassArg pa,
                                                   const TileOp* __restrict__ ops,
                                                   const TileDiag* __restrict__ diags,
                                                   const T* __restrict__ pool) {
  constexpr uint32_t N = 1u << TB;
  constexpr int A = amps_per_chunk<T>();        // amplitudes per 16-byte access
  constexpr uint32_t PER = N / (kThreads * A);  // chunks per thread
  static_assert(PER >= 1, "tile smaller than one chunk per thread");
  __shared__ T sr[N], si[N];

  // Chunk c of lane tid is tile-local index e = (tid + c NT) A, whose state offset is the
  // OR of a per-lane part (the bits of tid A) and a uniform part (the bits of c NT A).
  constexpr uint32_t LA = tile::ilog2c(kThreads * A);
  const uint64_t off_lane = tile::offset(threadIdx.x * A, pa.ts, LA);
  for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const uint64_t base = tile::insert_zero_bits(t, pa.tsort, TB);
    // A consecutive tile-local indices are A consecutive amplitudes (ts starts 0, 1, 2, 3).
    // All the lane's loads are issued before any of them is written to LDS: a chunk offset
    // reads its bit positions through an opaque statement, which the scheduler does not move
    // loads across, so a load-and-store per chunk would wait out each load in turn.
    {
      T r[PER][A], m[PER][A];
#pragma unroll
      for (uint32_t c = 0; c < PER; ++c)
        cload<T, A>(amp, base + (off_lane | tile::chunk_offset(c, pa.ts, LA)), r[c], m[c]);
#pragma unroll
      for (uint32_t c = 0; c < PER; ++c) {
        const uint32_t e = (threadIdx.x + kThreads * c) * A;
#pragma unroll
        for (int w = 0; w < A; ++w) {
          sr[e + w] = r[c][w];
          si[e + w] = m[c][w];
        }
      }
    }
    __syncthreads();
    tile::run_ops<T, kThreads, int(N / kThreads)>(sr, si, N, base, ops + pa.op_begin, pa.n_ops,
                                                  diags, pool);
    {  // likewise every LDS read before the stores
      T r[PER][A], m[PER][A];
#pragma unroll
      for (uint32_t c = 0; c < PER; ++c) {
        const uint32_t e = (threadIdx.x + kThreads * c) * A;
#pragma unroll
        for (int w = 0; w < A; ++w) {
          r[c][w] = sr[e + w];
          m[c][w] = si[e + w];
        }
      }
#pragma unroll
      for (uint32_t c = 0; c < PER; ++c)
        cstore<T, A>(amp, base + (off_lane | tile::chunk_offset(c, pa.ts, LA)), r[c], m[c]);
    }
    __syncthreads();
  }
}

template <typename T, int TB>
hipError_t launch_tb(const State& st, const TilePassArg& pa, const TileOp* ops,
                     const TileDiag* diags, const void* pool, int max_blocks, hipStream_t s) {
  const uint64_t n_tiles = st.dim() >> TB;
  // one workgroup per tile unless max_blocks caps the grid (then they loop), and fewer
  // than 2^32 work-items per dispatch
  uint64_t g = n_tiles;
  if (max_blocks > 0 && g > uint64_t(max_blocks)) g = max_blocks;
  if (g > (uint64_t(1) << 31) / kThreads) g = (uint64_t(1) << 31) / kThreads;
  k_tile<T, TB><<<unsigned(g), kThreads, 0, s>>>(static_cast<T*>(st.amp), n_tiles, pa, ops,
                                                diags, static_cast<const T*>(pool));
  return hipGetLastError();
}

// Tile sizes: 2^TB amplitudes of T in LDS (16 KiB .. 64 KiB) and at least one 16-byte
// vector per thread: complex128 TB = 10..12, complex64 TB = 11..13.
template <typename T>
hipError_t launch_t(const State& st, const TilePassArg& pa, const TileOp* ops,
                    const TileDiag* diags, const void* pool, int max_blocks, hipStream_t s) {
  if constexpr (sizeof(T) == 8) {
    switch (pa.tb) {
      case 10: return launch_tb<T, 10>(st, pa, ops, diags, pool, max_blocks, s);
      case 11: return launch_tb<T, 11>(st, pa, ops, diags, pool, max_blocks, s);
      case 12: return launch_tb<T, 12>(st, pa, ops, diags, pool, max_blocks, s);
      default: return hipErrorInvalidValue;
    }
  } else {
    switch (pa.tb) {
      case 11: return launch_tb<T, 11>(st, pa, ops, diags, pool, max_blocks, s);
      case 12: return launch_tb<T, 12>(st, pa, ops, diags, pool, max_blocks, s);
      case 13: return launch_tb<T, 13>(st, pa, ops, diags, pool, max_blocks, s);
      default: return hipErrorInvalidValue;
    }
  }
}

}  // namespace

hipError_t launch_tile(const State& st, const TilePassArg& pa, const TileOp* ops,
                       const TileDiag* diags, const void* pool, int max_blocks, hipStream_t s) {
  if (pa.tb > st.n) return hipErrorInvalidValue;
  return st.dtype == 0 ? launch_t<double>(st, pa, ops, diags, pool, max_blocks, s)
                       : launch_t<float>(st, pa, ops, diags, pool, max_blocks, s);
}

}  // namespace qsvh
