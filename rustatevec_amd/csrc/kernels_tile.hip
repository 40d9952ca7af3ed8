// kernels_tile.hip -- tile passes: many gates per HBM round trip, exact arithmetic.
//
// A pass owns a set of "tile bits" ts[0..TB) (ascending; always including the lowest bits
// so that a tile is made of contiguous runs of at least 16 amplitudes).  The state is cut
// into 2^(n-TB) tiles, each the 2^TB amplitudes sharing the values of all other bits.  A
// workgroup loads a tile into LDS, interprets the pass's op list on it, and stores it:
//
//  * kind 1q: a 2x2 gate on one tile bit (pairs in LDS), optionally conditioned on bits
//    outside the tile (the controls of a controlled gate whose target is in the tile);
//  * kind mq: a dense gate on 2 or 3 tile bits, optionally conditioned the same way;
//  * kind diag: a run of diagonal gates on any qubits -- each amplitude is read once,
//    multiplied by every gate's phase d[gather_bits(i, qs)] in order, and written once.
//
// Every op performs the operations of the per-gate kernels (kernels_plain.hip) in the
// same order, without contraction, so a pass is bit-identical to applying its ops one by
// one (crates/qsv-core/src/backend/cpu.rs:71-114,221-247).
#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace qsvh {
namespace {

constexpr int kThreads = 512;

template <typename T>
__device__ __forceinline__ void cmul(T& xr, T& xi, T dr, T di) {
  const T a = dr * xr - di * xi;
  const T b = dr * xi + di * xr;
  xr = a;
  xi = b;
}

__device__ __forceinline__ uint32_t insert_zero_bit32(uint32_t x, uint32_t b) {
  return ((x >> b) << (b + 1)) | (x & ((1u << b) - 1u));
}

// Dense M-qubit op on tile bits op.lq[0..M) (gate order): mq_block in LDS.
template <typename T, int M, uint32_t N>
__device__ __forceinline__ void tile_mq(T* sr, T* si, const TileOp& op, const T* pool) {
  constexpr uint32_t D = 1u << M;
  uint32_t srt[M];
#pragma unroll
  for (int j = 0; j < M; ++j) srt[j] = op.lq[j];
#pragma unroll
  for (int a = 1; a < M; ++a)
#pragma unroll
    for (int c = a; c > 0; --c)
      if (srt[c - 1] > srt[c]) {
        const uint32_t x = srt[c];
        srt[c] = srt[c - 1];
        srt[c - 1] = x;
      }
  const T* g = pool + 2 * op.off;
  for (uint32_t blk = threadIdx.x; blk < (N >> M); blk += kThreads) {
    uint32_t b0 = blk;
#pragma unroll
    for (int j = 0; j < M; ++j) b0 = insert_zero_bit32(b0, srt[j]);
    T ar[D], ai[D];
    uint32_t idx[D];
#pragma unroll
    for (uint32_t s = 0; s < D; ++s) {
      uint32_t i = b0;
#pragma unroll
      for (int j = 0; j < M; ++j) i |= ((s >> j) & 1u) << op.lq[j];
      idx[s] = i;
      ar[s] = sr[i];
      ai[s] = si[i];
    }
#pragma unroll
    for (uint32_t r = 0; r < D; ++r) {
      T accr = T(0), acci = T(0);
#pragma unroll
      for (uint32_t c = 0; c < D; ++c) {
        const T er = g[2 * (r * D + c)], ei = g[2 * (r * D + c) + 1];
        const T pr = er * ar[c] - ei * ai[c];
        const T pi = er * ai[c] + ei * ar[c];
        accr = accr + pr;
        acci = acci + pi;
      }
      sr[idx[r]] = accr;
      si[idx[r]] = acci;
    }
  }
}

template <typename T, int TB>
__global__ __launch_bounds__(kThreads) void k_tile(T* __restrict__ re, T* __restrict__ im,
                                                   uint64_t n_tiles, TilePassArg pa,
                                                   const TileOp* __restrict__ ops,
                                                   const TileDiag* __restrict__ diags,
                                                   const T* __restrict__ pool) {
  constexpr uint32_t N = 1u << TB;
  constexpr int W = 16 / int(sizeof(T));  // elements per 16-byte access
  constexpr uint32_t PER = N / (kThreads * W);  // vector chunks per thread
  static_assert(PER >= 1, "tile smaller than one vector per thread");
  __shared__ T sr[N], si[N];
  using V = T __attribute__((ext_vector_type(W)));

  // Global offset of tile-local index e: bit j of e is state bit ts[j].
  auto offset = [&](uint32_t e) {
    uint64_t off = 0;
#pragma unroll
    for (int j = 0; j < TB; ++j) off |= uint64_t((e >> j) & 1u) << pa.ts[j];
    return off;
  };

  for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    uint64_t base = t;
#pragma unroll
    for (int j = 0; j < TB; ++j) {
      const uint32_t b = pa.ts[j];
      base = ((base >> b) << (b + 1)) | (base & ((1ull << b) - 1));
    }
    // load: W consecutive tile-local indices are W consecutive addresses (ts[0..] = 0, 1)
#pragma unroll
    for (uint32_t c = 0; c < PER; ++c) {
      const uint32_t e = (threadIdx.x + kThreads * c) * W;
      const uint64_t a = base + offset(e);
      const V vr = *reinterpret_cast<const V*>(re + a);
      const V vi = *reinterpret_cast<const V*>(im + a);
#pragma unroll
      for (int w = 0; w < W; ++w) {
        sr[e + w] = vr[w];
        si[e + w] = vi[w];
      }
    }
    __syncthreads();
    for (uint32_t o = 0; o < pa.n_ops; ++o) {
      const TileOp op = ops[pa.op_begin + o];
      if ((base & op.cmask) != op.cval) continue;  // uniform over the workgroup
      if (op.kind == 0) {
        const T* g = pool + 2 * op.off;
        const T g0r = g[0], g0i = g[1], g1r = g[2], g1i = g[3];
        const T g2r = g[4], g2i = g[5], g3r = g[6], g3i = g[7];
        const uint32_t b = op.lq[0];
        for (uint32_t p = threadIdx.x; p < N / 2; p += kThreads) {
          const uint32_t i0 = insert_zero_bit32(p, b), i1 = i0 | (1u << b);
          const T x0r = sr[i0], x0i = si[i0], x1r = sr[i1], x1i = si[i1];
          const T ar = g0r * x0r - g0i * x0i, ai = g0r * x0i + g0i * x0r;
          const T br = g1r * x1r - g1i * x1i, bi = g1r * x1i + g1i * x1r;
          const T cr = g2r * x0r - g2i * x0i, ci = g2r * x0i + g2i * x0r;
          const T dr = g3r * x1r - g3i * x1i, di = g3r * x1i + g3i * x1r;
          sr[i0] = ar + br;
          si[i0] = ai + bi;
          sr[i1] = cr + dr;
          si[i1] = ci + di;
        }
      } else if (op.kind == 1) {
        for (uint32_t e = threadIdx.x; e < N; e += kThreads) {
          const uint64_t full = base | offset(e);
          T xr = sr[e], xi = si[e];
          for (uint32_t d = 0; d < op.n_diag; ++d) {
            const TileDiag dg = diags[op.off + d];
            uint32_t s = 0;
            for (uint32_t j = 0; j < dg.m; ++j) s |= uint32_t((full >> dg.q[j]) & 1u) << j;
            const T* tab = pool + 2 * (dg.tab + s);
            cmul(xr, xi, tab[0], tab[1]);
          }
          sr[e] = xr;
          si[e] = xi;
        }
      } else if (op.kind == 2) {
        tile_mq<T, 2, N>(sr, si, op, pool);
      } else {
        tile_mq<T, 3, N>(sr, si, op, pool);
      }
      __syncthreads();
    }
#pragma unroll
    for (uint32_t c = 0; c < PER; ++c) {
      const uint32_t e = (threadIdx.x + kThreads * c) * W;
      const uint64_t a = base + offset(e);
      V vr, vi;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        vr[w] = sr[e + w];
        vi[w] = si[e + w];
      }
      *reinterpret_cast<V*>(re + a) = vr;
      *reinterpret_cast<V*>(im + a) = vi;
    }
    __syncthreads();
  }
}

template <typename T, int TB>
hipError_t launch_tb(const State& st, const TilePassArg& pa, const TileOp* ops,
                     const TileDiag* diags, const void* pool, int max_blocks, hipStream_t s) {
  const uint64_t n_tiles = st.dim() >> TB;
  uint64_t g = n_tiles;
  if (g > uint64_t(max_blocks)) g = max_blocks;
  k_tile<T, TB><<<unsigned(g), kThreads, 0, s>>>(static_cast<T*>(st.re), static_cast<T*>(st.im),
                                                n_tiles, pa, ops, diags,
                                                static_cast<const T*>(pool));
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
