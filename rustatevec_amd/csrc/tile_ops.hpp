// tile_ops.hpp -- the op interpreter shared by the tile-pass kernel and the batched
// circuit kernel: applies a list of TileOps to one tile held in LDS (sr, si: N
// amplitudes of T; tile-local index e <-> state index base | offset(e)).
//
// Arithmetic is that of the per-gate kernels (kernels_plain.hip), operation for operation,
// without contraction (files including this header are built with -ffp-contract=off).
#pragma once

#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace qsvh {
namespace tile {

__device__ __forceinline__ uint32_t insert_zero_bit32(uint32_t x, uint32_t b) {
  return ((x >> b) << (b + 1)) | (x & ((1u << b) - 1u));
}

__device__ __forceinline__ uint64_t insert_zero_bits(uint64_t t, const uint32_t* ts, uint32_t tb) {
  for (uint32_t j = 0; j < tb; ++j) {
    const uint32_t b = ts[j];
    t = ((t >> b) << (b + 1)) | (t & ((1ull << b) - 1));
  }
  return t;
}

// State offset of tile-local index e: bit j of e is state bit ts[j].
__device__ __forceinline__ uint64_t offset(uint32_t e, const uint32_t* ts, uint32_t tb) {
  uint64_t off = 0;
  for (uint32_t j = 0; j < tb; ++j) off |= uint64_t((e >> j) & 1u) << ts[j];
  return off;
}

template <typename T>
__device__ __forceinline__ void cmul(T& xr, T& xi, T dr, T di) {
  const T a = dr * xr - di * xi;
  const T b = dr * xi + di * xr;
  xr = a;
  xi = b;
}

// Dense M-qubit op on tile bits op.lq[0..M) (gate order): mq_block in LDS.
template <typename T, int M, int NT>
__device__ __forceinline__ void apply_mq(T* sr, T* si, uint32_t N, const TileOp& op,
                                         const T* pool) {
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
  for (uint32_t blk = threadIdx.x; blk < (N >> M); blk += NT) {
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

// A run of diagonal gates.  Each lane keeps its KMAX amplitudes (e = tid + k NT) and their
// state indices in registers for the whole run; a gate's descriptor and (for m <= 2) its
// phases are uniform, so per amplitude the work is a bit extraction, a select and one
// complex multiply -- the same multiply, in the same order, as diag_chunk.
template <typename T, int NT, int KMAX>
__device__ __forceinline__ void apply_diag_run(T* sr, T* si, uint32_t N, uint64_t base,
                                               const uint32_t* ts, uint32_t tb,
                                               const TileOp& op,
                                               const TileDiag* __restrict__ diags,
                                               const T* __restrict__ pool) {
  T xr[KMAX], xi[KMAX];
  uint64_t f[KMAX];
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    const uint32_t e = threadIdx.x + uint32_t(k) * NT;
    if (e < N) {
      xr[k] = sr[e];
      xi[k] = si[e];
      f[k] = base | offset(e, ts, tb);
    }
  }
  for (uint32_t d = 0; d < op.n_diag; ++d) {
    // fields read one by one (no local copy of the descriptor: its qubit list would be
    // indexed dynamically and land in private memory)
    const TileDiag* dg = diags + op.off + d;
    const uint32_t m = dg->m;
    const T* tab = pool + 2 * dg->tab;
    if (m <= 2) {
      const uint32_t q0 = dg->q[0], q1 = m == 2 ? dg->q[1] : 0;
      const T t0r = tab[0], t0i = tab[1], t1r = tab[2], t1i = tab[3];
      const T t2r = m == 2 ? tab[4] : T(0), t2i = m == 2 ? tab[5] : T(0);
      const T t3r = m == 2 ? tab[6] : T(0), t3i = m == 2 ? tab[7] : T(0);
#pragma unroll
      for (int k = 0; k < KMAX; ++k) {
        if (threadIdx.x + uint32_t(k) * NT < N) {
          const bool b0 = (f[k] >> q0) & 1u;
          const bool b1 = m == 2 && ((f[k] >> q1) & 1u);
          const T dr = b1 ? (b0 ? t3r : t2r) : (b0 ? t1r : t0r);
          const T di = b1 ? (b0 ? t3i : t2i) : (b0 ? t1i : t0i);
          cmul(xr[k], xi[k], dr, di);
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < KMAX; ++k) {
        if (threadIdx.x + uint32_t(k) * NT < N) {
          uint32_t s = 0;
          for (uint32_t j = 0; j < m; ++j) s |= uint32_t((f[k] >> dg->q[j]) & 1u) << j;
          cmul(xr[k], xi[k], tab[2 * s], tab[2 * s + 1]);
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < KMAX; ++k) {
    const uint32_t e = threadIdx.x + uint32_t(k) * NT;
    if (e < N) {
      sr[e] = xr[k];
      si[e] = xi[k];
    }
  }
}

// Interpret ops[0..n_ops) on the tile (LDS); ends with a workgroup barrier.
template <typename T, int NT, int KMAX>
__device__ __forceinline__ void run_ops(T* sr, T* si, uint32_t N, uint64_t base,
                                        const uint32_t* ts, uint32_t tb,
                                        const TileOp* __restrict__ ops, uint32_t n_ops,
                                        const TileDiag* __restrict__ diags,
                                        const T* __restrict__ pool) {
  for (uint32_t o = 0; o < n_ops; ++o) {
    const TileOp op = ops[o];
    if ((base & op.cmask) != op.cval) continue;  // uniform over the workgroup
    if (op.kind == 0) {
      const T* g = pool + 2 * op.off;
      const T g0r = g[0], g0i = g[1], g1r = g[2], g1i = g[3];
      const T g2r = g[4], g2i = g[5], g3r = g[6], g3i = g[7];
      const uint32_t b = op.lq[0];
      for (uint32_t p = threadIdx.x; p < N / 2; p += NT) {
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
      apply_diag_run<T, NT, KMAX>(sr, si, N, base, ts, tb, op, diags, pool);
    } else if (op.kind == 2) {
      apply_mq<T, 2, NT>(sr, si, N, op, pool);
    } else {
      apply_mq<T, 3, NT>(sr, si, N, op, pool);
    }
    __syncthreads();
  }
}

}  // namespace tile
}  // namespace qsvh
