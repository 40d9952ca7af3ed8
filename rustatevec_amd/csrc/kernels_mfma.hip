// kernels_mfma.hip -- fused k-qubit groups (k = 4..6) on the gfx950 matrix cores.
//
// A fused group (the output of fusion::fuse, crates/qsv-core/src/fusion.rs:39-139) is a
// dense D x D complex matrix U, D = 2^k, acting on k qubits G.  Viewing the complex64 state
// as X[s][o] (s: the k group bits, o: all other bits) the update is the complex GEMM
// Y = U X, written in real arithmetic as the block product
//
//     [Yr]   [Ur  -Ui] [Xr]
//     [Yi] = [Ui   Ur] [Xi]        (2D x 2D) . (2D x N),  8 * 2^k * 2^n flops.
//
// The state is processed in tiles of 2048 amplitudes: the k group bits plus the C lowest
// other bits (C = 11 - k), so every tile covers runs of >= 2^C consecutive addresses and
// the loads and stores are coalesced whatever qubits the group acts on.  A tile is staged
// in LDS as the 2D x 2^C real block operand; the four waves of a workgroup each produce
// one 32 x 32 output block with MFMA, write it back through LDS, and the tile is stored.
// Global loads of the next tile are issued before the current tile's MFMA work.
//
// FP32 variant: v_mfma_f32_32x32x2_f32, the block matrix held in registers as A fragments.
// BF16x3 variant: v_mfma_f32_32x32x16_bf16 with both operands split into bf16 hi + lo
// (x = hi + lo, lo = bf16(x - hi)); Y = Ahi Bhi + Ahi Blo + Alo Bhi accumulated in FP32
// (the lo*lo term, ~2^-16 relative, is dropped).
#include <hip/hip_runtime.h>

#include "amp.hpp"
#include "engine.hpp"

namespace qsvh {
namespace {

constexpr int kBlock = 256;
constexpr int kTileBits = 11;  // 2048 amplitudes per tile

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Tile-local index t (kTileBits bits) <-> state bits: bit j of t is state bit ts[j]
// (ascending).  role[j] = 0x80 | i if ts[j] is the i-th group bit, else the column bit
// index.  Same table for every tile; a tile's base address inserts zeros at all ts.
struct TileMap {
  uint32_t ts[kTileBits];
  uint8_t role[kTileBits];
};

__device__ __forceinline__ uint64_t tile_base(uint64_t t, const TileMap& m) {
#pragma unroll
  for (int j = 0; j < kTileBits; ++j) {
    const uint32_t b = m.ts[j];
    t = ((t >> b) << (b + 1)) | (t & ((1ull << b) - 1));
  }
  return t;
}

__device__ __forceinline__ uint64_t tile_offset(uint32_t tl, const TileMap& m) {
  uint64_t off = 0;
#pragma unroll
  for (int j = 0; j < kTileBits; ++j) off |= uint64_t((tl >> j) & 1u) << m.ts[j];
  return off;
}

// (row s, column n) of tile-local index tl.
__device__ __forceinline__ void tile_rc(uint32_t tl, const TileMap& m, uint32_t& s,
                                        uint32_t& n) {
  s = 0;
  n = 0;
#pragma unroll
  for (int j = 0; j < kTileBits; ++j) {
    const uint32_t bit = (tl >> j) & 1u;
    const uint32_t r = m.role[j];
    if (r & 0x80u)
      s |= bit << (r & 0x7fu);
    else
      n |= bit << r;
  }
}

// Each lane moves four 16-byte chunks of two interleaved amplitudes per tile
// (2048 / 2 / 256 = 4); tile-local bit 0 is state bit 0, so a chunk is consecutive.
constexpr int kChunks = (1 << kTileBits) / 2 / kBlock;

// ------------------------------------------------------------------------- FP32
// LDS layout [row][col], row stride chosen = 32 (mod 64) floats so that the two half-waves
// of a B-fragment read (rows r and r+1) fall in disjoint banks.
template <int K>
struct F32Shape {
  static constexpr int D = 1 << K;
  static constexpr int NT = 1 << (kTileBits - K);  // columns per tile
  static constexpr int STRIDE = NT + (NT >= 64 ? 32 : 0);
  static constexpr int NTILE_COLS = NT / 32;       // 32-wide output blocks per row band
};

template <int K, bool NTA>
__global__ __launch_bounds__(kBlock) void k_fused_f32(float* __restrict__ amp,
                                                      uint64_t n_tiles, TileMap map,
                                                      const float* __restrict__ ablock,
                                                      int xcd) {
  using S = F32Shape<K>;
  constexpr int D = S::D, R = 2 * D, STRIDE = S::STRIDE;
  __shared__ float xr[D * STRIDE];
  __shared__ float xi[D * STRIDE];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int mt = wave / S::NTILE_COLS, nt = wave % S::NTILE_COLS;  // this wave's 32x32 block

  // A fragments of the block matrix rows [32 mt, 32 mt + 32): a[kk] = A[row][2 kk + lane/32].
  float a[D];
#pragma unroll
  for (int kk = 0; kk < D; ++kk)
    a[kk] = ablock[(32 * mt + (lane & 31)) * R + 2 * kk + (lane >> 5)];

  // Per-lane chunk offsets and LDS positions (the same for every tile).
  uint64_t off[kChunks];
  uint32_t pos[kChunks][2];
#pragma unroll
  for (int c = 0; c < kChunks; ++c) {
    const uint32_t tl = 2u * (threadIdx.x + kBlock * c);
    off[c] = tile_offset(tl, map);
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      uint32_t s, n;
      tile_rc(tl + e, map, s, n);
      pos[c][e] = s * STRIDE + n;
    }
  }

  floatx4 v[kChunks];
  uint64_t t = xcd_block(xcd);  // consecutive tiles stay on one XCD
  if (t < n_tiles) {
    const uint64_t base = tile_base(t, map);
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
      v[c] = vload<NTA>(reinterpret_cast<const floatx4*>(amp + 2 * (base + off[c])));
    }
  }
  for (; t < n_tiles; t += gridDim.x) {
    const uint64_t base = tile_base(t, map);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kChunks; ++c)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        xr[pos[c][e]] = v[c][2 * e];
        xi[pos[c][e]] = v[c][2 * e + 1];
      }
    __syncthreads();
    const uint64_t tn = t + gridDim.x;
    if (tn < n_tiles) {  // prefetch the next tile while the matrix cores work
      const uint64_t nb = tile_base(tn, map);
#pragma unroll
      for (int c = 0; c < kChunks; ++c) {
        v[c] = vload<NTA>(reinterpret_cast<const floatx4*>(amp + 2 * (nb + off[c])));
      }
    }
    floatx16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f,
                    0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const int col = 32 * nt + (lane & 31);
#pragma unroll
    for (int kk = 0; kk < D; ++kk) {
      const int row = 2 * kk + (lane >> 5);  // block row of the B operand
      const float b = row < D ? xr[row * STRIDE + col] : xi[(row - D) * STRIDE + col];
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[kk], b, acc, 0, 0, 0);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 32 * mt + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      if (row < D)
        xr[row * STRIDE + col] = acc[r];
      else
        xi[(row - D) * STRIDE + col] = acc[r];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
      floatx4 y;
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        y[2 * e] = xr[pos[c][e]];
        y[2 * e + 1] = xi[pos[c][e]];
      }
      vstore<NTA>(reinterpret_cast<floatx4*>(amp + 2 * (base + off[c])), y);
    }
  }
}

// ------------------------------------------------------------------------- BF16x3
// LDS layout [col][block row] (block rows 0..D-1 = Xr, D..2D-1 = Xi), so that the eight
// consecutive K values of a bf16 B fragment are contiguous.
template <int K>
struct B16Shape {
  static constexpr int D = 1 << K;
  static constexpr int R = 2 * D;
  static constexpr int NT = 1 << (kTileBits - K);
  static constexpr int CSTRIDE = R + 4;  // floats per column (padding against conflicts)
  static constexpr int NTILE_COLS = NT / 32;
  static constexpr int KSTEPS = R / 16;
};

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = static_cast<__bf16>(x);
  lo = static_cast<__bf16>(x - static_cast<float>(hi));
}

template <int K, bool NTA>
__global__ __launch_bounds__(kBlock) void k_fused_bf16x3(float* __restrict__ amp,
                                                         uint64_t n_tiles, TileMap map,
                                                         const __bf16* __restrict__ ahi,
                                                         const __bf16* __restrict__ alo,
                                                         int xcd) {
  using S = B16Shape<K>;
  constexpr int D = S::D, R = S::R, CS = S::CSTRIDE, NT = S::NT;
  __shared__ float xb[NT * CS];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int mt = wave / S::NTILE_COLS, nt = wave % S::NTILE_COLS;

  // A fragments: lane holds A[32 mt + lane % 32][16 ks + 8 (lane / 32) + j], j = 0..7.
  bf16x8 ah[S::KSTEPS], al[S::KSTEPS];
#pragma unroll
  for (int ks = 0; ks < S::KSTEPS; ++ks) {
    const int idx = (32 * mt + (lane & 31)) * R + 16 * ks + 8 * (lane >> 5);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      ah[ks][j] = ahi[idx + j];
      al[ks][j] = alo[idx + j];
    }
  }

  uint64_t off[kChunks];
  uint32_t pos[kChunks][2];
#pragma unroll
  for (int c = 0; c < kChunks; ++c) {
    const uint32_t tl = 2u * (threadIdx.x + kBlock * c);
    off[c] = tile_offset(tl, map);
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      uint32_t s, n;
      tile_rc(tl + e, map, s, n);
      pos[c][e] = n * CS + s;  // re at +0, im at +D
    }
  }

  floatx4 v[kChunks];
  uint64_t t = xcd_block(xcd);  // consecutive tiles stay on one XCD
  if (t < n_tiles) {
    const uint64_t base = tile_base(t, map);
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
      v[c] = vload<NTA>(reinterpret_cast<const floatx4*>(amp + 2 * (base + off[c])));
    }
  }
  for (; t < n_tiles; t += gridDim.x) {
    const uint64_t base = tile_base(t, map);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kChunks; ++c)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        xb[pos[c][e]] = v[c][2 * e];
        xb[pos[c][e] + D] = v[c][2 * e + 1];
      }
    __syncthreads();
    const uint64_t tn = t + gridDim.x;
    if (tn < n_tiles) {
      const uint64_t nb = tile_base(tn, map);
#pragma unroll
      for (int c = 0; c < kChunks; ++c) {
        v[c] = vload<NTA>(reinterpret_cast<const floatx4*>(amp + 2 * (nb + off[c])));
      }
    }
    floatx16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f,
                    0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const int col = 32 * nt + (lane & 31);
#pragma unroll
    for (int ks = 0; ks < S::KSTEPS; ++ks) {
      const float* src = xb + col * CS + 16 * ks + 8 * (lane >> 5);
      const floatx4 b0 = *reinterpret_cast<const floatx4*>(src);
      const floatx4 b1 = *reinterpret_cast<const floatx4*>(src + 4);
      bf16x8 bh, bl;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        __bf16 h, l;
        split_bf16(j < 4 ? b0[j] : b1[j - 4], h, l);
        bh[j] = h;
        bl[j] = l;
      }
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[ks], bh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[ks], bl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[ks], bh, acc, 0, 0, 0);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 32 * mt + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      xb[col * CS + row] = acc[r];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
      floatx4 y;
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        y[2 * e] = xb[pos[c][e]];
        y[2 * e + 1] = xb[pos[c][e] + D];
      }
      vstore<NTA>(reinterpret_cast<floatx4*>(amp + 2 * (base + off[c])), y);
    }
  }
}

template <int K, bool NT>
hipError_t launch_k(const State& st, const TileMap& map, const void* mats, int precision,
                    const Tuning& t, hipStream_t s) {
  const uint64_t n_tiles = st.dim() >> kTileBits;
  uint64_t g = n_tiles;
  if (g > uint64_t(t.mfma_blocks)) g = t.mfma_blocks;
  float* amp = static_cast<float*>(st.amp);
  if (precision == 1) {
    k_fused_f32<K, NT><<<unsigned(g), kBlock, 0, s>>>(amp, n_tiles, map,
                                                      static_cast<const float*>(mats), t.xcd);
  } else {
    constexpr size_t n = size_t(4) << (2 * K);  // (2D)^2 entries
    const __bf16* hi = static_cast<const __bf16*>(mats);
    k_fused_bf16x3<K, NT><<<unsigned(g), kBlock, 0, s>>>(amp, n_tiles, map, hi, hi + n,
                                                         t.xcd);
  }
  return hipGetLastError();
}

template <int K>
hipError_t launch_k(const State& st, const TileMap& map, const void* mats, int precision,
                    const Tuning& t, hipStream_t s) {
  return nontemporal_for(st, t, true) ? launch_k<K, true>(st, map, mats, precision, t, s)
                                      : launch_k<K, false>(st, map, mats, precision, t, s);
}

}  // namespace

// `q` are the group's qubits in ascending order; `mats` is the block matrix
// [[Ur, -Ui], [Ui, Ur]] (row-major, 2D x 2D) in f32 (precision 1) or as bf16 hi then lo
// arrays (precision 2).  Requires a complex64 state of at least 11 qubits.
hipError_t launch_fused_mfma(const State& st, uint32_t k, const uint32_t* q, const void* mats,
                             int precision, const Tuning& t, hipStream_t s) {
  if (st.dtype != 1 || st.n < uint32_t(kTileBits) || k < 4 || k > 6) return hipErrorInvalidValue;
  // Tile bits: the group bits plus the (11 - k) lowest other bits.
  uint32_t ts[kTileBits];
  uint32_t nts = 0;
  bool in_g[64] = {false};
  for (uint32_t j = 0; j < k; ++j) in_g[q[j]] = true;
  uint32_t ncol = 0;
  TileMap map;
  for (uint32_t b = 0; b < st.n && nts < uint32_t(kTileBits); ++b) {
    if (in_g[b]) {
      ts[nts++] = b;
    } else if (ncol < kTileBits - k) {
      ts[nts++] = b;
      ++ncol;
    }
  }
  if (nts != uint32_t(kTileBits)) return hipErrorInvalidValue;
  uint32_t gi = 0, ci = 0;
  for (uint32_t j = 0; j < nts; ++j) {
    map.ts[j] = ts[j];
    if (in_g[ts[j]]) {
      // q is ascending, so the i-th group bit met in ascending order is q[i]
      map.role[j] = uint8_t(0x80u | gi++);
    } else {
      map.role[j] = uint8_t(ci++);
    }
  }
  switch (k) {
    case 4: return launch_k<4>(st, map, mats, precision, t, s);
    case 5: return launch_k<5>(st, map, mats, precision, t, s);
    default: return launch_k<6>(st, map, mats, precision, t, s);
  }
}

}  // namespace qsvh
