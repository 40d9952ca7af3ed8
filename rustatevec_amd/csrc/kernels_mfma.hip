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

// LDS position of tile-local index tl: row s, column n at s * rs + n * cs.
__device__ __forceinline__ uint32_t tile_pos(uint32_t tl, const TileMap& m, uint32_t rs,
                                             uint32_t cs) {
  uint32_t s, n;
  tile_rc(tl, m, s, n);
  return s * rs + n * cs;
}

// Each lane moves four 16-byte chunks of two interleaved amplitudes per tile
// (2048 / 2 / 256 = 4); tile-local bit 0 is state bit 0, so a chunk is consecutive.
// Chunk c of a lane holds tile-local indices 2 (tid + 256 c) + {0, 1}: against chunk 0 it
// sets tile-local bits 9 and 10, its second amplitude bit 0.  So a chunk's state offset and
// LDS positions are the lane's chunk-0 values plus terms that are the same for every lane
// (kept in scalar registers); the lane keeps one offset and one position, not eight.
constexpr int kChunks = (1 << kTileBits) / 2 / kBlock;
constexpr uint32_t chunk_tl(int c, int e) { return uint32_t(2 * kBlock * c + e); }

// ------------------------------------------------------------------------- FP32
// The tile sits in LDS as the real 2D x NT block operand (block rows 0..D-1 = Xr,
// D..2D-1 = Xi), element (row, col) at row * RS + col * CS.  Row-major (RS = NT + pad,
// CS = 1) keeps the B-fragment reads conflict-free and suits tiles whose lowest bits
// (the ones consecutive lanes step through when staging) are column bits; column-major
// (RS = 1, CS = 2D + 33) suits tiles whose lowest bits are group (row) bits, which
// row-major would stage into a single bank.  CS = 2D + 33 puts the 32 columns of a
// half-wave's B read on distinct banks and the two half-waves (rows r, r + 1) on
// disjoint ones.
template <int K, bool CM>
struct F32Shape {
  static constexpr int D = 1 << K;
  static constexpr int NT = 1 << (kTileBits - K);  // columns per tile
  static constexpr int RS = CM ? 1 : NT + (NT >= 64 ? 32 : 0);
  static constexpr int CS = CM ? 2 * D + 33 : 1;
  static constexpr int WORDS = CM ? NT * CS : 2 * D * RS;
  static constexpr int NTILE_COLS = NT / 32;       // 32-wide output blocks per row band
};

template <int K, bool NTA, bool CM>
__global__ __launch_bounds__(kBlock) void k_fused_f32(float* __restrict__ amp,
                                                      uint64_t n_tiles, TileMap map,
                                                      const float* __restrict__ ablock,
                                                      int xcd) {
  using S = F32Shape<K, CM>;
  constexpr int D = S::D, R = 2 * D, RS = S::RS, CS = S::CS;
  __shared__ float xb[S::WORDS];

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int mt = wave / S::NTILE_COLS, nt = wave % S::NTILE_COLS;  // this wave's 32x32 block

  // A fragments of the block matrix rows [32 mt, 32 mt + 32): a[kk] = A[row][2 kk + lane/32].
  float a[D];
#pragma unroll
  for (int kk = 0; kk < D; ++kk)
    a[kk] = ablock[(32 * mt + (lane & 31)) * R + 2 * kk + (lane >> 5)];
  // Complete these loads here: left pending, their waits land inside the tile loop, where
  // (memory operations completing in order) they also wait for the next tile's prefetch.
#pragma unroll
  for (int kk = 0; kk < D; ++kk) asm volatile("" : "+v"(a[kk]));

  // The lane's chunk-0 offset and LDS position (re; im D rows further).
  uint64_t off0 = tile_offset(2u * threadIdx.x, map);
  uint32_t pos0 = tile_pos(2u * threadIdx.x, map, RS, CS);
  auto off = [&](int c) { return off0 | tile_offset(chunk_tl(c, 0), map); };
  auto pos = [&](int c, int e) { return pos0 + tile_pos(chunk_tl(c, e), map, RS, CS); };

  // Per tile: the next tile's loads are issued, the matrix cores work on this tile (staged
  // in LDS), the result is stored, and the next tile is staged.  Staging the first tile
  // ahead of the loop keeps one memory-wait pattern in the loop, so that each staging
  // waits for its own loads only, not for the stores issued after them.
  floatx4 v[kChunks];
  const uint64_t G = gridDim.x;
  auto load = [&](uint64_t tt) {
    const uint64_t b = tile_base(tt, map);
#pragma unroll
    for (int c = 0; c < kChunks; ++c)
      v[c] = vload<NTA>(reinterpret_cast<const floatx4*>(amp + 2 * (b + off(c))));
  };
  auto stage = [&]() {
#pragma unroll
    for (int c = 0; c < kChunks; ++c)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const uint32_t q = pos(c, e);
        xb[q] = v[c][2 * e];
        xb[q + D * RS] = v[c][2 * e + 1];
      }
  };
  uint64_t t = xcd_block(xcd);  // consecutive tiles stay on one XCD
  if (t < n_tiles) {
    load(t);
    stage();
  }
  __syncthreads();
  for (; t < n_tiles; t += G) {
    // opaque to the compiler, so that the chunk offsets are formed here from the one
    // offset and position instead of being held in registers across the loop
    asm volatile("" : "+v"(off0), "+v"(pos0));
    const uint64_t base = tile_base(t, map);
    const uint64_t tn = t + G;
    if (tn < n_tiles) load(tn);  // in flight while the matrix cores work
    floatx16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f,
                    0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const int col = 32 * nt + (lane & 31);
    const int bcol = (lane >> 5) * RS + col * CS;
#pragma unroll
    for (int kk = 0; kk < D; ++kk) {
      const float b = xb[bcol + 2 * kk * RS];  // B[2 kk + lane / 32][col]
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[kk], b, acc, 0, 0, 0);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 32 * mt + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      xb[row * RS + col * CS] = acc[r];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
      floatx4 y;
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const uint32_t q = pos(c, e);
        y[2 * e] = xb[q];
        y[2 * e + 1] = xb[q + D * RS];
      }
      vstore<NTA>(reinterpret_cast<floatx4*>(amp + 2 * (base + off(c))), y);
    }
    if (tn < n_tiles) {
      __syncthreads();
      stage();
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------------- BF16x3
// The tile is split into bf16 hi and lo parts once, as it is staged: LDS holds the hi and
// lo B operands [col][block row] (block rows 0..D-1 = Xr, D..2D-1 = Xi), so that the eight
// consecutive K values of a B fragment are one 16-byte read and the MFMA loop converts
// nothing.  The FP32 result goes back through the same LDS as [col][block row] floats.
template <int K>
struct B16Shape {
  static constexpr int D = 1 << K;
  static constexpr int R = 2 * D;
  static constexpr int NT = 1 << (kTileBits - K);
  static constexpr int HSTRIDE = R + 8;  // bf16 per column of a B operand (16-byte rows,
                                         // 4-dword skew: conflict-free fragment reads)
  static constexpr int CSTRIDE = R + 4;  // floats per column of the result
  static constexpr int NTILE_COLS = NT / 32;
  static constexpr int KSTEPS = R / 16;
  static constexpr int B_BYTES = 2 * NT * HSTRIDE * 2;
  static constexpr int Y_BYTES = NT * CSTRIDE * 4;
  static constexpr int LDS_BYTES = B_BYTES > Y_BYTES ? B_BYTES : Y_BYTES;
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
  constexpr int D = S::D, R = S::R, HS = S::HSTRIDE, CS = S::CSTRIDE, NT = S::NT;
  __shared__ __align__(16) unsigned char lds[S::LDS_BYTES];
  __bf16* bhs = reinterpret_cast<__bf16*>(lds);  // hi B operand, NT x HS
  __bf16* bls = bhs + NT * HS;                    // lo B operand
  float* yb = reinterpret_cast<float*>(lds);      // result, NT x CS

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
  // Complete these loads here: left pending, their waits land inside the tile loop, where
  // (memory operations completing in order) they also wait for the next tile's prefetch.
#pragma unroll
  for (int ks = 0; ks < S::KSTEPS; ++ks) asm volatile("" : "+v"(ah[ks]), "+v"(al[ks]));

  // The lane's chunk-0 offset and LDS positions (column-major, re at +0, im at +D) in
  // bf16 units for the B operands and in floats for the result.
  uint64_t off0 = tile_offset(2u * threadIdx.x, map);
  uint32_t hpos0 = tile_pos(2u * threadIdx.x, map, 1, HS);
  uint32_t ypos0 = tile_pos(2u * threadIdx.x, map, 1, CS);
  auto off = [&](int c) { return off0 | tile_offset(chunk_tl(c, 0), map); };
  auto hpos = [&](int c, int e) { return hpos0 + tile_pos(chunk_tl(c, e), map, 1, HS); };
  auto ypos = [&](int c, int e) { return ypos0 + tile_pos(chunk_tl(c, e), map, 1, CS); };

  // Per tile as in the FP32 kernel: next tile's loads, matrix-core work on this tile, and
  // the next tile staged (split into hi and lo as it goes to LDS) before this one's stores.
  floatx4 v[kChunks];
  const uint64_t G = gridDim.x;
  auto load = [&](uint64_t tt) {
    const uint64_t b = tile_base(tt, map);
#pragma unroll
    for (int c = 0; c < kChunks; ++c)
      v[c] = vload<NTA>(reinterpret_cast<const floatx4*>(amp + 2 * (b + off(c))));
  };
  auto stage = [&]() {
#pragma unroll
    for (int c = 0; c < kChunks; ++c)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const uint32_t h = hpos(c, e);
        __bf16 hr, lr, hi_, li;
        split_bf16(v[c][2 * e], hr, lr);
        split_bf16(v[c][2 * e + 1], hi_, li);
        bhs[h] = hr;
        bls[h] = lr;
        bhs[h + D] = hi_;
        bls[h + D] = li;
      }
  };
  uint64_t t = xcd_block(xcd);  // consecutive tiles stay on one XCD
  if (t < n_tiles) {
    load(t);
    stage();
  }
  __syncthreads();
  for (; t < n_tiles; t += G) {
    // opaque to the compiler, so that the chunk offsets are formed here from the one
    // offset and position instead of being held in registers across the loop
    asm volatile("" : "+v"(off0), "+v"(hpos0), "+v"(ypos0));
    const uint64_t base = tile_base(t, map);
    const uint64_t tn = t + G;
    if (tn < n_tiles) load(tn);  // in flight while the matrix cores work
    floatx16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f,
                    0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const int col = 32 * nt + (lane & 31);
#pragma unroll
    for (int ks = 0; ks < S::KSTEPS; ++ks) {
      const int at = col * HS + 16 * ks + 8 * (lane >> 5);
      const bf16x8 bh = *reinterpret_cast<const bf16x8*>(bhs + at);
      const bf16x8 bl = *reinterpret_cast<const bf16x8*>(bls + at);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[ks], bh, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[ks], bl, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[ks], bh, acc, 0, 0, 0);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int row = 32 * mt + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      yb[col * CS + row] = acc[r];
    }
    __syncthreads();
    // this tile's result into registers, the next tile staged (its loads complete), and
    // only then this tile's stores: measured a little faster than storing first
    floatx4 y[kChunks];
#pragma unroll
    for (int c = 0; c < kChunks; ++c)
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const uint32_t q = ypos(c, e);
        y[c][2 * e] = yb[q];
        y[c][2 * e + 1] = yb[q + D];
      }
    if (tn < n_tiles) {
      __syncthreads();
      stage();
    }
#pragma unroll
    for (int c = 0; c < kChunks; ++c)
      vstore<NTA>(reinterpret_cast<floatx4*>(amp + 2 * (base + off(c))), y[c]);
    if (tn < n_tiles) __syncthreads();
  }
}

template <int K, bool NT>
hipError_t launch_k(const State& st, const TileMap& map, const void* mats, int precision,
                    const Tuning& t, hipStream_t s) {
  const uint64_t n_tiles = st.dim() >> kTileBits;
  const int cus = st.ctx && st.ctx->cus > 0 ? st.ctx->cus : 256;
  const uint64_t cap = t.mfma_blocks > 0 ? uint64_t(t.mfma_blocks)
                                         : uint64_t(precision == 1 && K == 4 ? 1 : 2) * cus;
  const uint64_t g = n_tiles < cap ? n_tiles : cap;
  float* amp = static_cast<float*>(st.amp);
  if (precision == 1) {
    // Column-major staging when most of tile-local bits 1..5 (stepped through by the 32
    // lanes of a half-wave) are group bits; knob 1 / 2 forces row- / column-major.
    int lane_rows = 0;
    for (int j = 1; j <= 5; ++j) lane_rows += (map.role[j] & 0x80u) ? 1 : 0;
    const bool cm = t.mfma_layout == 2 || (t.mfma_layout == 0 && lane_rows >= 3);
    const float* a = static_cast<const float*>(mats);
    if (cm)
      k_fused_f32<K, NT, true><<<unsigned(g), kBlock, 0, s>>>(amp, n_tiles, map, a, t.xcd);
    else
      k_fused_f32<K, NT, false><<<unsigned(g), kBlock, 0, s>>>(amp, n_tiles, map, a, t.xcd);
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
