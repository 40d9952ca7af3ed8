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
// The next tile's loads are issued before the current tile's MFMA work.
//
// FP32 variant: v_mfma_f32_32x32x2_f32, the block matrix held in registers as A fragments.
// BF16x3 variant: v_mfma_f32_32x32x16_bf16 with both operands split into bf16 hi + lo
// (x = hi + lo, lo = bf16(x - hi)); Y = Ahi Bhi + Ahi Blo + Alo Bhi accumulated in FP32
// (the lo*lo term, ~2^-16 relative, is dropped).
//
// MFMA tile passes (k_mfma_tiles, end of the file): a run of fused groups per HBM round
// trip.  A 2^12-amplitude tile stays in LDS while the pass's dense groups (BF16x3 products
// reading their operands straight from the tile) and diagonal groups (phases, the qubits
// outside the tile fixed per tile) are applied to it in order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

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

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

// split_bf16 of two values at once (one packed conversion each for hi and lo): the same
// roundings, so the same bits.
__device__ __forceinline__ void split_bf16x2(float a, float b, bf16x2& hi, bf16x2& lo) {
  const floatx2 x = {a, b};
  hi = __builtin_convertvector(x, bf16x2);
  lo = __builtin_convertvector(x - __builtin_convertvector(hi, floatx2), bf16x2);
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
  // Split-BF16 at k = 4: one workgroup per tile (its A fragments are 4 KiB, cheap to load
  // per tile, and a looping workgroup's next loads wait out its previous stores): 10-21%
  // faster than the persistent grid at 30 qubits; at k = 5 (16 KiB of fragments) four
  // tiles per workgroup, 1-7% faster.  6-qubit groups (64 KiB) keep a persistent grid of
  // two workgroups per CU; FP32 one (4-qubit) or two per CU, except 6-qubit groups: one
  // workgroup per 64 tiles, 2-3% faster.
  const uint64_t persistent = uint64_t(precision == 1 && K == 4 ? 1 : 2) * cus;
  const uint64_t cap = t.mfma_blocks > 0 ? uint64_t(t.mfma_blocks)
                       : precision == 2 && K == 4 ? (uint64_t(1) << 31) / kBlock
                       : precision == 2 && K == 5 ? std::max(n_tiles / 4, persistent)
                       : precision == 1 && K == 6 ? std::max(n_tiles / 64, persistent)
                                                  : persistent;
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

// ======================================================================= MFMA tile passes
namespace {

// LDS slot (8 bytes: re, im) of tile-local amplitude e: the 5-bit groups of e above bit 4
// are XOR-folded into its low 5 bits.  The map is linear over GF(2), so the slot of a | b
// (disjoint bits) is the XOR of the slots of a and b, and 32 lanes stepping through 5 tile
// bits whose positions differ mod 5 hit 32 distinct slots of a 256-byte row of banks,
// wherever in the tile those bits are.
__device__ __forceinline__ uint32_t swz(uint32_t e) {
  return e ^ (((e >> 5) ^ (e >> 10)) & 31u);
}

// State offset of tile-local index tl (tile-local bit j = state bit ts[j]).
__device__ __forceinline__ uint64_t mt_offset(uint32_t tl, const MtPassArg& pa) {
  uint64_t o = 0;
#pragma unroll
  for (int j = 0; j < 13; ++j)
    if (j < int(pa.tb)) o |= uint64_t((tl >> j) & 1u) << pa.ts[j];
  return o;
}

__device__ __forceinline__ uint64_t mt_base(uint64_t t, const MtPassArg& pa) {
#pragma unroll
  for (int j = 0; j < 13; ++j)
    if (j < int(pa.tb)) {
      const uint32_t b = pa.ts[j];
      t = ((t >> b) << (b + 1)) | (t & ((1ull << b) - 1));
    }
  return t;
}

// Byte j of a packed bit-position array (compile-time j: one scalar load and shift).
__device__ __forceinline__ uint32_t op_byte(const uint32_t* w, int j) {
  return (w[j >> 2] >> (8 * (j & 3))) & 0xffu;
}

// XOR of the words U[i] of the set bits i of v (v < 2^K).
template <int K>
__device__ __forceinline__ uint32_t uxor(const uint32_t (&U)[K], uint32_t v) {
  uint32_t x = 0;
#pragma unroll
  for (int i = 0; i < K; ++i)
    if ((v >> i) & 1u) x ^= U[i];
  return x;
}

// One dense k-qubit group on the tile in LDS: Y = U X as the real block product, split-BF16
// x3 on v_mfma_f32_32x32x16_bf16 (as k_fused_bf16x3).  Row s, column n of X is the real
// (s < D) or imaginary (s >= D) part of tile-local amplitude dep(s mod D over the group
// bits) | dep(n over the other bits).  The tile is interleaved complex in LDS, so one
// 8-byte read gives a column's entry for a real K step and for the matching imaginary one.
// The block matrix comes from the host with its rows permuted: row band mt (32 rows) holds
// the real rows 16 mt .. 16 mt + 15, then the imaginary rows D + 16 mt .. D + 16 mt + 15, so
// a lane's accumulators pair up into the re, im of 8 output amplitudes per block.  The
// 2^(TB-9) output blocks of 32 x 32 are dealt 2^(TB-11) per wave, all of one row band, so
// the band's A fragments serve them all.  The caller puts a barrier between the last
// group's writes and any later read of the tile.
template <int K, int TB>
__device__ __forceinline__ void mt_dense(float2* buf, const MtOp& op,
                                         const unsigned char* pool) {
  constexpr int D = 1 << K, R = 2 * D, KSTEPS = R / 16, KH = KSTEPS / 2;
  constexpr int NBC = 1 << (TB - K - 5);  // 32-column blocks per row band
  constexpr int BPW = 1 << (TB - 11);     // output blocks per wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = lane >> 5, l = lane & 31;
  const int b0 = wave * BPW, mt = b0 / NBC, nt0 = b0 % NBC;

  const __bf16* ahi = reinterpret_cast<const __bf16*>(pool + op.mat);
  const __bf16* alo = ahi + R * R;
  bf16x8 ah[KSTEPS], al[KSTEPS];
#pragma unroll
  for (int ks = 0; ks < KSTEPS; ++ks) {
    const int idx = (32 * mt + l) * R + 16 * ks + 8 * h;
    ah[ks] = *reinterpret_cast<const bf16x8*>(ahi + idx);
    al[ks] = *reinterpret_cast<const bf16x8*>(alo + idx);
  }
  // the tile's previous op done by every wave (the A loads above, issued first, are in
  // flight meanwhile)
  __syncthreads();
  // LDS byte offsets (8 x the slots: the XOR structure carries over)
  uint32_t U[K];
#pragma unroll
  for (int i = 0; i < K; ++i) U[i] = 8u * swz(1u << op_byte(op.bits, i));
  uint32_t wl = 0;  // the lane's column l over the other bits
#pragma unroll
  for (int i = 0; i < 5; ++i) wl |= uint32_t((l >> i) & 1) << op_byte(op.obits, i);
  wl = 8u * swz(wl);
  auto colw = [&](int nt) {  // block nt's first column over the other bits (uniform)
    uint32_t u = 0;
#pragma unroll
    for (int i = 5; i < TB - K; ++i) u |= uint32_t((nt >> (i - 5)) & 1) << op_byte(op.obits, i);
    return 8u * swz(u);
  };
  auto slot = [&](uint32_t byte) -> float2& {
    return *reinterpret_cast<float2*>(reinterpret_cast<char*>(buf) + byte);
  };
  auto mfma3 = [](const bf16x8& a_hi, const bf16x8& a_lo, const float* x, floatx16 c) {
    bf16x2 hh[4], lo[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) split_bf16x2(x[2 * j], x[2 * j + 1], hh[j], lo[j]);
    const bf16x8 bh = __builtin_shufflevector(
        __builtin_shufflevector(hh[0], hh[1], 0, 1, 2, 3),
        __builtin_shufflevector(hh[2], hh[3], 0, 1, 2, 3), 0, 1, 2, 3, 4, 5, 6, 7);
    const bf16x8 bl = __builtin_shufflevector(
        __builtin_shufflevector(lo[0], lo[1], 0, 1, 2, 3),
        __builtin_shufflevector(lo[2], lo[3], 0, 1, 2, 3), 0, 1, 2, 3, 4, 5, 6, 7);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_hi, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_hi, bl, c, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a_lo, bh, c, 0, 0, 0);
  };

  floatx16 acc[BPW];
  uint32_t wb[BPW];  // B rows 16 ks + 8 h + j of block bi
  auto init = [&](int bi) {
    acc[bi] = floatx16{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f,
                       0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    wb[bi] = wl ^ colw(nt0 + bi) ^ (h ? U[3] : 0u);
  };
  // K steps kr (real rows 16 kr + 8 h + j) and kr + KH (the imaginary rows of the same
  // amplitudes) of block bi: one read of 8 amplitudes, two splits, six products
  auto steps = [&](int bi, int kr) {
    // opaque per step: the step's read addresses are formed here, not all held at once
    asm volatile("" : "+v"(wb[bi]));
    float xr[8], xi[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float2 v = slot(wb[bi] ^ uxor<K>(U, uint32_t(16 * kr + j)));
      xr[j] = v.x;
      xi[j] = v.y;
    }
    acc[bi] = mfma3(ah[kr], al[kr], xr, acc[bi]);
    acc[bi] = mfma3(ah[kr + KH], al[kr + KH], xi, acc[bi]);
    // the next step's LDS reads are issued behind these products: hoisted, all of a
    // block's reads would be live at once
    asm volatile("" ::: "memory");
  };
  // output amplitude 16 mt + 8 (r >> 2) + 4 h + (r & 3): re in acc[r], im in acc[r + 8]
  auto write = [&](int bi) {
    uint32_t wo = wl ^ colw(nt0 + bi) ^ (h ? U[2] : 0u);
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      if ((r & 3) == 0) asm volatile("" : "+v"(wo));
      const uint32_t a = uint32_t(16 * mt + 8 * (r >> 2) + (r & 3));
      slot(wo ^ uxor<K>(U, a)) = float2{acc[bi][r], acc[bi][r + 8]};
    }
  };
  if constexpr (R == 32) {
    // one row band (k = 4): a wave's output blocks are its own columns, which no other wave
    // reads, so each block is written as soon as it is formed -- no barrier between the
    // reads and the writes, one block's accumulators live at a time
#pragma unroll
    for (int bi = 0; bi < BPW; ++bi) {
      init(bi);
#pragma unroll
      for (int kr = 0; kr < KH; ++kr) steps(bi, kr);
      write(bi);
    }
  } else {
    // several row bands read every column: all waves read before any writes.  Block by
    // block (alternating the blocks' steps, to overlap their chains of dependent products,
    // measured no faster)
#pragma unroll
    for (int bi = 0; bi < BPW; ++bi) init(bi);
#pragma unroll
    for (int bi = 0; bi < BPW; ++bi)
#pragma unroll
      for (int kr = 0; kr < KH; ++kr) steps(bi, kr);
    __syncthreads();
#pragma unroll
    for (int bi = 0; bi < BPW; ++bi) write(bi);
  }
  // no barrier here: whatever reads the tile next starts with one (the next dense group
  // after its A-fragment loads, a diagonal run, the tile's store)
}

// A run of diagonal ops [i0, i1): the lane's amplitudes e = tid + 256 j stay in registers
// across the run.  A phase index is the bits of the op's qubits in the state index: those
// outside the tile come from the tile's base (the same for the whole tile), those in it
// from e, split into the lane's part and a per-j part.  A single diagonal op reads its
// phases from the pass's table; a merged run (head op, then members) first builds, per
// tile, the product of its base table and its members' phases over its <= 8 tile bits (one
// entry per thread, in one of two LDS tables of `grp` entries used in turn), then applies
// that table once.
template <int TB>
__device__ __forceinline__ void mt_diag_run(float2* buf, const float2* tabs, float2* tgroup,
                                            uint32_t grp, const MtOp* ops,
                                            const unsigned char* pool, uint32_t i0,
                                            uint32_t i1, uint64_t base) {
  constexpr int E = (1 << TB) / kBlock;
  // opaque, so that the amplitudes' LDS slots are formed here, not held across the tiles
  uint32_t tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  float2 v[E];
#pragma unroll
  for (int j = 0; j < E; ++j) v[j] = buf[swz(tid + kBlock * j)];
  // multiply every amplitude by tab[its bits at the k positions `bw` (tile-local bits where
  // in_mask is set, else state bits)]
  auto apply = [&](const float2* tab, uint32_t k, uint32_t in_mask, uint64_t bw) {
    uint32_t fix = 0, lp = 0;  // out-of-tile bits; the lane's bits (tile-local bits < 8)
    uint32_t g8[TB - 8];       // index bits of tile-local bit 8 + b
#pragma unroll
    for (int c = 0; c < TB - 8; ++c) g8[c] = 0;
    for (uint32_t q = 0; q < k; ++q) {
      const uint32_t b = uint32_t(bw >> (8 * q)) & 0xffu;
      if (!((in_mask >> q) & 1u)) {
        fix |= uint32_t((base >> b) & 1u) << q;
      } else if (b < 8) {
        lp |= ((tid >> b) & 1u) << q;
      } else {
#pragma unroll
        for (int c = 0; c < TB - 8; ++c)
          if (int(b) == 8 + c) g8[c] = 1u << q;
      }
    }
    lp |= fix;
#pragma unroll
    for (int j = 0; j < E; ++j) {
      if ((j & 3) == 0) asm volatile("" : "+v"(lp));  // indices formed 4 at a time
      uint32_t idx = lp;
#pragma unroll
      for (int c = 0; c < TB - 8; ++c)
        if ((j >> c) & 1) idx ^= g8[c];
      const float2 ph = tab[idx];
      const float a = v[j].x, b = v[j].y;
      v[j] = float2{a * ph.x - b * ph.y, a * ph.y + b * ph.x};
    }
  };
  auto words = [](const MtOp& o) { return uint64_t(o.bits[0]) | (uint64_t(o.bits[1]) << 32); };
  uint32_t g = 0;
  for (uint32_t i = i0; i < i1;) {
    const MtOp& op = ops[i];
    if (op.kind == 1) {
      apply(tabs + op.tab, op.k, op.in_mask, words(op));
      ++i;
      continue;
    }
    float2* T = tgroup + grp * (g++ & 1u);
    if (tid < (1u << op.k)) {  // entry tid: the head's bits r set where bit r of tid is
      float2 p = op.tab == kMtNoTable ? float2{1.f, 0.f} : tabs[op.tab + tid];
      // members four at a time from their records: the records, then the phases, fetched
      // together, then multiplied in order.  Index bit q of a member is bit rank[q] of tid
      // (in the tile) or bit sbit[q] of the tile's base (outside); an unused rank 15 or
      // state bit 63 reads 0.
      const uint32_t* rec = reinterpret_cast<const uint32_t*>(pool + op.mat);
      const uint32_t nm = op.members;
      for (uint32_t m0 = 0; m0 < nm; m0 += 4) {
        uint32_t at[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t* r = rec + 8 * (m0 + u);
          uint32_t idx = 0, fix = 0;
#pragma unroll
          for (int q = 0; q < int(kMaxGateQubits); ++q) {
            const uint32_t rk = (r[1 + q / 4] >> (8 * (q % 4))) & 0xffu;
            const uint32_t sb = (r[3 + q / 4] >> (8 * (q % 4))) & 0xffu;
            idx |= ((tid >> rk) & 1u) << q;
            fix |= uint32_t((base >> sb) & 1u) << q;
          }
          at[u] = r[0] + (idx | fix);
        }
        float2 ph[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) ph[u] = tabs[at[u]];
#pragma unroll
        for (int u = 0; u < 4; ++u)
          if (m0 + u < nm) p = float2{p.x * ph[u].x - p.y * ph[u].y, p.x * ph[u].y + p.y * ph[u].x};
      }
      T[tid] = p;
    }
    __syncthreads();  // the table is complete; the other one's last readers are done
    apply(T, op.k, (1u << op.k) - 1u, words(op));
    i += 1 + op.members;
  }
#pragma unroll
  for (int j = 0; j < E; ++j) buf[swz(tid + kBlock * j)] = v[j];
}

// Waves per SIMD asked of the compiler: four for passes of groups up to 5 qubits (128
// registers: a few spill around the merged diagonal tables, and the fourth workgroup per CU
// still gains 7-10% where the pass's phase table leaves it room in LDS), two with 6-qubit
// groups (asked for 3, that kernel spills).
template <int TB, int KMAX, bool NTA>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(KMAX == 6 ? 2 : 4)))
void k_mfma_tiles(float* __restrict__ amp, uint64_t n_tiles, MtPassArg pa,
                  const MtOp* __restrict__ ops, const unsigned char* __restrict__ pool,
                  int xcd) {
  constexpr int N = 1 << TB;
  constexpr int CH = N / 2 / kBlock;  // 16-byte chunks (2 amplitudes) per lane
  // the tile, interleaved complex: slot swz(e) holds amplitude e
  __shared__ float2 buf[N];
  // the pass's phase table, then the merged runs' two per-tile tables: dynamic LDS of their
  // own size (pa.tab_entries + 2 pa.grp_entries), so passes with few diagonal phases leave
  // room for another workgroup per CU
  extern __shared__ float2 tabs[];
  float2* const tgroup = tabs + pa.tab_entries;
  {  // the pass's phase table, once per workgroup
    const float2* src = reinterpret_cast<const float2*>(pool + pa.tab_off);
    for (uint32_t e = threadIdx.x; e < pa.tab_entries; e += kBlock) tabs[e] = src[e];
  }
  // chunk c of a lane holds tile-local indices 2 (tid + 256 c) + {0, 1}: against chunk 0 it
  // sets tile-local bits 9.., the same for every lane (tile-local bit 0 is state bit 0)
  uint64_t off0 = mt_offset(2u * threadIdx.x, pa);
  const MtOp* pops = ops + pa.op_begin;
  for (uint64_t t = xcd_block(xcd); t < n_tiles; t += gridDim.x) {
    asm volatile("" : "+v"(off0));
    const uint64_t base = mt_base(t, pa);
    __syncthreads();  // the previous tile's stores have read LDS
    {
      floatx4 v[CH];
#pragma unroll
      for (int c = 0; c < CH; ++c)
        v[c] = vload<NTA>(reinterpret_cast<const floatx4*>(
            amp + 2 * (base + (off0 | mt_offset(uint32_t(2 * kBlock * c), pa)))));
#pragma unroll
      for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int e = 0; e < 2; ++e)
          buf[swz(2u * threadIdx.x + uint32_t(2 * kBlock * c + e))] =
              float2{v[c][2 * e], v[c][2 * e + 1]};
    }
    for (uint32_t i = 0; i < pa.n_ops;) {
      const MtOp& op = pops[i];
      if (op.kind == 0) {
        // (mt_dense starts with the barrier after its A-fragment loads)
        // instantiated up to the pass's widest group only: a kernel holding the 6-qubit
        // product's registers runs at lower occupancy
        if (KMAX == 4 || op.k == 4) mt_dense<4, TB>(buf, op, pool);
        else if (KMAX == 5 || op.k == 5) mt_dense<KMAX == 4 ? 4 : 5, TB>(buf, op, pool);
        else mt_dense<KMAX, TB>(buf, op, pool);
        ++i;
      } else {
        const uint32_t j = i + op.run;
        __syncthreads();  // the diagonal run's amplitudes were staged by other lanes
        mt_diag_run<TB>(buf, tabs, tgroup, pa.grp_entries, pops, pool, i, j, base);
        i = j;
      }
    }
    __syncthreads();
    asm volatile("" : "+v"(off0));  // chunk offsets re-formed here, not held across the ops
#pragma unroll
    for (int c = 0; c < CH; ++c) {
      floatx4 y;
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const float2 a = buf[swz(2u * threadIdx.x + uint32_t(2 * kBlock * c + e))];
        y[2 * e] = a.x;
        y[2 * e + 1] = a.y;
      }
      vstore<NTA>(reinterpret_cast<floatx4*>(
                      amp + 2 * (base + (off0 | mt_offset(uint32_t(2 * kBlock * c), pa)))),
                  y);
    }
  }
}

template <int TB, int KMAX, bool NT>
hipError_t launch_mt(const State& st, const MtPassArg& pa, const MtOp* ops, const void* pool,
                     const Tuning& t, hipStream_t s) {
  const uint64_t n_tiles = st.dim() >> TB;
  const size_t entries = size_t(pa.tab_entries) + 2 * size_t(pa.grp_entries);
  const size_t dyn = (entries > 0 ? entries : 1) * sizeof(float2);
  // one workgroup per tile (tile_blocks 0) or tile_blocks workgroups looping over them;
  // one per tile measured 4-10% faster than as many as are resident at once
  uint64_t g = n_tiles;
  if (t.tile_blocks > 0 && g > uint64_t(t.tile_blocks)) g = uint64_t(t.tile_blocks);
  if (g > (uint64_t(1) << 31) / kBlock) g = (uint64_t(1) << 31) / kBlock;
  k_mfma_tiles<TB, KMAX, NT><<<unsigned(g), kBlock, dyn, s>>>(
      static_cast<float*>(st.amp), n_tiles, pa, ops,
      static_cast<const unsigned char*>(pool), t.xcd);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_mfma_tiles(const State& st, const MtPassArg& pa, const MtOp* ops,
                             const void* pool, const Tuning& t, hipStream_t s) {
  if (st.dtype != 1 || pa.tb != 12 || st.n < pa.tb || pa.ts[0] != 0 || pa.kmax < 4 ||
      pa.kmax > 6 || pa.tab_entries > kMtTabEntries || pa.grp_entries > 256)
    return hipErrorInvalidValue;
  const bool nt = nontemporal_for(st, t, true);
  auto by_k = [&](auto tb) {
    constexpr int TB = decltype(tb)::value;
    switch (pa.kmax) {
      case 4: return nt ? launch_mt<TB, 4, true>(st, pa, ops, pool, t, s)
                        : launch_mt<TB, 4, false>(st, pa, ops, pool, t, s);
      case 5: return nt ? launch_mt<TB, 5, true>(st, pa, ops, pool, t, s)
                        : launch_mt<TB, 5, false>(st, pa, ops, pool, t, s);
      default: return nt ? launch_mt<TB, 6, true>(st, pa, ops, pool, t, s)
                         : launch_mt<TB, 6, false>(st, pa, ops, pool, t, s);
    }
  };
  return by_k(std::integral_constant<int, 12>{});
}

}  // namespace qsvh
