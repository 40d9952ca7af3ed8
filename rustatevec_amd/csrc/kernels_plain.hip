// kernels_plain.hip -- the plain (unfused, exact-precision) gate kernels for gfx950.
//
// Amplitudes are interleaved complex pairs of T (double for complex128, float for
// complex64; amp.hpp).  Every streaming kernel moves 16-byte chunks (A = 1 complex128 or
// 2 complex64 amplitudes per access), so a wave64 load instruction covers 1 KiB of
// contiguous HBM, and orders its blocks XCD-contiguously (Tuning::xcd).
//
// Arithmetic is the reference's, operation for operation and without contraction (this
// file is compiled with -ffp-contract=off): complex products as (ac - bd, ad + bc), sums
// in the reference's order.  The complex128 results are therefore bit-identical to
// CpuBackend (crates/qsv-core/src/backend/cpu.rs:71-114,221-247).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "amp.hpp"
#include "engine.hpp"

namespace qsvh {
namespace {

constexpr int kBlock = 256;

// 2x2 gate, row-major entries g00 g01 g10 g11.
template <typename T>
struct G2 {
  T r[4], i[4];
};

// one_qubit_block body (cpu.rs:84-93): y0 = g00 x0 + g01 x1, y1 = g10 x0 + g11 x1.
template <typename T>
__device__ __forceinline__ void pair_update(T& x0r, T& x0i, T& x1r, T& x1i, const G2<T>& g) {
  const T ar = g.r[0] * x0r - g.i[0] * x0i, ai = g.r[0] * x0i + g.i[0] * x0r;
  const T br = g.r[1] * x1r - g.i[1] * x1i, bi = g.r[1] * x1i + g.i[1] * x1r;
  const T cr = g.r[2] * x0r - g.i[2] * x0i, ci = g.r[2] * x0i + g.i[2] * x0r;
  const T dr = g.r[3] * x1r - g.i[3] * x1i, di = g.r[3] * x1i + g.i[3] * x1r;
  x0r = ar + br;
  x0i = ai + bi;
  x1r = cr + dr;
  x1i = ci + di;
}

// Work item of unroll slot u in the grid-stride iteration whose first item for this thread
// is v0: strided (slot u is u * nthr items further, so a wave's loads stay 1 KiB contiguous
// per slot) or, with `contig`, the thread's U items adjacent.
__device__ __forceinline__ uint64_t item_of(uint64_t v0, uint64_t tid, int u, int U,
                                            uint64_t nthr, int contig) {
  return contig ? (v0 - tid) + tid * uint64_t(U) + uint64_t(u) : v0 + uint64_t(u) * nthr;
}

// ---------------------------------------------------------------------------- 1q
// Target q >= log2(A): each work item is A consecutive "lower" amplitudes, i.e. the
// nested-block walk of cpu.rs:148-175 with the 2^(n-1) pairs enumerated as
// lo = insert_zero_bit(x, q), hi = lo + 2^q (layout.rs:22-26).  Consecutive lanes read
// consecutive 16-byte chunks at lo and at hi: every access is coalesced for any q.
template <typename T, int A, int U, bool NT>
__global__ __launch_bounds__(kBlock) void k_1q_wide(T* __restrict__ amp, uint64_t n_items,
                                                    uint32_t q, G2<T> g, int contig, int xcd) {
  const uint64_t lowmask = (1ull << q) - 1;
  const uint64_t half = 1ull << q;
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  const uint64_t tid = xcd_block(xcd) * blockDim.x + threadIdx.x;
  for (uint64_t v0 = tid; v0 < n_items; v0 += nthr * U) {
    T ar[U][A], ai[U][A], br[U][A], bi[U][A];
    uint64_t lo[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t v = item_of(v0, tid, u, U, nthr, contig);
      const uint64_t x = v * A;
      lo[u] = ((x & ~lowmask) << 1) | (x & lowmask);
      if (v < n_items) {
        cload<T, A, NT>(amp, lo[u], ar[u], ai[u]);
        cload<T, A, NT>(amp, lo[u] + half, br[u], bi[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t v = item_of(v0, tid, u, U, nthr, contig);
      if (v < n_items) {
#pragma unroll
        for (int j = 0; j < A; ++j) pair_update(ar[u][j], ai[u][j], br[u][j], bi[u][j], g);
        cstore<T, A, NT>(amp, lo[u], ar[u], ai[u]);
        cstore<T, A, NT>(amp, lo[u] + half, br[u], bi[u]);
      }
    }
  }
}

// Same gate, one 16-byte chunk per lane: lanes 0..31 of a wave load the lower chunks of 32
// consecutive pair items, lanes 32..63 the matching upper chunks, and v_permlane32_swap
// hands each lane its partner chunk.  Every lane then does one load and one store (the
// access pattern of the diagonal kernel, 4-8% faster than k_1q_wide's two of each on
// HBM-resident states) and computes its own output row of pair_update with the same
// operations, so the results are bitwise those of k_1q_wide.  n_chunks must be a multiple
// of 64, so that whole waves stay together through the grid-stride loop.
template <typename T, int A>
__device__ __forceinline__ void partner_chunk(const T (&r)[A], const T (&m)[A], T (&pr)[A],
                                              T (&pm)[A], bool up) {
  static_assert(sizeof(T) * A == 8, "16-byte chunks");
  uint32_t w[4], o[4];
  __builtin_memcpy(w, r, 8);
  __builtin_memcpy(w + 2, m, 8);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    // lanes 0..31 receive lanes 32..63 of w in the second result, and vice versa
    const auto sw = __builtin_amdgcn_permlane32_swap(w[k], w[k], false, false);
    o[k] = up ? sw[0] : sw[1];
  }
  __builtin_memcpy(pr, o, 8);
  __builtin_memcpy(pm, o + 2, 8);
}

template <typename T, int A, bool NT>
__global__ __launch_bounds__(kBlock) void k_1q_lane(T* __restrict__ amp, uint64_t n_chunks,
                                                    uint32_t q, G2<T> g, int xcd) {
  const uint64_t lowmask = (1ull << q) - 1;
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  const uint32_t lane = threadIdx.x & 63;
  const bool up = lane >= 32;
  // this lane's output row of the gate: (g00, g01) for lower chunks, (g10, g11) for upper
  const T g0r = up ? g.r[2] : g.r[0], g0i = up ? g.i[2] : g.i[0];
  const T g1r = up ? g.r[3] : g.r[1], g1i = up ? g.i[3] : g.i[1];
  for (uint64_t t = xcd_block(xcd) * blockDim.x + threadIdx.x; t < n_chunks; t += nthr) {
    const uint64_t x = ((t >> 6) * 32 + (lane & 31)) * A;  // first lower amplitude of the item
    const uint64_t lo = ((x & ~lowmask) << 1) | (x & lowmask);
    const uint64_t me = up ? lo + (1ull << q) : lo;
    T xr[A], xi[A], pr[A], pi[A];
    cload<T, A, NT>(amp, me, xr, xi);
    partner_chunk<T, A>(xr, xi, pr, pi, up);
#pragma unroll
    for (int j = 0; j < A; ++j) {
      const T x0r = up ? pr[j] : xr[j], x0i = up ? pi[j] : xi[j];
      const T x1r = up ? xr[j] : pr[j], x1i = up ? xi[j] : pi[j];
      const T ar = g0r * x0r - g0i * x0i, ai = g0r * x0i + g0i * x0r;
      const T br = g1r * x1r - g1i * x1i, bi = g1r * x1i + g1i * x1r;
      xr[j] = ar + br;
      xi[j] = ai + bi;
    }
    cstore<T, A, NT>(amp, me, xr, xi);
  }
}

// Target q < log2(A) (complex64, q = 0): both partners of every pair sit inside one
// 16-byte chunk, so each work item is one chunk of A amplitudes holding A/2 complete
// pairs.  Q is the target qubit (a compile-time constant so the pairing stays in registers).
template <typename T, int A, int Q, int U, bool NT>
__global__ __launch_bounds__(kBlock) void k_1q_intra(T* __restrict__ amp, uint64_t n_items,
                                                     G2<T> g, int contig, int xcd) {
  constexpr int h = 1 << Q;
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  const uint64_t tid = xcd_block(xcd) * blockDim.x + threadIdx.x;
  for (uint64_t v0 = tid; v0 < n_items; v0 += nthr * U) {
    T xr[U][A], xi[U][A];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t v = item_of(v0, tid, u, U, nthr, contig);
      if (v < n_items) cload<T, A, NT>(amp, v * A, xr[u], xi[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t v = item_of(v0, tid, u, U, nthr, contig);
      if (v < n_items) {
#pragma unroll
        for (int j = 0; j < A; ++j)
          if (((j >> Q) & 1) == 0) pair_update(xr[u][j], xi[u][j], xr[u][j + h], xi[u][j + h], g);
        cstore<T, A, NT>(amp, v * A, xr[u], xi[u]);
      }
    }
  }
}

// ---------------------------------------------------------------------- diagonal
// diag_chunk (cpu.rs:99-114): psi_i <- d[gather_bits(i, qs)] * psi_i in one sequential
// pass (layout.rs:43-49).  The 2^m phases are staged in LDS.
template <typename T>
struct DiagArg {
  uint32_t m;
  uint32_t q[kMaxGateQubits];
  T dr[64], di[64];
};

template <typename T, int A, int U, bool NT>
__global__ __launch_bounds__(kBlock) void k_diag(T* __restrict__ amp, uint64_t n_items,
                                                 DiagArg<T> a, int contig, int xcd) {
  // each wave its own copy of the (up to 64-entry) table, written by its own lanes after
  // they issue their first loads: LDS operations of one wave complete in order, so no
  // workgroup barrier is needed before the reads (one per block measured 3-4% on 30-qubit
  // launches), and the table's kernel-argument loads overlap the amplitudes' loads
  __shared__ T wdr[kBlock / 64][64], wdi[kBlock / 64][64];
  const uint32_t sub = 1u << a.m, lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  T* const sdr = wdr[wv];
  T* const sdi = wdi[wv];
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  const uint64_t tid = xcd_block(xcd) * blockDim.x + threadIdx.x;
  T xr[U][A], xi[U][A];
  auto load = [&](uint64_t v0) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t v = item_of(v0, tid, u, U, nthr, contig);
      if (v < n_items) cload<T, A, NT>(amp, v * A, xr[u], xi[u]);
    }
  };
  load(tid);
  if (lane < sub) {
    sdr[lane] = a.dr[lane];
    sdi[lane] = a.di[lane];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  auto apply = [&](uint64_t v0) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t v = item_of(v0, tid, u, U, nthr, contig);
      if (v < n_items) {
#pragma unroll
        for (int j = 0; j < A; ++j) {
          const uint64_t i = v * A + j;
          uint32_t s = 0;
          for (uint32_t b = 0; b < a.m; ++b) s |= uint32_t((i >> a.q[b]) & 1u) << b;
          const T dr = sdr[s], di = sdi[s];
          const T x_r = xr[u][j], x_i = xi[u][j];
          xr[u][j] = dr * x_r - di * x_i;
          xi[u][j] = dr * x_i + di * x_r;
        }
        cstore<T, A, NT>(amp, v * A, xr[u], xi[u]);
      }
    }
  };
  apply(tid);
  for (uint64_t v0 = tid + nthr * U; v0 < n_items; v0 += nthr * U) {
    load(v0);
    apply(v0);
  }
}

// Diagonal gates on 1 or 2 qubits (RZ, CP, CZ, RZZ: almost every diagonal launch) take
// their phases as scalar arguments: no LDS table, no barrier, nothing per block but the
// loads.  Same multiply as k_diag.
template <typename T>
struct Diag4 {
  T r[4], i[4];
};

template <typename T, int A, int M, bool NT>
__global__ __launch_bounds__(kBlock) void k_diag_small(T* __restrict__ amp, uint64_t n_items,
                                                       uint32_t q0, uint32_t q1, Diag4<T> d,
                                                       int xcd) {
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t v = xcd_block(xcd) * blockDim.x + threadIdx.x; v < n_items; v += nthr) {
    T xr[A], xi[A];
    cload<T, A, NT>(amp, v * A, xr, xi);
#pragma unroll
    for (int j = 0; j < A; ++j) {
      const uint64_t i = v * A + j;
      const uint32_t b0 = uint32_t(i >> q0) & 1u;
      const uint32_t b1 = M == 2 ? uint32_t(i >> q1) & 1u : 0u;
      const T lo_r = b0 ? d.r[1] : d.r[0], lo_i = b0 ? d.i[1] : d.i[0];
      const T hi_r = b0 ? d.r[3] : d.r[2], hi_i = b0 ? d.i[3] : d.i[2];
      const T dr = b1 ? hi_r : lo_r, di = b1 ? hi_i : lo_i;
      const T x_r = xr[j], x_i = xi[j];
      xr[j] = dr * x_r - di * x_i;
      xi[j] = dr * x_i + di * x_r;
    }
    cstore<T, A, NT>(amp, v * A, xr, xi);
  }
}

// ---------------------------------------------------------------------- m-qubit// ---------------------------------------------------------------------- m-qubit
// insert_zero_bits over strictly ascending positions (layout.rs:70-87).
template <int M>
__device__ __forceinline__ uint64_t insert_zero_bits(uint64_t x, const uint32_t* sorted) {
#pragma unroll
  for (int j = 0; j < M; ++j) {
    const uint32_t b = sorted[j];
    x = ((x >> b) << (b + 1)) | (x & ((1ull << b) - 1));
  }
  return x;
}

template <int M>
struct MqIdx {
  uint32_t sorted[M];
  uint64_t off[1 << M];  // scatter_bits(s, qs) for every internal sub-index s
};

template <typename T, int M>
struct MatArg {
  T r[1 << (2 * M)], i[1 << (2 * M)];
};

// mq_block (cpu.rs:221-247) for m = 2, 3 with the matrix in kernel arguments (scalar
// registers).  A work item is VW consecutive blocks o..o+VW-1; when the lowest gate qubit
// is >= log2(VW) their amplitudes are consecutive, so every gather is one 16-byte chunk.
template <typename T, int M, int VW, bool NT>
__global__ __launch_bounds__(kBlock) void k_mq_reg(T* __restrict__ amp, uint64_t n_items,
                                                   MqIdx<M> idx, MatArg<T, M> g, int xcd) {
  constexpr int D = 1 << M;
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t v = xcd_block(xcd) * blockDim.x + threadIdx.x; v < n_items; v += nthr) {
    const uint64_t base = insert_zero_bits<M>(v * VW, idx.sorted);
    T ar[D][VW], ai[D][VW];
#pragma unroll
    for (int s = 0; s < D; ++s) cload<T, VW, NT>(amp, base + idx.off[s], ar[s], ai[s]);
#pragma unroll
    for (int r = 0; r < D; ++r) {
      T yr[VW], yi[VW];
#pragma unroll
      for (int w = 0; w < VW; ++w) {
        T accr = T(0), acci = T(0);
#pragma unroll
        for (int c = 0; c < D; ++c) {
          const T er = g.r[r * D + c], ei = g.i[r * D + c];
          const T pr = er * ar[c][w] - ei * ai[c][w];
          const T pi = er * ai[c][w] + ei * ar[c][w];
          accr = accr + pr;
          acci = acci + pi;
        }
        yr[w] = accr;
        yi[w] = acci;
      }
      cstore<T, VW, NT>(amp, base + idx.off[r], yr, yi);
    }
  }
}

// mq_block for m = 2, 3 with one 16-byte chunk per lane (the access pattern of the
// diagonal and k_1q_lane kernels: one load and one store per lane).  A wave covers 64 >> M
// consecutive work items (A consecutive blocks each, as in k_mq_reg: the lowest gate qubit
// is >= log2(A)); its lanes are 2^M groups of 64 >> M, group s loading member s (internal
// sub-index s, offset scatter_bits(s, q)) of every item.  A lane gathers the other members
// of its item from the lanes 8, 16 and 32 apart (DPP row rotate by 8, v_permlane16_swap,
// v_permlane32_swap) and computes its own output row s with k_mq_reg's operations in the
// same order, so results are bitwise those of k_mq_reg.
template <typename T>
__device__ __forceinline__ void swap_rows(uint32_t v, uint32_t& lo, uint32_t& hi, int bit) {
  // the value of lane (lane & ~2^bit) into lo and of lane (lane | 2^bit) into hi
  if (bit == 5) {
    const auto sw = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    lo = sw[0];
    hi = sw[1];
  } else {
    const auto sw = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    lo = sw[0];
    hi = sw[1];
  }
}

template <typename T, int M, int A, bool NT>
__global__ __launch_bounds__(kBlock) void k_mq_lane(T* __restrict__ amp, uint64_t n_waves,
                                                    MqIdx<M> idx, MatArg<T, M> g, int xcd) {
  constexpr int D = 1 << M;
  constexpr int GL = 64 >> M;          // lanes per member group = items per wave
  constexpr int W = 4;                 // 32-bit words per 16-byte chunk
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t s = lane / GL;        // this lane's member (and output row)
  // this lane's row of the gate (kernel arguments selected once per lane)
  T er[D], ei[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    T r = g.r[c], i = g.i[c];
#pragma unroll
    for (int rr = 1; rr < D; ++rr)
      if (s == uint32_t(rr)) {
        r = g.r[rr * D + c];
        i = g.i[rr * D + c];
      }
    er[c] = r;
    ei[c] = i;
  }
  uint64_t my_off = idx.off[0];
#pragma unroll
  for (int rr = 1; rr < D; ++rr)
    if (s == uint32_t(rr)) my_off = idx.off[rr];
  const uint64_t nw = (uint64_t(gridDim.x) * blockDim.x) >> 6;
  for (uint64_t w = (xcd_block(xcd) * blockDim.x + threadIdx.x) >> 6; w < n_waves; w += nw) {
    const uint64_t item = w * GL + (lane % GL);
    const uint64_t a = insert_zero_bits<M>(item * A, idx.sorted) + my_off;
    T xr[A], xi[A];
    cload<T, A, NT>(amp, a, xr, xi);
    uint32_t v[W];
    __builtin_memcpy(v, xr, 8);
    __builtin_memcpy(v + 2, xi, 8);
    // x[c][k]: word k of member c's chunk
    uint32_t x[D][W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
      uint32_t y[D];
      if constexpr (M == 3) {
        // lanes 8 apart (member bit 0): the row-rotated copy, ordered by the lane's bit 3
        const uint32_t o = uint32_t(__builtin_amdgcn_update_dpp(int(v[k]), int(v[k]), 0x128,
                                                                0xF, 0xF, true));
        const bool b3 = (lane >> 3) & 1u;
        const uint32_t e0 = b3 ? o : v[k], e1 = b3 ? v[k] : o;
        uint32_t a0, a1, b0, b1;
        swap_rows<T>(e0, a0, a1, 4);  // member bit 1
        swap_rows<T>(e1, b0, b1, 4);
        swap_rows<T>(a0, y[0], y[4], 5);  // member bit 2
        swap_rows<T>(b0, y[1], y[5], 5);
        swap_rows<T>(a1, y[2], y[6], 5);
        swap_rows<T>(b1, y[3], y[7], 5);
      } else {
        uint32_t e0, e1;
        swap_rows<T>(v[k], e0, e1, 4);  // member bit 0
        swap_rows<T>(e0, y[0], y[2], 5);  // member bit 1
        swap_rows<T>(e1, y[1], y[3], 5);
      }
#pragma unroll
      for (int c = 0; c < D; ++c) x[c][k] = y[c];
    }
    T yr[A], yi[A];
#pragma unroll
    for (int j = 0; j < A; ++j) {
      T accr = T(0), acci = T(0);
#pragma unroll
      for (int c = 0; c < D; ++c) {
        T cr[A], ci[A];
        __builtin_memcpy(cr, x[c], 8);
        __builtin_memcpy(ci, x[c] + 2, 8);
        const T pr = er[c] * cr[j] - ei[c] * ci[j];
        const T pi = er[c] * ci[j] + ei[c] * cr[j];
        accr = accr + pr;
        acci = acci + pi;
      }
      yr[j] = accr;
      yi[j] = acci;
    }
    cstore<T, A, NT>(amp, a, yr, yi);
  }
}

// mq_block for complex128 gates on a qubit below 6 (one block per thread gathers chunks
// 2^q0 x 16 bytes apart: below 3, half or a quarter of each 128-byte line per load; from 3
// to 5, lines in 256-1024-byte steps): a wave takes the 64
// consecutive amplitudes of a run (index bits 0..5 are its lanes) for each value of the
// gate's NH qubits at 6 and above (one load each), gathers every member of its blocks from
// the lanes holding them (the gate qubits below 6 are lane bits; ds_bpermute), and
// computes its own rows with k_mq_reg's operations in the same order, so the results are
// bitwise those of k_mq_reg.  The lane's rows (one per value of the high qubits) are
// selected from the kernel arguments once.  Launched for m = 2, NH = 1.
template <int M, int NH>
struct MqLow {
  uint32_t hsort[NH > 0 ? NH : 1];  // the high gate qubits, ascending
  uint64_t hoff[1 << NH];           // offset of high sub-index h (bit i: hsort[i])
  uint32_t q[M];                    // the gate's qubits in matrix-index order
  uint32_t hrank[M];                // of a high qubit q[j]: its bit in h
  uint32_t c_lane[1 << M];          // member c: its lane bits (the low gate qubits)
  uint32_t c_h[1 << M];             // member c: its high sub-index
  uint32_t glmask;                  // lane bits of the low gate qubits
};

template <int M, int NH, bool NT>
__global__ __launch_bounds__(kBlock) void k_mq_low(double* __restrict__ amp, uint64_t n_runs,
                                                   MqLow<M, NH> ix, MatArg<double, M> g,
                                                   int xcd) {
  constexpr int D = 1 << M, H = 1 << NH;
  const uint32_t lane = threadIdx.x & 63;
  // row of each of the lane's amplitudes: bit j from lane bit q[j] or from h
  double er[H][D], ei[H][D];
#pragma unroll
  for (int h = 0; h < H; ++h) {
    uint32_t r = 0;
#pragma unroll
    for (int j = 0; j < M; ++j)
      r |= (ix.q[j] < 6 ? (lane >> ix.q[j]) & 1u : (uint32_t(h) >> ix.hrank[j]) & 1u) << j;
#pragma unroll
    for (int c = 0; c < D; ++c) {
      double vr = g.r[c], vi = g.i[c];
#pragma unroll
      for (int rr = 1; rr < D; ++rr)
        if (r == uint32_t(rr)) {
          vr = g.r[rr * D + c];
          vi = g.i[rr * D + c];
        }
      er[h][c] = vr;
      ei[h][c] = vi;
    }
  }
  const uint32_t keep = lane & ~ix.glmask;
  const uint64_t nw = (uint64_t(gridDim.x) * blockDim.x) >> 6;
  for (uint64_t w = (xcd_block(xcd) * blockDim.x + threadIdx.x) >> 6; w < n_runs; w += nw) {
    uint64_t base = w << 6;
#pragma unroll
    for (int i = 0; i < NH; ++i) {
      const uint32_t b = ix.hsort[i];
      base = ((base >> b) << (b + 1)) | (base & ((uint64_t(1) << b) - 1));
    }
    base += lane;
    uint32_t v[H][4];  // the lane's amplitude for each high sub-index, as 32-bit words
#pragma unroll
    for (int h = 0; h < H; ++h) {
      double xr[1], xi[1];
      cload<double, 1, NT>(amp, base + ix.hoff[h], xr, xi);
      __builtin_memcpy(v[h], xr, 8);
      __builtin_memcpy(v[h] + 2, xi, 8);
    }
    // member c of each of the lane's blocks: held by lane keep | c_lane[c], in its v[c_h[c]]
    double cr[D], ci[D];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      const int src = int((keep | ix.c_lane[c]) << 2);
      uint32_t u[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        uint32_t val = v[0][k];
#pragma unroll
        for (int h = 1; h < H; ++h)
          if (ix.c_h[c] == uint32_t(h)) val = v[h][k];
        u[k] = uint32_t(__builtin_amdgcn_ds_bpermute(src, int(val)));
      }
      __builtin_memcpy(&cr[c], u, 8);
      __builtin_memcpy(&ci[c], u + 2, 8);
    }
#pragma unroll
    for (int h = 0; h < H; ++h) {
      double accr = 0.0, acci = 0.0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const double pr = er[h][c] * cr[c] - ei[h][c] * ci[c];
        const double pi = er[h][c] * ci[c] + ei[h][c] * cr[c];
        accr = accr + pr;
        acci = acci + pi;
      }
      const double yr[1] = {accr}, yi[1] = {acci};
      cstore<double, 1, NT>(amp, base + ix.hoff[h], yr, yi);
    }
  }
}

// mq_block for m = 4..6 (fused groups): a workgroup stages a tile of C consecutive blocks
// (2^m x C amplitudes) and the gate matrix in LDS, then each lane forms outputs
// out[s][c] = sum_j U[s][j] X[j][c] in ascending j (the reference's order).  Loads run
// along c, the direction in which the blocks' amplitudes are consecutive.
template <typename T, int M>
__global__ __launch_bounds__(kBlock) void k_mq_lds(T* __restrict__ amp, uint64_t n_blocks,
                                                   MqIdx<M> idx, const T* __restrict__ mat) {
  constexpr int D = 1 << M;
  constexpr int C = 1024 / D;
  __shared__ T ur[D * D], ui[D * D];
  __shared__ T xr[D * C], xi[D * C];
  for (int e = threadIdx.x; e < D * D; e += kBlock) {
    ur[e] = mat[2 * e];
    ui[e] = mat[2 * e + 1];
  }
  const uint64_t n_tiles = (n_blocks + C - 1) / C;
  for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    __syncthreads();
    for (int e = threadIdx.x; e < D * C; e += kBlock) {
      const int c = e % C, s = e / C;
      const uint64_t o = t * C + c;
      if (o < n_blocks) {
        const uint64_t a = insert_zero_bits<M>(o, idx.sorted) + idx.off[s];
        T r[1], m[1];
        cload<T, 1>(amp, a, r, m);
        xr[e] = r[0];
        xi[e] = m[0];
      }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < D * C; e += kBlock) {
      const int c = e % C, s = e / C;
      const uint64_t o = t * C + c;
      if (o >= n_blocks) continue;
      T accr = T(0), acci = T(0);
      for (int j = 0; j < D; ++j) {
        const T er = ur[s * D + j], ei = ui[s * D + j];
        const T x_r = xr[j * C + c], x_i = xi[j * C + c];
        const T pr = er * x_r - ei * x_i;
        const T pi = er * x_i + ei * x_r;
        accr = accr + pr;
        acci = acci + pi;
      }
      const uint64_t a = insert_zero_bits<M>(o, idx.sorted) + idx.off[s];
      const T r[1] = {accr}, m[1] = {acci};
      cstore<T, 1>(amp, a, r, m);
    }
  }
}

// ---------------------------------------------------------------------- init / reduce
template <typename T, int A>
__global__ __launch_bounds__(kBlock) void k_init_basis(T* __restrict__ amp, uint64_t n_items,
                                                       uint64_t basis, int xcd) {
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t v = xcd_block(xcd) * blockDim.x + threadIdx.x; v < n_items; v += nthr) {
    T zr[A], zi[A];
#pragma unroll
    for (int j = 0; j < A; ++j) {
      zr[j] = (v * A + j == basis) ? T(1) : T(0);
      zi[j] = T(0);
    }
    cstore<T, A>(amp, v * A, zr, zi);
  }
}

// StateVector::probabilities (state/mod.rs:130-134): p_i = re^2 + im^2 in T.
template <typename T, int A>
__global__ __launch_bounds__(kBlock) void k_probs(const T* __restrict__ amp,
                                                  T* __restrict__ p, uint64_t n_items,
                                                  int xcd) {
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t v = xcd_block(xcd) * blockDim.x + threadIdx.x; v < n_items; v += nthr) {
    T xr[A], xi[A];
    cload<T, A>(amp, v * A, xr, xi);
    VecT<T, A> y;
#pragma unroll
    for (int j = 0; j < A; ++j) y[j] = xr[j] * xr[j] + xi[j] * xi[j];
    *reinterpret_cast<VecT<T, A>*>(p + v * A) = y;
  }
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// f64 partial sums of sign(i & mask) * |psi_i|^2 (norm when mask == 0; the expectation
// of sample.rs:177-200 otherwise), one partial per block; fixed order, so repeatable.
template <typename T, int A>
__global__ __launch_bounds__(kBlock) void k_reduce(const T* __restrict__ amp, uint64_t n_items,
                                                   uint64_t mask, double* __restrict__ part) {
  __shared__ double ws[kBlock / 64];
  double acc = 0.0;
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < n_items; v += nthr) {
    T xr[A], xi[A];
    cload<T, A>(amp, v * A, xr, xi);
#pragma unroll
    for (int j = 0; j < A; ++j) {
      const double a = double(xr[j]), b = double(xi[j]);
      const double pj = a * a + b * b;
      acc += (__popcll((v * A + j) & mask) & 1) ? -pj : pj;
    }
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += ws[w];
    part[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(kBlock) void k_reduce_final(const double* __restrict__ part, int n,
                                                         double* __restrict__ out) {
  __shared__ double ws[kBlock / 64];
  double acc = 0.0;
  for (int i = threadIdx.x; i < n; i += kBlock) acc += part[i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += ws[w];
    *out = s;
  }
}

template <typename T, int A, int U, bool NT>
__global__ __launch_bounds__(kBlock) void k_copy(const T* __restrict__ src, T* __restrict__ dst,
                                                 uint64_t n_items, int contig, int xcd) {
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  const uint64_t tid = xcd_block(xcd) * blockDim.x + threadIdx.x;
  for (uint64_t v0 = tid; v0 < n_items; v0 += nthr * U) {
    T a[U][A], b[U][A];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t v = item_of(v0, tid, u, U, nthr, contig);
      if (v < n_items) cload<T, A, NT>(src, v * A, a[u], b[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint64_t v = item_of(v0, tid, u, U, nthr, contig);
      if (v < n_items) cstore<T, A, NT>(dst, v * A, a[u], b[u]);
    }
  }
}

// ---------------------------------------------------------------------- subspace
// Zero bits inserted at up to six ascending positions (x + (x & ~(2^p - 1)) inserts a zero
// at p; unused masks are 0), then the fixed bits set: the index of the x-th amplitude of a
// subspace with some bits held at given values.
struct SubIdx {
  uint64_t insm[kMaxGateQubits];
  uint64_t fix;
};

__device__ __forceinline__ uint64_t sub_index(uint64_t x, const SubIdx& a) {
#pragma unroll
  for (int j = 0; j < kMaxGateQubits; ++j) x += x & a.insm[j];
  return x | a.fix;
}

// A controlled 1-qubit gate on its control subspace: only the pairs (lo, lo + 2^t) whose
// control bits hold the required values are read and written; elsewhere the dense product
// (mq_block, cpu.rs:221-247) returns each amplitude unchanged, and on the subspace its two
// nonzero terms are the pair update's, in the same order.
template <typename T, int A>
__global__ __launch_bounds__(kBlock) void k_ctrl1q(T* __restrict__ amp, uint64_t n_items,
                                                   uint32_t t, SubIdx a, G2<T> g, int xcd) {
  const uint64_t half = 1ull << t;
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t v = xcd_block(xcd) * blockDim.x + threadIdx.x; v < n_items; v += nthr) {
    const uint64_t lo = sub_index(v * A, a);
    T ar[A], ai[A], br[A], bi[A];
    cload<T, A>(amp, lo, ar, ai);
    cload<T, A>(amp, lo + half, br, bi);
#pragma unroll
    for (int j = 0; j < A; ++j) pair_update(ar[j], ai[j], br[j], bi[j], g);
    cstore<T, A>(amp, lo, ar, ai);
    cstore<T, A>(amp, lo + half, br, bi);
  }
}

// SWAP on its |01>, |10> subspace: amplitudes i + 2^a and i + 2^b (i with both bits clear)
// trade places; the |00> and |11> amplitudes, which the dense product (mq_block,
// cpu.rs:221-247) returns unchanged, are not touched -- half the bytes of the full pass,
// the same values (each output is one input times exactly 1 plus exact zeros).
template <typename T, int A, bool NT>
__global__ __launch_bounds__(kBlock) void k_swap_sub(T* __restrict__ amp, uint64_t n_items,
                                                     SubIdx a, uint64_t da, uint64_t db,
                                                     int xcd) {
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t v = xcd_block(xcd) * blockDim.x + threadIdx.x; v < n_items; v += nthr) {
    const uint64_t i = sub_index(v * A, a);
    T ar[A], ai[A], br[A], bi[A];
    cload<T, A, NT>(amp, i + da, ar, ai);
    cload<T, A, NT>(amp, i + db, br, bi);
    cstore<T, A, NT>(amp, i + da, br, bi);
    cstore<T, A, NT>(amp, i + db, ar, ai);
  }
}

// The one non-unit phase of a diagonal gate on its amplitudes (diag_chunk's multiply,
// cpu.rs:99-114; the others would be multiplied by exactly 1 + 0i).
template <typename T, int A>
__global__ __launch_bounds__(kBlock) void k_diag_sub(T* __restrict__ amp, uint64_t n_items,
                                                     SubIdx a, T dr, T di, int xcd) {
  const uint64_t nthr = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t v = xcd_block(xcd) * blockDim.x + threadIdx.x; v < n_items; v += nthr) {
    const uint64_t i = sub_index(v * A, a);
    T xr[A], xi[A];
    cload<T, A>(amp, i, xr, xi);
#pragma unroll
    for (int j = 0; j < A; ++j) {
      const T x_r = xr[j], x_i = xi[j];
      xr[j] = dr * x_r - di * x_i;
      xi[j] = dr * x_i + di * x_r;
    }
    cstore<T, A>(amp, i, xr, xi);
  }
}

// ------------------------------------------------------------------------ bit swaps
// The state bits of disjoint pairs exchanged: the amplitude at index i moves to rho(i)
// (BitSwapArg).  rho is an involution, so amplitudes trade places in pairs and the pass
// runs in place.  A tile is the 2^(lo + n_hi) amplitudes over the low bits 0..lo-1 and the
// partners of those bits; rho maps the tile bits among themselves and the other bits among
// themselves, so the tile of base b goes to the tile of base rho(b).  One workgroup per
// pair of tiles loads both (runs of 2^lo consecutive amplitudes), then writes each,
// permuted through LDS, to the other's place; a tile that rho maps to itself is permuted
// in place.  Pure data movement: the values are copied, not computed.
__device__ __forceinline__ uint64_t swap_bit_pairs(uint64_t x, uint32_t n, const uint32_t* a,
                                                   const uint32_t* b) {
  for (uint32_t k = 0; k < n; ++k) {
    const uint64_t t = ((x >> a[k]) ^ (x >> b[k])) & 1u;
    x ^= (t << a[k]) | (t << b[k]);
  }
  return x;
}

template <typename T, int LO, bool NT>
__global__ __launch_bounds__(kBlock) void k_bitswap(T* __restrict__ amp, BitSwapArg a, int xcd) {
  constexpr int A = amps_per_chunk<T>();
  constexpr uint32_t kTile = 1u << (2 * LO);  // amplitudes of the largest tile
  constexpr int NI = int(kTile / A / kBlock);  // chunks per lane and tile, at most
  static_assert(NI >= 1, "a tile holds at least one chunk per lane");
  __shared__ T sbuf[2][kTile][2];
  const uint32_t chunks = (1u << (a.lo + a.n_hi)) / A;
  const uint32_t lomask = (1u << a.lo) - 1;
  // LDS slot of tile position p: the high position bits XORed into the low ones, so the
  // lanes of a wave meet distinct banks both in position order and in permuted order
  auto slot = [&](uint32_t p) { return p ^ ((p >> a.lo) & lomask); };
  // Pair w of tiles (the bases bu, bv = rho(bu)): the low n_rest bits of w are the other
  // bits without a partner; the rest, wl, numbers the values (x, y) of the pairs' bits:
  // wl < 2^p, the tiles with x = y everywhere (rho maps each to itself), x = wl; beyond,
  // d = x ^ y != 0 and x with its bit at the lowest set bit of d cleared (the other tile
  // of the pair has it set), so every pair of tiles has one number.
  auto pair_of = [&](uint64_t w, uint64_t& bu, uint64_t& bv) {
    const uint32_t p = a.n_op;
    const uint64_t wl = w >> a.n_rest;
    uint64_t d = 0, x = wl;
    if (p > 0 && wl >= (uint64_t(1) << p)) {
      const uint64_t r = wl - (uint64_t(1) << p);
      d = 1 + (r >> (p - 1));
      const uint64_t x1 = r & ((uint64_t(1) << (p - 1)) - 1);
      const uint32_t j = uint32_t(__builtin_ctzll(d));
      x = ((x1 >> j) << (j + 1)) | (x1 & ((uint64_t(1) << j) - 1));
    }
    uint64_t base = 0;
    for (uint32_t k = 0; k < a.n_rest; ++k) base |= ((w >> k) & 1u) << a.rest_bit[k];
    bu = base;
    bv = base;
    for (uint32_t k = 0; k < p; ++k) {
      const uint64_t xk = (x >> k) & 1u, yk = xk ^ ((d >> k) & 1u);
      bu |= (xk << a.op_a[k]) | (yk << a.op_b[k]);
      bv |= (yk << a.op_a[k]) | (xk << a.op_b[k]);
    }
  };
  // state offset of each of this lane's chunks within a tile (the low bits as they are,
  // the tile bits above them to their state bits)
  uint64_t off[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const uint32_t p = (threadIdx.x + uint32_t(i) * kBlock) * A;
    uint64_t o = p & lomask;
    for (uint32_t k = 0; k < a.n_hi; ++k) o |= uint64_t((p >> (a.lo + k)) & 1u) << a.hi_bit[k];
    off[i] = o;
  }
  T r0[NI][A], m0[NI][A], r1[NI][A], m1[NI][A];
  auto load = [&](uint64_t bu, uint64_t bv) {
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      if (threadIdx.x + uint32_t(i) * kBlock < chunks) {
        cload<T, A, NT>(amp, bu + off[i], r0[i], m0[i]);
        if (bv != bu) cload<T, A, NT>(amp, bv + off[i], r1[i], m1[i]);
      }
    }
  };
  uint64_t bu = 0, bv = 0;
  uint64_t w = xcd_block(xcd);
  if (w < a.n_items) {
    pair_of(w, bu, bv);
    load(bu, bv);
  }
  while (w < a.n_items) {
    const bool two = bv != bu;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const uint32_t c = threadIdx.x + uint32_t(i) * kBlock;
      if (c < chunks) {
#pragma unroll
        for (int j = 0; j < A; ++j) {
          const uint32_t s = slot(c * A + j);
          sbuf[0][s][0] = r0[i][j];
          sbuf[0][s][1] = m0[i][j];
          if (two) {
            sbuf[1][s][0] = r1[i][j];
            sbuf[1][s][1] = m1[i][j];
          }
        }
      }
    }
    __syncthreads();
    // the next pair's loads go out before this pair's stores (their tiles are other tiles)
    const uint64_t su = bu, sv = bv;
    w += gridDim.x;
    if (w < a.n_items) {
      pair_of(w, bu, bv);
      load(bu, bv);
    }
    // position p of the tile at sv receives position rho(p) of the tile at su, and the
    // tile at su position rho(p) of the tile at sv (of itself when sv == su)
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const uint32_t c = threadIdx.x + uint32_t(i) * kBlock;
      if (c < chunks) {
        uint32_t src[A];
#pragma unroll
        for (int j = 0; j < A; ++j)
          src[j] = slot(uint32_t(swap_bit_pairs(c * A + j, a.n_sp, a.sp_a, a.sp_b)));
        T r[A], m[A];
#pragma unroll
        for (int j = 0; j < A; ++j) {
          r[j] = sbuf[0][src[j]][0];
          m[j] = sbuf[0][src[j]][1];
        }
        cstore<T, A, NT>(amp, sv + off[i], r, m);
        if (two) {
#pragma unroll
          for (int j = 0; j < A; ++j) {
            r[j] = sbuf[1][src[j]][0];
            m[j] = sbuf[1][src[j]][1];
          }
          cstore<T, A, NT>(amp, su + off[i], r, m);
        }
      }
    }
    __syncthreads();  // the LDS is read before the next pair is written
  }
}

// ------------------------------------------------------------------- host helpers
// A dispatch holds fewer than 2^32 work-items (the packet's grid size is 32-bit), so
// grids are capped at 2^31 work-items -- a multiple of 8 blocks, keeping the XCD order --
// and the kernels' grid-stride loops cover the rest (complex128 at 32 qubits has 2^32
// one-chunk items).
constexpr uint64_t kMaxStreamBlocks = (uint64_t(1) << 31) / kBlock;

unsigned grid_for(uint64_t items, int unroll, const Tuning& t) {
  const uint64_t per = uint64_t(kBlock) * uint64_t(unroll);
  uint64_t g = (items + per - 1) / per;
  if (t.stream_max_blocks > 0 && g > uint64_t(t.stream_max_blocks)) g = t.stream_max_blocks;
  if (g < 1) g = 1;
  if (g > kMaxStreamBlocks) g = kMaxStreamBlocks;
  return unsigned(g);
}

// Call f(unroll, nontemporal) with the tuning's values as compile-time constants.
template <typename F>
hipError_t with_shape(const State& st, const Tuning& t, bool streams, F&& f) {
  using std::integral_constant;
  const bool nt = nontemporal_for(st, t, streams);
  switch (t.unroll) {
    case 1:
      return nt ? f(integral_constant<int, 1>{}, std::true_type{})
                : f(integral_constant<int, 1>{}, std::false_type{});
    case 4:
      return nt ? f(integral_constant<int, 4>{}, std::true_type{})
                : f(integral_constant<int, 4>{}, std::false_type{});
    default:
      return nt ? f(integral_constant<int, 2>{}, std::true_type{})
                : f(integral_constant<int, 2>{}, std::false_type{});
  }
}

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }

template <typename T, int A>
hipError_t launch_1q_t(const State& st, uint32_t q, const G2<T>& g, const Tuning& t,
                       hipStream_t s) {
  const uint64_t dim = st.dim();
  T* amp = static_cast<T*>(st.amp);
  const bool streams = (uint64_t(2 * st.elem()) << q) >= 128;  // pair stride, bytes
  if constexpr (sizeof(T) * A == 8) {
    if (t.unroll == 1 && !t.contig && int(q) >= ilog2(A) && dim % (64 * A) == 0) {
      const uint64_t chunks = dim / A;
      const unsigned grid = grid_for(chunks, 1, t);
      if (nontemporal_for(st, t, true))  // one access per chunk: pays at every target
        k_1q_lane<T, A, true><<<grid, kBlock, 0, s>>>(amp, chunks, q, g, t.xcd);
      else
        k_1q_lane<T, A, false><<<grid, kBlock, 0, s>>>(amp, chunks, q, g, t.xcd);
      return hipGetLastError();
    }
  }
  return with_shape(st, t, streams, [&](auto uc, auto ntc) {
    constexpr int U = decltype(uc)::value;
    constexpr bool NT = decltype(ntc)::value;
    if (int(q) >= ilog2(A)) {
      const uint64_t items = (dim >> 1) / A;
      k_1q_wide<T, A, U, NT><<<grid_for(items, U, t), kBlock, 0, s>>>(amp, items, q, g,
                                                                   t.contig, t.xcd);
    } else if constexpr (A >= 2) {
      const uint64_t items = dim / A;
      k_1q_intra<T, A, 0, U, NT><<<grid_for(items, U, t), kBlock, 0, s>>>(amp, items, g,
                                                                      t.contig, t.xcd);
    }
    return hipGetLastError();
  });
}

template <typename T>
hipError_t launch_1q_any(const State& st, uint32_t q, const double* gd, const Tuning& t,
                         hipStream_t s) {
  G2<T> g;
  for (int e = 0; e < 4; ++e) {
    g.r[e] = T(gd[2 * e]);
    g.i[e] = T(gd[2 * e + 1]);
  }
  constexpr int A = amps_per_chunk<T>();
  if (st.dim() >= uint64_t(A)) return launch_1q_t<T, A>(st, q, g, t, s);
  return launch_1q_t<T, 1>(st, q, g, t, s);
}

template <typename T>
hipError_t launch_diag_any(const State& st, uint32_t m, const uint32_t* q, const double* d,
                           const Tuning& t, hipStream_t s) {
  DiagArg<T> a;
  a.m = m;
  for (uint32_t j = 0; j < kMaxGateQubits; ++j) a.q[j] = j < m ? q[j] : 0;
  for (uint32_t e = 0; e < 64; ++e) {
    a.dr[e] = e < (1u << m) ? T(d[2 * e]) : T(0);
    a.di[e] = e < (1u << m) ? T(d[2 * e + 1]) : T(0);
  }
  T* amp = static_cast<T*>(st.amp);
  constexpr int A = amps_per_chunk<T>();
  if (m <= 2 && st.dim() >= uint64_t(A)) {
    Diag4<T> d4{};
    for (uint32_t e = 0; e < 4; ++e) {
      d4.r[e] = a.dr[e < (1u << m) ? e : 0];
      d4.i[e] = a.di[e < (1u << m) ? e : 0];
    }
    const uint64_t items = st.dim() / A;
    const uint32_t q1 = m == 2 ? q[1] : 0;
    const unsigned grid = grid_for(items, 1, t);
    const bool nt = nontemporal_for(st, t, true);
    if (m == 1 && nt)
      k_diag_small<T, A, 1, true><<<grid, kBlock, 0, s>>>(amp, items, q[0], q1, d4, t.xcd);
    else if (m == 1)
      k_diag_small<T, A, 1, false><<<grid, kBlock, 0, s>>>(amp, items, q[0], q1, d4, t.xcd);
    else if (nt)
      k_diag_small<T, A, 2, true><<<grid, kBlock, 0, s>>>(amp, items, q[0], q1, d4, t.xcd);
    else
      k_diag_small<T, A, 2, false><<<grid, kBlock, 0, s>>>(amp, items, q[0], q1, d4, t.xcd);
    return hipGetLastError();
  }
  return with_shape(st, t, false, [&](auto uc, auto ntc) {
    constexpr int U = decltype(uc)::value;
    constexpr bool NT = decltype(ntc)::value;
    if (st.dim() >= uint64_t(A)) {
      const uint64_t items = st.dim() / A;
      unsigned grid = grid_for(items, U, t);
      if (t.diag_blocks > 0 && grid > unsigned(t.diag_blocks)) grid = unsigned(t.diag_blocks);
      k_diag<T, A, U, NT><<<grid, kBlock, 0, s>>>(amp, items, a, t.contig, t.xcd);
    } else {
      const uint64_t items = st.dim();
      k_diag<T, 1, U, NT><<<grid_for(items, U, t), kBlock, 0, s>>>(amp, items, a, t.contig,
                                                                  t.xcd);
    }
    return hipGetLastError();
  });
}

template <int M>
MqIdx<M> make_idx(const uint32_t* q) {
  MqIdx<M> idx;
  uint32_t sorted[M];
  for (int j = 0; j < M; ++j) sorted[j] = q[j];
  for (int a = 1; a < M; ++a)
    for (int b = a; b > 0 && sorted[b - 1] > sorted[b]; --b) {
      const uint32_t x = sorted[b];
      sorted[b] = sorted[b - 1];
      sorted[b - 1] = x;
    }
  for (int j = 0; j < M; ++j) idx.sorted[j] = sorted[j];
  for (int sidx = 0; sidx < (1 << M); ++sidx) {
    uint64_t off = 0;
    for (int j = 0; j < M; ++j) off |= uint64_t((sidx >> j) & 1) << q[j];
    idx.off[sidx] = off;
  }
  return idx;
}

template <int M, int NH>
hipError_t launch_mq_low(const State& st, const uint32_t* q, const MatArg<double, M>& g,
                         const Tuning& t, hipStream_t s) {
  MqLow<M, NH> ix{};
  int nh = 0;
  for (int j = 0; j < M; ++j)
    if (q[j] >= 6) ix.hsort[nh++] = q[j];
  std::sort(ix.hsort, ix.hsort + NH);
  for (int h = 0; h < (1 << NH); ++h) {
    uint64_t o = 0;
    for (int i = 0; i < NH; ++i) o |= uint64_t((h >> i) & 1) << ix.hsort[i];
    ix.hoff[h] = o;
  }
  for (int j = 0; j < M; ++j) {
    ix.q[j] = q[j];
    ix.hrank[j] = 0;
    for (int i = 0; i < NH; ++i)
      if (ix.hsort[i] == q[j]) ix.hrank[j] = uint32_t(i);
    if (q[j] < 6) ix.glmask |= 1u << q[j];
  }
  for (int c = 0; c < (1 << M); ++c) {
    uint32_t ln = 0, h = 0;
    for (int j = 0; j < M; ++j) {
      const uint32_t bit = uint32_t(c >> j) & 1u;
      if (q[j] < 6) ln |= bit << q[j];
      else h |= bit << ix.hrank[j];
    }
    ix.c_lane[c] = ln;
    ix.c_h[c] = h;
  }
  const uint64_t runs = st.dim() >> (6 + NH);
  const unsigned grid = grid_for(runs * 64, 1, t);
  double* amp = static_cast<double*>(st.amp);
  if (nontemporal_for(st, t, true))  // whole 1 KiB runs per load
    k_mq_low<M, NH, true><<<grid, kBlock, 0, s>>>(amp, runs, ix, g, t.xcd);
  else
    k_mq_low<M, NH, false><<<grid, kBlock, 0, s>>>(amp, runs, ix, g, t.xcd);
  return hipGetLastError();
}

template <typename T, int M>
hipError_t launch_mq_reg(const State& st, const uint32_t* q, const double* gd,
                         const Tuning& t, hipStream_t s) {
  const MqIdx<M> idx = make_idx<M>(q);
  MatArg<T, M> g;
  for (int e = 0; e < (1 << (2 * M)); ++e) {
    g.r[e] = T(gd[2 * e]);
    g.i[e] = T(gd[2 * e + 1]);
  }
  T* amp = static_cast<T*>(st.amp);
  constexpr int A = amps_per_chunk<T>();
  const uint64_t blocks = st.dim() >> M;
  constexpr uint64_t GL = 64 >> M;
  // non-temporal only when every load of a wave covers whole 128-byte lines (the lowest gate
  // qubit at or above 128 B / amplitude size): on partial lines it cost 2-6x (CX(0, 1) at
  // 30 qubits: 4.7 -> 1.5 TB/s), on whole ones it gains 4-7%
  const bool nt = nontemporal_for(st, t, true) &&
                  (uint64_t(1) << idx.sorted[0]) * 2 * sizeof(T) >= 128;
  // complex128 2-qubit gates on a qubit below 6 and one at 6 or above: two whole 1 KiB runs
  // per wave, members across lanes (30 qubits: SWAP(0, 29) 5.25 -> 6.48 TB/s, SWAP(4, 25)
  // 5.80 -> 6.38, CX(4, 20) 5.67 -> 6.27, CX(3, 17) 6.35 -> 6.61).  With both qubits below
  // 6 (one load and 16 lane exchanges per amplitude) or three qubits below 6 (32) the
  // exchanges cost more than the partial lines (U(0, 1) 4.74 -> 4.20, U(0, 1, 2) 3.68 ->
  // 2.20 TB/s), so those keep one block per thread, as do gates with both qubits at 6 or
  // above (one block per thread already loads whole runs).
  if constexpr (sizeof(T) == 8 && M == 2) {
    const bool high = q[0] >= 6 || q[1] >= 6;
    if (t.mq_lane && idx.sorted[0] < 6 && high && st.n >= 7)
      return launch_mq_low<M, 1>(st, q, g, t, s);
  }
  // the lane form measured faster only for complex64 2-qubit gates (5.8-6.0 against
  // 5.4-5.7 TB/s at 30 qubits); for complex128 its exchanges cost more than they save
  if (t.mq_lane && sizeof(T) == 4 && M == 2 && int(idx.sorted[0]) >= ilog2(A) &&
      blocks % (A * GL) == 0 && blocks > 0) {
    const uint64_t waves = blocks / A / GL;
    const unsigned grid = grid_for(waves * 64, 1, t);
    if (nt)
      k_mq_lane<T, M, A, true><<<grid, kBlock, 0, s>>>(amp, waves, idx, g, t.xcd);
    else
      k_mq_lane<T, M, A, false><<<grid, kBlock, 0, s>>>(amp, waves, idx, g, t.xcd);
  } else if (int(idx.sorted[0]) >= ilog2(A) && blocks >= uint64_t(A)) {
    const uint64_t items = blocks / A;
    if (nt)
      k_mq_reg<T, M, A, true><<<grid_for(items, 1, t), kBlock, 0, s>>>(amp, items, idx, g, t.xcd);
    else
      k_mq_reg<T, M, A, false><<<grid_for(items, 1, t), kBlock, 0, s>>>(amp, items, idx, g, t.xcd);
  } else {
    if (nt)
      k_mq_reg<T, M, 1, true><<<grid_for(blocks, 1, t), kBlock, 0, s>>>(amp, blocks, idx, g, t.xcd);
    else
      k_mq_reg<T, M, 1, false><<<grid_for(blocks, 1, t), kBlock, 0, s>>>(amp, blocks, idx, g, t.xcd);
  }
  return hipGetLastError();
}

template <typename T, int M>
hipError_t launch_mq_lds(const State& st, const uint32_t* q, const void* dmat,
                         hipStream_t s) {
  const MqIdx<M> idx = make_idx<M>(q);
  constexpr int C = 1024 >> M;
  const uint64_t blocks = st.dim() >> M;
  uint64_t tiles = (blocks + C - 1) / C;
  if (tiles > 65536) tiles = 65536;
  k_mq_lds<T, M><<<unsigned(tiles), kBlock, 0, s>>>(static_cast<T*>(st.amp), blocks, idx,
                                                     static_cast<const T*>(dmat));
  return hipGetLastError();
}

template <typename T>
hipError_t launch_mq_any(const State& st, uint32_t m, const uint32_t* q, const double* g,
                         const void* dmat, const Tuning& t, hipStream_t s) {
  switch (m) {
    case 2: return launch_mq_reg<T, 2>(st, q, g, t, s);
    case 3: return launch_mq_reg<T, 3>(st, q, g, t, s);
    case 4: return launch_mq_lds<T, 4>(st, q, dmat, s);
    case 5: return launch_mq_lds<T, 5>(st, q, dmat, s);
    case 6: return launch_mq_lds<T, 6>(st, q, dmat, s);
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t launch_init_t(const State& st, uint64_t basis, hipStream_t s) {
  constexpr int A = amps_per_chunk<T>();
  T* amp = static_cast<T*>(st.amp);
  Tuning t;
  if (st.dim() >= uint64_t(A)) {
    const uint64_t items = st.dim() / A;
    k_init_basis<T, A><<<grid_for(items, 1, t), kBlock, 0, s>>>(amp, items, basis, t.xcd);
  } else {
    k_init_basis<T, 1><<<1, kBlock, 0, s>>>(amp, st.dim(), basis, 0);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_probs_t(const State& st, void* out, hipStream_t s) {
  constexpr int A = amps_per_chunk<T>();
  const T* amp = static_cast<const T*>(st.amp);
  Tuning t;
  if (st.dim() >= uint64_t(A)) {
    const uint64_t items = st.dim() / A;
    k_probs<T, A><<<grid_for(items, 1, t), kBlock, 0, s>>>(amp, static_cast<T*>(out), items,
                                                          t.xcd);
  } else {
    k_probs<T, 1><<<1, kBlock, 0, s>>>(amp, static_cast<T*>(out), st.dim(), 0);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_reduce_t(const State& st, uint64_t mask, double* part, int max_part,
                           double* result, hipStream_t s) {
  constexpr int A = amps_per_chunk<T>();
  const T* amp = static_cast<const T*>(st.amp);
  const bool vec = st.dim() >= uint64_t(A);
  const uint64_t items = vec ? st.dim() / A : st.dim();
  uint64_t g = (items + kBlock * 8 - 1) / (kBlock * 8);
  if (g < 1) g = 1;
  if (g > uint64_t(max_part)) g = max_part;
  if (vec)
    k_reduce<T, A><<<unsigned(g), kBlock, 0, s>>>(amp, items, mask, part);
  else
    k_reduce<T, 1><<<unsigned(g), kBlock, 0, s>>>(amp, items, mask, part);
  k_reduce_final<<<1, kBlock, 0, s>>>(part, int(g), result);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_copy_t(const State& st, void* dst, const Tuning& t, hipStream_t s) {
  constexpr int A = amps_per_chunk<T>();
  const T* src = static_cast<const T*>(st.amp);
  T* d = static_cast<T*>(dst);
  return with_shape(st, t, true, [&](auto uc, auto ntc) {
    constexpr int U = decltype(uc)::value;
    constexpr bool NT = decltype(ntc)::value;
    if (st.dim() >= uint64_t(A)) {
      const uint64_t items = st.dim() / A;
      k_copy<T, A, U, NT><<<grid_for(items, U, t), kBlock, 0, s>>>(src, d, items, t.contig,
                                                                  t.xcd);
    } else {
      k_copy<T, 1, U, NT><<<1, kBlock, 0, s>>>(src, d, st.dim(), 0, 0);
    }
    return hipGetLastError();
  });
}

// Subspace of the m gate qubits q (any order) with their bits held at fix.
SubIdx make_sub(uint32_t m, const uint32_t* q, uint64_t fix, uint32_t* lowest) {
  uint32_t srt[kMaxGateQubits];
  for (uint32_t j = 0; j < m; ++j) srt[j] = q[j];
  for (uint32_t a = 1; a < m; ++a)
    for (uint32_t b = a; b > 0 && srt[b - 1] > srt[b]; --b) {
      const uint32_t x = srt[b];
      srt[b] = srt[b - 1];
      srt[b - 1] = x;
    }
  SubIdx s{};
  for (uint32_t j = 0; j < kMaxGateQubits; ++j)
    s.insm[j] = j < m ? ~((uint64_t(1) << srt[j]) - 1) : 0;
  s.fix = fix;
  *lowest = srt[0];
  return s;
}

template <typename T>
hipError_t launch_ctrl1q_t(const State& st, uint32_t m, const uint32_t* q, uint64_t fix,
                           const double* gd, const Tuning& t, hipStream_t s) {
  G2<T> g;
  for (int e = 0; e < 4; ++e) {
    g.r[e] = T(gd[2 * e]);
    g.i[e] = T(gd[2 * e + 1]);
  }
  uint32_t lowest;
  const SubIdx a = make_sub(m, q, fix, &lowest);
  constexpr int A = amps_per_chunk<T>();
  const uint64_t items = st.dim() >> m;
  T* amp = static_cast<T*>(st.amp);
  if (int(lowest) >= ilog2(A) && items >= uint64_t(A))
    k_ctrl1q<T, A><<<grid_for(items / A, 1, t), kBlock, 0, s>>>(amp, items / A, q[0], a, g, t.xcd);
  else
    k_ctrl1q<T, 1><<<grid_for(items, 1, t), kBlock, 0, s>>>(amp, items, q[0], a, g, t.xcd);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_diag_sub_t(const State& st, uint32_t m, const uint32_t* q, uint64_t fix,
                             const double* ph, const Tuning& t, hipStream_t s) {
  uint32_t lowest;
  const SubIdx a = make_sub(m, q, fix, &lowest);
  constexpr int A = amps_per_chunk<T>();
  const uint64_t items = st.dim() >> m;
  T* amp = static_cast<T*>(st.amp);
  if (int(lowest) >= ilog2(A) && items >= uint64_t(A))
    k_diag_sub<T, A><<<grid_for(items / A, 1, t), kBlock, 0, s>>>(amp, items / A, a, T(ph[0]),
                                                                 T(ph[1]), t.xcd);
  else
    k_diag_sub<T, 1><<<grid_for(items, 1, t), kBlock, 0, s>>>(amp, items, a, T(ph[0]), T(ph[1]),
                                                             t.xcd);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_swap_sub_t(const State& st, const uint32_t* q, const Tuning& t,
                             hipStream_t s) {
  uint32_t lowest;
  const SubIdx a = make_sub(2, q, 0, &lowest);
  constexpr int A = amps_per_chunk<T>();
  const uint64_t items = st.dim() >> 2;
  const uint64_t da = uint64_t(1) << q[0], db = uint64_t(1) << q[1];
  T* amp = static_cast<T*>(st.amp);
  const bool nt = nontemporal_for(st, t, true) && (uint64_t(1) << lowest) * 2 * sizeof(T) >= 128;
  if (int(lowest) >= ilog2(A) && items >= uint64_t(A)) {
    if (nt)
      k_swap_sub<T, A, true><<<grid_for(items / A, 1, t), kBlock, 0, s>>>(amp, items / A, a, da,
                                                                          db, t.xcd);
    else
      k_swap_sub<T, A, false><<<grid_for(items / A, 1, t), kBlock, 0, s>>>(amp, items / A, a, da,
                                                                           db, t.xcd);
  } else {
    k_swap_sub<T, 1, false><<<grid_for(items, 1, t), kBlock, 0, s>>>(amp, items, a, da, db,
                                                                     t.xcd);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_bitswap_t(const State& st, const uint8_t* img, const Tuning& t,
                            hipStream_t s) {
  constexpr int LO = 5;  // tiles of up to 2^10 amplitudes (at least 2^5 in a run)
  const uint32_t n = st.n;
  for (uint32_t b = 0; b < n; ++b)
    if (img[b] >= n || img[img[b]] != b) return hipErrorInvalidValue;
  BitSwapArg a;
  // the most low bits whose tile (they and their partners above them) fits the LDS tile
  uint32_t lo = n < uint32_t(2 * LO) ? n : uint32_t(2 * LO);
  for (;; --lo) {
    uint32_t hi = 0;
    for (uint32_t b = 0; b < lo; ++b) hi += img[b] >= lo ? 1u : 0u;
    if (lo + hi <= uint32_t(2 * LO)) break;
  }
  a.lo = lo;
  uint32_t pos[64];  // tile position bit of each state bit, or ~0u outside the tile
  for (uint32_t b = 0; b < n; ++b) pos[b] = b < a.lo ? b : ~0u;
  for (uint32_t x = a.lo; x < n; ++x)  // ascending: the partners of low bits above them
    if (img[x] < a.lo) {
      a.hi_bit[a.n_hi] = x;
      pos[x] = a.lo + a.n_hi++;
    }
  for (uint32_t b = 0; b < n; ++b) {
    if (pos[b] != ~0u) {
      if (img[b] > b) {
        a.sp_a[a.n_sp] = pos[b];
        a.sp_b[a.n_sp++] = pos[img[b]];
      }
    } else if (img[b] == b) {
      a.rest_bit[a.n_rest++] = b;
    } else if (img[b] > b) {
      a.op_a[a.n_op] = b;
      a.op_b[a.n_op++] = img[b];
    }
  }
  if (a.n_sp == 0 && a.n_op == 0) return hipSuccess;
  const uint32_t p = a.n_op;
  a.n_items = (uint64_t(1) << a.n_rest) *
              ((uint64_t(1) << p) + (p ? ((uint64_t(1) << p) - 1) << (p - 1) : 0));
  // a workgroup per pair of tiles, or the tuning's grid-stride cap
  uint64_t g = a.n_items < kMaxStreamBlocks ? a.n_items : kMaxStreamBlocks;
  if (t.stream_max_blocks > 0 && g > uint64_t(t.stream_max_blocks)) g = t.stream_max_blocks;
  const unsigned grid = unsigned(g);
  T* amp = static_cast<T*>(st.amp);
  if (nontemporal_for(st, t, true))
    k_bitswap<T, LO, true><<<grid, kBlock, 0, s>>>(amp, a, t.xcd);
  else
    k_bitswap<T, LO, false><<<grid, kBlock, 0, s>>>(amp, a, t.xcd);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_bitswap(const State& st, const uint8_t* img, const Tuning& t,
                          hipStream_t s) {
  if (st.n > 64) return hipErrorInvalidValue;
  return st.dtype == 0 ? launch_bitswap_t<double>(st, img, t, s)
                       : launch_bitswap_t<float>(st, img, t, s);
}

hipError_t launch_swap_sub(const State& st, const uint32_t* q, const Tuning& t, hipStream_t s) {
  if (st.n < 2) return hipErrorInvalidValue;
  return st.dtype == 0 ? launch_swap_sub_t<double>(st, q, t, s)
                       : launch_swap_sub_t<float>(st, q, t, s);
}

hipError_t launch_ctrl1q(const State& st, uint32_t m, const uint32_t* q, uint64_t fix,
                         const double* g, const Tuning& t, hipStream_t s) {
  return st.dtype == 0 ? launch_ctrl1q_t<double>(st, m, q, fix, g, t, s)
                       : launch_ctrl1q_t<float>(st, m, q, fix, g, t, s);
}

hipError_t launch_diag_sub(const State& st, uint32_t m, const uint32_t* q, uint64_t fix,
                           const double* phase, const Tuning& t, hipStream_t s) {
  return st.dtype == 0 ? launch_diag_sub_t<double>(st, m, q, fix, phase, t, s)
                       : launch_diag_sub_t<float>(st, m, q, fix, phase, t, s);
}

hipError_t launch_init_basis(const State& st, uint64_t basis, hipStream_t s) {
  return st.dtype == 0 ? launch_init_t<double>(st, basis, s) : launch_init_t<float>(st, basis, s);
}

hipError_t launch_1q(const State& st, uint32_t q, const double* g, const Tuning& t,
                     hipStream_t s) {
  return st.dtype == 0 ? launch_1q_any<double>(st, q, g, t, s)
                       : launch_1q_any<float>(st, q, g, t, s);
}

hipError_t launch_diag(const State& st, uint32_t m, const uint32_t* q, const double* d,
                       const Tuning& t, hipStream_t s) {
  return st.dtype == 0 ? launch_diag_any<double>(st, m, q, d, t, s)
                       : launch_diag_any<float>(st, m, q, d, t, s);
}

hipError_t launch_mq(const State& st, uint32_t m, const uint32_t* q, const double* g,
                     const void* dmat, const Tuning& t, hipStream_t s) {
  return st.dtype == 0 ? launch_mq_any<double>(st, m, q, g, dmat, t, s)
                       : launch_mq_any<float>(st, m, q, g, dmat, t, s);
}

hipError_t launch_probs(const State& st, void* out, hipStream_t s) {
  return st.dtype == 0 ? launch_probs_t<double>(st, out, s) : launch_probs_t<float>(st, out, s);
}

hipError_t launch_reduce(const State& st, uint64_t mask, double* part, int max_part,
                         double* result, hipStream_t s) {
  return st.dtype == 0 ? launch_reduce_t<double>(st, mask, part, max_part, result, s)
                       : launch_reduce_t<float>(st, mask, part, max_part, result, s);
}

hipError_t launch_copy(const State& st, void* dst, const Tuning& t, hipStream_t s) {
  return st.dtype == 0 ? launch_copy_t<double>(st, dst, t, s)
                       : launch_copy_t<float>(st, dst, t, s);
}

}  // namespace qsvh
