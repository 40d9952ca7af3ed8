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
    // the bit positions are read at each use: masks derived from them and hoisted out of
    // a kernel's tile loop would hold two scalar registers per tile bit
    uint32_t b = ts[j];
    asm volatile("" : "+s"(b));
    t = ((t >> b) << (b + 1)) | (t & ((1ull << b) - 1));
  }
  return t;
}

// Uniform part of a chunk's state offset: bit j of c is tile-local bit lb + j (state bit
// ts[lb + j]).  The bit positions are read at each use (scalar work per tile): offsets
// hoisted out of a kernel's tile loop would hold registers across all the ops.
__device__ __forceinline__ uint64_t chunk_offset(uint32_t c, const uint32_t* ts, uint32_t lb) {
  uint64_t off = 0;
  for (uint32_t j = 0; (c >> j) != 0; ++j)
    if ((c >> j) & 1u) {
      uint32_t b = ts[lb + j];
      asm volatile("" : "+s"(b));
      off |= 1ull << b;
    }
  return off;
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

// cmul in place, for multiplies under a branch (a phase skipped when it is exactly 1): the
// result lands in the registers of x, so the join after the branch needs no copies (written
// in C, the product goes to new registers and both arms copy into them, two moves per
// amplitude whether or not it is multiplied).  The operations and their order are cmul's:
// dr xr - di xi, dr xi + di xr.  S: the phase is uniform (scalar registers).
template <bool S>
__device__ __forceinline__ void cmul_ip(double& xr, double& xi, double dr, double di) {
  double t0, t1;
  if constexpr (S)
    asm volatile(
        "v_mul_f64 %2, %4, %0\n\t"
        "v_mul_f64 %3, %5, %0\n\t"
        "v_mul_f64 %0, %5, %1\n\t"
        "v_mul_f64 %1, %4, %1\n\t"
        "v_add_f64 %0, %2, -%0\n\t"
        "v_add_f64 %1, %1, %3"
        : "+v"(xr), "+v"(xi), "=&v"(t0), "=&v"(t1)
        : "s"(dr), "s"(di));
  else
    asm volatile(
        "v_mul_f64 %2, %4, %0\n\t"
        "v_mul_f64 %3, %5, %0\n\t"
        "v_mul_f64 %0, %5, %1\n\t"
        "v_mul_f64 %1, %4, %1\n\t"
        "v_add_f64 %0, %2, -%0\n\t"
        "v_add_f64 %1, %1, %3"
        : "+v"(xr), "+v"(xi), "=&v"(t0), "=&v"(t1)
        : "v"(dr), "v"(di));
}
template <bool S>
__device__ __forceinline__ void cmul_ip(float& xr, float& xi, float dr, float di) {
  float t0, t1;
  if constexpr (S)
    asm volatile(
        "v_mul_f32 %2, %4, %0\n\t"
        "v_mul_f32 %3, %5, %0\n\t"
        "v_mul_f32 %0, %5, %1\n\t"
        "v_mul_f32 %1, %4, %1\n\t"
        "v_sub_f32 %0, %2, %0\n\t"
        "v_add_f32 %1, %1, %3"
        : "+v"(xr), "+v"(xi), "=&v"(t0), "=&v"(t1)
        : "s"(dr), "s"(di));
  else
    asm volatile(
        "v_mul_f32 %2, %4, %0\n\t"
        "v_mul_f32 %3, %5, %0\n\t"
        "v_mul_f32 %0, %5, %1\n\t"
        "v_mul_f32 %1, %4, %1\n\t"
        "v_sub_f32 %0, %2, %0\n\t"
        "v_add_f32 %1, %1, %3"
        : "+v"(xr), "+v"(xi), "=&v"(t0), "=&v"(t1)
        : "v"(dr), "v"(di));
}

// First tile-local index of block `blk` of a dense or permutation op: zeros inserted at
// the op's action and condition bits (ascending), condition bits set to their values.
__device__ __forceinline__ uint32_t op_base(uint32_t blk, const TileOp& op) {
  // x + (x & ~(2^p - 1)) inserts a zero bit at p (ascending p); a zero mask is a no-op.
  // Constant field indices only: a dynamically indexed op field would put the op in scratch.
  uint32_t b = blk;
  b += b & op.insm[0];
  b += b & op.insm[1];
  b += b & op.insm[2];
  return b | op.lcval;
}

// Tile-local index of sub-index s of the block at b0 (bit j of s is action bit lq[j]).
template <int M>
__device__ __forceinline__ uint32_t op_index(uint32_t b0, uint32_t s, const TileOp& op) {
  uint32_t i = b0;
#pragma unroll
  for (int j = 0; j < M; ++j) i |= ((s >> j) & 1u) << op.lq[j];
  return i;
}

// Dense M-qubit block in LDS: mq_block (cpu.rs:221-247), y_r = sum_c U[r][c] x_c from an
// exact zero in ascending c.  For M = 1 this equals one_qubit_block's y = g0 x0 + g1 x1
// (cpu.rs:84-93): 0 + a is a.  REAL: every entry has a zero imaginary part, whose products
// (exact zeros) are not formed.
template <typename T, int M, int NT, int KMAX, bool REAL>
__device__ __forceinline__ void apply_dense(T* sr, T* si, uint32_t N, const TileOp& op,
                                            const T* __restrict__ pool) {
  constexpr uint32_t D = 1u << M;
  constexpr int kRowUnroll = M == 1 ? 2 : 1;
  // blocks per thread without conditions (a compile-time trip count lets the loads of
  // all blocks issue together)
  constexpr int IT = (KMAX >> M) > 0 ? (KMAX >> M) : 1;
  const T* __restrict__ g = pool + 2 * op.off;
  const uint32_t nblk = N >> op.nins;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const uint32_t blk = threadIdx.x + uint32_t(it) * NT;
    if (blk >= nblk) continue;
    const uint32_t b0 = op_base(blk, op);
    T ar[D], ai[D];
#pragma unroll
    for (uint32_t s = 0; s < D; ++s) {
      const uint32_t i = op_index<M>(b0, s, op);
      ar[s] = sr[i];
      ai[s] = si[i];
    }
    // rows one at a time (M >= 2): the matrix is read row by row instead of being held
    // in registers for the whole block loop
#pragma unroll kRowUnroll
    for (uint32_t r = 0; r < D; ++r) {
      // the sum starts at the first term: the reference's 0 + t is t (up to a zero's sign)
      T accr, acci;
#pragma unroll
      for (uint32_t c = 0; c < D; ++c) {
        const T er = g[2 * (r * D + c)];
        T pr, pi;
        if constexpr (REAL) {
          pr = er * ar[c];
          pi = er * ai[c];
        } else {
          const T ei = g[2 * (r * D + c) + 1];
          pr = er * ar[c] - ei * ai[c];
          pi = er * ai[c] + ei * ar[c];
        }
        accr = c == 0 ? pr : accr + pr;
        acci = c == 0 ? pi : acci + pi;
      }
      const uint32_t i = op_index<M>(b0, r, op);
      sr[i] = accr;
      si[i] = acci;
    }
  }
}

// Permutation block: y_r = e_r x_{perm[r]}; a coefficient exactly 1 + 0i is a move.  The
// dense product would add exact zeros to the one nonzero term, so values are the same.
template <typename T, int M, int NT, int KMAX>
__device__ __forceinline__ void apply_perm(T* sr, T* si, uint32_t N, const TileOp& op,
                                           const T* __restrict__ pool) {
  constexpr uint32_t D = 1u << M;
  constexpr int IT = (KMAX >> M) > 0 ? (KMAX >> M) : 1;
  const T* __restrict__ e = pool + 2 * op.off;
  const uint32_t nblk = N >> op.nins;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const uint32_t blk = threadIdx.x + uint32_t(it) * NT;
    if (blk >= nblk) continue;
    const uint32_t b0 = op_base(blk, op);
    T yr[D], yi[D];
#pragma unroll
    for (uint32_t r = 0; r < D; ++r) {
      const uint32_t i = op_index<M>(b0, op.perm[r], op);
      yr[r] = sr[i];
      yi[r] = si[i];
    }
    if (op.pid != (1u << D) - 1u) {  // some coefficient is not 1: multiply those
#pragma unroll
      for (uint32_t r = 0; r < D; ++r)
        if (!((op.pid >> r) & 1u)) cmul(yr[r], yi[r], e[2 * r], e[2 * r + 1]);
    }
#pragma unroll
    for (uint32_t r = 0; r < D; ++r) {
      const uint32_t i = op_index<M>(b0, r, op);
      sr[i] = yr[r];
      si[i] = yi[r];
    }
  }
}

constexpr uint32_t ilog2c(uint32_t x) { return x <= 1 ? 0 : 1 + ilog2c(x / 2); }

// f(integral_constant<int, k>) for k = K0 .. K1-1.  Register arrays indexed only through
// it are indexed by constants from the start, so they become plain registers (a loop
// unrolled later leaves them as one vector value that control flow then copies whole).
// Every lambda passed here and to the other dispatchers is always_inline: an outlined
// lambda would take its captured register arrays by address and put them in scratch.
template <int K0, int K1, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (K0 < K1) {
    f(std::integral_constant<int, K0>{});
    static_for<K0 + 1, K1>(f);
  }
}

// Entry i (0..3) of a 4-entry table held in registers.  Selects on values, not on
// element references: a select of references becomes a select of addresses and puts the
// table in private memory.
template <typename T>
__device__ __forceinline__ T sel4(const T (&t)[4], uint32_t i) {
  const T t0 = t[0], t1 = t[1], t2 = t[2], t3 = t[3];
  const T lo = (i & 1u) ? t1 : t0;
  const T hi = (i & 1u) ? t3 : t2;
  return (i & 2u) ? hi : lo;
}

// One diagonal gate of m <= 2 qubits on the lane's KMAX amplitudes (tile-local index
// e = tid + k NT), lowered by the host (TileDiag): its outside bits pick one 4-entry phase
// table per tile, and the phase index c is made of at most two bits, each a lane bit (fixed
// for the lane) or a register bit (a bit of k).  Register-bit positions become compile-time
// constants (a switch over the few possible positions), so per amplitude the work is one
// complex multiply whose phase operand is known to the instruction.  An amplitude whose
// phase is exactly 1 + 0i is left as it is: the multiply would return it unchanged (up to
// the sign of a zero).
// Call f with integral_constant<int, p> for a runtime p < LK.
template <int LK, typename F>
__device__ __forceinline__ void with_kpos(uint32_t p, F&& f) {
  using std::integral_constant;
  if constexpr (LK > 0) if (p == 0) return f(integral_constant<int, 0>{});
  if constexpr (LK > 1) if (p == 1) return f(integral_constant<int, 1>{});
  if constexpr (LK > 2) if (p == 2) return f(integral_constant<int, 2>{});
  if constexpr (LK > 3) if (p == 3) return f(integral_constant<int, 3>{});
  if constexpr (LK > 4) if (p == 4) return f(integral_constant<int, 4>{});
  if constexpr (LK > 5) if (p == 5) return f(integral_constant<int, 5>{});
}

// Phases uniform over the workgroup: pr/pi[c] is the phase of an amplitude whose register
// bits at positions KA, KB (-1: none) read c = bit KA | bit KB << 1; bit c of `id` set:
// that phase is the identity.
template <typename T, int NT, int KMAX, int KA, int KB>
__device__ __forceinline__ void diag_uniform(T (&xr)[KMAX], T (&xi)[KMAX], uint32_t N,
                                             const T (&pr)[4], const T (&pi)[4], uint32_t id) {
  auto cidx = [](auto kc) __attribute__((always_inline)) {
    constexpr int k = decltype(kc)::value;
    return (KA >= 0 ? (k >> (KA < 0 ? 0 : KA)) & 1 : 0) |
           (KB >= 0 ? ((k >> (KB < 0 ? 0 : KB)) & 1) << 1 : 0);
  };
  if (id == 0) {  // no identity phase: no branches
    static_for<0, KMAX>([&](auto kc) __attribute__((always_inline)) {
      constexpr int k = kc.value;
      const int c = cidx(kc);
      if (threadIdx.x + uint32_t(k) * NT < N) cmul(xr[k], xi[k], pr[c], pi[c]);
    });
    return;
  }
  static_for<0, KMAX>([&](auto kc) __attribute__((always_inline)) {
    constexpr int k = kc.value;
    const int c = cidx(kc);
    // (the in-place asm multiply also keeps the uniform skip a branch: no speculated
    // multiply + select)
    if (!((id >> c) & 1u) && threadIdx.x + uint32_t(k) * NT < N)
      cmul_ip<true>(xr[k], xi[k], pr[c], pi[c]);
  });
}

// Phases per lane: p0 for register bit KA = 0 (or KA = -1: every k), p1 for bit KA = 1.
template <typename T, int NT, int KMAX, int KA>
__device__ __forceinline__ void diag_lane(T (&xr)[KMAX], T (&xi)[KMAX], uint32_t N, T p0r,
                                          T p0i, T p1r, T p1i, bool skip0, bool skip1) {
  if (!skip0 && !skip1) {  // no branches
    static_for<0, KMAX>([&](auto kc) __attribute__((always_inline)) {
      constexpr int k = kc.value;
      constexpr bool hi = KA >= 0 && ((k >> (KA < 0 ? 0 : KA)) & 1);
      if (threadIdx.x + uint32_t(k) * NT < N) cmul(xr[k], xi[k], hi ? p1r : p0r, hi ? p1i : p0i);
    });
    return;
  }
  static_for<0, KMAX>([&](auto kc) __attribute__((always_inline)) {
    constexpr int k = kc.value;
    constexpr bool hi = KA >= 0 && ((k >> (KA < 0 ? 0 : KA)) & 1);
    if (!(hi ? skip1 : skip0) && threadIdx.x + uint32_t(k) * NT < N)
      cmul_ip<false>(xr[k], xi[k], hi ? p1r : p0r, hi ? p1i : p0i);
  });
}

// The words of a TileDiag a gate's multiply needs, and its 4 phases for the tile (the
// table of its outside bits' values): one scalar load each.
struct DiagHdr {
  uint32_t m, code, oq, tab, idw, pad[3];
};
template <typename T>
struct DiagPh {
  T r[4], i[4];
};

__device__ __forceinline__ DiagHdr load_hdr(const TileDiag* __restrict__ dg) {
  return *reinterpret_cast<const DiagHdr*>(dg);
}

template <typename T>
__device__ __forceinline__ DiagPh<T> load_ph(const DiagHdr& h, uint64_t base,
                                             const T* __restrict__ pool) {
  const uint32_t nout = (h.code >> 16) & 3u;
  uint32_t so = 0;
  if (nout > 0) so = uint32_t((base >> (h.oq & 63u)) & 1u);
  if (nout > 1) so |= uint32_t((base >> ((h.oq >> 8) & 63u)) & 1u) << 1;
  const T* __restrict__ tab = pool + 2 * (h.tab + 4 * so);
  DiagPh<T> p;
  static_for<0, 4>([&](auto c) __attribute__((always_inline)) {
    p.r[c.value] = tab[2 * c.value];
    p.i[c.value] = tab[2 * c.value + 1];
  });
  return p;
}

template <typename T, int NT, int KMAX>
__device__ __forceinline__ void diag_gate_small(T (&xr)[KMAX], T (&xi)[KMAX], uint32_t N,
                                                uint64_t base, const DiagHdr& h,
                                                const DiagPh<T>& ph) {
  constexpr int LK = int(ilog2c(KMAX));
  const uint32_t nout = (h.code >> 16) & 3u;
  uint32_t so = 0;
  if (nout > 0) so = uint32_t((base >> (h.oq & 63u)) & 1u);
  if (nout > 1) so |= uint32_t((base >> ((h.oq >> 8) & 63u)) & 1u) << 1;
  const T(&er)[4] = ph.r;
  const T(&ei)[4] = ph.i;
  const uint32_t id = (h.idw >> (4 * so)) & 15u;
  const uint32_t v = h.code & 255u;
  const uint32_t kind0 = v & 15u, kind1 = v >> 4;
  const uint32_t lp0 = (h.code >> 8) & 15u, lp1 = (h.code >> 12) & 15u;
  if (kind0 == 0) {  // one phase for the whole tile
    diag_uniform<T, NT, KMAX, -1, -1>(xr, xi, N, er, ei, id);
  } else if (kind0 >= 2) {  // register bits only: phases uniform for each k
    if (kind1 == 0) {
      with_kpos<LK>(kind0 - 2, [&](auto A) __attribute__((always_inline)) {
        diag_uniform<T, NT, KMAX, A.value, -1>(xr, xi, N, er, ei, id);
      });
    } else {
      with_kpos<LK>(kind0 - 2, [&](auto A) __attribute__((always_inline)) {
        with_kpos<LK>(kind1 - 2, [&](auto B) __attribute__((always_inline)) {
          if constexpr (A.value < B.value)
            diag_uniform<T, NT, KMAX, A.value, B.value>(xr, xi, N, er, ei, id);
        });
      });
    }
  } else {  // index bit 0 is a lane bit
    const bool l0 = (threadIdx.x >> lp0) & 1u;
    if (kind1 == 0) {
      if ((id & 3u) != 3u) {
        const T pr = l0 ? er[1] : er[0], pi = l0 ? ei[1] : ei[0];
        diag_lane<T, NT, KMAX, -1>(xr, xi, N, pr, pi, pr, pi, false, false);
      }
    } else if (kind1 == 1) {
      if (id != 15u) {
        const uint32_t c = uint32_t(l0) | (((threadIdx.x >> lp1) & 1u) << 1);
        const T pr = sel4(er, c), pi = sel4(ei, c);
        diag_lane<T, NT, KMAX, -1>(xr, xi, N, pr, pi, pr, pi, false, false);
      }
    } else {
      const T p0r = l0 ? er[1] : er[0], p0i = l0 ? ei[1] : ei[0];
      const T p1r = l0 ? er[3] : er[2], p1i = l0 ? ei[3] : ei[2];
      const bool skip0 = (id & 3u) == 3u, skip1 = (id >> 2) == 3u;
      with_kpos<LK>(kind1 - 2, [&](auto A) __attribute__((always_inline)) {
        diag_lane<T, NT, KMAX, A.value>(xr, xi, N, p0r, p0i, p1r, p1i, skip0, skip1);
      });
    }
  }
}

// A diagonal gate on m > 2 qubits: sub-index from the tile-local index bit by bit.
template <typename T, int NT, int KMAX>
__device__ __forceinline__ void diag_gate_wide(T (&xr)[KMAX], T (&xi)[KMAX], uint32_t N,
                                               uint64_t base, const TileDiag* __restrict__ dg,
                                               const T* __restrict__ pool) {
  const uint32_t m = dg->m;
  const T* __restrict__ tab = pool + 2 * dg->tab;
  static_for<0, KMAX>([&](auto kc) __attribute__((always_inline)) {
    constexpr int k = kc.value;
    const uint32_t e = threadIdx.x + uint32_t(k) * NT;
    if (e < N) {
      uint32_t s = 0;
      for (uint32_t j = 0; j < m; ++j) {
        const uint32_t loc = dg->loc[j];
        const uint32_t bit = loc == kOutside ? uint32_t((base >> dg->q[j]) & 1u) : (e >> loc) & 1u;
        s |= bit << j;
      }
      cmul(xr[k], xi[k], tab[2 * s], tab[2 * s + 1]);
    }
  });
}

// Diagonal gates dgs[0 .. n) in order on the lane's registers: per gate one scalar load of
// its words, then one of its phases.  (Issuing a gate's loads one gate ahead was measured
// slower: QFT(30) complex128 tile passes 232 -> 283 ms.)
template <typename T, int NT, int KMAX>
__device__ __forceinline__ void diag_gates(T (&xr)[KMAX], T (&xi)[KMAX], uint32_t N,
                                           uint64_t base, const TileDiag* __restrict__ dgs,
                                           uint32_t n, const T* __restrict__ pool) {
  for (uint32_t d = 0; d < n; ++d) {
    const DiagHdr h = load_hdr(dgs + d);
    if (h.m <= 2)
      diag_gate_small<T, NT, KMAX>(xr, xi, N, base, h, load_ph<T>(h, base, pool));
    else
      diag_gate_wide<T, NT, KMAX>(xr, xi, N, base, dgs + d, pool);
  }
}

// Kind 5 lane-run member: n phases, each controlled by one outside bit (diags[off + d].oq)
// on the run index bit b: where the tile's base has bit oq set, the amplitudes whose bit b is
// 1 are multiplied by pool[ins[0] + d] -- the one phase of the gate's table that is not
// exactly 1 (the identities' multiplies are left out, as in diag_gate_small).  Each gate
// costs a bit test and the multiplies; its two loads do not depend on each other, and four
// gates' loads are issued together.
template <typename T, int NT, int KMAX>
__device__ __forceinline__ void ctrl_phases(T (&xr)[KMAX], T (&xi)[KMAX], uint32_t N,
                                            uint64_t base, const TileOp& g,
                                            const TileDiag* __restrict__ diags,
                                            const T* __restrict__ pool) {
  constexpr int LK = int(ilog2c(KMAX));
  const TileDiag* __restrict__ dg = diags + g.off;
  const T* __restrict__ ph = pool + 2 * size_t(g.ins[0]);
  const uint32_t n = g.n_diag, b = g.lq[0];
  auto each = [&](auto&& mul) __attribute__((always_inline)) {
    uint32_t d = 0;
    for (; d + 4 <= n; d += 4) {
      uint32_t oq[4];
      T pr[4], pi[4];
      static_for<0, 4>([&](auto u) __attribute__((always_inline)) {
        oq[u.value] = dg[d + u.value].oq;
        pr[u.value] = ph[2 * (d + u.value)];
        pi[u.value] = ph[2 * (d + u.value) + 1];
      });
      static_for<0, 4>([&](auto u) __attribute__((always_inline)) {
        if ((base >> (oq[u.value] & 63u)) & 1u) mul(pr[u.value], pi[u.value]);
      });
    }
    for (; d < n; ++d) {
      const uint32_t oq = dg[d].oq;
      const T pr = ph[2 * d], pi = ph[2 * d + 1];
      if ((base >> (oq & 63u)) & 1u) mul(pr, pi);
    }
  };
  if (b >= kTileLaneBits) {
    if constexpr (LK > 0)
      with_kpos<LK>(b - kTileLaneBits, [&](auto P) __attribute__((always_inline)) {
        each([&](T pr, T pi) __attribute__((always_inline)) {
          static_for<0, KMAX>([&](auto kc) __attribute__((always_inline)) {
            constexpr int k = kc.value;
            if constexpr ((k >> P.value) & 1)
              if (threadIdx.x + uint32_t(k) * NT < N) cmul_ip<true>(xr[k], xi[k], pr, pi);
          });
        });
      });
  } else {
    const bool on = (threadIdx.x >> b) & 1u;
    each([&](T pr, T pi) __attribute__((always_inline)) {
      if (on)
        static_for<0, KMAX>([&](auto kc) __attribute__((always_inline)) {
          constexpr int k = kc.value;
          if (threadIdx.x + uint32_t(k) * NT < N) cmul_ip<true>(xr[k], xi[k], pr, pi);
        });
    });
  }
}

// Kind 6 lane-run member: n phases on two run index bits, the stretch's register bit b and
// each gate's own bit a (diags[off + d].oq): the amplitudes whose bits a and b are both 1 are
// multiplied by pool[ins[0] + d] (the identities' multiplies left out, as above).
template <typename T, int NT, int KMAX>
__device__ __forceinline__ void pair_phases(T (&xr)[KMAX], T (&xi)[KMAX], uint32_t N,
                                            const TileOp& g, const TileDiag* __restrict__ diags,
                                            const T* __restrict__ pool) {
  constexpr int LK = int(ilog2c(KMAX));
  const TileDiag* __restrict__ dg = diags + g.off;
  const T* __restrict__ ph = pool + 2 * size_t(g.ins[0]);
  const uint32_t n = g.n_diag, b = g.lq[0];
  uint32_t tid = threadIdx.x;
  // mul(a, pr, pi): the gate with other bit a and phase pr + i pi
  auto each = [&](auto&& mul) __attribute__((always_inline)) {
#pragma unroll 1
    for (uint32_t d = 0; d < n; ++d) {
      const uint32_t a = dg[d].oq;
      mul(a, ph[2 * d], ph[2 * d + 1]);
    }
  };
  // registers k that have bit p set (compile-time) and whose lane or register bit a is 1
  auto regs = [&](auto P, uint32_t a, T pr, T pi) __attribute__((always_inline)) {
    if (a >= kTileLaneBits) {
      const uint32_t q = a - kTileLaneBits;
      static_for<0, KMAX>([&](auto kc) __attribute__((always_inline)) {
        constexpr int k = kc.value;
        if constexpr ((k >> P.value) & 1)
          if (((uint32_t(k) >> q) & 1u) && threadIdx.x + uint32_t(k) * NT < N)
            cmul_ip<true>(xr[k], xi[k], pr, pi);
      });
    } else if ((tid >> a) & 1u) {
      static_for<0, KMAX>([&](auto kc) __attribute__((always_inline)) {
        constexpr int k = kc.value;
        if constexpr ((k >> P.value) & 1)
          if (threadIdx.x + uint32_t(k) * NT < N) cmul_ip<true>(xr[k], xi[k], pr, pi);
      });
    }
  };
  // (the host makes b a register bit)
  if constexpr (LK > 0)
    with_kpos<LK>(b - kTileLaneBits, [&](auto P) __attribute__((always_inline)) {
      each([&](uint32_t a, T pr, T pi) __attribute__((always_inline)) { regs(P, a, pr, pi); });
    });
}

// A run of diagonal gates.  Each lane keeps its KMAX amplitudes (e = tid + k NT) in
// registers for the whole run and multiplies them by every gate's phase in order -- the
// multiply of diag_chunk, in the same order.
template <typename T, int NT, int KMAX>
__device__ __forceinline__ void apply_diag_run(T* sr, T* si, uint32_t N, uint64_t base,
                                               const TileOp& op,
                                               const TileDiag* __restrict__ diags,
                                               const T* __restrict__ pool) {
  T xr[KMAX], xi[KMAX];
  static_for<0, KMAX>([&](auto kc) __attribute__((always_inline)) {
    const uint32_t e = threadIdx.x + uint32_t(kc.value) * NT;
    if (e < N) {
      xr[kc.value] = sr[e];
      xi[kc.value] = si[e];
    }
  });
  diag_gates<T, NT, KMAX>(xr, xi, N, base, diags + op.off, op.n_diag, pool);
  static_for<0, KMAX>([&](auto kc) __attribute__((always_inline)) {
    const uint32_t e = threadIdx.x + uint32_t(kc.value) * NT;
    if (e < N) {
      sr[e] = xr[kc.value];
      si[e] = xi[kc.value];
    }
  });
}

// Math of a 1-qubit op on the pair (x0, x1): one_qubit_block's y0 = g00 x0 + g01 x1,
// y1 = g10 x0 + g11 x1 (cpu.rs:84-93) in the same operation order.  V = 0: general;
// 1: real matrix (the products of its zero imaginary parts are not formed); 2:
// anti-diagonal (y0 = g01 x1, y1 = g10 x0: the zero entries' products are not formed);
// 3: a swap (anti-diagonal with unit entries).  Same values as the dense product: only
// exact zeros are left out (up to the sign of a zero).
constexpr uint32_t kPairGeneral = 0, kPairReal = 1, kPairAnti = 2, kPairSwap = 3;

template <typename T, uint32_t V>
__device__ __forceinline__ void pair_math(T& x0r, T& x0i, T& x1r, T& x1i,
                                          const T* __restrict__ m) {
  if constexpr (V == kPairSwap) {
    const T tr = x0r, ti = x0i;
    x0r = x1r;
    x0i = x1i;
    x1r = tr;
    x1i = ti;
  } else if constexpr (V == kPairAnti) {
    const T y0r = x1r, y0i = x1i, y1r = x0r, y1i = x0i;
    x0r = y0r;
    x0i = y0i;
    x1r = y1r;
    x1i = y1i;
    cmul(x0r, x0i, m[2], m[3]);  // g01
    cmul(x1r, x1i, m[4], m[5]);  // g10
  } else if constexpr (V == kPairReal) {
    const T a = m[0], b = m[2], c = m[4], d = m[6];
    const T y0r = a * x0r + b * x1r, y0i = a * x0i + b * x1i;
    const T y1r = c * x0r + d * x1r, y1i = c * x0i + d * x1i;
    x0r = y0r;
    x0i = y0i;
    x1r = y1r;
    x1i = y1i;
  } else {
    const T g0r = m[0], g0i = m[1], g1r = m[2], g1i = m[3];
    const T g2r = m[4], g2i = m[5], g3r = m[6], g3i = m[7];
    const T ar = g0r * x0r - g0i * x0i, ai = g0r * x0i + g0i * x0r;
    const T br = g1r * x1r - g1i * x1i, bi = g1r * x1i + g1i * x1r;
    const T cr = g2r * x0r - g2i * x0i, ci = g2r * x0i + g2i * x0r;
    const T dr = g3r * x1r - g3i * x1i, di = g3r * x1i + g3i * x1r;
    x0r = ar + br;
    x0i = ai + bi;
    x1r = cr + dr;
    x1i = ci + di;
  }
}

// Call f(integral_constant<uint32_t, V>) for a 1-qubit op's pair variant (uniform).
template <typename F>
__device__ __forceinline__ void with_pair_variant(uint32_t v, F&& f) {
  using std::integral_constant;
  if (v == kPairSwap) return f(integral_constant<uint32_t, kPairSwap>{});
  if (v == kPairAnti) return f(integral_constant<uint32_t, kPairAnti>{});
  if (v == kPairReal) return f(integral_constant<uint32_t, kPairReal>{});
  return f(integral_constant<uint32_t, kPairGeneral>{});
}

// A standalone 1-qubit op (kind 0, m = 1) on action bit lq[0] with conditions.
template <typename T, uint32_t V, int NT, int KMAX>
__device__ __forceinline__ void apply_1q(T* sr, T* si, uint32_t N, const TileOp& op,
                                         const T* __restrict__ pool) {
  constexpr int IT = (KMAX >> 1) > 0 ? (KMAX >> 1) : 1;
  const T* __restrict__ m = pool + 2 * op.off;
  const uint32_t nblk = N >> op.nins;
  const uint32_t h = 1u << op.lq[0];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const uint32_t blk = threadIdx.x + uint32_t(it) * NT;
    if (blk >= nblk) continue;
    const uint32_t i0 = op_base(blk, op), i1 = i0 | h;
    T x0r = sr[i0], x0i = si[i0], x1r = sr[i1], x1i = si[i1];
    pair_math<T, V>(x0r, x0i, x1r, x1i, m);
    sr[i0] = x0r;
    si[i0] = x0i;
    sr[i1] = x1r;
    si[i1] = x1i;
  }
}

// Register block (kind 3): each lane loads blocks of 2^NB amplitudes (the block's bits
// lq[0..NB)), applies the op's members in order -- 1-qubit ops on block bit j, skipped for
// sub-indices whose condition bits differ -- and stores them: one LDS round trip and one
// barrier for the whole run instead of one per gate.
template <typename T, int NB, int NT, int KMAX>
__device__ __forceinline__ void apply_regblock(T* sr, T* si, uint32_t N, uint64_t base,
                                               const TileOp& op,
                                               const TileOp* __restrict__ mem,
                                               const T* __restrict__ pool) {
  constexpr uint32_t D = 1u << NB;
  constexpr int IT = (KMAX >> NB) > 0 ? (KMAX >> NB) : 1;
  const uint32_t nblk = N >> NB;
  const uint32_t nmem = op.n_diag;
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const uint32_t blk = threadIdx.x + uint32_t(it) * NT;
    if (blk >= nblk) continue;
    const uint32_t b0 = op_base(blk, op);
    T xr[D], xi[D];
    static_for<0, int(D)>([&](auto sc) __attribute__((always_inline)) {
      const uint32_t i = op_index<NB>(b0, uint32_t(sc.value), op);
      xr[sc.value] = sr[i];
      xi[sc.value] = si[i];
    });
    for (uint32_t u = 0; u < nmem; ++u) {
      const TileOp& g = mem[u];
      if ((base & g.cmask) != g.cval) continue;
      const uint32_t cm = g.insm[0], cv = g.lcval;
      const T* __restrict__ mat = pool + 2 * g.off;
      with_pair_variant(g.real, [&](auto vc) __attribute__((always_inline)) {
        constexpr uint32_t V = vc.value;
        with_kpos<NB>(g.lq[0], [&](auto jc) __attribute__((always_inline)) {
          constexpr int J = jc.value;
          static_for<0, int(D)>([&](auto sc) __attribute__((always_inline)) {
            constexpr uint32_t S = uint32_t(sc.value);
            if constexpr (((S >> J) & 1u) == 0) {
              constexpr uint32_t S1 = S | (1u << J);
              if ((S & cm) == cv) pair_math<T, V>(xr[S], xi[S], xr[S1], xi[S1], mat);
            }
          });
        });
      });
    }
    static_for<0, int(D)>([&](auto sc) __attribute__((always_inline)) {
      const uint32_t i = op_index<NB>(b0, uint32_t(sc.value), op);
      sr[i] = xr[sc.value];
      si[i] = xi[sc.value];
    });
  }
}

// ------------------------------------------------------------------ lane runs (kind 4)
// A lane run holds the tile in the register layout of the diagonal runs -- lane tid keeps
// the KMAX amplitudes e = tid + k NT -- and applies its members in order: runs of diagonal
// gates (member kind 1: TileDiags off .. off + n_diag) and 1-qubit ops (member kind 0)
// whose target is a lane bit (tile-local bit < 6: the pair partner is in the same wave,
// lane ^ 2^p) or a register bit (bit >= log2 NT: both amplitudes of a pair are in the
// lane's registers).
// Only the run's ends touch LDS, and one barrier follows it, where the LDS form of the same
// ops takes a round trip and a barrier per op.  Member conditions: outside bits cmask/cval
// (uniform per tile), tile-local bits insm[0] / lcval (per amplitude).

// Cross-lane exchange for lane bit P < 4: the value of lane (lane ^ 2^P) by DPP (VALU
// moves within a row of 16 lanes): quad_perm for P = 0, 1, row_ror:8 for P = 3, and for
// P = 2 row_half_mirror (lane ^ 7 within 8) followed by quad_perm [3,2,1,0] (lane ^ 3).
template <int P>
__device__ __forceinline__ uint32_t dpp_xor_u32(uint32_t v) {
  const int x = int(v);
  if constexpr (P == 0) return uint32_t(__builtin_amdgcn_update_dpp(x, x, 0xB1, 0xF, 0xF, true));
  if constexpr (P == 1) return uint32_t(__builtin_amdgcn_update_dpp(x, x, 0x4E, 0xF, 0xF, true));
  if constexpr (P == 3) return uint32_t(__builtin_amdgcn_update_dpp(x, x, 0x128, 0xF, 0xF, true));
  if constexpr (P == 2) {
    const int h = __builtin_amdgcn_update_dpp(x, x, 0x141, 0xF, 0xF, true);
    return uint32_t(__builtin_amdgcn_update_dpp(h, h, 0x1B, 0xF, 0xF, true));
  }
  return v;
}

template <int P, typename T>
__device__ __forceinline__ T dpp_xor(T v) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "float or double");
  uint32_t w[sizeof(T) / 4];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < int(sizeof(T) / 4); ++i) w[i] = dpp_xor_u32<P>(w[i]);
  T r;
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}

// Lane bit P = 4 or 5: v_permlane16_swap / v_permlane32_swap of v with a copy of itself
// leaves in the first result the value of lane (lane & ~2^P) and in the second that of
// lane (lane | 2^P) -- the pair's lower and upper amplitudes, in every lane.
template <int P, typename T>
__device__ __forceinline__ void permlane_pair(T v, T& lo, T& hi) {
  static_assert(P == 4 || P == 5, "permlane swaps exchange rows of 16 or halves of 32");
  uint32_t w[sizeof(T) / 4], a[sizeof(T) / 4], b[sizeof(T) / 4];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < int(sizeof(T) / 4); ++i) {
    const auto sw = P == 5 ? __builtin_amdgcn_permlane32_swap(w[i], w[i], false, false)
                           : __builtin_amdgcn_permlane16_swap(w[i], w[i], false, false);
    a[i] = sw[0];
    b[i] = sw[1];
  }
  __builtin_memcpy(&lo, a, sizeof(T));
  __builtin_memcpy(&hi, b, sizeof(T));
}

// A value that is the same in every lane, read into scalar registers: a select between two
// such values stays a select of values (a select between m[a] and m[b] would otherwise
// become a per-lane load from a selected address).
template <typename T>
__device__ __forceinline__ T uniform(T v) {
  uint32_t w[sizeof(T) / 4];
  __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
  for (int i = 0; i < int(sizeof(T) / 4); ++i)
    w[i] = uint32_t(__builtin_amdgcn_readfirstlane(int(w[i])));
  __builtin_memcpy(&v, w, sizeof(T));
  return v;
}

// Row `up` of pair_math<V> (y0 for the lower amplitude, y1 for the upper) for one lane:
// the lane's coefficients are picked once per op (c0 multiplies x0, c1 multiplies x1;
// anti-diagonal: c0 multiplies the partner), then every amplitude takes the operations
// pair_math performs for that row, in its order.
template <typename T, uint32_t V>
struct PairRow {
  T c0r, c0i, c1r, c1i;  // multiply x0 and x1 (anti-diagonal: c0 multiplies the partner)
  T csr, csi, cpr, cpi;  // multiply the lane's own amplitude and its partner
  __device__ __forceinline__ PairRow(const T* __restrict__ m, bool up) {
    // all entries are read before the selects (reads inside a select's arms would become
    // a divergent branch)
    T e[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) e[i] = uniform(m[i]);
    if constexpr (V == kPairAnti) {  // y0 = g01 x1, y1 = g10 x0
      c0r = up ? e[4] : e[2];
      c0i = up ? e[5] : e[3];
      c1r = c1i = T(0);
      csr = csi = T(0);
      cpr = c0r;
      cpi = c0i;
    } else {
      c0r = up ? e[4] : e[0];
      c0i = up ? e[5] : e[1];
      c1r = up ? e[6] : e[2];
      c1i = up ? e[7] : e[3];
      csr = up ? e[6] : e[0];  // g11 for the upper amplitude, g00 for the lower
      csi = up ? e[7] : e[1];
      cpr = up ? e[4] : e[2];  // g10 / g01
      cpi = up ? e[5] : e[3];
    }
  }
  // The same row as own term + partner term: pair_math adds the x0 term to the x1 term,
  // and IEEE addition is commutative, so for either row this is its value bit for bit.
  __device__ __forceinline__ void apply_sp(T& yr, T& yi, T xr, T xi, T pr, T pi) const {
    if constexpr (V == kPairSwap) {
      yr = pr;
      yi = pi;
    } else if constexpr (V == kPairAnti) {
      T ar = pr, ai = pi;
      cmul(ar, ai, cpr, cpi);
      yr = ar;
      yi = ai;
    } else if constexpr (V == kPairReal) {
      yr = csr * xr + cpr * pr;
      yi = csr * xi + cpr * pi;
    } else {
      const T ar = csr * xr - csi * xi, ai = csr * xi + csi * xr;
      const T br = cpr * pr - cpi * pi, bi = cpr * pi + cpi * pr;
      yr = ar + br;
      yi = ai + bi;
    }
  }
  // The row from the pair's lower and upper amplitudes (x0, x1), held in every lane.
  __device__ __forceinline__ void apply_x01(T& yr, T& yi, T x0r, T x0i, T x1r, T x1i,
                                            bool up) const {
    if constexpr (V == kPairSwap || V == kPairAnti) {  // only the partner enters
      apply_sp(yr, yi, T(0), T(0), up ? x0r : x1r, up ? x0i : x1i);
    } else if constexpr (V == kPairReal) {
      yr = c0r * x0r + c1r * x1r;
      yi = c0r * x0i + c1r * x1i;
    } else {
      const T ar = c0r * x0r - c0i * x0i, ai = c0r * x0i + c0i * x0r;
      const T br = c1r * x1r - c1i * x1i, bi = c1r * x1i + c1i * x1r;
      yr = ar + br;
      yi = ai + bi;
    }
  }
};

// 1-qubit member on lane bit P.
template <typename T, uint32_t V, int P, int NT, int KMAX>
__device__ __forceinline__ void run_1q_lane(T (&xr)[KMAX], T (&xi)[KMAX], uint32_t N,
                                            const T* __restrict__ m, uint32_t cm, uint32_t cv) {
  // recomputed per member (one VALU op): a lane mask per bit position hoisted to the kernel
  // entry would hold scalar registers for the whole kernel
  uint32_t tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const bool up = (tid >> P) & 1u;
  const PairRow<T, V> row(m, up);
  static_for<0, KMAX>([&](auto kc) __attribute__((always_inline)) {
    constexpr int k = kc.value;
    if (uint32_t(k) * NT < N) {  // uniform: whole waves are in or out for N >= 64
      T yr, yi;
      if constexpr (P < 4) {
        const T pr = dpp_xor<P>(xr[k]), pi = dpp_xor<P>(xi[k]);
        row.apply_sp(yr, yi, xr[k], xi[k], pr, pi);
      } else {
        T x0r, x1r, x0i, x1i;
        permlane_pair<P>(xr[k], x0r, x1r);
        permlane_pair<P>(xi[k], x0i, x1i);
        row.apply_x01(yr, yi, x0r, x0i, x1r, x1i, up);
      }
      if (cm == 0) {
        xr[k] = yr;
        xi[k] = yi;
      } else {
        const bool c = ((tid + uint32_t(k) * NT) & cm) == cv;
        xr[k] = c ? yr : xr[k];
        xi[k] = c ? yi : xi[k];
      }
    }
  });
}

// 1-qubit member on register bit J (tile-local bit log2(NT) + J).
template <typename T, uint32_t V, int J, int NT, int KMAX>
__device__ __forceinline__ void run_1q_reg(T (&xr)[KMAX], T (&xi)[KMAX], uint32_t N,
                                           const T* __restrict__ m, uint32_t cm, uint32_t cv) {
  static_for<0, KMAX>([&](auto kc) __attribute__((always_inline)) {
    constexpr int k = kc.value;
    if constexpr (((k >> J) & 1) == 0) {
      constexpr int k1 = k | (1 << J);
      if (uint32_t(k1) * NT < N) {
        if (cm == 0) {
          pair_math<T, V>(xr[k], xi[k], xr[k1], xi[k1], m);
        } else {
          T a0r = xr[k], a0i = xi[k], a1r = xr[k1], a1i = xi[k1];
          pair_math<T, V>(a0r, a0i, a1r, a1i, m);
          const bool c = ((threadIdx.x + uint32_t(k) * NT) & cm) == cv;
          xr[k] = c ? a0r : xr[k];
          xi[k] = c ? a0i : xi[k];
          xr[k1] = c ? a1r : xr[k1];
          xi[k1] = c ? a1i : xi[k1];
        }
      }
    }
  });
}

// FREE: the run may have its own register bits (op.m = 1) and kind 5 and 6 members; a
// kernel short of registers leaves that code out, and its host lowers with the fixed
// layout and kind 1 members only (TileBuilder::free_runs).
template <typename T, int NT, int KMAX, bool FREE = true>
__device__ __forceinline__ void apply_lane_run(T* sr, T* si, uint32_t N, uint64_t base,
                                               const TileOp& op,
                                               const TileOp* __restrict__ mem,
                                               const TileDiag* __restrict__ diags,
                                               const T* __restrict__ pool) {
  constexpr int LB = int(ilog2c(NT));    // lane + wave bits
  constexpr int LK = int(ilog2c(KMAX));  // register bits
  // Run coordinates u = tid + k NT: with the run's own register bits (op.m = 1; tile-local
  // positions perm[0..lq[0]), ascending, lq[0] <= LK register bits, N = NT 2^lq[0]) register
  // bit j of k is tile-local bit perm[j] and the lane's bits fill the other positions in
  // order; registers k >= 2^lq[0] hold nothing (u >= N).
  const bool own = FREE && op.m != 0;
  const uint32_t nreg = op.lq[0];
  uint32_t lane_e = threadIdx.x;
  if (own) {
    static_for<0, LK>([&](auto jc) __attribute__((always_inline)) {
      const uint32_t p = op.perm[jc.value];
      if (uint32_t(jc.value) < nreg) lane_e += lane_e & ~((1u << p) - 1u);
    });
  }
  auto tile_index = [&](auto kc) __attribute__((always_inline)) {
    constexpr int k = decltype(kc)::value;
    if (!own) return threadIdx.x + uint32_t(k) * NT;
    uint32_t r = 0;
    static_for<0, LK>([&](auto jc) __attribute__((always_inline)) {
      if constexpr ((k >> jc.value) & 1)
        if (uint32_t(jc.value) < nreg) r |= 1u << op.perm[jc.value];
    });
    return lane_e | r;
  };
  T xr[KMAX], xi[KMAX];
  static_for<0, KMAX>([&](auto kc) __attribute__((always_inline)) {
    const uint32_t e = tile_index(kc);
    xr[kc.value] = T(0);
    xi[kc.value] = T(0);
    if (threadIdx.x + uint32_t(kc.value) * NT < N) {
      xr[kc.value] = sr[e];
      xi[kc.value] = si[e];
    }
  });
  const uint32_t nmem = op.n_diag;
  for (uint32_t u = 0; u < nmem; ++u) {
    const TileOp& g = mem[u];
    if ((base & g.cmask) != g.cval) continue;
    if (g.kind == 1) {  // consecutive diagonal gates diags[off .. off + n_diag)
      diag_gates<T, NT, KMAX>(xr, xi, N, base, diags + g.off, g.n_diag, pool);
      continue;
    }
    if constexpr (FREE) {
      if (g.kind == 5) {  // phases controlled by outside bits
        ctrl_phases<T, NT, KMAX>(xr, xi, N, base, g, diags, pool);
        continue;
      }
      if (g.kind == 6) {  // phases on two in-tile bits
        pair_phases<T, NT, KMAX>(xr, xi, N, g, diags, pool);
        continue;
      }
    }
    const T* __restrict__ mat = pool + 2 * g.off;
    const uint32_t p = g.lq[0], cm = g.insm[0], cv = g.lcval;
    with_pair_variant(g.real, [&](auto vc) __attribute__((always_inline)) {
      constexpr uint32_t V = vc.value;
      if (p < 6u) {
        with_kpos<6>(p, [&](auto pc) __attribute__((always_inline)) {
          run_1q_lane<T, V, pc.value, NT, KMAX>(xr, xi, N, mat, cm, cv);
        });
      } else if constexpr (LK > 0) {
        with_kpos<LK>(p - uint32_t(LB), [&](auto jc) __attribute__((always_inline)) {
          run_1q_reg<T, V, jc.value, NT, KMAX>(xr, xi, N, mat, cm, cv);
        });
      }
    });
  }
  static_for<0, KMAX>([&](auto kc) __attribute__((always_inline)) {
    const uint32_t e = tile_index(kc);
    if (threadIdx.x + uint32_t(kc.value) * NT < N) {
      sr[e] = xr[kc.value];
      si[e] = xi[kc.value];
    }
  });
}

// Interpret ops[0..n_ops) on the tile (LDS); ends with a workgroup barrier.
template <typename T, int NT, int KMAX, bool FREE = true>
__device__ __forceinline__ void run_ops(T* sr, T* si, uint32_t N, uint64_t base,
                                        const TileOp* __restrict__ ops, uint32_t n_ops,
                                        const TileDiag* __restrict__ diags,
                                        const T* __restrict__ pool) {
  for (uint32_t o = 0; o < n_ops; ++o) {
    const TileOp op = ops[o];
    if (op.kind == 4) {  // lane run; its op.n_diag members follow it
      apply_lane_run<T, NT, KMAX, FREE>(sr, si, N, base, op, ops + o + 1, diags, pool);
      o += op.n_diag;
      __syncthreads();
      continue;
    }
    if (op.kind == 3) {  // register block; its op.n_diag members follow it
      if ((base & op.cmask) != op.cval) {  // the members' common outside condition
        o += op.n_diag;
        continue;
      }
      if (op.m == 1)
        apply_regblock<T, 1, NT, KMAX>(sr, si, N, base, op, ops + o + 1, pool);
      else if (op.m == 2)
        apply_regblock<T, 2, NT, KMAX>(sr, si, N, base, op, ops + o + 1, pool);
      else
        apply_regblock<T, 3, NT, KMAX>(sr, si, N, base, op, ops + o + 1, pool);
      o += op.n_diag;
      __syncthreads();
      continue;
    }
    if ((base & op.cmask) != op.cval) continue;  // uniform over the workgroup
    if (op.kind == 1) {
      apply_diag_run<T, NT, KMAX>(sr, si, N, base, op, diags, pool);
    } else if (op.kind == 0) {
      if (op.m == 1) {
        with_pair_variant(op.real, [&](auto vc) __attribute__((always_inline)) {
          apply_1q<T, vc.value, NT, KMAX>(sr, si, N, op, pool);
        });
      } else if (op.m == 2) {
        if (op.real)
          apply_dense<T, 2, NT, KMAX, true>(sr, si, N, op, pool);
        else
          apply_dense<T, 2, NT, KMAX, false>(sr, si, N, op, pool);
      } else {
        apply_dense<T, 3, NT, KMAX, false>(sr, si, N, op, pool);
      }
    } else {
      if (op.m == 2)
        apply_perm<T, 2, NT, KMAX>(sr, si, N, op, pool);
      else
        apply_perm<T, 3, NT, KMAX>(sr, si, N, op, pool);
    }
    __syncthreads();
  }
}

}  // namespace tile
}  // namespace qsvh
