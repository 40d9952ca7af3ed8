// plan.cpp -- launch-list passes of a program: routing wide dense gates to the matrix
// cores (MFMA precisions on complex64 states), the exact LDS tile passes, and the MFMA tile
// passes (runs of fused groups per 2^12-amplitude tile).
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/qsvh.h"
#include "program.hpp"

namespace qsvh {

namespace {

constexpr uint32_t kMfmaMinQubits = 4;   // 2D = 32 block rows: one 32x32 MFMA tile
constexpr uint32_t kMfmaTileBits = 11;   // the MFMA kernels' tile (2048 amplitudes)

// Round-to-nearest-even f32 -> bf16 bits (finite inputs).
uint16_t bf16_bits(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}
float bf16_value(uint16_t b) {
  const uint32_t u = uint32_t(b) << 16;
  float x;
  std::memcpy(&x, &u, 4);
  return x;
}

// The 2D x 2D real block matrix [[Ur, -Ui], [Ui, Ur]] of a gate whose internal index is
// re-expressed over its qubits in ascending order.
std::vector<float> block_matrix(const Launch& l, uint32_t* sorted) {
  const uint32_t m = l.m, D = 1u << m;
  std::copy(l.q, l.q + m, sorted);
  std::sort(sorted, sorted + m);
  // bit j of the original index is qubit l.q[j]; it moves to position rank(l.q[j]).
  uint32_t rank[kMaxGateQubits];
  for (uint32_t j = 0; j < m; ++j)
    rank[j] = uint32_t(std::find(sorted, sorted + m, l.q[j]) - sorted);
  auto remap = [&](uint32_t s) {
    uint32_t t = 0;
    for (uint32_t j = 0; j < m; ++j) t |= ((s >> j) & 1u) << rank[j];
    return t;
  };
  std::vector<float> a(size_t(4) * D * D);
  const uint32_t R = 2 * D;
  for (uint32_t r = 0; r < D; ++r)
    for (uint32_t c = 0; c < D; ++c) {
      const float ur = float(l.mat[2 * (size_t(r) * D + c)]);
      const float ui = float(l.mat[2 * (size_t(r) * D + c) + 1]);
      const uint32_t i = remap(r), j = remap(c);
      a[size_t(i) * R + j] = ur;
      a[size_t(i) * R + D + j] = -ui;
      a[size_t(D + i) * R + j] = ui;
      a[size_t(D + i) * R + D + j] = ur;
    }
  return a;
}

}  // namespace

// ----------------------------------------------------------------------- tile passes
namespace {

constexpr uint32_t kTileLowBits = 4;  // bits 0..3 always in a tile: 16-amplitude runs
// A 1-qubit op on a bit that the diagonal gates after it use this many times as an index
// bit starts a lane run with that bit among the register bits (a lane-bit index costs the
// multiply on every amplitude of the lane's registers, a register-bit one on those it picks).
constexpr uint32_t kHeavyDiag = 4;

uint32_t popcount64(uint64_t x) { return uint32_t(__builtin_popcountll(x)); }

// Qubits a gate actually acts on: those w.r.t. which its matrix is not block-diagonal
// (a control qubit c has U[r][c'] == 0 whenever bit c of r and c' differ).  A diagonal
// gate acts on none; CX(c, t) acts on t only.
uint64_t action_mask(const Launch& l) {
  if (l.kind == Launch::kDiag) return 0;
  const uint32_t D = 1u << l.m;
  uint64_t mask = 0;
  for (uint32_t j = 0; j < l.m; ++j) {
    bool control = true;
    for (uint32_t r = 0; r < D && control; ++r)
      for (uint32_t c = 0; c < D; ++c)
        if (((r ^ c) >> j) & 1u) {
          const size_t e = size_t(r) * D + c;
          if (l.mat[2 * e] != 0.0 || l.mat[2 * e + 1] != 0.0) {
            control = false;
            break;
          }
        }
    if (!control) mask |= 1ull << l.q[j];
  }
  return mask;
}

uint64_t qubit_mask(const Launch& l) {
  uint64_t m = 0;
  for (uint32_t j = 0; j < l.m; ++j) m |= 1ull << l.q[j];
  return m;
}


}  // namespace

void lower_exact_pass(const std::vector<Launch>& in, const std::vector<uint64_t>& take,
                      uint32_t n, uint32_t tb, uint64_t tmask, TileBuilder& bld,
                      std::vector<TilePassArg>& passes, std::vector<Launch>& out,
                      std::vector<uint64_t>& order) {
  const uint64_t low = (1ull << std::min(kTileLowBits, tb)) - 1;
  // Tile-local positions: the low bits first (in order: contiguous runs), then the other
  // tile qubits by their activity in this pass, least active first.  Positions from
  // kTileLaneBits up are register bits (a 1-qubit op there pairs amplitudes in a lane's
  // registers) and positions 4 .. kTileLaneBits - 1 need an LDS round trip, so the
  // qubits the pass's non-diagonal ops act on most take the register positions.
  // Activity: 2 per non-diagonal op acting on the qubit, 1 per diagonal gate on it.
  uint32_t busy[64] = {0};
  for (uint64_t i : take) {
    const uint64_t act = action_mask(in[i]);
    for (uint32_t j = 0; j < in[i].m; ++j) {
      const uint32_t q = in[i].q[j];
      if (in[i].kind == Launch::kDiag)
        busy[q] += (tmask >> q) & 1u;
      else if ((act >> q) & 1u)
        busy[q] += 2;
    }
  }
  std::vector<uint32_t> others;
  for (uint32_t b = 0; b < 64; ++b)
    if (((tmask & ~low) >> b) & 1u) others.push_back(b);
  std::stable_sort(others.begin(), others.end(),
                   [&](uint32_t a, uint32_t c) { return busy[a] < busy[c]; });
  TilePassArg pa{};
  pa.tb = tb;
  uint32_t pos[64] = {0};
  uint32_t k = 0, ks = 0;
  for (uint32_t b = 0; b < 64; ++b) {
    if ((tmask >> b) & 1u) pa.tsort[ks++] = b;
    if ((low >> b) & 1u) {
      pa.ts[k] = b;
      pos[b] = k++;
    }
  }
  for (uint32_t b : others) {
    pa.ts[k] = b;
    pos[b] = k++;
  }
  (void)n;
  pa.op_begin = uint32_t(bld.ops.size());
  bld.pass_begin = bld.ops.size();
  // Lane runs choose their register bits where the kernel allows it (TileBuilder::free_runs;
  // a tile pass kernel's tile is always full, tb bits of the kernel's tb).
  bld.tile_bits = tb;
  bld.run_lk = bld.free_runs && tb > kTileLaneBits ? std::min<uint32_t>(tb - kTileLaneBits, 4) : 0;
  for (size_t t = 0; t < take.size(); ++t) {
    const uint64_t i = take[t];
    // the tile bits that the diagonal gates up to the next other op use kHeavyDiag times
    bld.heavy = 0;
    if (in[i].kind != Launch::kDiag && bld.run_lk > 0) {
      uint32_t use[64] = {0};
      for (size_t u = t + 1; u < take.size() && in[take[u]].kind == Launch::kDiag; ++u)
        for (uint32_t j = 0; j < in[take[u]].m; ++j) {
          const uint32_t q = in[take[u]].q[j];
          if ((tmask >> q) & 1u && ++use[pos[q]] == kHeavyDiag) bld.heavy |= 1ull << pos[q];
        }
    }
    bld.lower(in[i], tmask, pos);
    order.push_back(i);
  }
  bld.close_block();
  bld.close_run();
  pa.n_ops = uint32_t(bld.ops.size()) - pa.op_begin;
  Launch l;
  l.kind = Launch::kTile;
  l.arena_off = passes.size();
  passes.push_back(pa);
  out.push_back(std::move(l));
}

void schedule_tiles(const std::vector<Launch>& in, uint32_t n, uint32_t tb, TileBuilder& bld,
                    std::vector<TilePassArg>& passes, std::vector<Launch>& out,
                    std::vector<uint64_t>& order) {
  std::vector<uint64_t> act(in.size()), qm(in.size());
  std::vector<bool> tileable(in.size());
  for (size_t i = 0; i < in.size(); ++i) {
    qm[i] = qubit_mask(in[i]);
    act[i] = action_mask(in[i]);
    tileable[i] = in[i].kind == Launch::kDiag || (popcount64(act[i]) <= 3 && in[i].m <= 3);
  }
  const uint64_t low = (1ull << std::min(kTileLowBits, tb)) - 1;

  auto flush = [&](std::vector<uint64_t> rem) {
    while (!rem.empty()) {
      // choose the tile bits: the low bits, then the action qubits of the gates that
      // come next, in order, skipping gates behind a blocked qubit
      uint64_t tmask = low, blocked = 0;
      for (uint64_t i : rem) {
        if (qm[i] & blocked) {
          blocked |= qm[i];
          continue;
        }
        const uint64_t need = act[i] & ~tmask;
        if (popcount64(tmask) + popcount64(need) <= tb)
          tmask |= need;
        else
          blocked |= qm[i];
        if (popcount64(tmask) == tb) break;
      }
      for (uint32_t b = 0; popcount64(tmask) < tb && b < n; ++b) tmask |= 1ull << b;
      // the pass takes every gate whose action qubits are in the tile and that is not
      // behind a blocked qubit
      std::vector<uint64_t> take, rest;
      blocked = 0;
      for (uint64_t i : rem) {
        if ((qm[i] & blocked) || (act[i] & ~tmask)) {
          blocked |= qm[i];
          rest.push_back(i);
          continue;
        }
        take.push_back(i);
      }
      lower_exact_pass(in, take, n, tb, tmask, bld, passes, out, order);
      rem.swap(rest);
    }
  };

  std::vector<uint64_t> seg;
  for (uint64_t i = 0; i < in.size(); ++i) {
    if (tileable[i]) {
      seg.push_back(i);
      continue;
    }
    flush(seg);
    seg.clear();
    out.push_back(in[i]);
    order.push_back(i);
  }
  flush(seg);
}

uint64_t TileBuilder::add_matrix(const double* g, size_t n_complex) {
    const uint64_t off = pool.size() / 2;
    pool.insert(pool.end(), g, g + 2 * n_complex);
    return off;
  }

void TileBuilder::lower(const Launch& l, uint64_t tmask, const uint32_t* pos) {
    if (l.kind == Launch::kDiag) {
      close_block();
      TileDiag d{};
      d.m = l.m;
      for (uint32_t j = 0; j < l.m; ++j) {
        d.q[j] = l.q[j];
        d.loc[j] = ((tmask >> l.q[j]) & 1u) ? pos[l.q[j]] : kOutside;
      }
      if (l.m > 2) {
        d.tab = uint32_t(add_matrix(l.mat.data(), size_t(1) << l.m));
        diag_phases.emplace_back();
      } else {  // its table waits for the run's coordinates
        diag_phases.emplace_back(l.mat.begin(), l.mat.begin() + 2 * (size_t(1) << l.m));
      }
      diags.push_back(d);
      if (open_run >= 0 && ops.back().kind == 1 && ops.size() > size_t(open_run) + 1 &&
          ops.back().off + ops.back().n_diag + 1 == diags.size()) {
        ops.back().n_diag++;  // extends the run's trailing diagonal member
        return;
      }
      TileOp mem{};
      mem.kind = 1;
      mem.off = diags.size() - 1;
      mem.n_diag = 1;
      run_member(mem);
      return;
    }
    const uint32_t D = 1u << l.m;
    uint32_t in_j[3], out_j[3];
    uint32_t nin = 0, nout = 0;
    for (uint32_t j = 0; j < l.m; ++j) {
      if ((tmask >> l.q[j]) & 1u)
        in_j[nin++] = j;
      else
        out_j[nout++] = j;
    }
    uint64_t cmask = 0;
    for (uint32_t k = 0; k < nout; ++k) cmask |= 1ull << l.q[out_j[k]];
    const uint32_t Din = 1u << nin;
    std::vector<double> sub(2 * size_t(Din) * Din);
    for (uint32_t v = 0; v < (1u << nout); ++v) {
      // sub-block for outside value v over the inside bits (in gate order)
      auto full = [&](uint32_t s) {
        uint32_t r = 0;
        for (uint32_t k = 0; k < nin; ++k) r |= ((s >> k) & 1u) << in_j[k];
        for (uint32_t k = 0; k < nout; ++k) r |= ((v >> k) & 1u) << out_j[k];
        return r;
      };
      for (uint32_t r = 0; r < Din; ++r)
        for (uint32_t c = 0; c < Din; ++c) {
          const size_t e = size_t(full(r)) * D + full(c);
          sub[2 * (size_t(r) * Din + c)] = l.mat[2 * e];
          sub[2 * (size_t(r) * Din + c) + 1] = l.mat[2 * e + 1];
        }
      uint64_t cval = 0;
      for (uint32_t k = 0; k < nout; ++k) cval |= uint64_t((v >> k) & 1u) << l.q[out_j[k]];
      lower_block(sub.data(), nin, in_j, l.q, pos, cmask, cval);
    }
  }

// Phase table and lowering variant of a diagonal gate of m <= 2 qubits whose loc[] are in
// its run's coordinates (lane bits below kTileLaneBits, register bits from there up).
void TileBuilder::finish_small_diag(TileDiag& d, const std::vector<double>& mat) {
  // index bits: lane bits first, then register bits by ascending k bit
  struct Bit { uint32_t j, kind, lp; };
  Bit ib[2];
  uint32_t ni = 0, oj[2], no = 0;
  for (uint32_t j = 0; j < d.m; ++j) {
    if (d.loc[j] == kOutside) {
      oj[no++] = j;
    } else if (d.loc[j] < kTileLaneBits) {
      ib[ni++] = Bit{j, 1, d.loc[j]};
    } else {
      ib[ni++] = Bit{j, 2 + d.loc[j] - kTileLaneBits, 0};
    }
  }
  if (ni == 2 && (ib[0].kind > ib[1].kind)) std::swap(ib[0], ib[1]);
  const uint32_t variant = (ni > 0 ? ib[0].kind : 0) | ((ni > 1 ? ib[1].kind : 0) << 4);
  const uint32_t lp0 = ni > 0 && ib[0].kind == 1 ? ib[0].lp : 0;
  const uint32_t lp1 = ni > 1 && ib[1].kind == 1 ? ib[1].lp : 0;
  d.code = variant | lp0 << 8 | lp1 << 12 | no << 16;
  d.oq = (no > 0 ? d.q[oj[0]] : 0) | (no > 1 ? d.q[oj[1]] : 0) << 8;
  d.idw = 0;
  // 4 phases per outside value so, indexed by c; unreachable c hold zeros
  std::vector<double> tab(2 * 4 * (size_t(1) << no), 0.0);
  for (uint32_t so = 0; so < (1u << no); ++so)
    for (uint32_t c = 0; c < 4; ++c) {
      if ((c >> ni) != 0) continue;
      uint32_t s = 0;
      for (uint32_t i = 0; i < ni; ++i) s |= ((c >> i) & 1u) << ib[i].j;
      for (uint32_t o = 0; o < no; ++o) s |= ((so >> o) & 1u) << oj[o];
      const size_t at = 2 * (4 * size_t(so) + c);
      tab[at] = mat[2 * s];
      tab[at + 1] = mat[2 * s + 1];
      if (tab[at] == 1.0 && tab[at + 1] == 0.0) d.idw |= 1u << (4 * so + c);
    }
  d.tab = uint32_t(add_matrix(tab.data(), tab.size() / 2));
}

// Emit the ops of one sub-block (gate order over the inside bits in_j, nin of them, with
// the outside condition cmask/cval).  Inside bits the block is block-diagonal in (no
// nonzero entry links a row and a column that differ in the bit: controls) become
// conditions on the tile-local index; each non-identity block over the remaining action
// bits becomes a permutation op (every row and column holding exactly one nonzero) or a
// dense op.  Skipping exact zero and exact identity factors changes no value: the
// multiplies skipped would add or multiply by exact 0 or 1 (up to the sign of a zero).
void TileBuilder::lower_block(const double* sub, uint32_t nin, const uint32_t* in_j,
                              const uint32_t* q, const uint32_t* pos, uint64_t cmask,
                              uint64_t cval) {
  const uint32_t Din = 1u << nin;
  auto nz = [&](uint32_t r, uint32_t c) {
    const size_t e = 2 * (size_t(r) * Din + c);
    return sub[e] != 0.0 || sub[e + 1] != 0.0;
  };
  bool is_identity = true;
  for (uint32_t r = 0; r < Din; ++r)
    for (uint32_t c = 0; c < Din; ++c) {
      const size_t e = 2 * (size_t(r) * Din + c);
      if (sub[e] != (r == c ? 1.0 : 0.0) || sub[e + 1] != 0.0) is_identity = false;
    }
  if (is_identity) return;
  uint32_t ctrl = 0;  // sub-index bits that are conditions
  for (uint32_t k = 0; k < nin; ++k) {
    bool blockdiag = true;
    for (uint32_t r = 0; r < Din; ++r)
      for (uint32_t c = 0; c < Din; ++c)
        if ((((r ^ c) >> k) & 1u) && nz(r, c)) blockdiag = false;
    if (blockdiag) ctrl |= 1u << k;
  }
  if (ctrl == Din - 1) ctrl = 0;  // diagonal block (not from a diagonal gate): keep it dense
  uint32_t act[3], na = 0, cb[3], nc = 0;
  for (uint32_t k = 0; k < nin; ++k) {
    if ((ctrl >> k) & 1u)
      cb[nc++] = k;
    else
      act[na++] = k;
  }
  const uint32_t Da = 1u << na;
  std::vector<double> blk(2 * size_t(Da) * Da);
  for (uint32_t w = 0; w < (1u << nc); ++w) {
    auto idx = [&](uint32_t a) {
      uint32_t s = 0;
      for (uint32_t i = 0; i < na; ++i) s |= ((a >> i) & 1u) << act[i];
      for (uint32_t i = 0; i < nc; ++i) s |= ((w >> i) & 1u) << cb[i];
      return s;
    };
    bool ident = true, real = true;
    for (uint32_t r = 0; r < Da; ++r)
      for (uint32_t c = 0; c < Da; ++c) {
        const size_t e = 2 * (size_t(idx(r)) * Din + idx(c));
        const size_t o = 2 * (size_t(r) * Da + c);
        blk[o] = sub[e];
        blk[o + 1] = sub[e + 1];
        if (blk[o] != (r == c ? 1.0 : 0.0) || blk[o + 1] != 0.0) ident = false;
        if (blk[o + 1] != 0.0) real = false;
      }
    if (ident) continue;
    TileOp op{};
    op.m = na;
    for (uint32_t i = 0; i < na; ++i) op.lq[i] = pos[q[in_j[act[i]]]];
    uint32_t ins[3], ni = 0;
    for (uint32_t i = 0; i < na; ++i) ins[ni++] = op.lq[i];
    for (uint32_t i = 0; i < nc; ++i) {
      const uint32_t b = pos[q[in_j[cb[i]]]];
      ins[ni++] = b;
      op.lcval |= ((w >> i) & 1u) << b;
    }
    std::sort(ins, ins + ni);
    op.nins = ni;
    for (uint32_t i = 0; i < ni; ++i) {
      op.ins[i] = ins[i];
      op.insm[i] = ~((1u << ins[i]) - 1u);
    }
    op.cmask = cmask;
    op.cval = cval;
    // permutation: exactly one nonzero per row and per column
    bool perm = true;
    uint32_t col_used = 0;
    for (uint32_t r = 0; r < Da && perm; ++r) {
      int found = -1;
      for (uint32_t c = 0; c < Da; ++c) {
        const size_t o = 2 * (size_t(r) * Da + c);
        if (blk[o] != 0.0 || blk[o + 1] != 0.0) {
          if (found >= 0) perm = false;
          found = int(c);
        }
      }
      if (found < 0 || ((col_used >> found) & 1u)) perm = false;
      if (perm) {
        col_used |= 1u << found;
        op.perm[r] = uint8_t(found);
      }
    }
    if (na == 1) {
      // 1-qubit blocks keep the dense 2x2; `real` names the pair variant of the kernel
      op.kind = 0;
      const bool anti = perm && op.perm[0] == 1;
      const bool unit = anti && blk[2] == 1.0 && blk[3] == 0.0 && blk[4] == 1.0 && blk[5] == 0.0;
      op.real = unit ? 3u : anti ? 2u : real ? 1u : 0u;
      op.off = add_matrix(blk.data(), 4);
    } else if (perm) {
      op.kind = 2;
      std::vector<double> coef(2 * Da);
      for (uint32_t r = 0; r < Da; ++r) {
        const size_t o = 2 * (size_t(r) * Da + op.perm[r]);
        coef[2 * r] = blk[o];
        coef[2 * r + 1] = blk[o + 1];
        if (blk[o] == 1.0 && blk[o + 1] == 0.0) op.pid |= 1u << r;
      }
      op.off = add_matrix(coef.data(), Da);
    } else {
      op.kind = 0;
      op.real = real ? 1u : 0u;
      op.off = add_matrix(blk.data(), size_t(Da) * Da);
    }
    emit(op);
  }
}

// Register blocks take 1-qubit ops on tile-local bits >= kBlockMinBit only: a block on
// bits 0..2 has every lane 8 amplitudes apart in LDS (16-way bank conflicts), where a
// standalone pair op keeps consecutive lanes on consecutive amplitudes.  All members of a
// block share its outside condition, so a tile that fails it skips the whole block.
constexpr uint32_t kBlockMinBit = 3;

// Lane-run members: 1-qubit ops on tile-local bits below kRunLaneBits (lanes of one row of
// 16: the pair partner is one or two DPP moves away) or from kTileLaneBits up (register
// bits).  Bits 4 and 5 are lanes too, but their exchange (permlane swaps) costs more than
// sharing a register block's LDS round trip with the neighbouring bits, as bits 6..8 do.
constexpr uint32_t kRunLaneBits = 4;

void TileBuilder::run_member(const TileOp& m) {
  if (open_run < 0) {
    TileOp h{};
    h.kind = 4;
    open_run = int64_t(ops.size());
    ops.push_back(h);
  }
  ops[size_t(open_run)].n_diag++;
  ops.push_back(m);
}

// Position of tile-local bit b among the bits that are not the open run's register bits
// (its lane position once those are its register bits).
static uint32_t lane_rank(uint32_t b, const uint32_t* reg, uint32_t nreg) {
  uint32_t r = b;
  for (uint32_t j = 0; j < nreg; ++j) r -= reg[j] < b ? 1u : 0u;
  return r;
}

void TileBuilder::emit(const TileOp& op) {
  if (op.kind == 0 && op.m == 1 && run_lk > 0) {
    // a run of free register bits: the op joins the open run or starts one
    close_block();
    const uint32_t b = op.lq[0];
    bool in = false;
    // (a run whose register bits are among tile bits 0..3 loads and stores its amplitudes
    // 8 or 16 apart in LDS, with bank conflicts; measured on QFT(30) c128, keeping those
    // bits lane bits instead cost more: 110 -> 122 ms)
    if (open_run >= 0) {
      for (uint32_t j = 0; j < run_nreg; ++j) in = in || run_reg[j] == b;
      if (!in && run_nreg < run_lk) {
        run_reg[run_nreg++] = b;
        in = true;
      }
      if (!in && lane_rank(b, run_reg, run_nreg) < kRunLaneBits && !((heavy >> b) & 1u))
        in = true;
    }
    if (!in) {
      close_run();
      run_reg[0] = b;
      run_nreg = 1;
    }
    TileOp mem = op;
    uint32_t cm = 0;  // condition bits: the op's inserted bits other than its target
    for (uint32_t i = 0; i < op.nins; ++i)
      if (op.ins[i] != op.lq[0]) cm |= 1u << op.ins[i];
    mem.insm[0] = cm;
    run_member(mem);
    return;
  }
  if (op.kind == 0 && op.m == 1 && (op.lq[0] < kRunLaneBits || op.lq[0] >= kTileLaneBits)) {
    close_block();
    TileOp mem = op;
    uint32_t cm = 0;  // condition bits: the op's inserted bits other than its target
    for (uint32_t i = 0; i < op.nins; ++i)
      if (op.ins[i] != op.lq[0]) cm |= 1u << op.ins[i];
    mem.insm[0] = cm;
    run_member(mem);
    return;
  }
  close_run();
  uint32_t bits = 0;
  for (uint32_t i = 0; i < op.nins; ++i) bits |= 1u << op.ins[i];
  if (op.m == 1 && op.kind == 0 && (bits & ((1u << kBlockMinBit) - 1u)) == 0) {
    if (open_block >= 0 && __builtin_popcount(block_bits | bits) <= 3 &&
        ops[size_t(open_block)].cmask == op.cmask && ops[size_t(open_block)].cval == op.cval) {
      block_bits |= bits;
      ops[size_t(open_block)].n_diag++;
      ops.push_back(op);
      return;
    }
    close_block();
    TileOp b{};
    b.kind = 3;
    b.n_diag = 1;
    b.cmask = op.cmask;
    b.cval = op.cval;
    open_block = int64_t(ops.size());
    block_bits = bits;
    ops.push_back(b);
    ops.push_back(op);
    return;
  }
  close_block();
  ops.push_back(op);
}

// Close the open lane run: choose its register bits (those its 1-qubit members claimed,
// then the in-tile bits its diagonal members use most, then the highest tile bits), record
// them in the head op (m = 1, lq[0] of them, perm[j] = tile-local position of register bit
// j, ascending)
// and rewrite the members in the run's coordinates: lane position (rank among the other
// tile bits) for the rest, kTileLaneBits + j for register bit j.  With fixed register bits
// the coordinates are the tile-local bits themselves.
void TileBuilder::close_run() {
  if (open_run < 0) return;
  const size_t h = size_t(open_run);
  open_run = -1;
  const uint32_t nmem = ops[h].n_diag;
  uint32_t map[32];
  for (uint32_t b = 0; b < 32; ++b) map[b] = b;
  if (run_lk > 0) {
    uint32_t use[32] = {0};
    for (size_t k = h + 1; k <= h + nmem; ++k)
      if (ops[k].kind == 1)
        for (uint32_t d = 0; d < ops[k].n_diag; ++d) {
          const TileDiag& g = diags[ops[k].off + d];
          for (uint32_t j = 0; j < g.m; ++j)
            if (g.loc[j] != kOutside) use[g.loc[j]]++;
        }
    auto taken = [&](uint32_t b) {
      for (uint32_t j = 0; j < run_nreg; ++j)
        if (run_reg[j] == b) return true;
      return false;
    };
    while (run_nreg < run_lk) {
      uint32_t best = 32;
      for (uint32_t b = 0; b < tile_bits; ++b)
        if (!taken(b) && use[b] > 0 && (best == 32 || use[b] > use[best])) best = b;
      if (best == 32)  // the highest free tile bit
        for (uint32_t b = tile_bits; b-- > 0;)
          if (!taken(b)) {
            best = b;
            break;
          }
      run_reg[run_nreg++] = best;
    }
    std::sort(run_reg, run_reg + run_nreg);
    for (uint32_t b = 0; b < tile_bits; ++b) map[b] = lane_rank(b, run_reg, run_nreg);
    for (uint32_t j = 0; j < run_nreg; ++j) map[run_reg[j]] = kTileLaneBits + j;
    ops[h].m = 1;
    ops[h].lq[0] = run_nreg;
    for (uint32_t j = 0; j < run_nreg; ++j) ops[h].perm[j] = uint8_t(run_reg[j]);
  }
  auto remap = [&](uint32_t mask) {
    uint32_t r = 0;
    for (uint32_t b = 0; b < 32; ++b)
      if ((mask >> b) & 1u) r |= 1u << map[b];
    return r;
  };
  for (size_t k = h + 1; k <= h + nmem; ++k) {
    TileOp& g = ops[k];
    if (g.kind == 1) {
      for (uint32_t d = 0; d < g.n_diag; ++d) {
        TileDiag& dg = diags[g.off + d];
        for (uint32_t j = 0; j < dg.m; ++j)
          if (dg.loc[j] != kOutside) dg.loc[j] = map[dg.loc[j]];
        if (dg.m <= 2) finish_small_diag(dg, diag_phases[g.off + d]);
      }
    } else {
      g.lq[0] = map[g.lq[0]];
      g.insm[0] = remap(g.insm[0]);
      g.lcval = remap(g.lcval);
    }
  }
  run_nreg = 0;
  // Stretches of diagonal gates that are one phase controlled by one outside bit, on the
  // same index bit (a controlled phase whose control is outside the tile: QFT's ladders)
  // become kind 5 members: no per-gate lowering variant to decode, the phases side by side.
  if (!free_runs) return;
  const std::vector<TileOp> mem(ops.begin() + std::ptrdiff_t(h) + 1, ops.end());
  ops.resize(h + 1);
  for (const TileOp& g : mem) {
    if (g.kind != 1) {
      ops.push_back(g);
      continue;
    }
    uint32_t d = 0;
    auto special = [&](uint32_t i) {
      uint32_t x, y;
      return ctrl_phase_bit(diags[g.off + i]) >= 0 ||
             (ctrl_pair_bits(diags[g.off + i], &x, &y) &&
              (x >= kTileLaneBits || y >= kTileLaneBits));
    };
    while (d < g.n_diag) {
      uint32_t e = d;
      while (e < g.n_diag && !special(e)) ++e;
      if (e > d) {
        TileOp t = g;
        t.off = g.off + d;
        t.n_diag = e - d;
        ops.push_back(t);
      }
      if (e == g.n_diag) break;
      std::vector<double> ph;
      uint32_t f = e;
      TileOp t{};
      const int bit = ctrl_phase_bit(diags[g.off + e]);
      if (bit >= 0) {
        for (; f < g.n_diag && ctrl_phase_bit(diags[g.off + f]) == bit; ++f) {
          const size_t at = 2 * (size_t(diags[g.off + f].tab) + 5);  // so = 1, c = 1
          ph.push_back(pool[at]);
          ph.push_back(pool[at + 1]);
        }
        t.kind = 5;
        t.lq[0] = uint32_t(bit);
      } else {
        // both bits in the tile: the stretch keeps a register bit its gates share (QFT's
        // ladder target); each gate's other bit goes into its oq
        uint32_t x, y;
        ctrl_pair_bits(diags[g.off + e], &x, &y);
        if (x < kTileLaneBits) std::swap(x, y);  // x: a register bit
        uint32_t b = x, nx, ny;
        if (y >= kTileLaneBits && e + 1 < g.n_diag &&
            ctrl_pair_bits(diags[g.off + e + 1], &nx, &ny) && (y == nx || y == ny) &&
            x != nx && x != ny)
          b = y;
        for (; f < g.n_diag && ctrl_pair_bits(diags[g.off + f], &nx, &ny) && (nx == b || ny == b);
             ++f) {
          TileDiag& dg = diags[g.off + f];
          dg.oq = nx == b ? ny : nx;
          const size_t at = 2 * (size_t(dg.tab) + 3);  // so = 0, c = 3
          ph.push_back(pool[at]);
          ph.push_back(pool[at + 1]);
        }
        t.kind = 6;
        t.lq[0] = b;
      }
      t.off = g.off + e;
      t.n_diag = f - e;
      t.ins[0] = uint32_t(add_matrix(ph.data(), f - e));
      t.cmask = g.cmask;
      t.cval = g.cval;
      ops.push_back(t);
      d = f;
    }
  }
  ops[h].n_diag = uint32_t(ops.size() - h - 1);
}

// A lowered diagonal gate that multiplies by one phase exactly the amplitudes whose two
// in-tile bits are both 1 (every other phase exactly 1 + 0i): true, and its bits in run
// coordinates.
bool TileBuilder::ctrl_pair_bits(const TileDiag& d, uint32_t* x, uint32_t* y) const {
  if (d.m != 2 || ((d.code >> 16) & 3u) != 0u) return false;
  const uint32_t kind0 = d.code & 15u, kind1 = (d.code >> 4) & 15u;
  if (kind0 == 0 || kind1 == 0) return false;
  if ((d.idw & 0xFu) != 0x7u) return false;  // identities at c = 0, 1, 2, not at c = 3
  *x = kind0 == 1 ? (d.code >> 8) & 15u : kTileLaneBits + kind0 - 2;
  *y = kind1 == 1 ? (d.code >> 12) & 15u : kTileLaneBits + kind1 - 2;
  return true;
}

// The index bit (run coordinates) of a lowered diagonal gate that multiplies by one phase
// exactly the amplitudes whose one outside bit and one in-tile bit are both 1 (every other
// phase of its table exactly 1 + 0i), else -1.
int TileBuilder::ctrl_phase_bit(const TileDiag& d) const {
  if (d.m > 2 || ((d.code >> 16) & 3u) != 1u) return -1;
  const uint32_t kind0 = d.code & 15u, kind1 = (d.code >> 4) & 15u;
  if (kind0 == 0 || kind1 != 0) return -1;
  // identities: so = 0 at c = 0, 1; so = 1 at c = 0, not at c = 1
  if ((d.idw & 0x33u) != 0x13u) return -1;
  return kind0 == 1 ? int((d.code >> 8) & 15u) : int(kTileLaneBits + kind0 - 2);
}

// Finalise the open register block: its bits (ascending) become lq / ins of the block
// op; each member's target becomes an index j into them and its conditions a mask and
// value over the block's sub-index.  A block of one member is just that member.
void TileBuilder::close_block() {
  if (open_block < 0) return;
  const size_t h = size_t(open_block);
  open_block = -1;
  TileOp& b = ops[h];
  if (b.n_diag == 1) {
    ops.erase(ops.begin() + std::ptrdiff_t(h));
    return;
  }
  uint32_t nb = 0, bit[3];
  for (uint32_t t = 0; t < 32; ++t)
    if ((block_bits >> t) & 1u) bit[nb++] = t;
  b.m = nb;
  b.nins = nb;
  for (uint32_t j = 0; j < nb; ++j) {
    b.lq[j] = bit[j];
    b.ins[j] = bit[j];
    b.insm[j] = ~((1u << bit[j]) - 1u);
  }
  auto index_of = [&](uint32_t t) {
    for (uint32_t j = 0; j < nb; ++j)
      if (bit[j] == t) return j;
    return 0u;
  };
  for (size_t k = h + 1; k < h + 1 + b.n_diag; ++k) {
    TileOp& g = ops[k];
    const uint32_t target = g.lq[0];
    uint32_t cm = 0, cv = 0;
    for (uint32_t i = 0; i < g.nins; ++i) {
      const uint32_t t = g.ins[i];
      if (t == target) continue;
      const uint32_t j = index_of(t);
      cm |= 1u << j;
      cv |= ((g.lcval >> t) & 1u) << j;
    }
    g.lq[0] = index_of(target);
    g.insm[0] = cm;
    g.lcval = cv;
  }
}

int plan_tiles(Program* p, int tile_qubits, PermTail* pt) {
  State& st = *p->st;
  p->order.clear();
  if (tile_qubits <= 0) {
    for (uint64_t i = 0; i < p->launches.size(); ++i) p->order.push_back(i);
    return kOk;
  }
  const uint32_t lo = st.dtype == 0 ? 10 : 11, hi = st.dtype == 0 ? 12 : 13;
  if (uint32_t(tile_qubits) < lo || uint32_t(tile_qubits) > hi)
    return set_err(st.ctx, kInvalidArg,
                   st.dtype == 0 ? "tile_qubits must be 10..12 for complex128"
                                 : "tile_qubits must be 11..13 for complex64");
  const uint32_t tb = uint32_t(tile_qubits);
  if (st.n < tb) {  // the state is smaller than one tile: plain launches
    for (uint64_t i = 0; i < p->launches.size(); ++i) p->order.push_back(i);
    return kOk;
  }
  const std::vector<Launch> in = std::move(p->launches);
  p->launches.clear();
  TileBuilder tbld;
  if (pt && !pt->swaps.empty()) {  // launches before k, the bit swaps, the rest
    const std::vector<Launch> pre(in.begin(), in.begin() + pt->k);
    const std::vector<Launch> suf(in.begin() + pt->k, in.end());
    schedule_tiles(pre, st.n, tb, tbld, p->tile_passes, p->launches, p->order);
    for (Launch& l : pt->swaps) p->launches.push_back(std::move(l));
    pt->swaps.clear();
    const size_t o0 = p->order.size();
    schedule_tiles(suf, st.n, tb, tbld, p->tile_passes, p->launches, p->order);
    for (size_t i = o0; i < p->order.size(); ++i) p->order[i] += pt->k;
  } else {
    schedule_tiles(in, st.n, tb, tbld, p->tile_passes, p->launches, p->order);
  }
  if (tbld.ops.empty()) return kOk;
  // upload ops, diagonal descriptors and the matrix pool (state precision)
  auto upload = [&](const void* src, size_t bytes, void** dst) {
    hipError_t e = hipMalloc(dst, bytes ? bytes : 16);
    if (e != hipSuccess) return hip_err(st.ctx, e, "tile buffers alloc");
    p->owned.push_back(*dst);
    if (bytes) e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    return hip_err(st.ctx, e, "tile buffers upload");
  };
  int rc = upload(tbld.ops.data(), tbld.ops.size() * sizeof(TileOp),
                  reinterpret_cast<void**>(&p->tile_ops));
  if (rc != kOk) return rc;
  rc = upload(tbld.diags.data(), tbld.diags.size() * sizeof(TileDiag),
              reinterpret_cast<void**>(&p->tile_diags));
  if (rc != kOk) return rc;
  if (st.dtype == 0) {
    rc = upload(tbld.pool.data(), tbld.pool.size() * sizeof(double), &p->tile_pool);
  } else {
    std::vector<float> f(tbld.pool.begin(), tbld.pool.end());
    rc = upload(f.data(), f.size() * sizeof(float), &p->tile_pool);
  }
  return rc;
}

// Subspace forms of per-gate launches (qsvh_program_opts::subspace): a diagonal gate with
// one entry other than exactly 1 + 0i touches only that entry's amplitudes; an m = 2, 3
// gate that is a 1-qubit gate on one qubit under controls on the others (one control value
// with a non-identity block, identity blocks elsewhere) touches only that control subspace.
// Exactly the SWAP matrix (gate.rs:160-168): ones at (0,0), (1,2), (2,1), (3,3), zeros
// elsewhere (its qubit order does not matter).
static bool is_swap(const Launch& l) {
  for (uint32_t r = 0; r < 4; ++r)
    for (uint32_t c = 0; c < 4; ++c) {
      const bool one = (r == 0 && c == 0) || (r == 3 && c == 3) || (r == 1 && c == 2) ||
                       (r == 2 && c == 1);
      const size_t e = 2 * (size_t(r) * 4 + c);
      if (l.mat[e] != (one ? 1.0 : 0.0) || l.mat[e + 1] != 0.0) return false;
    }
  return true;
}

void plan_subspace(Program* p) {
  for (Launch& l : p->launches) {
    if (l.kind == Launch::kDiag) {
      const uint32_t d = 1u << l.m;
      int one = -1, count = 0;
      for (uint32_t s = 0; s < d; ++s)
        if (l.mat[2 * s] != 1.0 || l.mat[2 * s + 1] != 0.0) {
          one = int(s);
          ++count;
        }
      if (count != 1) continue;
      uint64_t fix = 0;
      for (uint32_t j = 0; j < l.m; ++j) fix |= uint64_t((uint32_t(one) >> j) & 1u) << l.q[j];
      const double pr = l.mat[2 * one], pi = l.mat[2 * one + 1];
      l.kind = Launch::kDiagSub;
      l.fix = fix;
      l.mat = {pr, pi};
    } else if (l.kind == Launch::kMq && l.m == 2 && is_swap(l)) {
      l.kind = Launch::kSwapSub;
      l.mat.clear();
    } else if (l.kind == Launch::kMq && l.m <= 3) {
      const uint32_t D = 1u << l.m;
      auto nz = [&](uint32_t r, uint32_t c) {
        const size_t e = 2 * (size_t(r) * D + c);
        return l.mat[e] != 0.0 || l.mat[e + 1] != 0.0;
      };
      int target = -1;
      bool ok = true;
      for (uint32_t j = 0; j < l.m && ok; ++j) {
        bool control = true;
        for (uint32_t r = 0; r < D && control; ++r)
          for (uint32_t c = 0; c < D; ++c)
            if ((((r ^ c) >> j) & 1u) && nz(r, c)) {
              control = false;
              break;
            }
        if (!control) {
          if (target >= 0) ok = false;  // two action qubits
          target = int(j);
        }
      }
      if (!ok || target < 0) continue;
      // control values w (over the other bits, in gate order) with a non-identity block
      int hit = -1, hits = 0;
      std::vector<double> blk(8);
      for (uint32_t w = 0; w < (D >> 1); ++w) {
        auto idx = [&](uint32_t bit) {  // sub-index with target bit `bit`, controls = w
          const uint32_t lowmask = (1u << target) - 1u;
          return ((w & ~lowmask) << 1) | (w & lowmask) | (bit << target);
        };
        bool ident = true;
        for (uint32_t r = 0; r < 2; ++r)
          for (uint32_t c = 0; c < 2; ++c) {
            const size_t e = 2 * (size_t(idx(r)) * D + idx(c));
            if (l.mat[e] != (r == c ? 1.0 : 0.0) || l.mat[e + 1] != 0.0) ident = false;
          }
        if (!ident) {
          hit = int(w);
          ++hits;
          for (uint32_t r = 0; r < 2; ++r)
            for (uint32_t c = 0; c < 2; ++c) {
              const size_t e = 2 * (size_t(idx(r)) * D + idx(c));
              blk[2 * (2 * r + c)] = l.mat[e];
              blk[2 * (2 * r + c) + 1] = l.mat[e + 1];
            }
        }
      }
      if (hits != 1) continue;
      uint32_t q[kMaxGateQubits];
      uint64_t fix = 0;
      q[0] = l.q[target];
      uint32_t k = 1;
      const uint32_t lowmask = (1u << target) - 1u;
      const uint32_t full = ((uint32_t(hit) & ~lowmask) << 1) | (uint32_t(hit) & lowmask);
      for (uint32_t j = 0; j < l.m; ++j) {
        if (int(j) == target) continue;
        q[k++] = l.q[j];
        fix |= uint64_t((full >> j) & 1u) << l.q[j];
      }
      std::copy(q, q + l.m, l.q);
      l.kind = Launch::kCtrl1q;
      l.fix = fix;
      l.mat = blk;
    }
  }
}

int plan_mfma(Program* p) {
  if (p->precision == QSVH_PREC_EXACT) return kOk;
  if (p->precision != QSVH_PREC_MFMA_F32 && p->precision != QSVH_PREC_MFMA_BF16X3)
    return set_err(p->st->ctx, kInvalidArg, "unknown precision");
  State& st = *p->st;
  if (st.n < kMfmaTileBits) return kOk;  // too small for one MFMA tile: VALU kernels
  std::vector<uint8_t> blob;
  for (Launch& l : p->launches) {
    if (l.kind != Launch::kMq || l.m < kMfmaMinQubits) continue;
    uint32_t sorted[kMaxGateQubits];
    const std::vector<float> a = block_matrix(l, sorted);
    blob.resize((blob.size() + 255) & ~size_t(255));
    l.mfma_off = blob.size();
    if (p->precision == QSVH_PREC_MFMA_F32) {
      const size_t bytes = a.size() * sizeof(float);
      blob.resize(blob.size() + bytes);
      std::memcpy(blob.data() + l.mfma_off, a.data(), bytes);
    } else {
      std::vector<uint16_t> hi(a.size()), lo(a.size());
      for (size_t e = 0; e < a.size(); ++e) {
        hi[e] = bf16_bits(a[e]);
        lo[e] = bf16_bits(a[e] - bf16_value(hi[e]));
      }
      const size_t bytes = a.size() * sizeof(uint16_t);
      blob.resize(blob.size() + 2 * bytes);
      std::memcpy(blob.data() + l.mfma_off, hi.data(), bytes);
      std::memcpy(blob.data() + l.mfma_off + bytes, lo.data(), bytes);
    }
    std::copy(sorted, sorted + l.m, l.q);
    l.kind = Launch::kFusedMfma;
  }
  if (blob.empty()) return kOk;
  void* d = nullptr;
  hipError_t e = hipMalloc(&d, blob.size());
  if (e != hipSuccess) return hip_err(st.ctx, e, "mfma matrices alloc");
  p->owned.push_back(d);
  p->mfma = d;
  e = hipMemcpy(d, blob.data(), blob.size(), hipMemcpyHostToDevice);
  return hip_err(st.ctx, e, "mfma matrices upload");
}

// ----------------------------------------------------------------------- MFMA tile passes
namespace {

// A dense gate on m < 4 qubits widened to 4 (identity on the added qubits `extra`): matrix
// index bit j <-> q[j], the gate's own bits lowest.
Launch widen_to_4(const Launch& l, const uint32_t* extra) {
  Launch w;
  w.kind = Launch::kMq;
  w.m = 4;
  std::copy(l.q, l.q + l.m, w.q);
  std::copy(extra, extra + (4 - l.m), w.q + l.m);
  const uint32_t d = 1u << l.m, D = 16;
  w.mat.assign(2 * D * D, 0.0);
  for (uint32_t r = 0; r < D; ++r)
    for (uint32_t c = 0; c < D; ++c)
      if ((r >> l.m) == (c >> l.m)) {
        const size_t src = 2 * (size_t(r & (d - 1)) * d + (c & (d - 1)));
        w.mat[2 * (size_t(r) * D + c)] = l.mat[src];
        w.mat[2 * (size_t(r) * D + c) + 1] = l.mat[src + 1];
      }
  return w;
}

}  // namespace

// A pass takes ops in circuit order while the qubits of its dense ops (with bits 0..3) fit
// in tb bits; diagonal ops join any pass (their qubits outside the tile are fixed per tile).
// An op that does not fit is left for a later pass, and the scan goes on past it: a later
// op joins the pass only if it commutes with every op left behind before it -- a dense op
// shares no qubit with them, a diagonal op none with the dense ones among them (diagonal
// ops commute with each other).  Moving such ops ahead leaves the product unchanged, and
// each pass applies its ops in circuit order.
constexpr size_t kMtLookahead = 1024;

// Emit one MFMA tile pass holding the launches `cur` (in order): the tile is the bits of S
// (the dense ops' qubits and the low bits), filled up to tb bits with the lowest others.
static uint32_t op_bit(const MtOp& o, uint32_t j) {
  return (o.bits[j >> 2] >> (8 * (j & 3))) & 0xffu;
}

void emit_mt_pass(const std::vector<Launch>& in, const std::vector<size_t>& cur, uint32_t n,
                  uint32_t tb, uint64_t S, MtPlan* out) {
  std::vector<MtOp>& ops = out->ops;
  std::vector<uint8_t>& pool = out->pool;
  auto put = [&](const void* src, size_t bytes) {
    pool.resize((pool.size() + 255) & ~size_t(255));
    const uint64_t at = pool.size();
    const uint8_t* b = static_cast<const uint8_t*>(src);
    pool.insert(pool.end(), b, b + bytes);
    return at;
  };
  if (cur.empty()) return;
  uint64_t tm = S;
  for (uint32_t b = 0; b < n && popcount64(tm) < tb; ++b) tm |= 1ull << b;
  MtPassArg pa{};
  pa.tb = tb;
  pa.kmax = 4;
  uint32_t pos[64] = {0};
  uint32_t nb = 0;
  for (uint32_t b = 0; b < n; ++b)
    if ((tm >> b) & 1) {
      pa.ts[nb] = b;
      pos[b] = nb++;
    }
  pa.op_begin = uint32_t(ops.size());
  std::vector<const Launch*> src;  // the launch of each of this pass's ops
  for (size_t li : cur) {
    const Launch& l = in[li];
    out->order.push_back(li);
    src.push_back(&l);
    MtOp op{};
    if (l.kind == Launch::kDiag) {
      op.kind = 1;
      op.k = l.m;
      for (uint32_t j = 0; j < l.m; ++j) {
        const uint32_t q = l.q[j];
        if ((tm >> q) & 1) {
          op.in_mask |= 1u << j;
          op.set_bit(j, pos[q]);
        } else {
          op.set_bit(j, q);
        }
      }
    } else {
      Launch g = l;
      if (l.m < 4) {
        uint32_t extra[4], ne = 0;
        for (uint32_t b = 0; b < n && ne < 4 - l.m; ++b)
          if (((tm >> b) & 1) && std::find(l.q, l.q + l.m, b) == l.q + l.m) extra[ne++] = b;
        g = widen_to_4(l, extra);
      }
      uint32_t sorted[kMaxGateQubits];
      const std::vector<float> a = block_matrix(g, sorted);
      op.kind = 0;
      op.k = g.m;
      pa.kmax = std::max(pa.kmax, g.m);
      uint32_t gm = 0;
      for (uint32_t j = 0; j < g.m; ++j) {
        op.set_bit(j, pos[sorted[j]]);
        gm |= 1u << pos[sorted[j]];
      }
      // the other bits, the first 5 (a wave's 32 columns) chosen with distinct residues
      // mod 5 where the group leaves them: tile-local bit b lands in LDS slot bit b mod 5
      // (the tile's XOR fold), so such columns hit 32 slots of a row of banks
      uint32_t o = 0, res = 0, used = gm;
      for (uint32_t b = 0; b < tb && o < 5; ++b)
        if (!((used >> b) & 1) && !((res >> (b % 5)) & 1)) {
          op.set_obit(o++, b);
          res |= 1u << (b % 5);
          used |= 1u << b;
        }
      for (uint32_t b = 0; b < tb; ++b)
        if (!((used >> b) & 1)) op.set_obit(o++, b);
      // rows in the kernel's band order: band b (32 rows) holds the real rows 16 b ..
      // 16 b + 15, then the imaginary rows D + 16 b .. D + 16 b + 15 (the identity at k = 4)
      const uint32_t D = 1u << g.m, R = 2 * D;
      std::vector<uint16_t> hl(2 * a.size());
      for (uint32_t row = 0; row < R; ++row) {
        const uint32_t band = row / 32, i = row % 32;
        const uint32_t sr = i < 16 ? 16 * band + i : D + 16 * band + (i - 16);
        for (uint32_t col = 0; col < R; ++col) {
          const float x = a[size_t(sr) * R + col];
          const size_t e = size_t(row) * R + col;
          hl[e] = bf16_bits(x);
          hl[a.size() + e] = bf16_bits(x - bf16_value(hl[e]));
        }
      }
      op.mat = put(hl.data(), hl.size() * sizeof(uint16_t));
    }
    ops.push_back(op);
  }
  // The pass's phase table (staged in LDS once per workgroup): every diagonal op's phases,
  // rounded to f32 once.
  std::vector<float> tabs;
  auto add_table = [&](const double* ph, uint32_t entries) {
    const uint32_t at = uint32_t(tabs.size() / 2);
    for (uint32_t e = 0; e < 2 * entries; ++e) tabs.push_back(float(ph[e]));
    return at;
  };
  // Runs of diagonal ops: consecutive ones (they commute) whose in-tile qubits fit 8 tile
  // bits share one table over those bits (the head op's).  The members with every qubit in
  // the tile compose into the head's base table here, once, in f64; the others (a qubit
  // outside the tile, fixed per tile) stay member ops, multiplied into a per-tile copy of
  // the base table by the kernel.  A run without such members becomes one single diagonal
  // op over its tile bits.  A run keeps all its members per tile where its base table
  // would take the pass's tables past kMtTabEntries, or past kMtDynEntriesFour when they
  // fit it without base tables (the fourth workgroup per CU is worth more).
  {
    struct Run {
      size_t e, f;       // ops [e, f)
      uint32_t U, k;     // tile-local bits, their count
      uint32_t own_static, n_dyn;  // entries of the in-tile members' own tables; the others
      bool fold;
    };
    std::vector<Run> runs;
    uint32_t tab = 0;  // entries of the ops' own tables (the callers keep this in bounds)
    for (size_t e = pa.op_begin; e < ops.size();) {
      if (ops[e].kind != 1) {
        ++e;
        continue;
      }
      Run r{e, e, 0, 0, 0, 0, false};
      for (; r.f < ops.size() && ops[r.f].kind == 1; ++r.f) {
        uint32_t u = r.U;
        for (uint32_t j = 0; j < ops[r.f].k; ++j)
          if ((ops[r.f].in_mask >> j) & 1) u |= 1u << op_bit(ops[r.f], j);
        if (popcount64(u) > 8) break;
        r.U = u;
      }
      for (size_t i = r.e; i < r.f; ++i) tab += 1u << ops[i].k;
      if (r.f - r.e >= 2) {
        r.k = uint32_t(popcount64(r.U));
        for (size_t i = r.e; i < r.f; ++i) {
          if (ops[i].in_mask == (1u << ops[i].k) - 1) r.own_static += 1u << ops[i].k;
          else ++r.n_dyn;
        }
        runs.push_back(r);
      }
      e = r.f;
    }
    uint32_t grp_dyn = 0;  // per-tile table entries the runs with outside qubits need anyway
    for (const Run& r : runs)
      if (r.n_dyn) grp_dyn = std::max(grp_dyn, 1u << r.k);
    uint32_t grp_static = 0;  // the same for the all-inside runs kept as heads
    for (const Run& r : runs) grp_static = std::max(grp_static, r.n_dyn ? 0u : 1u << r.k);
    const uint32_t cap = tab + 2 * std::max(grp_dyn, grp_static) <= kMtDynEntriesFour
                             ? kMtDynEntriesFour
                             : kMtTabEntries + 2 * 256;
    grp_static = 0;
    for (Run& r : runs) {
      if (!r.own_static) continue;
      const uint32_t t2 = tab - r.own_static + (1u << r.k);
      // (the later all-inside runs are counted as folded)
      const uint32_t g2 = std::max(grp_dyn, grp_static);
      r.fold = t2 <= kMtTabEntries && t2 + 2 * g2 <= cap;
      if (r.fold) tab = t2;
      else if (!r.n_dyn) grp_static = std::max(grp_static, 1u << r.k);
    }

    std::vector<MtOp> merged(ops.begin(), ops.begin() + pa.op_begin);
    auto own_table = [&](size_t i) {
      const Launch& l = *src[i - pa.op_begin];
      return add_table(l.mat.data(), 1u << l.m);
    };
    size_t ri = 0;
    uint32_t grp = 0;
    for (size_t e = pa.op_begin; e < ops.size();) {
      if (ri < runs.size() && runs[ri].e == e) {
        const Run& r = runs[ri++];
        MtOp head{};
        head.kind = 2;
        uint32_t rank[32] = {0};
        for (uint32_t b = 0; b < 32; ++b)
          if ((r.U >> b) & 1) {
            rank[b] = head.k;
            head.set_bit(head.k++, b);
          }
        const uint32_t entries = 1u << head.k;
        head.tab = kMtNoTable;
        if (r.fold) {
          // base table: entry x (the head's bits j set where bit j of x is) = product of
          // the in-tile members' phases
          std::vector<double> base(2 * entries);
          for (uint32_t x = 0; x < entries; ++x) {
            double pr = 1.0, pi = 0.0;
            for (size_t i = r.e; i < r.f; ++i) {
              if (ops[i].in_mask != (1u << ops[i].k) - 1) continue;
              const Launch& l = *src[i - pa.op_begin];
              uint32_t idx = 0;
              for (uint32_t j = 0; j < ops[i].k; ++j)
                idx |= ((x >> rank[op_bit(ops[i], j)]) & 1u) << j;
              const double a = l.mat[2 * idx], b = l.mat[2 * idx + 1];
              const double re = pr * a - pi * b;
              pi = pr * b + pi * a;
              pr = re;
            }
            base[2 * x] = pr;
            base[2 * x + 1] = pi;
          }
          head.tab = add_table(base.data(), entries);
          if (!r.n_dyn) {  // one diagonal op over the run's tile bits
            head.kind = 1;
            head.in_mask = entries - 1;
            merged.push_back(head);
            e = r.f;
            continue;
          }
        }
        grp = std::max(grp, entries);
        // members: byte j of the bits is the head's rank of qubit j's tile bit (in_mask
        // bit j set) or its state bit.  The kernel reads them as records in the pool, 8
        // words each, padded to a multiple of 4 (with copies of the last): the table
        // offset; per qubit j (bytes) the head's rank of its tile bit or 15; per qubit j
        // its state bit or 63 (a rank 15 or state bit 63 reads 0).
        const size_t hi = merged.size();
        merged.push_back(head);
        std::vector<uint32_t> rec;
        for (size_t i = r.e; i < r.f; ++i) {
          if (r.fold && ops[i].in_mask == (1u << ops[i].k) - 1) continue;
          MtOp m = ops[i];
          m.kind = 3;
          m.bits[0] = m.bits[1] = 0;
          uint32_t w[8] = {0, 0x0f0f0f0fu, 0x0f0f0f0fu, 0x3f3f3f3fu, 0x3f3f3f3fu, 0, 0, 0};
          for (uint32_t j = 0; j < m.k; ++j) {
            const uint32_t b = op_bit(ops[i], j);
            const bool in_tile = (m.in_mask >> j) & 1;
            m.set_bit(j, in_tile ? rank[b] : b);
            uint32_t& word = w[(in_tile ? 1 : 3) + j / 4];
            const uint32_t sh = 8 * (j % 4);
            word = (word & ~(0xffu << sh)) | ((in_tile ? rank[b] : b) << sh);
          }
          m.tab = w[0] = own_table(i);
          merged.push_back(m);
          rec.insert(rec.end(), w, w + 8);
        }
        while ((rec.size() / 8) % 4) rec.insert(rec.end(), rec.end() - 8, rec.end());
        merged[hi].mat = put(rec.data(), rec.size() * sizeof(uint32_t));
        merged[hi].members = uint32_t(merged.size() - hi - 1);
        e = r.f;
        continue;
      }
      merged.push_back(ops[e]);
      if (ops[e].kind == 1) merged.back().tab = own_table(e);
      ++e;
    }
    ops.swap(merged);
    pa.grp_entries = grp;
  }
  pa.n_ops = uint32_t(ops.size()) - pa.op_begin;
  pa.tab_entries = uint32_t(tabs.size() / 2);
  pa.tab_off = tabs.empty() ? 0 : put(tabs.data(), tabs.size() * sizeof(float));
  for (size_t e = ops.size(); e-- > pa.op_begin;)
    if (ops[e].kind != 0)
      ops[e].run = 1 + (e + 1 < ops.size() && ops[e + 1].kind != 0 ? ops[e + 1].run : 0);
  Launch L;
  L.kind = Launch::kMfmaTile;
  L.arena_off = out->passes.size();
  out->passes.push_back(pa);
  out->launches.push_back(std::move(L));
}

void build_mt_plan(const std::vector<Launch>& in, uint32_t n, uint32_t tb, MtPlan* out) {
  std::vector<size_t> cur;
  const uint64_t low = (1ull << kTileLowBits) - 1;
  uint64_t S = low;
  uint32_t cur_tab = 0;  // phase entries of the pass's diagonal ops so far
  auto close = [&]() {
    emit_mt_pass(in, cur, n, tb, S, out);
    cur.clear();
    S = low;
    cur_tab = 0;
  };
  auto is_diag = [](const Launch& l) { return l.kind == Launch::kDiag && l.m <= 6; };
  auto is_dense = [](const Launch& l) {
    return (l.kind == Launch::k1q || l.kind == Launch::kMq) && l.m <= 6;
  };
  const uint64_t all = n >= 64 ? ~0ull : (1ull << n) - 1;
  std::vector<char> taken(in.size(), 0);
  size_t first = 0;  // first op not yet placed
  while (first < in.size()) {
    if (!is_diag(in[first]) && !is_dense(in[first])) {  // runs on its own, in order
      out->order.push_back(first);
      out->launches.push_back(in[first++]);
      continue;
    }
    uint64_t left_dense = 0, left_diag = 0;  // qubits of the ops left behind so far
    // the scan ends once the ops left behind block every qubit, or after a bounded number
    // of them (a long circuit on few qubits would otherwise be rescanned whole per pass)
    size_t left = 0;
    for (size_t i = first; i < in.size() && left_dense != all && left < kMtLookahead; ++i) {
      if (taken[i]) continue;
      const Launch& l = in[i];
      if (!is_diag(l) && !is_dense(l)) break;  // no op moves ahead of it
      uint64_t Q = 0;
      for (uint32_t j = 0; j < l.m; ++j) Q |= 1ull << l.q[j];
      bool take;
      if (is_diag(l))
        take = !(Q & left_dense) && cur_tab + (1u << l.m) <= kMtTabEntries;
      else
        take = !(Q & (left_dense | left_diag)) && popcount64(S | Q) <= tb;
      if (!take) {
        (is_diag(l) ? left_diag : left_dense) |= Q;
        ++left;
        continue;
      }
      if (is_diag(l))
        cur_tab += 1u << l.m;
      else
        S |= Q;
      cur.push_back(i);
      taken[i] = 1;
    }
    close();
    while (first < in.size() && taken[first]) ++first;
  }
}

int plan_mfma_tiles(Program* p, int tb, PermTail* pt) {
  State& st = *p->st;
  if (p->precision != QSVH_PREC_MFMA_BF16X3)
    return set_err(st.ctx, kInvalidArg, "mfma_tiles runs split-BF16 groups (MFMA_BF16X3)");
  if (tb != 12) return set_err(st.ctx, kInvalidArg, "mfma_tiles must be 12");
  if (st.n < uint32_t(tb)) return plan_mfma(p);  // smaller than one tile
  MtPlan plan;
  if (pt && !pt->swaps.empty()) {  // launches before k, the bit swaps, the rest
    const std::vector<Launch> pre(p->launches.begin(), p->launches.begin() + pt->k);
    const std::vector<Launch> suf(p->launches.begin() + pt->k, p->launches.end());
    build_mt_plan(pre, st.n, uint32_t(tb), &plan);
    for (Launch& l : pt->swaps) plan.launches.push_back(std::move(l));
    pt->swaps.clear();
    const size_t o0 = plan.order.size();
    build_mt_plan(suf, st.n, uint32_t(tb), &plan);
    for (size_t i = o0; i < plan.order.size(); ++i) plan.order[i] += pt->k;
  } else {
    build_mt_plan(p->launches, st.n, uint32_t(tb), &plan);
  }
  // p->order is the identity here (tile_qubits excludes mfma_tiles): launch i is op i
  p->order = std::move(plan.order);
  p->launches = std::move(plan.launches);
  p->mt_passes = std::move(plan.passes);
  const std::vector<MtOp>& ops = plan.ops;
  const std::vector<uint8_t>& pool = plan.pool;
  if (ops.empty()) return kOk;
  auto upload = [&](const void* src, size_t bytes, void** dst) {
    hipError_t e = hipMalloc(dst, bytes ? bytes : 16);
    if (e != hipSuccess) return hip_err(st.ctx, e, "mfma tile buffers alloc");
    p->owned.push_back(*dst);
    if (bytes) e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
    return hip_err(st.ctx, e, "mfma tile buffers upload");
  };
  void* d = nullptr;
  int rc = upload(ops.data(), ops.size() * sizeof(MtOp), &d);
  if (rc != kOk) return rc;
  p->mt_ops = static_cast<MtOp*>(d);
  rc = upload(pool.data(), pool.size(), &p->mt_pool);
  return rc;
}

// Where each bit of the sub-index goes when l's matrix is a permutation of the basis that
// moves bits: every column c has one entry exactly 1 + 0i, at row sigma(c) = the sum of
// bit j of c shifted to tau[j], and exact zeros elsewhere (y[sigma(c)] = x[c]).
static bool qubit_perm_of(const Launch& l, uint32_t* tau) {
  if (l.kind != Launch::kMq || l.m < 1 || l.m > kMaxGateQubits) return false;
  const uint32_t d = 1u << l.m;
  std::vector<uint32_t> sigma(d);
  for (uint32_t c = 0; c < d; ++c) {
    uint32_t ones = 0;
    for (uint32_t r = 0; r < d; ++r) {
      const double re = l.mat[2 * (size_t(r) * d + c)], im = l.mat[2 * (size_t(r) * d + c) + 1];
      if (im != 0.0 || (re != 0.0 && re != 1.0)) return false;
      if (re == 1.0) {
        sigma[c] = r;
        ++ones;
      }
    }
    if (ones != 1) return false;
  }
  uint32_t used = 0;
  for (uint32_t j = 0; j < l.m; ++j) {
    const uint32_t v = sigma[1u << j];
    if (v == 0 || (v & (v - 1)) || (used & v)) return false;
    used |= v;
    tau[j] = uint32_t(__builtin_ctz(v));
  }
  for (uint32_t c = 0; c < d; ++c) {
    uint32_t want = 0;
    for (uint32_t j = 0; j < l.m; ++j) want |= ((c >> j) & 1u) << tau[j];
    if (sigma[c] != want) return false;
  }
  return true;
}

uint64_t strip_perm_suffix(std::vector<Launch>& ls, uint32_t n, std::vector<Launch>* out,
                           std::vector<uint32_t>* dest_out) {
  uint64_t first = ls.size();
  uint32_t tau[kMaxGateQubits];
  while (first > 0 && qubit_perm_of(ls[first - 1], tau)) --first;
  if (first == ls.size()) return first;
  // dest[b]: the state bit that the data of bit b ends at, the launches applied in order
  std::vector<uint32_t> dest(n);
  for (uint32_t b = 0; b < n; ++b) dest[b] = b;
  for (uint64_t i = first; i < ls.size(); ++i) {
    const Launch& l = ls[i];
    qubit_perm_of(l, tau);
    std::vector<uint32_t> g(n);
    for (uint32_t b = 0; b < n; ++b) g[b] = b;
    for (uint32_t j = 0; j < l.m; ++j) g[l.q[j]] = l.q[tau[j]];
    for (uint32_t b = 0; b < n; ++b) dest[b] = g[dest[b]];
  }
  ls.resize(first);
  if (dest_out) *dest_out = dest;
  // dest = beta . alpha with alpha, beta involutions: on each cycle c_0 -> c_1 -> ... of
  // length k, alpha(c_i) = c_{-i mod k} and beta(c_i) = c_{1-i mod k}
  std::vector<uint8_t> alpha(n), beta(n);
  std::vector<bool> seen(n, false);
  for (uint32_t b = 0; b < n; ++b) {
    if (seen[b]) continue;
    std::vector<uint32_t> cyc;
    for (uint32_t x = b; !seen[x]; x = dest[x]) {
      seen[x] = true;
      cyc.push_back(x);
    }
    const size_t k = cyc.size();
    for (size_t i = 0; i < k; ++i) {
      alpha[cyc[i]] = uint8_t(cyc[(k - i) % k]);
      beta[cyc[i]] = uint8_t(cyc[(k + 1 - i) % k]);
    }
  }
  for (const std::vector<uint8_t>* inv : {&alpha, &beta}) {
    bool ident = true;
    for (uint32_t b = 0; b < n; ++b) ident = ident && (*inv)[b] == b;
    if (ident) continue;
    Launch l;
    l.kind = Launch::kBitSwap;
    l.img = *inv;
    out->push_back(std::move(l));
  }
  return first;
}

uint64_t hoist_perm(std::vector<Launch>& body, const std::vector<uint32_t>& dest,
                    const std::function<size_t(const std::vector<Launch>&)>& count) {
  const size_t m = body.size();
  std::vector<Launch> moved = body;
  for (Launch& l : moved)
    for (uint32_t j = 0; j < l.m; ++j) l.q[j] = dest[l.q[j]];
  // qubits the body acts on from k on
  std::vector<uint64_t> sup(m + 1, 0);
  for (size_t k = m; k-- > 0;) sup[k] = sup[k + 1] | qubit_mask(body[k]);
  uint64_t best_k = m;
  size_t best = count(body);
  for (size_t k = 0; k < m; ++k) {
    if (k > 0 && popcount64(sup[k]) == popcount64(sup[k - 1])) continue;
    const std::vector<Launch> pre(body.begin(), body.begin() + k);
    const std::vector<Launch> suf(moved.begin() + k, moved.end());
    const size_t c = (k ? count(pre) : 0) + count(suf);
    if (c < best) {
      best = c;
      best_k = k;
    }
  }
  for (size_t k = best_k; k < m; ++k) body[k] = std::move(moved[k]);
  return best_k;
}

size_t mt_launch_count(const std::vector<Launch>& in, uint32_t n) {
  MtPlan plan;
  build_mt_plan(in, n, 12, &plan);
  return plan.launches.size();
}

void release_extra(Program* p) {
  for (void* b : p->owned) hipFree(b);
  p->owned.clear();
  p->mfma = nullptr;
  p->mt_ops = nullptr;
  p->mt_pool = nullptr;
  p->tile_ops = nullptr;
  p->tile_diags = nullptr;
  p->tile_pool = nullptr;
}

}  // namespace qsvh
