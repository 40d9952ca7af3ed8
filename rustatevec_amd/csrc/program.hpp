// program.hpp -- a circuit bound to a device state: the launch list the executor
// replays, the device-resident matrices it reads, and the optional captured graph.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

#include "engine.hpp"

namespace qsvh {

struct Program {
  State* st = nullptr;
  std::vector<Launch> launches;
  void* arena = nullptr;  // device matrices (state precision, interleaved re/im)
  size_t arena_bytes = 0;
  bool use_graph = false;
  hipGraphExec_t graph = nullptr;
  int precision = 0;
  void* mfma = nullptr;      // block matrices of the MFMA group launches
  std::vector<void*> owned;  // further device buffers freed with the program
  // Tile passes: Launch::kTile launches index `tile_passes` through arena_off.
  std::vector<TilePassArg> tile_passes;
  TileOp* tile_ops = nullptr;
  TileDiag* tile_diags = nullptr;
  void* tile_pool = nullptr;
  // MFMA tile passes: Launch::kMfmaTile launches index `mt_passes` through arena_off.
  std::vector<MtPassArg> mt_passes;
  MtOp* mt_ops = nullptr;
  void* mt_pool = nullptr;
  // Original op index of each applied gate, in execution order (tile passes reorder
  // gates that act on disjoint qubits).
  std::vector<uint64_t> order;
};

int set_err(Ctx* ctx, int code, const std::string& msg);
int hip_err(Ctx* ctx, hipError_t e, const char* where);

// Lowered tile-pass ops of one or more circuits, with their matrices (f64, interleaved).
struct TileBuilder {
  size_t pass_begin = 0;  // first op of the pass being built
  std::vector<TileOp> ops;
  std::vector<TileDiag> diags;
  std::vector<double> pool;  // complex entries, interleaved
  uint64_t add_matrix(const double* g, size_t n_complex);
  // Lower launch l into ops for a pass whose tile bits are `tmask` (pos[q] = index of
  // state bit q among the tile bits).
  void lower(const Launch& l, uint64_t tmask, const uint32_t* pos);
  void lower_block(const double* sub, uint32_t nin, const uint32_t* in_j, const uint32_t* q,
                   const uint32_t* pos, uint64_t cmask, uint64_t cval);
  // Consecutive 1-qubit ops whose target and condition bits fit in 3 tile bits are
  // gathered under one register-block op (kind 3), whose members follow it.
  int64_t open_block = -1;  // index of the open register-block op, or -1
  uint32_t block_bits = 0;  // tile-local bits of the open block
  void emit(const TileOp& op);
  void close_block();
  // Diagonal gates and 1-qubit ops on lane bits (< 4) or register bits are gathered, in
  // order, under one lane-run op (kind 4) whose members follow it.  A lane run may choose
  // its own register bits (run_lk of them; 0: the tile-local bits from kTileLaneBits up):
  // a 1-qubit op on a bit outside them joins while there is room, or as a lane member when
  // its lane position is below 4 and the coming diagonal gates do not lean on that bit
  // (`heavy`); otherwise it starts a new run.  Members are lowered in tile-local bits and
  // rewritten in the run's coordinates (lane bits 0..8, register bits from kTileLaneBits
  // up) when the run closes.
  int64_t open_run = -1;  // index of the open lane-run op, or -1
  bool free_runs = true;   // runs may choose register bits and get kind 5, 6 members (the
                           // kernel runs run_ops<..., true>)
  uint32_t run_lk = 0;     // register bits the current pass's runs choose (0: fixed)
  uint32_t tile_bits = 0;  // tile bits of the current pass
  uint32_t run_reg[4] = {0, 0, 0, 0};  // the open run's register bits (tile-local)
  uint32_t run_nreg = 0;
  uint64_t heavy = 0;  // tile-local bits the diagonal gates after the current gate lean on
  // the phases of each diagonal gate of <= 2 qubits (parallel to diags), lowered to its
  // phase table when its run closes
  std::vector<std::vector<double>> diag_phases;
  void run_member(const TileOp& m);
  void close_run();
  void finish_small_diag(TileDiag& d, const std::vector<double>& mat);
  int ctrl_phase_bit(const TileDiag& d) const;
  bool ctrl_pair_bits(const TileDiag& d, uint32_t* x, uint32_t* y) const;
};

// Tile schedule of one circuit's classified launches on n qubits with tb tile bits
// (tb <= n): appends lowered ops to `bld`, passes to `passes`, the resulting launch list
// to `out` (kTile launches index `passes` through arena_off; launches that cannot be
// tiled -- dense gates on more than 3 qubits -- are kept as they are) and the applied
// order of the input launches to `order`.
void schedule_tiles(const std::vector<Launch>& in, uint32_t n, uint32_t tb, TileBuilder& bld,
                    std::vector<TilePassArg>& passes, std::vector<Launch>& out,
                    std::vector<uint64_t>& order);

// A trailing qubit permutation taken out of a circuit and the point k at which its bit
// swaps run: before launch k of the rest, whose launches from k on act on the permuted
// qubits (the data of state bit b is at bit dest[b] after the swaps).  The same values
// result: each moved gate does its arithmetic on the same amplitudes, at other indices.
struct PermTail {
  std::vector<Launch> swaps;   // the kBitSwap launches; emptied once a planner placed them
  std::vector<uint32_t> dest;
  uint64_t k = 0;
};

// Passes that rewrite the plain launch list (plan.cpp).  A tile planner given a PermTail
// with swaps schedules the launches before k and from k separately, the swaps between
// (k at the end: the swaps after the passes).
int plan_tiles(Program* p, int tile_qubits, PermTail* pt = nullptr);
void plan_subspace(Program* p);
int plan_mfma(Program* p);
// Runs of fused groups as MFMA tile passes of 2^tb amplitudes (split-BF16 precision).
int plan_mfma_tiles(Program* p, int tb, PermTail* pt = nullptr);
// Trailing qubit permutations (dense launches whose matrix only moves the bits of the
// sub-index, such as the SWAP network that ends a QFT, alone or fused into groups) are
// removed from `ls`; their product, a permutation of the state bits (dest, if given), is
// appended to `out` as at most two kBitSwap launches.  Returns the index of the first
// launch removed.
uint64_t strip_perm_suffix(std::vector<Launch>& ls, uint32_t n, std::vector<Launch>* out,
                           std::vector<uint32_t>* dest = nullptr);
// The point at which a trailing permutation (dest) runs among `body`: the end, the start,
// or a point where the qubits the rest of the body acts on shrink -- whichever gives the
// fewest launches by `count` (prefix plus moved suffix; the end on ties).  The body's
// launches from the returned k on are moved to the permuted qubits.
uint64_t hoist_perm(std::vector<Launch>& body, const std::vector<uint32_t>& dest,
                    const std::function<size_t(const std::vector<Launch>&)>& count);
// Launches of the MFMA tile form of `in` (12-bit tiles).
size_t mt_launch_count(const std::vector<Launch>& in, uint32_t n);
// The host side of plan_mfma_tiles: the passes, their ops and matrix pool, and the launch
// sequence (kMfmaTile launches and the launches no pass holds).
struct MtPlan {
  std::vector<MtPassArg> passes;
  std::vector<MtOp> ops;
  std::vector<uint8_t> pool;
  std::vector<Launch> launches;
  std::vector<uint64_t> order;  // indices into `in` in the order the plan applies them
};
void build_mt_plan(const std::vector<Launch>& in, uint32_t n, uint32_t tb, MtPlan* out);
// One MFMA tile pass of the launches `cur` (in order) on the tile of the bits of S filled up
// to tb bits with the lowest others; appended to `out`.
void emit_mt_pass(const std::vector<Launch>& in, const std::vector<size_t>& cur, uint32_t n,
                  uint32_t tb, uint64_t S, MtPlan* out);
// One exact tile pass of the launches `take` (in order; their action qubits in tmask, a set
// of tb bits holding bits 0..3), appended to bld / passes / out / order.
void lower_exact_pass(const std::vector<Launch>& in, const std::vector<uint64_t>& take,
                      uint32_t n, uint32_t tb, uint64_t tmask, TileBuilder& bld,
                      std::vector<TilePassArg>& passes, std::vector<Launch>& out,
                      std::vector<uint64_t>& order);
void release_extra(Program* p);
// The launches of a circuit's recorded pass plan (passes.cpp).
int plan_planned_passes(Program* p, const Circuit& c);

}  // namespace qsvh
