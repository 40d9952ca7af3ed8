// engine.hpp -- internal types shared by the C ABI (api.cpp) and the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "ir.hpp"

namespace qsvh {

// Launch-shape knobs; defaults chosen by measurement on MI355X, overridable per context
// (the autotuner writes them).
struct Tuning {
  int stream_block = 256;      // threads per block for the streaming kernels
  int stream_max_blocks = 0;   // 0: one thread per work item; >0: grid-stride cap
  int unroll = 1;              // independent vector items per thread per iteration
  int nontemporal = 2;         // 1: non-temporal (streaming) loads and stores; 2: where
                               //    they pay (streaming kernels, states beyond the MALL)
  int mfma_blocks = 0;         // workgroups of the persistent MFMA group kernels; 0: by
                               // form (measured: 1 per CU for FP32 at k = 4, else 2 per
                               // CU; the HBM rate drops with more tiles in flight)
  // workgroups of the tile-pass kernels (exact and MFMA): 0, one per tile; n > 0: n
  // workgroups looping over the tiles.  One per tile measured 6-10% faster than a
  // resident-sized loop (whose next tile's loads wait out the previous tile's stores:
  // the vector-memory wait counts stores too).
  int tile_blocks = 0;
  int mfma_layout = 0;         // FP32 group kernel's LDS layout: 0 by the group's qubits,
                               // 1 row-major, 2 column-major
  int contig = 0;              // 1: a thread's unrolled items are adjacent (else strided)
  // 1: blocks b and b + 1 of a streaming launch take neighbouring address ranges of one
  // XCD (the hardware deals blocks to the 8 XCDs round robin, b -> XCD b % 8).
  int xcd = 1;
  // 1: dense 2- and 3-qubit gates with one 16-byte chunk per lane (k_mq_lane); 0: one
  // block of 2^m amplitudes per thread (k_mq_reg)
  int mq_lane = 1;
  // workgroups of the LDS-table diagonal kernel (k_diag, m >= 3): 0 one thread per work
  // item (the phase table staged by every workgroup); n > 0 a grid-stride loop over n
  int diag_blocks = 0;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  Tuning tune;
  int cus = 0;
  // Reduction scratch: per-block f64 partials and the final value.
  double* d_partials = nullptr;
  double* d_result = nullptr;
  double* h_result = nullptr;  // pinned
  static constexpr int kMaxPartials = 4096;
  // Device address of the allocation-free sampler's scratch on this device (null when the
  // lookup failed; sample_scratch_err holds why).
  double* sample_scratch = nullptr;
  hipError_t sample_scratch_err = hipSuccess;
};

// A state vector in HBM: 2^n complex amplitudes, interleaved (re, im) pairs of T --
// amplitude i at amp[2 i], amp[2 i + 1].
struct State {
  Ctx* ctx = nullptr;
  uint32_t n = 0;
  int dtype = 0;  // 0 = c128 (f64 pairs), 1 = c64 (f32 pairs)
  void* amp = nullptr;
  uint64_t dim() const { return 1ull << n; }
  size_t elem() const { return dtype == 0 ? sizeof(double) : sizeof(float); }  // one part
  size_t bytes() const { return size_t(dim()) * 2 * elem(); }
};

// Non-temporal accesses: knob 0 never, 1 always, 2 (auto) where they were measured to pay
// on MI355X -- for a kernel that touches each cache line once (`streams`) and a state
// larger than the 256 MiB infinity cache.  A smaller state stays cache-resident between
// gates, which non-temporal accesses forfeit; the pair-per-thread 1q kernel where its pair
// halves share 128-byte lines, the in-chunk 1q kernel and the LDS-table diagonal ran slower
// with them.
inline bool nontemporal_for(const State& st, const Tuning& t, bool streams) {
  if (t.nontemporal == 2) return streams && st.bytes() > (size_t(256) << 20);
  return t.nontemporal != 0;
}

// ---- gate descriptors passed by value to the kernels -------------------------------
// Qubit index data for an m-qubit gate: sorted positions (for insert_zero_bits) and
// the offset of each internal sub-index s, scatter_bits(s, q) (layout.rs:54-60).
struct MqIndex {
  uint32_t m;
  uint32_t sorted[kMaxGateQubits];
  uint64_t off[64];
};

// One kernel launch of the plain executor.  kCtrl1q: 2x2 matrix `mat` on q[0], applied
// where the state bits q[1..m) equal `fix`'s bits (a controlled 1-qubit gate on its control
// subspace); kDiagSub: the phase mat[0..1] on the amplitudes whose bits q[0..m) equal
// `fix`'s (the one non-unit entry of a diagonal gate); kBitSwap: the amplitude at index i
// moved to the index whose bit img[b] is bit b of i, img an involution of the state bits.
struct Launch {
  enum Kind {
    k1q, kDiag, kMq, kFusedMfma, kTile, kCtrl1q, kDiagSub, kMfmaTile, kSwapSub, kBitSwap
  } kind;
  uint32_t m = 0;
  uint32_t q[kMaxGateQubits] = {0, 0, 0, 0, 0, 0};
  // matrix in f64 (interleaved re, im), row-major; for diagonal launches the 2^m phases.
  std::vector<double> mat;
  uint64_t arena_off = 0;  // element offset of the matrix in the device arena
  uint64_t mfma_off = 0;   // byte offset of the block matrix in the program's MFMA buffer
  uint64_t fix = 0;        // kCtrl1q / kDiagSub: required values of the fixed state bits
  std::vector<uint8_t> img;  // kBitSwap: image of each state bit
};

// The kernel argument of a kBitSwap launch: the involution's bit pairs split into those
// among a tile's bits (the low bits 0..lo-1 and their partners) and the others.
struct BitSwapArg {
  uint32_t lo = 0, n_hi = 0;  // low tile bits; tile bits above them (partners of low bits)
  uint32_t hi_bit[8] = {};    // state bits of tile position bits lo, lo + 1, ... (ascending)
  uint32_t n_sp = 0;          // pairs among the tile bits, as tile position bits
  uint32_t sp_a[8] = {}, sp_b[8] = {};
  uint32_t n_op = 0;          // pairs among the other bits, as state bits (a < b)
  uint32_t op_a[32] = {}, op_b[32] = {};
  uint32_t n_rest = 0;        // the other bits without a partner (32-bit entries: read by
  uint32_t rest_bit[64] = {}; // scalar loads)
  uint64_t n_items = 0;       // pairs of tiles: 2^n_rest (2^p + (2^p - 1) 2^(p-1)), p = n_op
};

// ---- launchers (kernels_plain.hip) ---------------------------------------------------
hipError_t launch_init_basis(const State& st, uint64_t basis, hipStream_t s);
hipError_t launch_1q(const State& st, uint32_t q, const double* g /*8: re,im interleaved*/,
                     const Tuning& t, hipStream_t s);
hipError_t launch_diag(const State& st, uint32_t m, const uint32_t* q,
                       const double* d /*2*2^m interleaved*/, const Tuning& t,
                       hipStream_t s);
// m-qubit dense gate; `dmat` is the device copy of the matrix in the state's precision
// (interleaved re, im), needed for m >= 4 (smaller gates travel as kernel arguments).
hipError_t launch_mq(const State& st, uint32_t m, const uint32_t* q, const double* g,
                     const void* dmat, const Tuning& t, hipStream_t s);
// Subspace launches (kCtrl1q, kDiagSub; see Launch).
hipError_t launch_ctrl1q(const State& st, uint32_t m, const uint32_t* q, uint64_t fix,
                         const double* g, const Tuning& t, hipStream_t s);
hipError_t launch_diag_sub(const State& st, uint32_t m, const uint32_t* q, uint64_t fix,
                           const double* phase, const Tuning& t, hipStream_t s);
// kSwapSub: SWAP(q[0], q[1]) on its |01>, |10> amplitudes only.
hipError_t launch_swap_sub(const State& st, const uint32_t* q, const Tuning& t, hipStream_t s);
// kBitSwap (img: the involution's image of each of the state's n bits; no launch when it
// is the identity, hipErrorInvalidValue when it is not an involution).
hipError_t launch_bitswap(const State& st, const uint8_t* img, const Tuning& t, hipStream_t s);
hipError_t launch_probs(const State& st, void* out, hipStream_t s);
// f64 sum of |psi_i|^2 * sign(i & mask) (mask 0: the norm) into *d_result.
hipError_t launch_reduce(const State& st, uint64_t mask, double* d_partials, int max_partials,
                         double* d_result, hipStream_t s);
hipError_t launch_copy(const State& src, void* dst, const Tuning& t, hipStream_t s);
// Synchronous copies between a state and host SoA arrays re[dim], im[dim] of the state's
// precision (api.cpp).
hipError_t download_split(const State& s, void* re, void* im);
hipError_t upload_join(const State& s, const void* re, const void* im);
// kernels_util.hip: amplitudes [first, first + count) of the state to / from separate device
// re / im arrays.
hipError_t launch_split(const State& st, uint64_t first, uint64_t count, void* re, void* im,
                        hipStream_t s);
hipError_t launch_join(const State& st, uint64_t first, uint64_t count, const void* re,
                       const void* im, hipStream_t s);

// ---- kernels_mfma.hip ------------------------------------------------------------------
// Fused k-qubit group (k = 4..6, qubits `q` ascending) on a complex64 state of >= 11
// qubits; `mats` is the 2D x 2D real block matrix in f32 (precision 1) or bf16 hi, lo
// (precision 2).
hipError_t launch_fused_mfma(const State& st, uint32_t k, const uint32_t* q, const void* mats,
                             int precision, const Tuning& t, hipStream_t s);

// MFMA tile passes: a run of fused groups applied to a tile of 2^12 amplitudes (complex64)
// held in LDS as f32, one HBM round trip for the run.  Dense groups
// (k = 4..6 qubits, all in the tile) are split-BF16x3 MFMA products; diagonal groups (up to
// 6 qubits anywhere: the qubits outside the tile are fixed per tile) scale each amplitude.
// Bit positions are bytes packed in dword words (byte j of `bits` at bits[j / 4] >> 8 (j % 4)),
// so that a kernel reads every field with scalar (uniform) loads.
// Kinds: 0 a dense group; 1 a diagonal group applied on its own; 2 the head of a merged run
// of diagonal groups -- `members` kind-3 ops follow it, whose in-tile qubits span at most 8
// tile bits (the head's `bits`, k of them): per tile, the product of the head's base table
// (the run's ops with every qubit in the tile, composed on the host) and the members'
// phases over those bits is built as one table, then applied once.
struct MtOp {
  uint32_t kind;
  uint32_t k;         // dense: group qubits (4..6); diagonal / member: m (1..6); head: its
                      // in-tile bits (<= 8)
  uint32_t in_mask;   // diagonal / member: bit j set when qubit j is in the tile
  uint32_t run;       // kinds 1..3: the consecutive non-dense ops from this one to the
                      // pass's next dense op or end (this one included)
  uint32_t tab;       // diagonal / member: first entry of its phases in the pass's table;
                      // head: its base table (the product of the run's ops with every
                      // qubit in the tile) or kMtNoTable
  uint32_t members;   // head: the member ops that follow it
  uint32_t bits[2];   // byte j: dense: tile-local bit j of the group, ascending; diagonal:
                      // tile-local bit of qubit j (in the tile) or its state bit (outside);
                      // member: the head's index of qubit j's tile bit, or its state bit;
                      // head: its tile-local bits, ascending
  uint32_t obits[3];  // byte i: dense: the tb - k other tile-local bits, ascending
  uint64_t mat;       // dense: byte offset in the pool of the bf16 block matrix hi, then lo
                      // (2D x 2D each); head: of its members' records (8 words each: the
                      // table offset, per qubit the head's rank of its tile bit or 15, per
                      // qubit its state bit or 63; padded to a multiple of 4 records)
  void set_bit(uint32_t j, uint32_t b) { bits[j >> 2] |= (b & 0xffu) << (8 * (j & 3)); }
  void set_obit(uint32_t i, uint32_t b) { obits[i >> 2] |= (b & 0xffu) << (8 * (i & 3)); }
};
constexpr uint32_t kMtTabEntries = 2048;  // phase entries of a pass's diagonal ops (LDS)
constexpr uint32_t kMtNoTable = 0xffffffffu;  // a head op without a base table
struct MtPassArg {
  uint32_t tb;
  uint32_t kmax;     // widest dense group of the pass (4..6; 4 when it has none)
  uint32_t ts[13];   // state bits of the tile, ascending (tile-local bit j = state bit ts[j])
  uint32_t op_begin, n_ops;
  uint32_t tab_entries;  // the pass's phase table: complex f32 entries at pool + tab_off
  uint32_t grp_entries;  // entries of a merged run's per-tile table (2^ its widest head's
                         // bits; 0 without heads), two of them after the phase table in LDS
  uint64_t tab_off;
};
// Phase-table entries (pass table plus the two per-tile tables) up to which an MFMA tile
// pass keeps four workgroups per CU: 32 KiB of tile plus 8 KiB of tables each.
constexpr uint32_t kMtDynEntriesFour = 1024;
hipError_t launch_mfma_tiles(const State& st, const MtPassArg& pa, const MtOp* ops,
                             const void* pool, const Tuning& t, hipStream_t s);

// ---- kernels_tile.hip ------------------------------------------------------------------
// One op of a tile pass, applied to the tile held in LDS; it runs on a tile only when
// (tile base & cmask) == cval (controls outside the tile).
//  kind 1: a run of n_diag diagonal gates starting at diags[off].
//  kind 0 (dense) and 2 (permutation): a block on m = 1..3 tile-local action bits lq[0..m)
//  (gate order), applied to the amplitudes whose tile-local condition bits (controls
//  inside the tile) equal lcval; ins[0..nins) are the action and condition bits ascending.
//  Dense: row-major 4^m complex matrix at pool[off]; `real`: for m >= 2, 1 when every
//  entry has a zero imaginary part; for m = 1 the pair variant (0 general, 1 real, 2
//  anti-diagonal, 3 swap).  Permutation (m >= 2): output row r is pool[off + r] times
//  input perm[r]; bit r of pid set: that coefficient is exactly 1 + 0i (a move).
//  kind 3: a register block on bits lq[0..m); its n_diag members (1-qubit dense ops whose
//  lq[0] is an index into the block's bits and insm[0] / lcval a condition on the block's
//  sub-index) follow it in the op list.
//  kind 4: a lane run; its n_diag members follow it (kind 1 diagonal members, kind 0
//  1-qubit members with a condition insm[0] / lcval on the run index u = tid + k NT) in the
//  run's coordinates: u's bits below kTileLaneBits are the lane's, the ones above the
//  register index k's.  m = 0: u is the tile-local index; m = 1: register bit j < lq[0] of
//  k is tile-local bit perm[j] (ascending) and the lane's bits fill the others in order.
//  kind 5 (lane-run member): n_diag phases controlled by one outside bit on the run index
//  bit lq[0]: gate d (diags[off + d], outside bit oq) multiplies the amplitudes whose bit
//  lq[0] is 1 by pool[ins[0] + d] in tiles whose base has bit oq set.
//  kind 6 (lane-run member): n_diag phases on two run index bits: gate d multiplies the
//  amplitudes whose bits lq[0] and diags[off + d].oq are both 1 by pool[ins[0] + d].
struct TileOp {
  uint32_t kind;
  uint32_t m;
  uint32_t lq[3];
  uint32_t n_diag;
  uint32_t nins;
  uint32_t ins[3];
  uint32_t insm[3];  // ~((1 << ins[j]) - 1) for j < nins, else 0: x + (x & insm) inserts a 0
  uint32_t lcval;
  uint32_t real;
  uint32_t pid;
  uint8_t perm[8];
  uint64_t cmask, cval;
  uint64_t off;
};
// Tile kernels run 2^kTileLaneBits threads; tile-local index e = tid + k * 2^kTileLaneBits,
// so tile-local bits below kTileLaneBits are lane bits and the others register bits.
constexpr uint32_t kTileLaneBits = 9;
constexpr uint32_t kOutside = 0xffffffffu;
// A diagonal gate inside a diag run, lowered for one pass (its first 8 words are what a
// kernel reads per gate: one scalar load).
//  m <= 2: the gate's bits are sorted into outside bits (state bits oq0 = oq & 63, oq1 =
//  oq >> 8 of the tile base: so = bit oq0 | bit oq1 << 1, nout of them) and up to two
//  "index bits" forming c (0..3), each absent, a lane bit (tid bit lp0 / lp1) or a
//  register bit (k bit p); code = variant | lp0 << 8 | lp1 << 12 | nout << 16, variant =
//  kind0 | kind1 << 4 with kind 0 absent, 1 lane, 2 + p register bit p.  The phase of an
//  amplitude is pool[tab + 4 so + c]; bit 4 so + c of idw: it is exactly 1 + 0i.
//  m > 2: phases pool[tab + s], s = gather of bits q[j] (tile-local bit loc[j] or, for
//  kOutside, the base bit).
struct TileDiag {
  uint32_t m;
  uint32_t code;
  uint32_t oq;
  uint32_t tab;
  uint32_t idw;
  uint32_t pad[3];
  uint32_t q[kMaxGateQubits];
  uint32_t loc[kMaxGateQubits];
};
struct TilePassArg {
  uint32_t tb;         // tile bits
  uint32_t ts[16];     // state bit of tile-local bit j: bits 0..3 first, then the scheduler's order
  uint32_t tsort[16];  // the same state bits ascending (insertion of a tile's base)
  uint32_t op_begin, n_ops;
};
hipError_t launch_tile(const State& st, const TilePassArg& pa, const TileOp* ops,
                       const TileDiag* diags, const void* pool, int max_blocks, hipStream_t s);

// ---- kernels_batch.hip -----------------------------------------------------------------
constexpr uint32_t kBatchInit = 1;    // build |0...0> in LDS instead of loading the tile
constexpr uint32_t kBatchExpZ0 = 2;   // write the tile's partial <Z_0> to partials
// One pass of one circuit of a batch: its tiles are work items
// [tile_begin, tile_begin + n_tiles) of the launch.
struct BatchPass {
  uint64_t state_off;  // amplitude offset of the circuit's state in the pool
  uint64_t tile_begin;
  uint64_t n_tiles;
  uint32_t n, tb;
  uint32_t ts[16], tsort[16];  // as in TilePassArg
  uint32_t op_begin, n_ops;
  uint32_t flags;
  uint32_t part_off;  // first partial slot (kBatchExpZ0)
};
hipError_t launch_batch_pass(int dtype, const BatchPass* passes, uint32_t n_passes,
                             uint64_t total_work, unsigned long long* counter, void* amp_pool,
                             const TileOp* ops, const TileDiag* diags,
                             const void* pool, double* partials, int blocks, hipStream_t s);
hipError_t launch_batch_reduce(const uint32_t* part_off, const uint32_t* n_part,
                               uint32_t n_circ, const double* partials, double* out,
                               hipStream_t s);

// ---- kernels_sample.hip ----------------------------------------------------------------
// Inverse-CDF sampling of `shots` outcomes with SplitMix64(seed) draws.  Buffers: local
// (dim doubles), totals (sample_chunks(dim)), offset (sample_chunks(dim) + 1), bsum
// (sample_blocks(dim)), out (shots).
uint64_t sample_chunks(uint64_t dim);
uint64_t sample_blocks(uint64_t dim);
// The allocation-free sampler's scratch: a static device array for states of up to
// kNoAllocMaxQubits qubits (sample_noalloc_doubles() doubles), its address on `device` (the
// current device) looked up once per device.
constexpr uint32_t kNoAllocMaxQubits = 22;
uint64_t sample_noalloc_doubles();
hipError_t sample_noalloc_scratch(int device, double** ptr);
hipError_t launch_sample(const State& st, uint64_t shots, uint64_t seed, double* local,
                         double* totals, double* offset, double* bsum, uint64_t* out,
                         hipStream_t s);

// ---- kernels_util.hip ------------------------------------------------------------------
// Random normalised start state (counter-based SplitMix64 + Box-Muller); `d_norm` receives
// the pre-normalisation <psi|psi>.
hipError_t launch_init_random(const State& st, uint64_t seed, double* part, int max_part,
                              double* d_norm, hipStream_t s);
// out[0..5) = max |a-b|, Re<a|b>, Im<a|b>, <a|a>, <b|b>; `part` holds 5 * 512 doubles.
hipError_t launch_compare(const State& a, const State& b, double* part, double* out,
                          hipStream_t s);

// ---- fusion (kernels_fused.hip) ------------------------------------------------------
// Compose every group's product in f64 on the device.  `ops` lists, per group, the
// member gates (local qubit positions + matrices).  Output: row-major 4^k entries per
// group at the given offsets (interleaved re, im).
struct ComposeGate {
  uint32_t m;
  uint32_t local[kMaxGateQubits];  // positions within the group's sorted qubit list
  uint64_t mat_off;                // offset (complex entries) into the packed matrices
};
struct ComposeGroup {
  uint32_t k;
  uint32_t first_gate, n_gates;
  uint64_t out_off;  // offset (complex entries) into the output
};
int compose_groups(Ctx* ctx, const std::vector<ComposeGroup>& groups,
                   const std::vector<ComposeGate>& gates, const std::vector<double>& mats,
                   std::vector<double>* out);

}  // namespace qsvh
