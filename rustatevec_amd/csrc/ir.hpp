// ir.hpp -- host-side circuit IR of the engine: gate library, validated circuits and the
// workload generators.  Mirrors crates/qsv-core/src/{gate.rs,circuit.rs,circuits.rs}
// (nez0b/rustatevec) behind the C ABI in include/qsvh.h.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace qsvh {

constexpr uint32_t kMaxGateQubits = 6;  // crates/qsv-core/src/backend/cpu.rs:34

// One gate application: row-major 2^m x 2^m complex matrix on the ordered qubits q[0..m)
// (crates/qsv-core/src/circuit.rs:12-27).
struct Op {
  uint32_t m = 0;
  uint32_t q[kMaxGateQubits] = {0, 0, 0, 0, 0, 0};
  std::vector<double> re, im;  // 4^m entries each

  uint32_t dim() const { return 1u << m; }
  // DenseGate::is_diagonal (crates/qsv-core/src/gate.rs:56-67): exact zero test.
  bool is_diagonal() const;
};

// One pass of a planned circuit (qsvh_plan_passes): ops [op_begin, op_begin + n_ops),
// applied to one LDS tile of the state bits in tile_mask per HBM round trip -- exactly
// (form 0: the ops are the circuit's own gates) or as an MFMA tile pass (form k > 0: the
// ops are groups fused up to k qubits, dense groups on split-BF16 matrix cores).
struct PlanPass {
  int form = 0;
  uint64_t op_begin = 0, n_ops = 0;
  uint64_t tile_mask = 0;
};

struct Circuit {
  uint32_t n = 0;
  std::vector<Op> ops;
  std::vector<PlanPass> plan;  // empty unless the circuit came from the pass planner
};

// Status values shared with include/qsvh.h.
enum Status : int {
  kOk = 0,
  kInvalidArg = 1,
  kInvalidQubit = 2,
  kRepeatedQubit = 3,
  kArity = 4,
  kTooWide = 5,
  kOom = 6,
  kHip = 7,
  kDtype = 8,
  kNoDevice = 9,
  kParse = 10,
};

// Circuit::push validation (crates/qsv-core/src/circuit.rs:54-78) as a status code.
int validate_qubits(uint32_t n, uint32_t m, const uint32_t* q);

// Gate library lookup (crates/qsv-core/src/gate.rs:96-180, plus u3/ccx/ccz).
// Fills `op` (m, matrix); qubits are left to the caller.
int make_gate(const std::string& name, const double* params, uint32_t n_params, Op* op);

int push(Circuit* c, uint32_t m, const uint32_t* q, const double* re, const double* im);
int push_gate(Circuit* c, const char* name, const uint32_t* q, const double* params,
              uint32_t n_params);

// SplitMix64 (crates/qsv-core/src/circuits.rs:12-47).
struct SplitMix64 {
  uint64_t state;
  explicit SplitMix64(uint64_t seed) : state(seed) {}
  uint64_t next_u64() {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  uint64_t below(uint64_t n) { return next_u64() % n; }
  double unit() { return static_cast<double>(next_u64() >> 11) / static_cast<double>(1ull << 53); }
  double angle() { return unit() * 6.28318530717958647692; }
};

// Generators.
Circuit gen_ghz(uint32_t n);                                    // circuits.rs:50-57
Circuit gen_ghz_chain(uint32_t n);                              // H(0); CX(i, i+1)
Circuit gen_qft(uint32_t n);                                    // circuits.rs:61-74
Circuit gen_random(uint32_t n, uint64_t n_gates, uint64_t seed);  // circuits.rs:79-113
Circuit gen_layered(uint32_t n, uint32_t depth, uint64_t seed);
Circuit gen_hea(uint32_t n, uint32_t depth, uint64_t seed);
Circuit gen_mix(uint32_t n, uint64_t n_gates, uint64_t seed);

// Greedy fusion grouping (crates/qsv-core/src/fusion.rs:39-99).  Each group is a run of
// consecutive ops [begin, end) and the qubit set it acts on: sorted when the group holds
// more than one op, the op's own order otherwise.
struct Group {
  uint64_t begin = 0, end = 0;
  uint32_t m = 0;
  uint32_t q[kMaxGateQubits] = {0, 0, 0, 0, 0, 0};
};
std::vector<Group> greedy_groups(const Circuit& c, uint32_t max_qubits);

// OpenQASM 3 / 2.0 program text -> circuit (qasm.cpp).  kParse with "line L: reason" in
// *err on failure.
int parse_qasm(const std::string& text, Circuit* out, std::string* err);

}  // namespace qsvh
