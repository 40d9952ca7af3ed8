/*
 * qsv_oracle.h -- CPU restatement of the qsv statevector algorithm, used ONLY as test
 * infrastructure: the checker that the HIP engine's results are compared against, and
 * the CPU baseline timed beside the GPU. Product code never links or calls it.
 *
 * Every function cites the file:line of the reference (nez0b/rustatevec) it restates.
 * Conventions (crates/qsv-core/src/gate.rs:8-12): a gate on the ordered qubit list qs has
 * internal row/column index whose bit j corresponds to global qubit qs[j]; controlled
 * gates put the control first.  States are Structure-of-Arrays: re[2^n], im[2^n]
 * (crates/qsv-core/src/state/mod.rs:25-29).
 */
#ifndef QSV_ORACLE_H
#define QSV_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- SplitMix64 (crates/qsv-core/src/circuits.rs:12-47) ---- */
uint64_t or_sm64_next(uint64_t* state);
uint64_t or_sm64_below(uint64_t* state, uint64_t n);
double or_sm64_unit(uint64_t* state);
double or_sm64_angle(uint64_t* state);

/* ---- index arithmetic (crates/qsv-core/src/state/layout.rs:22-87) ---- */
uint64_t or_insert_zero_bit(uint64_t index, uint32_t bit);
uint64_t or_gather_bits(uint64_t index, const uint32_t* qs, uint32_t m);
uint64_t or_scatter_bits(uint64_t sub, const uint32_t* qs, uint32_t m);
uint64_t or_insert_zero_bits(uint64_t index, const uint32_t* sorted, uint32_t m);

/* ---- gate library (crates/qsv-core/src/gate.rs:70-180) ----
 * Gate kinds.  U3, CCX and CCZ are not in the reference's gate library; CCX follows the
 * test helper at crates/qsv-core/tests/equivalence.rs:182-193, U3 the OpenQASM 3 U gate. */
enum {
  OR_H = 0, OR_X, OR_Y, OR_Z, OR_S, OR_T, OR_SX,
  OR_RX, OR_RY, OR_RZ, OR_P,
  OR_CX, OR_CZ, OR_CP, OR_SWAP, OR_RZZ,
  OR_U3, OR_CCX, OR_CCZ,
  OR_NKINDS
};
/* Arity of a gate kind, or 0 if unknown. */
uint32_t or_gate_arity(int kind);
/* Write the row-major 2^m x 2^m matrix of `kind` (params: up to 3 angles) into re/im
 * (capacity 64 entries).  Returns m, or 0 for an unknown kind. */
uint32_t or_gate_matrix(int kind, const double* params, double* re, double* im);
/* DenseGate::is_diagonal (crates/qsv-core/src/gate.rs:56-67): exact zero test. */
int or_is_diagonal(uint32_t m, const double* re, const double* im);

/* ---- circuit generators (crates/qsv-core/src/circuits.rs:50-113 and the workloads the
 * engine is measured on).  A generated circuit is a list of "gate records":
 * kind, q[3], param[3].  Each function writes at most `cap` records and returns the
 * number it needs (call with cap = 0 to size the buffers). ---- */
size_t or_gen_ghz_star(uint32_t n, int32_t* kind, uint32_t* q, double* p, size_t cap);
size_t or_gen_ghz_chain(uint32_t n, int32_t* kind, uint32_t* q, double* p, size_t cap);
size_t or_gen_qft(uint32_t n, int32_t* kind, uint32_t* q, double* p, size_t cap);
size_t or_gen_random(uint32_t n, size_t n_gates, uint64_t seed,
                     int32_t* kind, uint32_t* q, double* p, size_t cap);
size_t or_gen_layered(uint32_t n, uint32_t depth, uint64_t seed,
                      int32_t* kind, uint32_t* q, double* p, size_t cap);
size_t or_gen_hea(uint32_t n, uint32_t depth, uint64_t seed,
                  int32_t* kind, uint32_t* q, double* p, size_t cap);
size_t or_gen_mix(uint32_t n, size_t n_gates, uint64_t seed,
                  int32_t* kind, uint32_t* q, double* p, size_t cap);

/* ---- backends.  A circuit is passed flat: op i has arity[i] qubits starting at
 * qubits[qoff[i]] and a row-major matrix starting at mre/mim[moff[i]]. ---- */

/* RefBackend::apply (crates/qsv-core/src/backend/reference.rs:35-62): out-of-place
 * gather/scatter per output amplitude.  Returns 0 or a negative error. */
int or_ref_apply(uint32_t n, double* re, double* im, uint32_t m, const uint32_t* qs,
                 const double* gre, const double* gim);

/* CpuBackend::apply (crates/qsv-core/src/backend/cpu.rs:281-290) with its diagonal,
 * nested-block 1q and block mq kernels, thresholds and chunking (cpu.rs:44-247).
 * `parallel` selects CpuBackend::parallel() (OpenMP over all threads) vs serial(). */
int or_cpu_apply_f64(uint32_t n, double* re, double* im, uint32_t m, const uint32_t* qs,
                     const double* gre, const double* gim, int parallel);
int or_cpu_apply_f32(uint32_t n, float* re, float* im, uint32_t m, const uint32_t* qs,
                     const double* gre, const double* gim, int parallel);

/* StateVector::zeros + first_touch + basis (crates/qsv-core/src/state/mod.rs:33-73). */
void or_init_basis_f64(uint32_t n, double* re, double* im, uint64_t basis, int parallel);
void or_init_basis_f32(uint32_t n, float* re, float* im, uint64_t basis, int parallel);

/* Backend::execute default loop (crates/qsv-core/src/backend/mod.rs:66-73) without the
 * init: applies ops in order.  backend: 0 = RefBackend, 1 = CpuBackend::serial,
 * 2 = CpuBackend::parallel.  Returns 0 or a negative error. */
int or_run_f64(int backend, uint32_t n, double* re, double* im, size_t n_ops,
               const uint32_t* arity, const uint64_t* qoff, const uint32_t* qubits,
               const uint64_t* moff, const double* mre, const double* mim);
int or_run_f32(int parallel, uint32_t n, float* re, float* im, size_t n_ops,
               const uint32_t* arity, const uint64_t* qoff, const uint32_t* qubits,
               const uint64_t* moff, const double* mre, const double* mim);

/* ---- fusion (crates/qsv-core/src/fusion.rs:39-139) ----
 * or_fuse_groups: greedy grouping.  For each output op writes the index of the first
 * and one-past-last input op (grp_begin/grp_end) and the group's qubit list: sorted when
 * the group holds more than one op, the op's own order otherwise (fusion.rs:84-96).
 * Group g's qubits are written at gq[8*g .. 8*g+gm[g]).  Returns the number of groups. */
size_t or_fuse_groups(size_t n_ops, const uint32_t* arity, const uint64_t* qoff,
                      const uint32_t* qubits, uint32_t max_qubits,
                      uint64_t* grp_begin, uint64_t* grp_end, uint32_t* gm, uint32_t* gq);
/* build_fused (fusion.rs:102-139): compose ops [b, e) onto the sorted qubits q[0..k)
 * in f64, writing the 2^k x 2^k row-major product to ore/oim. */
int or_build_fused(uint32_t k, const uint32_t* q, uint64_t b, uint64_t e,
                   const uint32_t* arity, const uint64_t* qoff, const uint32_t* qubits,
                   const uint64_t* moff, const double* mre, const double* mim,
                   double* ore, double* oim);

/* ---- reductions (crates/qsv-core/src/state/mod.rs:112-134, sample.rs:177-200) ---- */
double or_norm_sqr_f64(uint64_t dim, const double* re, const double* im);
void or_probabilities_f64(uint64_t dim, const double* re, const double* im, double* out);
double or_expectation_z_f64(uint64_t dim, const double* probs, uint64_t mask);

/* ---- sampling (crates/qsv-core/src/sample.rs:33-173) ---- */
/* parallel_prefix_sum's serial path (sample.rs:114-160): out[i] = sum_{k<=i} p[k] in f64. */
void or_prefix_sum_f64(uint64_t dim, const double* p, double* out);
/* cdf_sample (sample.rs:140-153): draw k of SplitMix64(*state) gives u = unit() * total and
 * outcome min(partition_point(cdf < u), dim - 1).  Advances *state by `shots` draws. */
void or_cdf_sample(uint64_t dim, const double* p, uint64_t shots, uint64_t* state,
                   uint64_t* out);
/* AliasTable::build (sample.rs:38-82, Vose; the same small/large stack discipline) and
 * draw (sample.rs:89-97: bucket from one unit(), accept test on the next). */
void or_alias_build(uint64_t dim, const double* p, double* prob, uint64_t* alias);
void or_alias_sample(uint64_t dim, const double* prob, const uint64_t* alias, uint64_t shots,
                     uint64_t* state, uint64_t* out);

/* Number of OpenMP threads the parallel paths use. */
int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
