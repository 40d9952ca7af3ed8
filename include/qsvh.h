/*
 * qsvh.h -- C ABI of the MI355X (gfx950) statevector engine.
 *
 * This is the drop-in boundary for the reference's `Backend<R>` trait
 * (nez0b/rustatevec crates/qsv-core/src/backend/mod.rs:31-74): a Rust `qsv-hip` crate
 * implements `Backend` by calling these functions (see INTEGRATION.md for the binding).
 * All signatures use plain pointers and sizes; no exceptions or aborts cross the ABI --
 * every call returns a qsvh_status.
 *
 * Conventions (crates/qsv-core/src/gate.rs:8-12): a gate acting on the ordered qubit
 * list qs[0..m) is a row-major 2^m x 2^m complex matrix whose internal index has bit j
 * equal to global qubit qs[j]; controlled gates list the control first.  Amplitudes cross
 * the ABI (download, upload) as Structure-of-Arrays like the reference's StateVector
 * (crates/qsv-core/src/state/mod.rs:25-29): re[2^n], im[2^n] of f64 (QSVH_C128) or f32
 * (QSVH_C64); in HBM the engine keeps them interleaved (re, im pairs), which halves the
 * address streams of an in-place gate.  Matrices always cross the ABI in f64 and are rounded to
 * the state's precision on the device side, as `DenseGate<R>` is built with
 * `Cplx::from_f64` (crates/qsv-core/src/complex.rs:43-50).
 *
 * Threading (crates/qsv-cuda/src/lib.rs:59,199 for the reference's CUDA backend): one HIP
 * stream per context; calls on one context must be serialised by the caller; separate
 * contexts are independent.  Work is asynchronous until a call that returns host data
 * (amplitude, probabilities, norm_sqr, expectation_z, download) or qsvh_synchronize.
 */
#ifndef QSVH_H
#define QSVH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QSVH_ABI_VERSION 1
#define QSVH_MAX_GATE_QUBITS 6 /* apply_mq limit, crates/qsv-core/src/backend/cpu.rs:34,180 */

/* Status codes.  The reference panics where these are returned:
 * Circuit::push asserts (crates/qsv-core/src/circuit.rs:55-71), apply_mq's m <= 6 assert
 * (crates/qsv-core/src/backend/cpu.rs:180), `.expect` on device calls
 * (crates/qsv-cuda/src/lib.rs:90-227). */
typedef enum qsvh_status {
  QSVH_OK = 0,
  QSVH_ERR_INVALID_ARG = 1,     /* null pointer, unknown enum/name, bad size          */
  QSVH_ERR_INVALID_QUBIT = 2,   /* qubit >= n_qubits               (circuit.rs:60-66) */
  QSVH_ERR_REPEATED_QUBIT = 3,  /* same qubit twice in one gate    (circuit.rs:67-71) */
  QSVH_ERR_ARITY = 4,           /* matrix size != 2^m x 2^m        (circuit.rs:55-59) */
  QSVH_ERR_TOO_WIDE = 5,        /* gate on more than 6 qubits      (cpu.rs:180)       */
  QSVH_ERR_OOM = 6,             /* device (or host) allocation failed                 */
  QSVH_ERR_HIP = 7,             /* a HIP runtime call failed; see qsvh_last_error     */
  QSVH_ERR_DTYPE = 8,           /* operation not available for this state dtype       */
  QSVH_ERR_NO_DEVICE = 9,       /* no HIP device with that ordinal                    */
  QSVH_ERR_PARSE = 10           /* OpenQASM text rejected; see the message buffer     */
} qsvh_status;

typedef enum qsvh_dtype { QSVH_C128 = 0, QSVH_C64 = 1 } qsvh_dtype;

/* How fused (non-diagonal, k >= 2) groups are applied by qsvh_program_*.
 * EXACT: the state's own precision on the vector ALUs (f64 for C128: the plain path).
 * MFMA_F32: C64 states; k-qubit groups as FP32 matrix-core contractions.
 * MFMA_BF16X3: C64 states; split-BF16 (hi+lo) 3-product matrix-core contractions
 * accumulated in FP32. */
typedef enum qsvh_precision {
  QSVH_PREC_EXACT = 0,
  QSVH_PREC_MFMA_F32 = 1,
  QSVH_PREC_MFMA_BF16X3 = 2
} qsvh_precision;

typedef struct qsvh_ctx qsvh_ctx;         /* device + stream (CudaBackend, lib.rs:41-55)  */
typedef struct qsvh_state qsvh_state;     /* device state (Backend::State, mod.rs:33)     */
typedef struct qsvh_circuit qsvh_circuit; /* host circuit (Circuit<R>, circuit.rs:30-34)  */
typedef struct qsvh_program qsvh_program; /* a circuit bound to a state, ready to replay   */

/* ------------------------------------------------------------------ context */
/* CudaBackend::new(ordinal) (crates/qsv-cuda/src/lib.rs:57-77). */
int qsvh_ctx_create(int device, qsvh_ctx** out);
/* Every state, program, batch and model created on ctx must be destroyed first. */
void qsvh_ctx_destroy(qsvh_ctx* ctx);
/* Wait until all work queued on the context's stream has finished. */
int qsvh_synchronize_ctx(qsvh_ctx* ctx);
/* Text of the last error recorded on ctx ("" if none).  No reference counterpart:
 * the reference panics with the message instead. */
const char* qsvh_last_error(const qsvh_ctx* ctx);
/* Device name, compute-unit count and HBM bytes (for reports). */
int qsvh_device_info(qsvh_ctx* ctx, char* name, size_t name_cap, int* compute_units,
                     uint64_t* hbm_bytes);
int qsvh_abi_version(void);
/* Launch-shape knobs of the streaming (1q, diagonal, copy) kernels, per context:
 * "max_blocks" (0: one thread per work item; n > 0: grid-stride loop over n blocks),
 * "unroll" (1, 2 or 4 independent 16-byte items per thread per iteration) and
 * "nontemporal" (0/1: streaming load/store hints; 2: only for states larger than the
 * 256 MiB infinity cache), "contig" (0: a thread's unrolled items
 * are a grid apart; 1: adjacent), "xcd" (1: consecutive blocks of a streaming kernel take
 * neighbouring address ranges within one XCD; 0: launch order); "mfma_blocks"
 * (workgroups of the matrix-core group kernels; 0, the default: one per tile for split-BF16
 * 4-qubit groups, one per 4 tiles for split-BF16 5-qubit groups, one per 64 tiles for FP32
 * 6-qubit groups, 1 per CU for FP32 4-qubit groups, else 2 per CU); "tile_blocks" (workgroups of the tile-pass
 * kernels, exact and MFMA; 0, the default: one per tile; n > 0: n looping over tiles);
 * "mfma_layout" (LDS layout of the FP32 group kernel: 0 chosen from the group's qubits,
 * 1 row-major, 2 column-major).
 * No reference counterpart: the
 * reference's kernels have a fixed 256-thread, one-item-per-thread shape
 * (crates/qsv-cuda/src/lib.rs:32,80-87). */
int qsvh_tune_set(qsvh_ctx* ctx, const char* key, int64_t value);
int qsvh_tune_get(qsvh_ctx* ctx, const char* key, int64_t* value);

/* ------------------------------------------------------------------ gates */
/* The gate library (crates/qsv-core/src/gate.rs:96-180): name is one of h x y z s t sx
 * rx ry rz p|phase cx cz cp|cphase swap rzz, plus u3 (OpenQASM 3 U(theta,phi,lambda)),
 * ccx and ccz (crates/qsv-core/tests/equivalence.rs:182-193).  Writes the row-major
 * matrix into re/im (capacity 64 entries each) and its arity into *m. */
int qsvh_gate_matrix(const char* name, const double* params, uint32_t n_params, double* re,
                     double* im, uint32_t* m);

/* ------------------------------------------------------------------ circuits */
/* Circuit::new (crates/qsv-core/src/circuit.rs:37-42). */
int qsvh_circuit_create(uint32_t n_qubits, qsvh_circuit** out);
void qsvh_circuit_destroy(qsvh_circuit* c);
/* Circuit::push (circuit.rs:54-78): validated, returns ARITY/INVALID_QUBIT/
 * REPEATED_QUBIT/TOO_WIDE instead of panicking.  re/im: row-major 2^m x 2^m. */
int qsvh_circuit_push(qsvh_circuit* c, uint32_t m, const uint32_t* qubits, const double* re,
                      const double* im);
/* The fluent builders (circuit.rs:81-130): push the named library gate. */
int qsvh_circuit_push_gate(qsvh_circuit* c, const char* name, const uint32_t* qubits,
                           const double* params, uint32_t n_params);
uint32_t qsvh_circuit_n_qubits(const qsvh_circuit* c);
size_t qsvh_circuit_len(const qsvh_circuit* c);
/* Circuit::ops()[i] (circuit.rs:48-51): arity, qubits (capacity 6), matrix (capacity 64
 * complex entries when the gate has m <= 3; 4^m entries in general). */
int qsvh_circuit_op(const qsvh_circuit* c, size_t i, uint32_t* m, uint32_t* qubits,
                    double* re, double* im);

/* Generators (crates/qsv-core/src/circuits.rs:50-113) and the measured workloads. */
int qsvh_circuit_ghz(uint32_t n, qsvh_circuit** out);       /* circuits.rs:50-57        */
int qsvh_circuit_ghz_chain(uint32_t n, qsvh_circuit** out); /* H(0); CX(i,i+1)          */
int qsvh_circuit_qft(uint32_t n, qsvh_circuit** out);       /* circuits.rs:61-74        */
int qsvh_circuit_random(uint32_t n, uint64_t n_gates, uint64_t seed,
                        qsvh_circuit** out);                /* circuits.rs:79-113       */
/* Random layered circuit: per layer U3 on every qubit, RZ on every qubit, CZ brickwork. */
int qsvh_circuit_layered(uint32_t n, uint32_t depth, uint64_t seed, qsvh_circuit** out);
/* Hardware-efficient ansatz: per layer RY, RZ on every qubit, then CX(i,i+1). */
int qsvh_circuit_hea(uint32_t n, uint32_t depth, uint64_t seed, qsvh_circuit** out);
/* Random gate mix: 50% U3, 30% CX/CZ, 15% RZZ, 5% CCX/CCZ on random qubits. */
int qsvh_circuit_mix(uint32_t n, uint64_t n_gates, uint64_t seed, qsvh_circuit** out);

/* OpenQASM 3 (or 2.0) program text -> circuit: the QASM3 front end the reference's CLI
 * names as future work (crates/qsv-cli/src/main.rs:1-2, README.md:63).  Registers are
 * concatenated in declaration order (q[0] of the first register = qubit 0, the least
 * significant index bit); operands map to the gate's qubits in order, control first
 * (gate.rs:8-12).  Standard gates (stdgates.inc / qelib1.inc names), gate definitions,
 * inv / pow(int) / ctrl / negctrl modifiers, gphase, broadcast and ranges are read;
 * barrier and terminal measure are skipped; anything else is QSVH_ERR_PARSE with
 * "line L: reason" in err (NUL-terminated, truncated to err_cap; err may be NULL). */
int qsvh_circuit_from_qasm(const char* text, size_t len, qsvh_circuit** out, char* err,
                           size_t err_cap);

/* fusion::fuse (crates/qsv-core/src/fusion.rs:39-77) with FusionConfig{max_qubits}:
 * greedy consecutive grouping on the host; every multi-op group's matrix is composed in
 * f64 on the device (same product order as build_fused, fusion.rs:102-139) and the group
 * is emitted on its sorted qubits (fusion.rs:91-94).  *out is a new circuit. */
int qsvh_fuse(qsvh_ctx* ctx, const qsvh_circuit* in, uint32_t max_qubits,
              qsvh_circuit** out);

/* The greedy grouping of fusion::fuse alone, on the host (fusion.rs:39-99): for output
 * op g, the input op range [begin[g], end[g]), its arity m[g] and its qubits at
 * qubits[6*g ..] (sorted when the group holds more than one op).  Writes at most `cap`
 * groups and the total count to *n_groups. */
int qsvh_fuse_groups(const qsvh_circuit* in, uint32_t max_qubits, uint64_t* begin,
                     uint64_t* end, uint32_t* m, uint32_t* qubits, size_t cap,
                     size_t* n_groups);

/* ------------------------------------------------------------------ state (Backend) */
/* Backend::alloc (crates/qsv-core/src/backend/mod.rs:36-37); contents unspecified. */
int qsvh_state_alloc(qsvh_ctx* ctx, uint32_t n_qubits, int dtype, qsvh_state** out);
/* Destroy a state's programs (qsvh_program_destroy) before the state: a program keeps a
 * reference to its state and must not outlive it. */
void qsvh_state_free(qsvh_state* st);
uint32_t qsvh_state_n_qubits(const qsvh_state* st);
int qsvh_state_dtype(const qsvh_state* st);
/* Backend::init_basis (mod.rs:39-40). */
int qsvh_init_basis(qsvh_state* st, uint64_t basis);
/* Backend::apply (mod.rs:42-43): dispatch diagonal -> 1q -> mq as CpuBackend::apply
 * (crates/qsv-core/src/backend/cpu.rs:281-290). */
int qsvh_apply(qsvh_state* st, uint32_t m, const uint32_t* qubits, const double* re,
               const double* im);
/* Backend::amplitude (mod.rs:45-46): single-element device->host copy. */
int qsvh_amplitude(qsvh_state* st, uint64_t index, double* re, double* im);
/* Backend::probabilities (mod.rs:48-49): |psi_i|^2 in the state's precision (f64 for
 * C128, f32 for C64) into host_out[2^n]. */
int qsvh_probabilities(qsvh_state* st, void* host_out);
/* Device-side probabilities into a caller-owned device buffer (no host copy). */
int qsvh_probabilities_device(qsvh_state* st, void* device_out);
/* <psi|psi> as an f64 tree reduction (StateVector::norm_sqr, state/mod.rs:112-117). */
int qsvh_norm_sqr(qsvh_state* st, double* out);
/* Backend::expectation_z (mod.rs:57-59, sample.rs:177-200): sum_i |psi_i|^2 *
 * (-1)^popcount(i & mask), f64 accumulation on the device. */
int qsvh_expectation_z(qsvh_state* st, uint64_t mask, double* out);
/* Backend::sample (mod.rs:52-54, which a device backend overrides "to sample on-device (no
 * full copy back)"): `shots` outcomes in [0, 2^n) drawn from |psi_i|^2 on the device by
 * inverse CDF as in cdf_sample (sample.rs:140-153): an f64 prefix sum of the probabilities,
 * then for draw k (k = 1..shots) of SplitMix64::new(seed) u = unit() * total and
 * outcome = min(first i with cdf_i >= u, 2^n - 1) -- the outcomes of
 * cdf_sample(probabilities, shots, &mut SplitMix64::new(seed)), except where u lies within
 * rounding of a CDF boundary (the device adds the CDF in another order).
 * Random-stream use differs from the trait: the default Backend::sample (alias_sample,
 * sample.rs:108-111) takes 2 draws per shot from the caller's generator and leaves it
 * advanced by 2 * shots; this call takes a seed instead and touches no host generator.  A
 * binding that seeds it with one rng.next_u64() (SplitMix64's state is private to qsv-core)
 * leaves the caller's generator advanced by one draw, so a caller that goes on drawing from
 * rng sees a different stream than after the default sampler, and the outcomes are a
 * different sequence of the same distribution.  Writes `shots` u64 to host out.  Temporary
 * device memory: 8 bytes per amplitude plus 8 per shot; at most 2^31 shots per call and 34
 * qubits. */
int qsvh_sample(qsvh_state* st, uint64_t shots, uint64_t seed, uint64_t* out);
/* qsvh_sample without allocating: the same kernels, the same outcomes for the same state
 * and seed, written to `device_out` (caller-owned device memory of `shots` u64) instead of a
 * host array.  No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops that sample every iteration,
 * harnesses that forbid hipMalloc).  It calls no allocator and does not synchronise: the
 * outcomes are ready when the context's stream is (qsvh_synchronize).  Its scratch is one
 * static device array per GPU (about 32 MiB, held until the process ends) that fits the CDF
 * tables of at most 22 qubits: larger states are rejected with QSVH_ERR_INVALID_ARG ("at
 * most 22 qubits"), as are more than 2^31 shots.  All calls on one GPU share that array, so
 * calls that overlap on one GPU (two contexts or streams at once) race: serialise them. */
int qsvh_sample_noalloc(qsvh_state* st, uint64_t shots, uint64_t seed, uint64_t* device_out);
/* Backend::download (mod.rs:61-62): the full SoA state to host re/im (2^n each). */
int qsvh_download(qsvh_state* st, void* re, void* im);
/* Host -> device copy of a full state.  No reference counterpart (the trait has no
 * upload); needed to start from a prepared state such as a random normalised vector. */
int qsvh_upload(qsvh_state* st, const void* re, const void* im);
/* Random normalised start state, generated on the device: psi_i = (a_i + i b_i) / norm
 * with a_i, b_i ~ N(0,1) by Box-Muller from outputs 2i and 2i+1 of SplitMix64(seed)
 * (crates/qsv-core/src/circuits.rs:12-47): u = (out >> 11) * 2^-53,
 * a + ib = sqrt(-2 ln(1 - u1)) * e^{2 pi i u2}.  No reference counterpart (the trait has
 * no upload); it is the start state of the QFT workload. */
int qsvh_init_random_normal(qsvh_state* st, uint64_t seed);
/* Compare two states of equal size (any dtypes): *max_abs_err = max_i |a_i - b_i|,
 * *fidelity = |<a|b>|^2 / (<a|a><b|b>), both accumulated in f64 on the device. */
int qsvh_compare(qsvh_state* a, qsvh_state* b, double* max_abs_err, double* fidelity);
/* Wait until all work queued on the state's context has finished. */
int qsvh_synchronize(qsvh_state* st);

/* Backend::execute (mod.rs:66-73): alloc, init_basis(0), apply every op; one submission
 * (no per-gate synchronisation). */
int qsvh_execute(qsvh_ctx* ctx, const qsvh_circuit* c, int dtype, qsvh_state** out);

/* ------------------------------------------------------------------ programs */
/* A program binds a circuit to a state: gate classification, device-resident matrix
 * arena (uploaded once) and, optionally, a captured HIP graph.  No reference
 * counterpart: it is how the default execute loop (mod.rs:66-73) becomes one submission
 * that can be replayed. */
typedef struct qsvh_program_opts {
  int precision;    /* qsvh_precision; EXACT is the plain path                          */
  int use_graph;    /* capture the launches in a hipGraph and replay it                */
  int tile_qubits;  /* >0: apply runs of gates on a 2^tile_qubits LDS tile per HBM pass */
  int subspace;     /* 1: per-gate launches touch only the amplitudes a gate changes: a
                       controlled 1-qubit gate (CX, CCX, ...) its control subspace, a
                       diagonal gate with one non-unit phase (CP, CZ, CCZ) that phase's
                       amplitudes.  Values equal the full pass (the skipped products are
                       exact zeros and ones); the bytes moved drop to 1/2 .. 1/8.        */
  int mfma_tiles;   /* 12 (precision MFMA_BF16X3): runs of ops -- fused groups of up to 6
                       qubits, dense (split-BF16 MFMA) or diagonal -- applied to a
                       2^12-amplitude LDS tile per HBM pass instead of one pass each;
                       consecutive ops share a pass while its dense ops' qubits and qubits
                       0..3 fit the tile and its diagonal ops' phases fit 2048 entries.
                       Excludes tile_qubits and subspace.                                */
  int planned;      /* 1: run the pass plan the circuit carries (qsvh_plan_passes): each of
                       its passes as one exact tile pass or one MFMA tile pass (precision
                       MFMA_BF16X3 when any pass is MFMA); complex64 states of >= 12 qubits.
                       Excludes tile_qubits, mfma_tiles and subspace.                    */
  int reserved[2];
} qsvh_program_opts;

int qsvh_program_create(qsvh_state* st, const qsvh_circuit* c, const qsvh_program_opts* opts,
                        qsvh_program** out);
/* Apply the program's gates to the state's current contents (asynchronous). */
int qsvh_program_launch(qsvh_program* p);
/* Launch once and return the device time of that launch in milliseconds, measured with
 * HIP events recorded on the context's stream. */
int qsvh_program_time(qsvh_program* p, double* ms);
/* Number of device kernels one launch issues (passes over the state). */
int qsvh_program_num_kernels(const qsvh_program* p, uint64_t* out);
/* Launch once with a HIP event between consecutive kernels: ms[j] = device milliseconds of
 * kernel j (at most cap written; *n = the program's kernel count).  Measurement helper for
 * per-pass timings (the pass planner's training data). */
int qsvh_program_time_kernels(qsvh_program* p, double* ms, size_t cap, size_t* n);
/* The circuit's op indices in the order the program applies them (tile passes move
 * gates past gates on disjoint qubits).  Writes at most cap entries; *n = total. */
int qsvh_program_order(const qsvh_program* p, uint64_t* order, size_t cap, size_t* n);
void qsvh_program_destroy(qsvh_program* p);

/* ------------------------------------------------------------------ batches */
/* Many independent circuits (uneven qubit counts, 4..30 qubits each) run together: each
 * circuit's gates are scheduled into LDS tile passes and launch j runs pass j of every
 * circuit in one persistent kernel (work items handed out by an atomic counter); the
 * launches are captured in one HIP graph.  A circuit that fits one tile never leaves LDS
 * until its final state is written.  Every run starts each circuit from |0...0> and ends
 * with its <Z_0>.  No reference counterpart: the reference executes one circuit per
 * Backend::execute call (crates/qsv-core/src/backend/mod.rs:66-73).  Gates must act on
 * at most 3 qubits (or be diagonal). */
typedef struct qsvh_batch qsvh_batch;
int qsvh_batch_create(qsvh_ctx* ctx, const qsvh_circuit* const* circuits, size_t count,
                      int dtype, qsvh_batch** out);
void qsvh_batch_destroy(qsvh_batch* b);
/* Re-plan with a tile size (complex128: 10..12 bits, complex64: 11..13) and a persistent
 * grid (0: two workgroups per compute unit). */
int qsvh_batch_configure(qsvh_batch* b, int tile_qubits, int blocks);
/* Time every (tile size, grid of 1, 2, 4, 8 workgroups per CU) candidate on this batch
 * (best of three runs each), keep the fastest and report it. */
int qsvh_batch_autotune(qsvh_batch* b, int* tile_qubits, int* blocks, double* best_ms);
/* Run every circuit (graph replay; the graph is captured on first use). */
int qsvh_batch_run(qsvh_batch* b);
/* Run once and return the device time in milliseconds (HIP events on the stream). */
int qsvh_batch_time(qsvh_batch* b, double* ms);
int qsvh_batch_info(const qsvh_batch* b, size_t* count, int* tile_qubits, int* blocks,
                    uint64_t* n_launches, uint64_t* total_amplitudes);
/* Passes of circuit i (the launches that run it; the first builds |0...0> in LDS instead
 * of reading the state, so a run moves (2 passes - 1) x 2^n_i amplitudes through HBM). */
int qsvh_batch_passes(const qsvh_batch* b, size_t i, uint32_t* passes);
/* <Z_0> of every circuit after the last run (out[count]). */
int qsvh_batch_expectation_z0(qsvh_batch* b, double* out);
/* Final state / probabilities of circuit i (host buffers of 2^n_i elements). */
int qsvh_batch_download(qsvh_batch* b, size_t i, void* re, void* im);
int qsvh_batch_probabilities(qsvh_batch* b, size_t i, void* host_out);
/* Circuit i's op indices in the order the batch applies them. */
int qsvh_batch_order(const qsvh_batch* b, size_t i, uint64_t* order, size_t cap, size_t* n);
/* The variational-sweep workload: `count` hardware-efficient-ansatz circuits; with
 * rng = SplitMix64(seed), circuit c has n_c = 10 + rng.below(13) qubits and is
 * qsvh_circuit_hea(n_c, depth, rng.next_u64()).  Writes count handles to out. */
int qsvh_circuit_hea_batch(size_t count, uint32_t depth, uint64_t seed, qsvh_circuit** out);

/* ------------------------------------------------------------------ learned planner */
/* Cost model: a fully connected network n_features -> hidden -> hidden -> 1 with ReLU,
 * FP32, trained on the device by the library's own forward, backward and Adam kernels.
 * No reference counterpart: the reference fuses greedily with a fixed max_qubits
 * (crates/qsv-core/src/fusion.rs:24-77). */
typedef struct qsvh_mlp qsvh_mlp;
int qsvh_mlp_create(qsvh_ctx* ctx, uint32_t n_features, uint32_t hidden, uint64_t seed,
                    qsvh_mlp** out);
void qsvh_mlp_destroy(qsvh_mlp* m);
int qsvh_mlp_num_params(const qsvh_mlp* m, size_t* n);
/* Flat parameters: W1[F x H], b1[H], W2[H x H], b2[H], W3[H x 1], b3 (row-major). */
int qsvh_mlp_get_params(qsvh_mlp* m, float* out);
int qsvh_mlp_set_params(qsvh_mlp* m, const float* in);
/* Raw network outputs for rows x n_features inputs. */
int qsvh_mlp_forward(qsvh_mlp* m, const float* x, size_t rows, float* out);
/* Mean squared error of the raw outputs against y and its gradient w.r.t. every parameter
 * (same layout as the parameters). */
int qsvh_mlp_gradients(qsvh_mlp* m, const float* x, const float* y, size_t rows, float* grads,
                       float* loss);
/* Full-batch Adam on the MSE; targets are standardised internally (mean/stddev kept for
 * qsvh_mlp_predict).  *final_loss: MSE on the standardised targets after training. */
int qsvh_mlp_train(qsvh_mlp* m, const float* x, const float* y, size_t rows, uint32_t epochs,
                   float lr, float* final_loss);
/* Outputs in target units (de-standardised). */
int qsvh_mlp_predict(qsvh_mlp* m, const float* x, size_t rows, float* out);
int qsvh_mlp_target_scale(const qsvh_mlp* m, float* mean, float* stddev);
/* The 12 cost-model features of a fused group on an n-qubit state: k/6, n/32, diagonal
 * flag, density (nonzeros / 4^k), (q_i + 1)/n for the ascending qubits (6 slots, 0 when
 * absent), min(lowest qubit, 8)/8, matrix-core flag (non-diagonal, k >= 4). */
int qsvh_group_features(uint32_t n, uint32_t k, const uint32_t* qubits, int diagonal,
                        double density, float* out);
/* Plan the fusion of `in` with the cost model (targets: log2 of device milliseconds per
 * group at 2^30 amplitudes): dynamic programming over contiguous groups of <= max_qubits
 * qubits and <= window gates (0: 48), scored against the greedy plans of fusion::fuse at
 * every max_qubits (fusion.rs:39-77), the cheapest predicted plan kept.  Groups are
 * composed in f64 on the device.  *chosen: 0 for the dynamic programme, m for greedy at
 * max_qubits m. */
int qsvh_plan_fusion(qsvh_ctx* ctx, qsvh_mlp* model, const qsvh_circuit* in,
                     uint32_t max_qubits, uint32_t window, qsvh_circuit** out,
                     double* predicted_ms, int* chosen);
/* Cost-model features of every tile pass the tile scheduler builds for `c` on a state of
 * `dtype` with `tile_qubits` tile bits (10..12 for QSVH_C128, 11..13 for QSVH_C64):
 * QSVH_TILE_FEATURES floats per pass into out (at most max_passes rows); *n_passes: passes
 * built, *n_other: launches left outside passes (gates a tile cannot hold).  Features:
 * dtype, tile bits/13, n/32, then counts /16 of diagonal gates (phase uniform per tile,
 * lane-bit, register-bit, m > 2), lane-run 1q members (lane / register target x general /
 * cheap pair variant), register blocks, their members (general / cheap), LDS ops, permutation
 * ops, lane runs, conditioned lane-run members.  No reference counterpart: the tile passes
 * and their cost model are this engine's. */
#define QSVH_TILE_FEATURES 18
int qsvh_tile_pass_features(const qsvh_circuit* c, int dtype, uint32_t tile_qubits, float* out,
                            size_t max_passes, size_t* n_passes, size_t* n_other);
/* Cost-model features of every MFMA tile pass (program option mfma_tiles = 12, split-BF16)
 * built for the fused circuit `c` (>= 12 qubits): QSVH_MT_FEATURES floats per pass into out
 * (at most max_passes rows); *n_passes, *n_other as qsvh_tile_pass_features.  Features:
 * counts of dense groups of 4, 5, 6 qubits (/16), single diagonal ops (/32), merged diagonal
 * runs (/16) and their members (/64), per-tile table entries of the merged runs and the
 * pass's phase-table entries (/2048), widest group 5 / 6 flags, n/32 (states that fit the
 * 256 MiB infinity cache run faster per amplitude).  No reference counterpart. */
#define QSVH_MT_FEATURES 11
int qsvh_mt_pass_features(const qsvh_circuit* c, float* out, size_t max_passes,
                          size_t* n_passes, size_t* n_other);
/* Learned choice of the execution form for `in` on a `dtype` state: the fused plan of
 * qsvh_plan_fusion (group_model, 12 inputs) against tile passes at every allowed tile size
 * (tile_model, QSVH_TILE_FEATURES inputs, trained on log2 of pass milliseconds at 2^30
 * amplitudes; gates outside passes priced by group_model) and, when mt_model is given
 * (QSVH_MT_FEATURES inputs, trained likewise; complex64, >= 12 qubits), against greedy
 * fusion at max_qubits 4 and 5 run as MFMA tile passes; the cheapest predicted kept.
 * Fused: *out is the fused circuit (run with an MFMA precision), *tile_qubits = 0,
 * *mfma_tiles = 0, *chosen as qsvh_plan_fusion.  Tiled: *out is `in` unchanged (run exact
 * with program option tile_qubits), *tile_qubits > 0, *chosen = -tile_qubits.  MFMA tiles:
 * *out is the fused circuit (run MFMA_BF16X3 with program option mfma_tiles = *mfma_tiles =
 * 12), *chosen = 100 + max_qubits.  Extends the fusion planner (fusion.rs:39-77 is the
 * greedy fallback) to the choice between the forms. */
int qsvh_plan_form(qsvh_ctx* ctx, qsvh_mlp* group_model, qsvh_mlp* tile_model,
                   qsvh_mlp* mt_model, const qsvh_circuit* in, int dtype, uint32_t max_qubits,
                   uint32_t window, qsvh_circuit** out, int* tile_qubits, int* mfma_tiles,
                   double* predicted_ms, int* chosen);

/* Learned choice of the form of every pass.  The circuit (>= 12 qubits) is cut into passes:
 * each takes gates in circuit order while its dense gates' qubits, with qubits 0..3, fit one
 * tile of 12 qubits (diagonal gates join on any qubits within a phase budget), a gate moving
 * ahead only past gates it commutes with.  Each pass then runs exactly (its gates one by one
 * on the tile, complex64 arithmetic) or as one MFMA tile pass of its gates fused greedily up
 * to k = 3, 4 or 5 qubits (fusion.rs:39-77 within the pass; a group is diagonal or inside
 * the tile), priced per pass by tile_model (QSVH_TILE_FEATURES inputs) and mt_model
 * (QSVH_MT_FEATURES inputs), the cheapest predicted form kept.  form: -1 learned (both
 * models needed), 0 every pass exact, 3..5 every pass MFMA at that width (a pass a forced
 * form cannot run takes the cheapest other).  *out: the circuit in the order it is applied,
 * exact passes' gates as given and MFMA passes' fused groups composed in f64 on the device,
 * carrying the pass plan (run it with program option `planned`, precision MFMA_BF16X3).
 * *predicted_ms: the models' total (0 without models).  No reference counterpart: the
 * reference fuses with one fixed width (fusion.rs:24-77). */
int qsvh_plan_passes(qsvh_ctx* ctx, qsvh_mlp* tile_model, qsvh_mlp* mt_model,
                     const qsvh_circuit* in, int form, qsvh_circuit** out,
                     double* predicted_ms);
/* Passes of a planned circuit (0 for any other circuit). */
int qsvh_circuit_plan_len(const qsvh_circuit* c, size_t* n_passes);
/* Pass i of a planned circuit: its form (0 exact, k MFMA at width k), op range, tile bits
 * and, when features is not NULL, its cost-model feature row for its form
 * (QSVH_TILE_FEATURES floats for an exact pass, QSVH_MT_FEATURES for an MFMA pass). */
int qsvh_circuit_plan_pass(const qsvh_circuit* c, size_t i, int* form, uint64_t* op_begin,
                           uint64_t* n_ops, uint64_t* tile_mask, float* features);

/* Apply one gate `reps` times back to back and return the mean device time per launch
 * (HIP events on the context stream).  Measurement helper for single-kernel rooflines. */
int qsvh_time_gate(qsvh_state* st, uint32_t m, const uint32_t* qubits, const double* re,
                   const double* im, uint32_t reps, double* ms_per_launch);
/* Device-to-device copy of the whole state, timed like qsvh_time_gate: the measured
 * bandwidth ceiling (2 * bytes(state) moved per launch). */
int qsvh_time_copy(qsvh_state* st, uint32_t reps, double* ms_per_launch);

#ifdef __cplusplus
}
#endif
#endif /* QSVH_H */
