/*
 * qsv_oracle.c -- CPU restatement of the qsv algorithm (nez0b/rustatevec), kept as TEST
 * INFRASTRUCTURE only: the checker for the HIP engine and the timed CPU baseline.  Only
 * tests/, the smoke check and bench.py's cpu_baseline leg may load it.
 *
 * Arithmetic mirrors the reference's non-contracted Rust complex arithmetic
 * (crates/qsv-core/src/complex.rs:95-105): build with -ffp-contract=off.  The parallel
 * paths use OpenMP where the reference uses rayon, with the same decomposition
 * (crates/qsv-core/src/backend/cpu.rs:44-247).
 *
 * Parity pins (what checks this file, see tests/test_oracle.py): SplitMix64 against the
 * published generator's first outputs; index arithmetic against the reference's own
 * unit-test examples (layout.rs:131-193); gate matrices against closed forms and
 * scipy.linalg.expm; the backends against numpy einsum / Kronecker golden states
 * (tests/golden) and against each other (Ref vs Cpu); fusion by fused == unfused.
 * The exact RNG->gate mapping of or_gen_random and the group boundaries of
 * or_fuse_groups are checked only for self-consistency ("parity unpinned" beyond the
 * invariants): the reference cannot be built here (no Rust toolchain).
 */
#define _GNU_SOURCE
#include "qsv_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAX_SUB 64 /* cpu.rs:34  -- up to 6-qubit gates in the block kernel */
#define MIN_PAIRS_LIGHT (1ull << 16) /* cpu.rs:44 */
#define MIN_PAIRS_HEAVY (1ull << 12) /* cpu.rs:45 */
#define FIRST_TOUCH_MIN (1ull << 18) /* state/mod.rs:21 */

/* sin and cos of one angle through glibc's sincos, so that the compiler's choice between
 * separate calls and a fused sincos call (which can differ in the last bit) is fixed. */
static void sc(double x, double* s, double* c) { sincos(x, s, c); }

static const double OR_PI = 3.14159265358979323846;
static const double OR_TAU = 6.28318530717958647692;
static const double OR_FRAC_1_SQRT_2 = 0.70710678118654752440;

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ SplitMix64 */
/* circuits.rs:21-28 */
uint64_t or_sm64_next(uint64_t* s) {
  *s += 0x9E3779B97F4A7C15ull;
  uint64_t z = *s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
/* circuits.rs:30-34 */
uint64_t or_sm64_below(uint64_t* s, uint64_t n) { return or_sm64_next(s) % n; }
/* circuits.rs:36-40 */
double or_sm64_unit(uint64_t* s) {
  return (double)(or_sm64_next(s) >> 11) / (double)(1ull << 53);
}
/* circuits.rs:42-46 */
double or_sm64_angle(uint64_t* s) { return or_sm64_unit(s) * OR_TAU; }

/* ------------------------------------------------------------ index arithmetic */
/* layout.rs:22-26 */
uint64_t or_insert_zero_bit(uint64_t index, uint32_t bit) {
  uint64_t left = (index >> bit) << bit;
  uint64_t right = index - left;
  return (left << 1) | right;
}
/* layout.rs:43-49 */
uint64_t or_gather_bits(uint64_t index, const uint32_t* qs, uint32_t m) {
  uint64_t g = 0;
  for (uint32_t j = 0; j < m; j++) g |= ((index >> qs[j]) & 1ull) << j;
  return g;
}
/* layout.rs:54-60 */
uint64_t or_scatter_bits(uint64_t sub, const uint32_t* qs, uint32_t m) {
  uint64_t x = 0;
  for (uint32_t j = 0; j < m; j++) x |= ((sub >> j) & 1ull) << qs[j];
  return x;
}
/* layout.rs:70-87 (portable path; the BMI2 variant is an x86 micro-optimisation) */
uint64_t or_insert_zero_bits(uint64_t index, const uint32_t* sorted, uint32_t m) {
  uint64_t x = index;
  for (uint32_t j = 0; j < m; j++) x = or_insert_zero_bit(x, sorted[j]);
  return x;
}

/* ------------------------------------------------------------------ gates */
uint32_t or_gate_arity(int kind) {
  switch (kind) {
    case OR_H: case OR_X: case OR_Y: case OR_Z: case OR_S: case OR_T: case OR_SX:
    case OR_RX: case OR_RY: case OR_RZ: case OR_P: case OR_U3:
      return 1;
    case OR_CX: case OR_CZ: case OR_CP: case OR_SWAP: case OR_RZZ:
      return 2;
    case OR_CCX: case OR_CCZ:
      return 3;
    default:
      return 0;
  }
}

static void set1(double* re, double* im, double a, double b, double c, double d,
                 double e, double f, double g, double h) {
  re[0] = a; im[0] = b; re[1] = c; im[1] = d; re[2] = e; im[2] = f; re[3] = g; im[3] = h;
}

/* controlled_1q (gate.rs:72-87): control-0 subspace identity, control-1 applies u. */
static void controlled(const double* ure, const double* uim, double* re, double* im) {
  memset(re, 0, 16 * sizeof(double));
  memset(im, 0, 16 * sizeof(double));
  re[0] = 1.0;
  re[2 * 4 + 2] = 1.0;
  for (int to = 0; to < 2; to++)
    for (int ti = 0; ti < 2; ti++) {
      int row = 1 + 2 * to, col = 1 + 2 * ti;
      re[row * 4 + col] = ure[to * 2 + ti];
      im[row * 4 + col] = uim[to * 2 + ti];
    }
}

uint32_t or_gate_matrix(int kind, const double* p, double* re, double* im) {
  double s, c;
  double ure[4], uim[4];
  switch (kind) {
    case OR_H: { /* gate.rs:96-99 */
      double r = OR_FRAC_1_SQRT_2;
      set1(re, im, r, 0, r, 0, r, 0, -r, 0);
      return 1;
    }
    case OR_X: set1(re, im, 0, 0, 1, 0, 1, 0, 0, 0); return 1;   /* gate.rs:100-102 */
    case OR_Y: set1(re, im, 0, 0, 0, -1, 0, 1, 0, 0); return 1;  /* gate.rs:103-105 */
    case OR_Z: set1(re, im, 1, 0, 0, 0, 0, 0, -1, 0); return 1;  /* gate.rs:106-108 */
    case OR_S: set1(re, im, 1, 0, 0, 0, 0, 0, 0, 1); return 1;   /* gate.rs:109-111 */
    case OR_T: { /* gate.rs:112-115 */
      double r = OR_FRAC_1_SQRT_2;
      set1(re, im, 1, 0, 0, 0, 0, 0, r, r);
      return 1;
    }
    case OR_SX: set1(re, im, 0.5, 0.5, 0.5, -0.5, 0.5, -0.5, 0.5, 0.5); return 1; /* :117-122 */
    case OR_RX: /* gate.rs:126-129 */
      sc(p[0] * 0.5, &s, &c);
      set1(re, im, c, 0, 0, -s, 0, -s, c, 0);
      return 1;
    case OR_RY: /* gate.rs:130-133 */
      sc(p[0] * 0.5, &s, &c);
      set1(re, im, c, 0, -s, 0, s, 0, c, 0);
      return 1;
    case OR_RZ: /* gate.rs:134-138 */
      sc(p[0] * 0.5, &s, &c);
      set1(re, im, c, -s, 0, 0, 0, 0, c, s);
      return 1;
    case OR_P: /* gate.rs:140-143 */
      sc(p[0], &s, &c);
      set1(re, im, 1, 0, 0, 0, 0, 0, c, s);
      return 1;
    case OR_U3: {
      /* OpenQASM 3 U(theta, phi, lambda) =
       *   [[cos(t/2), -e^{i l} sin(t/2)], [e^{i p} sin(t/2), e^{i(p+l)} cos(t/2)]] */
      double st, ct, sp, cp, sl, cl, spl, cpl;
      sc(p[0] * 0.5, &st, &ct);
      sc(p[1], &sp, &cp);
      sc(p[2], &sl, &cl);
      sc(p[1] + p[2], &spl, &cpl);
      set1(re, im, ct, 0, -cl * st, -sl * st, cp * st, sp * st, cpl * ct, spl * ct);
      return 1;
    }
    case OR_CX: /* gate.rs:148-150 */
      or_gate_matrix(OR_X, p, ure, uim);
      controlled(ure, uim, re, im);
      return 2;
    case OR_CZ: /* gate.rs:152-154 */
      or_gate_matrix(OR_Z, p, ure, uim);
      controlled(ure, uim, re, im);
      return 2;
    case OR_CP: /* gate.rs:156-158 */
      or_gate_matrix(OR_P, p, ure, uim);
      controlled(ure, uim, re, im);
      return 2;
    case OR_SWAP: /* gate.rs:160-168 */
      memset(re, 0, 16 * sizeof(double));
      memset(im, 0, 16 * sizeof(double));
      re[0] = re[6] = re[9] = re[15] = 1.0;
      return 2;
    case OR_RZZ: /* gate.rs:170-180 */
      memset(re, 0, 16 * sizeof(double));
      memset(im, 0, 16 * sizeof(double));
      for (int idx = 0; idx < 4; idx++) {
        int parity = (idx & 1) ^ ((idx >> 1) & 1);
        double zz = parity == 0 ? 1.0 : -1.0;
        double a = -0.5 * p[0] * zz;
        sc(a, &im[idx * 4 + idx], &re[idx * 4 + idx]);
      }
      return 2;
    case OR_CCX: /* tests/equivalence.rs:182-193: swap internal indices 3 and 7 */
      memset(re, 0, 64 * sizeof(double));
      memset(im, 0, 64 * sizeof(double));
      for (int i = 0; i < 8; i++) {
        int t = i == 3 ? 7 : (i == 7 ? 3 : i);
        re[i * 8 + t] = 1.0;
      }
      return 3;
    case OR_CCZ: /* diag(1,...,1,-1): phase -1 when all three bits are set */
      memset(re, 0, 64 * sizeof(double));
      memset(im, 0, 64 * sizeof(double));
      for (int i = 0; i < 8; i++) re[i * 8 + i] = i == 7 ? -1.0 : 1.0;
      return 3;
    default:
      return 0;
  }
}

/* gate.rs:56-67 */
int or_is_diagonal(uint32_t m, const double* re, const double* im) {
  uint64_t dim = 1ull << m;
  for (uint64_t r = 0; r < dim; r++)
    for (uint64_t c = 0; c < dim; c++)
      if (r != c && (re[r * dim + c] != 0.0 || im[r * dim + c] != 0.0)) return 0;
  return 1;
}

/* ------------------------------------------------------------------ generators */
#define EMIT(K, A, B, C, P0, P1, P2)          \
  do {                                        \
    if (cnt < cap) {                          \
      kind[cnt] = (K);                        \
      q[3 * cnt] = (A);                       \
      q[3 * cnt + 1] = (B);                   \
      q[3 * cnt + 2] = (C);                   \
      p[3 * cnt] = (P0);                      \
      p[3 * cnt + 1] = (P1);                  \
      p[3 * cnt + 2] = (P2);                  \
    }                                         \
    cnt++;                                    \
  } while (0)

/* ghz (circuits.rs:50-57): H(0); CX(0, q) for q = 1..n */
size_t or_gen_ghz_star(uint32_t n, int32_t* kind, uint32_t* q, double* p, size_t cap) {
  size_t cnt = 0;
  EMIT(OR_H, 0, 0, 0, 0, 0, 0);
  for (uint32_t t = 1; t < n; t++) EMIT(OR_CX, 0, t, 0, 0, 0, 0);
  return cnt;
}

/* GHZ as the chain H(0); CX(i, i+1) for i = 0..n-2 (the GHZ(20) workload). */
size_t or_gen_ghz_chain(uint32_t n, int32_t* kind, uint32_t* q, double* p, size_t cap) {
  size_t cnt = 0;
  EMIT(OR_H, 0, 0, 0, 0, 0, 0);
  for (uint32_t i = 0; i + 1 < n; i++) EMIT(OR_CX, i, i + 1, 0, 0, 0, 0);
  return cnt;
}

/* qft (circuits.rs:61-74) */
size_t or_gen_qft(uint32_t n, int32_t* kind, uint32_t* q, double* p, size_t cap) {
  size_t cnt = 0;
  for (uint32_t j = 0; j < n; j++) {
    EMIT(OR_H, j, 0, 0, 0, 0, 0);
    for (uint32_t k = j + 1; k < n; k++) {
      double lambda = OR_PI / (double)(1ull << (k - j));
      EMIT(OR_CP, k, j, 0, lambda, 0, 0);
    }
  }
  for (uint32_t j = 0; j < n / 2; j++) EMIT(OR_SWAP, j, n - 1 - j, 0, 0, 0, 0);
  return cnt;
}

/* random_circuit (circuits.rs:79-113) */
size_t or_gen_random(uint32_t n, size_t n_gates, uint64_t seed, int32_t* kind, uint32_t* q,
                     double* p, size_t cap) {
  uint64_t rng = seed;
  size_t cnt = 0;
  for (size_t g = 0; g < n_gates; g++) {
    uint32_t a = (uint32_t)or_sm64_below(&rng, n);
    uint64_t c = or_sm64_below(&rng, 15);
    if (c <= 6) {
      static const int fixed[7] = {OR_H, OR_X, OR_Y, OR_Z, OR_S, OR_T, OR_SX};
      EMIT(fixed[c], a, 0, 0, 0, 0, 0);
    } else if (c <= 10) {
      static const int rot[4] = {OR_RX, OR_RY, OR_RZ, OR_P};
      double ang = or_sm64_angle(&rng);
      EMIT(rot[c - 7], a, 0, 0, ang, 0, 0);
    } else {
      uint32_t b = (uint32_t)or_sm64_below(&rng, n);
      while (b == a) b = (uint32_t)or_sm64_below(&rng, n);
      uint64_t w = or_sm64_below(&rng, 4);
      if (w == 0) EMIT(OR_CX, a, b, 0, 0, 0, 0);
      else if (w == 1) EMIT(OR_CZ, a, b, 0, 0, 0, 0);
      else if (w == 2) EMIT(OR_SWAP, a, b, 0, 0, 0, 0);
      else {
        double ang = or_sm64_angle(&rng);
        EMIT(OR_RZZ, a, b, 0, ang, 0, 0);
      }
    }
  }
  return cnt;
}

/* Random layered circuit: per layer l, U3(theta, phi, lambda) on every qubit (angles drawn
 * in that order, qubit by qubit), RZ(angle) on every qubit, then CZ on brickwork pairs:
 * (0,1),(2,3),... for even l and (1,2),(3,4),... for odd l. */
size_t or_gen_layered(uint32_t n, uint32_t depth, uint64_t seed, int32_t* kind, uint32_t* q,
                      double* p, size_t cap) {
  uint64_t rng = seed;
  size_t cnt = 0;
  for (uint32_t l = 0; l < depth; l++) {
    for (uint32_t j = 0; j < n; j++) {
      double t = or_sm64_angle(&rng), f = or_sm64_angle(&rng), la = or_sm64_angle(&rng);
      EMIT(OR_U3, j, 0, 0, t, f, la);
    }
    for (uint32_t j = 0; j < n; j++) {
      double a = or_sm64_angle(&rng);
      EMIT(OR_RZ, j, 0, 0, a, 0, 0);
    }
    for (uint32_t j = (l & 1); j + 1 < n; j += 2) EMIT(OR_CZ, j, j + 1, 0, 0, 0, 0);
  }
  return cnt;
}

/* Hardware-efficient ansatz: per layer, RY(angle) on every qubit, RZ(angle) on every
 * qubit, then the CX ladder CX(i, i+1) for i = 0..n-2. */
size_t or_gen_hea(uint32_t n, uint32_t depth, uint64_t seed, int32_t* kind, uint32_t* q,
                  double* p, size_t cap) {
  uint64_t rng = seed;
  size_t cnt = 0;
  for (uint32_t l = 0; l < depth; l++) {
    for (uint32_t j = 0; j < n; j++) EMIT(OR_RY, j, 0, 0, or_sm64_angle(&rng), 0, 0);
    for (uint32_t j = 0; j < n; j++) EMIT(OR_RZ, j, 0, 0, or_sm64_angle(&rng), 0, 0);
    for (uint32_t j = 0; j + 1 < n; j++) EMIT(OR_CX, j, j + 1, 0, 0, 0, 0);
  }
  return cnt;
}

/* distinct random qubit other than the ones given (rejection, as circuits.rs:99-102) */
static uint32_t distinct(uint64_t* rng, uint32_t n, uint32_t a, uint32_t b, int nb) {
  for (;;) {
    uint32_t x = (uint32_t)or_sm64_below(rng, n);
    if (x != a && (nb < 2 || x != b)) return x;
  }
}

/* Gate mix for the planner workload: per gate r = below(100):
 *   r < 50: U3(theta, phi, lambda) on q = below(n)
 *   r < 80: CX (below(2) == 0) or CZ on a random distinct pair
 *   r < 95: RZZ(angle) on a random distinct pair
 *   else  : CCX (below(2) == 0) or CCZ on a random distinct triple */
size_t or_gen_mix(uint32_t n, size_t n_gates, uint64_t seed, int32_t* kind, uint32_t* q,
                  double* p, size_t cap) {
  uint64_t rng = seed;
  size_t cnt = 0;
  for (size_t g = 0; g < n_gates; g++) {
    uint64_t r = or_sm64_below(&rng, 100);
    uint32_t a = (uint32_t)or_sm64_below(&rng, n);
    if (r < 50) {
      double t = or_sm64_angle(&rng), f = or_sm64_angle(&rng), la = or_sm64_angle(&rng);
      EMIT(OR_U3, a, 0, 0, t, f, la);
    } else if (r < 80) {
      uint32_t b = distinct(&rng, n, a, 0, 1);
      int k = or_sm64_below(&rng, 2) == 0 ? OR_CX : OR_CZ;
      EMIT(k, a, b, 0, 0, 0, 0);
    } else if (r < 95) {
      uint32_t b = distinct(&rng, n, a, 0, 1);
      EMIT(OR_RZZ, a, b, 0, or_sm64_angle(&rng), 0, 0);
    } else {
      uint32_t b = distinct(&rng, n, a, 0, 1);
      uint32_t c = distinct(&rng, n, a, b, 2);
      int k = or_sm64_below(&rng, 2) == 0 ? OR_CCX : OR_CCZ;
      EMIT(k, a, b, c, 0, 0, 0);
    }
  }
  return cnt;
}

/* ------------------------------------------------------------------ validation */
/* Circuit::push checks (circuit.rs:54-78) restated as error codes. */
static int check_qubits(uint32_t n, uint32_t m, const uint32_t* qs) {
  if (m == 0 || m > 6) return -5;
  for (uint32_t i = 0; i < m; i++) {
    if (qs[i] >= n) return -2;
    for (uint32_t j = i + 1; j < m; j++)
      if (qs[i] == qs[j]) return -3;
  }
  return 0;
}

/* ------------------------------------------------------------------ RefBackend */
/* reference.rs:35-62 */
int or_ref_apply(uint32_t n, double* re, double* im, uint32_t m, const uint32_t* qs,
                 const double* gre, const double* gim) {
  int err = check_qubits(n, m, qs);
  if (err) return err;
  uint64_t dim = 1ull << n;
  uint64_t qmask = 0;
  for (uint32_t j = 0; j < m; j++) qmask |= 1ull << qs[j];
  uint64_t nonmask = ~qmask & (dim - 1);
  uint64_t gdim = 1ull << m;
  double* ore = (double*)malloc(dim * sizeof(double));
  double* oim = (double*)malloc(dim * sizeof(double));
  if (!ore || !oim) {
    free(ore);
    free(oim);
    return -6;
  }
  for (uint64_t r = 0; r < dim; r++) {
    uint64_t r_fixed = r & nonmask;
    uint64_t gr = or_gather_bits(r, qs, m);
    double ar = 0.0, ai = 0.0;
    for (uint64_t gc = 0; gc < gdim; gc++) {
      uint64_t col = r_fixed | or_scatter_bits(gc, qs, m);
      double er = gre[gr * gdim + gc], ei = gim[gr * gdim + gc];
      double xr = re[col], xi = im[col];
      double pr = er * xr - ei * xi;
      double pi = er * xi + ei * xr;
      ar = ar + pr;
      ai = ai + pi;
    }
    ore[r] = ar;
    oim[r] = ai;
  }
  memcpy(re, ore, dim * sizeof(double));
  memcpy(im, oim, dim * sizeof(double));
  free(ore);
  free(oim);
  return 0;
}

/* ------------------------------------------------------------------ CpuBackend */
/* The kernels are written once as macros over the real type R so the f64 and f32
 * instances share one text (the reference is generic over Real, real.rs:10-36). */

#define DEFINE_CPU(R, SUF)                                                                  \
  /* one_qubit_block (cpu.rs:71-94) */                                                      \
  static void one_qubit_block_##SUF(R* rl, R* il, R* rh, R* ih, uint64_t len,               \
                                    const R* g) {                                           \
    for (uint64_t k = 0; k < len; k++) {                                                    \
      R x0r = rl[k], x0i = il[k], x1r = rh[k], x1i = ih[k];                                 \
      R a_r = g[0] * x0r - g[1] * x0i, a_i = g[0] * x0i + g[1] * x0r;                       \
      R b_r = g[2] * x1r - g[3] * x1i, b_i = g[2] * x1i + g[3] * x1r;                       \
      R c_r = g[4] * x0r - g[5] * x0i, c_i = g[4] * x0i + g[5] * x0r;                       \
      R d_r = g[6] * x1r - g[7] * x1i, d_i = g[6] * x1i + g[7] * x1r;                       \
      rl[k] = a_r + b_r;                                                                    \
      il[k] = a_i + b_i;                                                                    \
      rh[k] = c_r + d_r;                                                                    \
      ih[k] = c_i + d_i;                                                                    \
    }                                                                                       \
  }                                                                                         \
  /* diag_chunk (cpu.rs:99-114) */                                                          \
  static void diag_chunk_##SUF(uint64_t base, R* rc, R* ic, uint64_t len, const R* d,       \
                               const uint32_t* qs, uint32_t m) {                            \
    for (uint64_t o = 0; o < len; o++) {                                                    \
      uint64_t s = or_gather_bits(base + o, qs, m);                                         \
      R dr = d[2 * s], di = d[2 * s + 1];                                                   \
      R xr = rc[o], xi = ic[o];                                                             \
      rc[o] = dr * xr - di * xi;                                                            \
      ic[o] = dr * xi + di * xr;                                                            \
    }                                                                                       \
  }                                                                                         \
  /* mq_block (cpu.rs:221-247) */                                                           \
  static void mq_block_##SUF(R* re, R* im, const R* g, const uint32_t* qs, uint32_t m,      \
                             uint64_t base) {                                               \
    uint64_t sub = 1ull << m;                                                               \
    R ar[MAX_SUB], ai[MAX_SUB];                                                             \
    for (uint64_t s = 0; s < sub; s++) {                                                    \
      uint64_t idx = base | or_scatter_bits(s, qs, m);                                      \
      ar[s] = re[idx];                                                                      \
      ai[s] = im[idx];                                                                      \
    }                                                                                       \
    for (uint64_t r = 0; r < sub; r++) {                                                    \
      R accr = 0, acci = 0;                                                                 \
      const R* row = g + 2 * r * sub;                                                       \
      for (uint64_t c = 0; c < sub; c++) {                                                  \
        R er = row[2 * c], ei = row[2 * c + 1];                                             \
        R pr = er * ar[c] - ei * ai[c];                                                     \
        R pi = er * ai[c] + ei * ar[c];                                                     \
        accr = accr + pr;                                                                   \
        acci = acci + pi;                                                                   \
      }                                                                                     \
      uint64_t idx = base | or_scatter_bits(r, qs, m);                                      \
      re[idx] = accr;                                                                       \
      im[idx] = acci;                                                                       \
    }                                                                                       \
  }                                                                                         \
  static int cpu_apply_##SUF(uint32_t n, R* re, R* im, uint32_t m, const uint32_t* qs,      \
                             const double* gre, const double* gim, int parallel) {          \
    int err = check_qubits(n, m, qs);                                                       \
    if (err) return err;                                                                    \
    uint64_t dim = 1ull << n, sub = 1ull << m;                                              \
    /* gate entries in R precision, interleaved (the gate is DenseGate<R>) */              \
    R* g = (R*)malloc(2 * sub * sub * sizeof(R));                                           \
    if (!g) return -6;                                                                      \
    for (uint64_t i = 0; i < sub * sub; i++) {                                              \
      g[2 * i] = (R)gre[i];                                                                 \
      g[2 * i + 1] = (R)gim[i];                                                             \
    }                                                                                       \
    if (or_is_diagonal(m, gre, gim)) { /* cpu.rs:117-146 */                                 \
      R d[2 * MAX_SUB];                                                                     \
      for (uint64_t s = 0; s < sub; s++) {                                                  \
        d[2 * s] = g[2 * (s * sub + s)];                                                    \
        d[2 * s + 1] = g[2 * (s * sub + s) + 1];                                            \
      }                                                                                     \
      if (parallel && dim >= (MIN_PAIRS_LIGHT << 1)) {                                      \
        uint64_t nth = (uint64_t)or_num_threads();                                          \
        uint64_t chunk = dim / (4 * nth);                                                   \
        if (chunk < 1024) chunk = 1024;                                                     \
        int64_t nchunks = (int64_t)((dim + chunk - 1) / chunk);                             \
        _Pragma("omp parallel for schedule(dynamic, 1)")                                    \
        for (int64_t c = 0; c < nchunks; c++) {                                             \
          uint64_t b0 = (uint64_t)c * chunk;                                                \
          uint64_t len = dim - b0 < chunk ? dim - b0 : chunk;                               \
          diag_chunk_##SUF(b0, re + b0, im + b0, len, d, qs, m);                            \
        }                                                                                   \
      } else {                                                                              \
        diag_chunk_##SUF(0, re, im, dim, d, qs, m);                                         \
      }                                                                                     \
    } else if (m == 1) { /* cpu.rs:148-175 */                                               \
      uint64_t half = 1ull << qs[0], block = half << 1;                                     \
      int64_t nblocks = (int64_t)(dim / block);                                             \
      if (parallel && (dim >> 1) >= MIN_PAIRS_LIGHT) {                                      \
        /* rayon splits the chunk list into contiguous ranges: a static schedule */        \
        _Pragma("omp parallel for schedule(static)")                                        \
        for (int64_t b = 0; b < nblocks; b++) {                                             \
          uint64_t o = (uint64_t)b * block;                                                 \
          one_qubit_block_##SUF(re + o, im + o, re + o + half, im + o + half, half, g);     \
        }                                                                                   \
      } else {                                                                              \
        for (int64_t b = 0; b < nblocks; b++) {                                             \
          uint64_t o = (uint64_t)b * block;                                                 \
          one_qubit_block_##SUF(re + o, im + o, re + o + half, im + o + half, half, g);     \
        }                                                                                   \
      }                                                                                     \
    } else { /* cpu.rs:177-213 */                                                           \
      uint32_t sorted[6];                                                                   \
      memcpy(sorted, qs, m * sizeof(uint32_t));                                             \
      for (uint32_t i = 1; i < m; i++)                                                      \
        for (uint32_t j = i; j > 0 && sorted[j - 1] > sorted[j]; j--) {                     \
          uint32_t t = sorted[j];                                                           \
          sorted[j] = sorted[j - 1];                                                        \
          sorted[j - 1] = t;                                                                \
        }                                                                                   \
      int64_t blocks = (int64_t)(dim >> m);                                                 \
      if (parallel && (dim >> 1) >= MIN_PAIRS_HEAVY && blocks > 1) {                        \
        _Pragma("omp parallel for schedule(static)")                                        \
        for (int64_t o = 0; o < blocks; o++)                                                \
          mq_block_##SUF(re, im, g, qs, m, or_insert_zero_bits((uint64_t)o, sorted, m));    \
      } else {                                                                              \
        for (int64_t o = 0; o < blocks; o++)                                                \
          mq_block_##SUF(re, im, g, qs, m, or_insert_zero_bits((uint64_t)o, sorted, m));    \
      }                                                                                     \
    }                                                                                       \
    free(g);                                                                                \
    return 0;                                                                               \
  }                                                                                         \
  /* StateVector::zeros (+ first_touch) and basis (state/mod.rs:33-73) */                   \
  static void init_basis_##SUF(uint32_t n, R* re, R* im, uint64_t basis, int parallel) {    \
    int64_t dim = (int64_t)(1ull << n);                                                     \
    if (parallel && (uint64_t)dim >= FIRST_TOUCH_MIN) {                                     \
      _Pragma("omp parallel for schedule(static)")                                          \
      for (int64_t i = 0; i < dim; i++) {                                                   \
        re[i] = 0;                                                                          \
        im[i] = 0;                                                                          \
      }                                                                                     \
    } else {                                                                                \
      memset(re, 0, (size_t)dim * sizeof(R));                                               \
      memset(im, 0, (size_t)dim * sizeof(R));                                               \
    }                                                                                       \
    re[basis] = (R)1;                                                                       \
  }

DEFINE_CPU(double, f64)
DEFINE_CPU(float, f32)

int or_cpu_apply_f64(uint32_t n, double* re, double* im, uint32_t m, const uint32_t* qs,
                     const double* gre, const double* gim, int parallel) {
  return cpu_apply_f64(n, re, im, m, qs, gre, gim, parallel);
}
int or_cpu_apply_f32(uint32_t n, float* re, float* im, uint32_t m, const uint32_t* qs,
                     const double* gre, const double* gim, int parallel) {
  return cpu_apply_f32(n, re, im, m, qs, gre, gim, parallel);
}
void or_init_basis_f64(uint32_t n, double* re, double* im, uint64_t basis, int parallel) {
  init_basis_f64(n, re, im, basis, parallel);
}
void or_init_basis_f32(uint32_t n, float* re, float* im, uint64_t basis, int parallel) {
  init_basis_f32(n, re, im, basis, parallel);
}

/* backend/mod.rs:66-73 (the op loop; init is separate so uploaded states can be run) */
int or_run_f64(int backend, uint32_t n, double* re, double* im, size_t n_ops,
               const uint32_t* arity, const uint64_t* qoff, const uint32_t* qubits,
               const uint64_t* moff, const double* mre, const double* mim) {
  for (size_t i = 0; i < n_ops; i++) {
    int err;
    if (backend == 0)
      err = or_ref_apply(n, re, im, arity[i], qubits + qoff[i], mre + moff[i], mim + moff[i]);
    else
      err = cpu_apply_f64(n, re, im, arity[i], qubits + qoff[i], mre + moff[i],
                          mim + moff[i], backend == 2);
    if (err) return err;
  }
  return 0;
}

int or_run_f32(int parallel, uint32_t n, float* re, float* im, size_t n_ops,
               const uint32_t* arity, const uint64_t* qoff, const uint32_t* qubits,
               const uint64_t* moff, const double* mre, const double* mim) {
  for (size_t i = 0; i < n_ops; i++) {
    int err = cpu_apply_f32(n, re, im, arity[i], qubits + qoff[i], mre + moff[i],
                            mim + moff[i], parallel);
    if (err) return err;
  }
  return 0;
}

/* ------------------------------------------------------------------ fusion */
static int contains(const uint32_t* xs, uint32_t n, uint32_t v) {
  for (uint32_t i = 0; i < n; i++)
    if (xs[i] == v) return 1;
  return 0;
}

/* fuse (fusion.rs:39-77) + the group bookkeeping of flush_group (fusion.rs:79-99) */
size_t or_fuse_groups(size_t n_ops, const uint32_t* arity, const uint64_t* qoff,
                      const uint32_t* qubits, uint32_t max_qubits, uint64_t* grp_begin,
                      uint64_t* grp_end, uint32_t* gm, uint32_t* gq) {
  uint32_t max = max_qubits < 1 ? 1 : max_qubits;
  size_t ng = 0;
  uint32_t cur[64];
  uint32_t ncur = 0;
  uint64_t begin = 0, count = 0;
#define FLUSH()                                                                  \
  do {                                                                           \
    if (count > 0) {                                                             \
      grp_begin[ng] = begin;                                                     \
      grp_end[ng] = begin + count;                                               \
      gm[ng] = ncur;                                                             \
      memcpy(gq + 8 * ng, cur, ncur * sizeof(uint32_t));                         \
      if (count > 1) { /* fused groups land on sorted qubits (fusion.rs:91-94) */ \
        uint32_t* s = gq + 8 * ng;                                               \
        for (uint32_t a = 1; a < ncur; a++)                                      \
          for (uint32_t b = a; b > 0 && s[b - 1] > s[b]; b--) {                  \
            uint32_t t = s[b];                                                   \
            s[b] = s[b - 1];                                                     \
            s[b - 1] = t;                                                        \
          }                                                                      \
      }                                                                          \
      ng++;                                                                      \
    }                                                                            \
    ncur = 0;                                                                    \
    count = 0;                                                                   \
  } while (0)
  for (size_t i = 0; i < n_ops; i++) {
    const uint32_t* qs = qubits + qoff[i];
    uint32_t m = arity[i];
    if (m > max) { /* fusion.rs:49-54: wider than the cap, emitted standalone */
      FLUSH();
      begin = i;
      count = 1;
      ncur = m;
      memcpy(cur, qs, m * sizeof(uint32_t));
      FLUSH();
      continue;
    }
    uint32_t uni[64];
    uint32_t nu = ncur;
    memcpy(uni, cur, ncur * sizeof(uint32_t));
    for (uint32_t j = 0; j < m; j++)
      if (!contains(uni, nu, qs[j])) uni[nu++] = qs[j];
    if (count == 0) {
      begin = i;
      count = 1;
      ncur = m;
      memcpy(cur, qs, m * sizeof(uint32_t));
    } else if (nu <= max) {
      count++;
      ncur = nu;
      memcpy(cur, uni, nu * sizeof(uint32_t));
    } else {
      FLUSH();
      begin = i;
      count = 1;
      ncur = m;
      memcpy(cur, qs, m * sizeof(uint32_t));
    }
  }
  FLUSH();
#undef FLUSH
  return ng;
}

/* build_fused (fusion.rs:102-139): mat = E(g_last) ... E(g_first) * I, in f64, skipping
 * zero gate entries so that products of diagonal gates stay exactly diagonal. */
int or_build_fused(uint32_t k, const uint32_t* q, uint64_t b, uint64_t e,
                   const uint32_t* arity, const uint64_t* qoff, const uint32_t* qubits,
                   const uint64_t* moff, const double* mre, const double* mim, double* ore,
                   double* oim) {
  uint64_t dim = 1ull << k;
  double* nr = (double*)calloc(dim * dim, sizeof(double));
  double* ni = (double*)calloc(dim * dim, sizeof(double));
  if (!nr || !ni) {
    free(nr);
    free(ni);
    return -6;
  }
  memset(ore, 0, dim * dim * sizeof(double));
  memset(oim, 0, dim * dim * sizeof(double));
  for (uint64_t i = 0; i < dim; i++) ore[i * dim + i] = 1.0;
  for (uint64_t op = b; op < e; op++) {
    uint32_t m = arity[op];
    const uint32_t* qs = qubits + qoff[op];
    const double* gr = mre + moff[op];
    const double* gi = mim + moff[op];
    uint32_t local[6];
    for (uint32_t j = 0; j < m; j++) {
      local[j] = 0xFFFFFFFFu;
      for (uint32_t t = 0; t < k; t++)
        if (q[t] == qs[j]) local[j] = t;
      if (local[j] == 0xFFFFFFFFu) {
        free(nr);
        free(ni);
        return -2;
      }
    }
    uint64_t nsub = 1ull << m;
    memset(nr, 0, dim * dim * sizeof(double));
    memset(ni, 0, dim * dim * sizeof(double));
    for (uint64_t r = 0; r < dim; r++) {
      uint64_t r_sub = 0;
      for (uint32_t j = 0; j < m; j++) r_sub |= ((r >> local[j]) & 1ull) << j;
      for (uint64_t ks = 0; ks < nsub; ks++) {
        double er = gr[r_sub * nsub + ks], ei = gi[r_sub * nsub + ks];
        if (er == 0.0 && ei == 0.0) continue;
        uint64_t kk = r;
        for (uint32_t j = 0; j < m; j++) {
          uint64_t bit = (ks >> j) & 1ull;
          kk = (kk & ~(1ull << local[j])) | (bit << local[j]);
        }
        for (uint64_t c = 0; c < dim; c++) {
          double xr = ore[kk * dim + c], xi = oim[kk * dim + c];
          double pr = er * xr - ei * xi;
          double pi = er * xi + ei * xr;
          nr[r * dim + c] = nr[r * dim + c] + pr;
          ni[r * dim + c] = ni[r * dim + c] + pi;
        }
      }
    }
    memcpy(ore, nr, dim * dim * sizeof(double));
    memcpy(oim, ni, dim * dim * sizeof(double));
  }
  free(nr);
  free(ni);
  return 0;
}

/* ------------------------------------------------------------------ reductions */
/* StateVector::norm_sqr (state/mod.rs:112-117): serial accumulation */
double or_norm_sqr_f64(uint64_t dim, const double* re, const double* im) {
  double s = 0.0;
  for (uint64_t i = 0; i < dim; i++) s = s + re[i] * re[i] + im[i] * im[i];
  return s;
}
/* StateVector::probabilities (state/mod.rs:130-134) */
void or_probabilities_f64(uint64_t dim, const double* re, const double* im, double* out) {
  for (uint64_t i = 0; i < dim; i++) out[i] = re[i] * re[i] + im[i] * im[i];
}
/* sample::expectation_z (sample.rs:177-200), serial path */
double or_expectation_z_f64(uint64_t dim, const double* probs, uint64_t mask) {
  double s = 0.0;
  for (uint64_t i = 0; i < dim; i++) {
    double v = probs[i];
    s += (__builtin_popcountll(i & mask) & 1) ? -v : v;
  }
  return s;
}

/* ---- sampling (sample.rs:33-173) ------------------------------------------------- */
void or_prefix_sum_f64(uint64_t dim, const double* p, double* out) {
  double acc = 0.0;
  for (uint64_t i = 0; i < dim; i++) {
    acc += p[i];
    out[i] = acc;
  }
}

void or_cdf_sample(uint64_t dim, const double* p, uint64_t shots, uint64_t* state,
                   uint64_t* out) {
  double* cdf = (double*)malloc(sizeof(double) * (dim ? dim : 1));
  or_prefix_sum_f64(dim, p, cdf);
  const double total = dim ? cdf[dim - 1] : 0.0;
  const uint64_t last = dim ? dim - 1 : 0;
  for (uint64_t k = 0; k < shots; k++) {
    const double u = or_sm64_unit(state) * total;
    /* partition_point(|c| c < u): first index whose cdf is >= u */
    uint64_t lo = 0, hi = dim;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (cdf[mid] < u)
        lo = mid + 1;
      else
        hi = mid;
    }
    out[k] = lo < last ? lo : last;
  }
  free(cdf);
}

void or_alias_build(uint64_t dim, const double* p, double* prob, uint64_t* alias) {
  double total = 0.0;
  for (uint64_t i = 0; i < dim; i++) total += p[i];
  const double scale = total > 0.0 ? (double)dim / total : 0.0;
  double* scaled = (double*)malloc(sizeof(double) * dim);
  uint64_t* small = (uint64_t*)malloc(sizeof(uint64_t) * dim);
  uint64_t* large = (uint64_t*)malloc(sizeof(uint64_t) * dim);
  uint64_t ns = 0, nl = 0;
  for (uint64_t i = 0; i < dim; i++) {
    scaled[i] = p[i] * scale;
    prob[i] = 0.0;
    alias[i] = 0;
    if (scaled[i] < 1.0)
      small[ns++] = i;
    else
      large[nl++] = i;
  }
  while (ns > 0 && nl > 0) {
    const uint64_t sm = small[--ns], lg = large[--nl];
    prob[sm] = scaled[sm];
    alias[sm] = lg;
    scaled[lg] = (scaled[lg] + scaled[sm]) - 1.0;
    if (scaled[lg] < 1.0)
      small[ns++] = lg;
    else
      large[nl++] = lg;
  }
  while (nl > 0) prob[large[--nl]] = 1.0;
  while (ns > 0) prob[small[--ns]] = 1.0;
  free(scaled);
  free(small);
  free(large);
}

void or_alias_sample(uint64_t dim, const double* prob, const uint64_t* alias, uint64_t shots,
                     uint64_t* state, uint64_t* out) {
  for (uint64_t k = 0; k < shots; k++) {
    uint64_t i = (uint64_t)(or_sm64_unit(state) * (double)dim);
    if (i > dim - 1) i = dim - 1;
    out[k] = or_sm64_unit(state) < prob[i] ? i : alias[i];
  }
}
