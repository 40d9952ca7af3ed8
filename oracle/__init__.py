"""ctypes view of the CPU oracle (``qsv_oracle.c``) -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may import
this package, and only as the checker or the timed CPU baseline; the HIP engine never
routes through it.

Circuits are described here as "gate records" ``(kind, qubits, params)`` with the kind
numbering of ``qsv_oracle.h``; :func:`flatten` turns a record list into the flat arrays
the C backends take.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build

KINDS = ["h", "x", "y", "z", "s", "t", "sx", "rx", "ry", "rz", "p",
         "cx", "cz", "cp", "swap", "rzz", "u3", "ccx", "ccz"]
KIND = {k: i for i, k in enumerate(KINDS)}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_build.LIB):
            _build.build()
        L = C.CDLL(_build.LIB)
        u32, u64, i32, f64 = C.c_uint32, C.c_uint64, C.c_int32, C.c_double
        P = C.c_void_p
        L.or_sm64_next.restype = u64
        L.or_sm64_next.argtypes = [P]
        L.or_sm64_below.restype = u64
        L.or_sm64_below.argtypes = [P, u64]
        L.or_sm64_unit.restype = f64
        L.or_sm64_unit.argtypes = [P]
        L.or_sm64_angle.restype = f64
        L.or_sm64_angle.argtypes = [P]
        L.or_insert_zero_bit.restype = u64
        L.or_insert_zero_bit.argtypes = [u64, u32]
        L.or_gather_bits.restype = u64
        L.or_gather_bits.argtypes = [u64, P, u32]
        L.or_scatter_bits.restype = u64
        L.or_scatter_bits.argtypes = [u64, P, u32]
        L.or_insert_zero_bits.restype = u64
        L.or_insert_zero_bits.argtypes = [u64, P, u32]
        L.or_gate_arity.restype = u32
        L.or_gate_arity.argtypes = [C.c_int]
        L.or_gate_matrix.restype = u32
        L.or_gate_matrix.argtypes = [C.c_int, P, P, P]
        L.or_is_diagonal.restype = C.c_int
        L.or_is_diagonal.argtypes = [u32, P, P]
        for name in ["or_gen_ghz_star", "or_gen_ghz_chain", "or_gen_qft"]:
            getattr(L, name).restype = C.c_size_t
            getattr(L, name).argtypes = [u32, P, P, P, C.c_size_t]
        for name in ["or_gen_random", "or_gen_mix"]:
            getattr(L, name).restype = C.c_size_t
            getattr(L, name).argtypes = [u32, C.c_size_t, u64, P, P, P, C.c_size_t]
        for name in ["or_gen_layered", "or_gen_hea"]:
            getattr(L, name).restype = C.c_size_t
            getattr(L, name).argtypes = [u32, u32, u64, P, P, P, C.c_size_t]
        L.or_ref_apply.restype = C.c_int
        L.or_ref_apply.argtypes = [u32, P, P, u32, P, P, P]
        L.or_cpu_apply_f64.restype = C.c_int
        L.or_cpu_apply_f64.argtypes = [u32, P, P, u32, P, P, P, C.c_int]
        L.or_cpu_apply_f32.restype = C.c_int
        L.or_cpu_apply_f32.argtypes = [u32, P, P, u32, P, P, P, C.c_int]
        L.or_init_basis_f64.restype = None
        L.or_init_basis_f64.argtypes = [u32, P, P, u64, C.c_int]
        L.or_init_basis_f32.restype = None
        L.or_init_basis_f32.argtypes = [u32, P, P, u64, C.c_int]
        L.or_run_f64.restype = C.c_int
        L.or_run_f64.argtypes = [C.c_int, u32, P, P, C.c_size_t, P, P, P, P, P, P]
        L.or_run_f32.restype = C.c_int
        L.or_run_f32.argtypes = [C.c_int, u32, P, P, C.c_size_t, P, P, P, P, P, P]
        L.or_fuse_groups.restype = C.c_size_t
        L.or_fuse_groups.argtypes = [C.c_size_t, P, P, P, u32, P, P, P, P]
        L.or_build_fused.restype = C.c_int
        L.or_build_fused.argtypes = [u32, P, u64, u64, P, P, P, P, P, P, P, P]
        L.or_norm_sqr_f64.restype = f64
        L.or_norm_sqr_f64.argtypes = [u64, P, P]
        L.or_probabilities_f64.restype = None
        L.or_probabilities_f64.argtypes = [u64, P, P, P]
        L.or_expectation_z_f64.restype = f64
        L.or_expectation_z_f64.argtypes = [u64, P, u64]
        L.or_prefix_sum_f64.argtypes = [u64, P, P]
        L.or_cdf_sample.argtypes = [u64, P, u64, P, P]
        L.or_alias_build.argtypes = [u64, P, P, P]
        L.or_alias_sample.argtypes = [u64, P, P, u64, P, P]
        L.or_num_threads.restype = C.c_int
        L.or_num_threads.argtypes = []
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------- RNG / index math
class SplitMix64:
    """crates/qsv-core/src/circuits.rs:12-47 via the C restatement."""

    def __init__(self, seed: int):
        self.s = np.array([seed], dtype=np.uint64)

    def next_u64(self) -> int:
        return int(lib().or_sm64_next(_p(self.s)))

    def below(self, n: int) -> int:
        return int(lib().or_sm64_below(_p(self.s), n))

    def unit(self) -> float:
        return float(lib().or_sm64_unit(_p(self.s)))

    def angle(self) -> float:
        return float(lib().or_sm64_angle(_p(self.s)))


def insert_zero_bit(i, b):
    return int(lib().or_insert_zero_bit(i, b))


def gather_bits(i, qs):
    a = np.asarray(qs, dtype=np.uint32)
    return int(lib().or_gather_bits(i, _p(a), len(a)))


def scatter_bits(s, qs):
    a = np.asarray(qs, dtype=np.uint32)
    return int(lib().or_scatter_bits(s, _p(a), len(a)))


def insert_zero_bits(i, sorted_qs):
    a = np.asarray(sorted_qs, dtype=np.uint32)
    return int(lib().or_insert_zero_bits(i, _p(a), len(a)))


# ---------------------------------------------------------------- gates / circuits
def gate_matrix(kind, params=()):
    """Row-major complex128 matrix of a gate kind (name or number)."""
    k = KIND[kind] if isinstance(kind, str) else int(kind)
    p = np.zeros(3)
    p[: len(params)] = params
    re = np.zeros(64)
    im = np.zeros(64)
    m = lib().or_gate_matrix(k, _p(p), _p(re), _p(im))
    if m == 0:
        raise ValueError(f"unknown gate kind {kind}")
    d = 1 << m
    return (re[: d * d] + 1j * im[: d * d]).reshape(d, d)


def is_diagonal(mat: np.ndarray) -> bool:
    m = int(np.log2(mat.shape[0]))
    re = np.ascontiguousarray(mat.real).ravel()
    im = np.ascontiguousarray(mat.imag).ravel()
    return bool(lib().or_is_diagonal(m, _p(re), _p(im)))


def _gen(fn, *args):
    L = lib()
    f = getattr(L, fn)
    n = f(*args, None, None, None, 0)
    kind = np.zeros(n, dtype=np.int32)
    q = np.zeros(3 * n, dtype=np.uint32)
    p = np.zeros(3 * n, dtype=np.float64)
    f(*args, _p(kind), _p(q), _p(p), n)
    recs = []
    for i in range(n):
        k = int(kind[i])
        ar = int(L.or_gate_arity(k))
        recs.append((KINDS[k], tuple(int(x) for x in q[3 * i: 3 * i + ar]),
                     tuple(float(x) for x in p[3 * i: 3 * i + 3])))
    return recs


def gen_ghz_star(n):
    return _gen("or_gen_ghz_star", n)


def gen_ghz_chain(n):
    return _gen("or_gen_ghz_chain", n)


def gen_qft(n):
    return _gen("or_gen_qft", n)


def gen_random(n, n_gates, seed):
    return _gen("or_gen_random", n, n_gates, seed)


def gen_layered(n, depth, seed):
    return _gen("or_gen_layered", n, depth, seed)


def gen_hea(n, depth, seed):
    return _gen("or_gen_hea", n, depth, seed)


def gen_mix(n, n_gates, seed):
    return _gen("or_gen_mix", n, n_gates, seed)


class Flat:
    """A circuit as flat arrays (the oracle's input format)."""

    def __init__(self, ops):
        """ops: list of (matrix (2^m x 2^m complex), qubits)."""
        self.n_ops = len(ops)
        self.arity = np.array([len(q) for _, q in ops], dtype=np.uint32)
        self.qoff = np.zeros(self.n_ops, dtype=np.uint64)
        self.moff = np.zeros(self.n_ops, dtype=np.uint64)
        qs, mr, mi = [], [], []
        qo = mo = 0
        for i, (mat, q) in enumerate(ops):
            self.qoff[i] = qo
            self.moff[i] = mo
            qs.extend(q)
            qo += len(q)
            mat = np.asarray(mat, dtype=np.complex128)
            mr.append(mat.real.ravel())
            mi.append(mat.imag.ravel())
            mo += mat.size
        self.qubits = np.array(qs if qs else [0], dtype=np.uint32)
        self.mre = np.ascontiguousarray(np.concatenate(mr) if mr else np.zeros(1))
        self.mim = np.ascontiguousarray(np.concatenate(mi) if mi else np.zeros(1))
        self.ops = ops

    def args(self):
        return (self.n_ops, _p(self.arity), _p(self.qoff), _p(self.qubits),
                _p(self.moff), _p(self.mre), _p(self.mim))


def records_to_ops(recs):
    return [(gate_matrix(k, p), q) for k, q, p in recs]


def flatten(ops_or_recs) -> Flat:
    if ops_or_recs and isinstance(ops_or_recs[0][0], str):
        ops_or_recs = records_to_ops(ops_or_recs)
    return Flat(list(ops_or_recs))


# ---------------------------------------------------------------- backends
REF, CPU_SERIAL, CPU_PARALLEL = 0, 1, 2


def basis(n, index=0, dtype=np.float64, parallel=True):
    re = np.empty(1 << n, dtype=dtype)
    im = np.empty(1 << n, dtype=dtype)
    f = lib().or_init_basis_f64 if dtype == np.float64 else lib().or_init_basis_f32
    f(n, _p(re), _p(im), index, 1 if parallel else 0)
    return re, im


def run(backend, n, flat: Flat, re, im):
    """Apply the circuit in place (Backend::execute's op loop)."""
    if re.dtype == np.float64:
        err = lib().or_run_f64(backend, n, _p(re), _p(im), *flat.args())
    else:
        err = lib().or_run_f32(1 if backend == CPU_PARALLEL else 0, n, _p(re), _p(im),
                               *flat.args())
    if err:
        raise ValueError(f"oracle run failed: {err}")
    return re, im


def execute(backend, n, ops_or_recs, dtype=np.float64, start=None):
    """Backend::execute (crates/qsv-core/src/backend/mod.rs:66-73): |0> (or the given
    start state) then every op."""
    flat = flatten(ops_or_recs)
    if start is None:
        re, im = basis(n, 0, dtype)
    else:
        start = np.asarray(start)
        re = np.ascontiguousarray(start.real, dtype=dtype).copy()
        im = np.ascontiguousarray(start.imag, dtype=dtype).copy()
    run(backend, n, flat, re, im)
    return re + 1j * im


def apply(backend, n, state: np.ndarray, mat, qubits):
    re = np.ascontiguousarray(state.real).astype(np.float64)
    im = np.ascontiguousarray(state.imag).astype(np.float64)
    mat = np.asarray(mat, dtype=np.complex128)
    q = np.asarray(qubits, dtype=np.uint32)
    mr = np.ascontiguousarray(mat.real).ravel()
    mi = np.ascontiguousarray(mat.imag).ravel()
    if backend == REF:
        err = lib().or_ref_apply(n, _p(re), _p(im), len(q), _p(q), _p(mr), _p(mi))
    else:
        err = lib().or_cpu_apply_f64(n, _p(re), _p(im), len(q), _p(q), _p(mr), _p(mi),
                                     1 if backend == CPU_PARALLEL else 0)
    if err:
        raise ValueError(f"oracle apply failed: {err}")
    return re + 1j * im


# ---------------------------------------------------------------- fusion
def fuse_groups(flat: Flat, max_qubits: int):
    """fusion::fuse grouping: list of (begin, end, qubits)."""
    n = flat.n_ops
    gb = np.zeros(max(n, 1), dtype=np.uint64)
    ge = np.zeros(max(n, 1), dtype=np.uint64)
    gm = np.zeros(max(n, 1), dtype=np.uint32)
    gq = np.zeros(8 * max(n, 1), dtype=np.uint32)
    ng = lib().or_fuse_groups(n, _p(flat.arity), _p(flat.qoff), _p(flat.qubits), max_qubits,
                              _p(gb), _p(ge), _p(gm), _p(gq))
    return [(int(gb[g]), int(ge[g]), tuple(int(x) for x in gq[8 * g: 8 * g + gm[g]]))
            for g in range(ng)]


def build_fused(flat: Flat, begin, end, qubits):
    k = len(qubits)
    q = np.asarray(qubits, dtype=np.uint32)
    d = 1 << k
    ore = np.zeros(d * d)
    oim = np.zeros(d * d)
    err = lib().or_build_fused(k, _p(q), begin, end, _p(flat.arity), _p(flat.qoff),
                               _p(flat.qubits), _p(flat.moff), _p(flat.mre), _p(flat.mim),
                               _p(ore), _p(oim))
    if err:
        raise ValueError(f"build_fused failed: {err}")
    return (ore + 1j * oim).reshape(d, d)


def fuse(flat: Flat, max_qubits: int):
    """fusion::fuse (crates/qsv-core/src/fusion.rs:39-77): the fused op list."""
    out = []
    for b, e, q in fuse_groups(flat, max_qubits):
        if e - b == 1:
            out.append(flat.ops[b])
        else:
            out.append((build_fused(flat, b, e, q), q))
    return out


# ---------------------------------------------------------------- reductions
def norm_sqr(state):
    re = np.ascontiguousarray(state.real)
    im = np.ascontiguousarray(state.imag)
    return float(lib().or_norm_sqr_f64(len(re), _p(re), _p(im)))


def probabilities(state):
    re = np.ascontiguousarray(state.real)
    im = np.ascontiguousarray(state.imag)
    out = np.empty(len(re))
    lib().or_probabilities_f64(len(re), _p(re), _p(im), _p(out))
    return out


def expectation_z(probs, mask):
    p = np.ascontiguousarray(probs, dtype=np.float64)
    return float(lib().or_expectation_z_f64(len(p), _p(p), mask))


def prefix_sum(p):
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.empty_like(p)
    lib().or_prefix_sum_f64(len(p), _p(p), _p(out))
    return out


def cdf_sample(p, shots, seed):
    """cdf_sample (sample.rs:140-153) with SplitMix64(seed); returns (outcomes, end state)."""
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.empty(shots, dtype=np.uint64)
    st = C.c_uint64(seed)
    lib().or_cdf_sample(len(p), _p(p), shots, C.byref(st), _p(out))
    return out, st.value


def alias_build(p):
    p = np.ascontiguousarray(p, dtype=np.float64)
    prob = np.empty(len(p))
    alias = np.empty(len(p), dtype=np.uint64)
    lib().or_alias_build(len(p), _p(p), _p(prob), _p(alias))
    return prob, alias


def alias_sample(p, shots, seed):
    """alias_sample (sample.rs:102-105) with SplitMix64(seed); returns (outcomes, end state)."""
    prob, alias = alias_build(p)
    out = np.empty(shots, dtype=np.uint64)
    st = C.c_uint64(seed)
    lib().or_alias_sample(len(prob), _p(prob), _p(alias), shots, C.byref(st), _p(out))
    return out, st.value


def num_threads():
    return int(lib().or_num_threads())
