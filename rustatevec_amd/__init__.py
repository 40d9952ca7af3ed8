"""rustatevec_amd -- Python view of the MI355X statevector engine (libqsvh.so).

The engine itself is C++/HIP behind the C ABI in ``include/qsvh.h``; this module binds
that ABI with ctypes and mirrors the reference's host interface for the path
(nez0b/rustatevec, crates/qsv-core):

* :class:`Circuit` -- ``Circuit<R>`` and its fluent builders (circuit.rs:30-131), plus
  the generators of circuits.rs:50-113 and the measured workloads;
* :func:`fuse` -- ``fusion::fuse`` (fusion.rs:39-77) with device-side f64 composition;
* :class:`HipBackend` -- the ``Backend<R>`` trait (backend/mod.rs:31-74): ``alloc``,
  ``init_basis``, ``apply``, ``amplitude``, ``probabilities``, ``expectation_z``,
  ``download``, ``execute``; plus ``upload``, ``norm_sqr`` and replayable programs.

There is no CPU fallback: if the library is missing, importing the backend raises.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from typing import Iterable, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqsvh.so")

C128, C64 = 0, 1
PREC_EXACT, PREC_MFMA_F32, PREC_MFMA_BF16X3 = 0, 1, 2

STATUS = {
    0: "OK", 1: "INVALID_ARG", 2: "INVALID_QUBIT", 3: "REPEATED_QUBIT", 4: "ARITY",
    5: "TOO_WIDE", 6: "OOM", 7: "HIP_ERROR", 8: "DTYPE", 9: "NO_DEVICE", 10: "PARSE",
}


class QsvhError(RuntimeError):
    """A non-OK status from the C ABI; ``.status`` holds the name."""

    def __init__(self, code: int, msg: str = ""):
        self.code = code
        self.status = STATUS.get(code, str(code))
        super().__init__(f"{self.status}: {msg}" if msg else self.status)


class _ProgramOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("use_graph", C.c_int), ("tile_qubits", C.c_int),
                ("subspace", C.c_int), ("mfma_tiles", C.c_int), ("planned", C.c_int),
                ("reserved", C.c_int * 2)]


_lib = None


def lib() -> C.CDLL:
    """Load libqsvh.so (built by ``rustatevec_amd/build.py``).  Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with rustatevec_amd/build.py "
                          "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P, u32, u64, i32, dbl = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_double
    PP = C.POINTER(C.c_void_p)
    sig = {
        "qsvh_abi_version": (i32, []),
        "qsvh_ctx_create": (i32, [i32, PP]),
        "qsvh_ctx_destroy": (None, [P]),
        "qsvh_synchronize_ctx": (i32, [P]),
        "qsvh_last_error": (C.c_char_p, [P]),
        "qsvh_tune_set": (i32, [P, C.c_char_p, C.c_int64]),
        "qsvh_tune_get": (i32, [P, C.c_char_p, C.POINTER(C.c_int64)]),
        "qsvh_device_info": (i32, [P, C.c_char_p, C.c_size_t, C.POINTER(C.c_int),
                                   C.POINTER(C.c_uint64)]),
        "qsvh_gate_matrix": (i32, [C.c_char_p, P, u32, P, P, C.POINTER(C.c_uint32)]),
        "qsvh_circuit_create": (i32, [u32, PP]),
        "qsvh_circuit_destroy": (None, [P]),
        "qsvh_circuit_push": (i32, [P, u32, P, P, P]),
        "qsvh_circuit_push_gate": (i32, [P, C.c_char_p, P, P, u32]),
        "qsvh_circuit_from_qasm": (i32, [C.c_char_p, C.c_size_t, P, C.c_char_p, C.c_size_t]),
        "qsvh_circuit_n_qubits": (u32, [P]),
        "qsvh_circuit_len": (C.c_size_t, [P]),
        "qsvh_circuit_op": (i32, [P, C.c_size_t, C.POINTER(C.c_uint32), P, P, P]),
        "qsvh_circuit_ghz": (i32, [u32, PP]),
        "qsvh_circuit_ghz_chain": (i32, [u32, PP]),
        "qsvh_circuit_qft": (i32, [u32, PP]),
        "qsvh_circuit_random": (i32, [u32, u64, u64, PP]),
        "qsvh_circuit_layered": (i32, [u32, u32, u64, PP]),
        "qsvh_circuit_hea": (i32, [u32, u32, u64, PP]),
        "qsvh_circuit_mix": (i32, [u32, u64, u64, PP]),
        "qsvh_fuse": (i32, [P, P, u32, PP]),
        "qsvh_fuse_groups": (i32, [P, u32, P, P, P, P, C.c_size_t, C.POINTER(C.c_size_t)]),
        "qsvh_state_alloc": (i32, [P, u32, i32, PP]),
        "qsvh_state_free": (None, [P]),
        "qsvh_state_n_qubits": (u32, [P]),
        "qsvh_state_dtype": (i32, [P]),
        "qsvh_init_basis": (i32, [P, u64]),
        "qsvh_apply": (i32, [P, u32, P, P, P]),
        "qsvh_amplitude": (i32, [P, u64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "qsvh_probabilities": (i32, [P, P]),
        "qsvh_probabilities_device": (i32, [P, P]),
        "qsvh_norm_sqr": (i32, [P, C.POINTER(C.c_double)]),
        "qsvh_expectation_z": (i32, [P, u64, C.POINTER(C.c_double)]),
        "qsvh_download": (i32, [P, P, P]),
        "qsvh_sample": (i32, [P, u64, u64, P]),
        "qsvh_sample_noalloc": (i32, [P, u64, u64, P]),
        "qsvh_upload": (i32, [P, P, P]),
        "qsvh_synchronize": (i32, [P]),
        "qsvh_init_random_normal": (i32, [P, u64]),
        "qsvh_compare": (i32, [P, P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "qsvh_execute": (i32, [P, P, i32, PP]),
        "qsvh_program_create": (i32, [P, P, C.POINTER(_ProgramOpts), PP]),
        "qsvh_program_launch": (i32, [P]),
        "qsvh_program_time": (i32, [P, C.POINTER(C.c_double)]),
        "qsvh_program_num_kernels": (i32, [P, C.POINTER(C.c_uint64)]),
        "qsvh_program_time_kernels": (i32, [P, P, C.c_size_t, C.POINTER(C.c_size_t)]),
        "qsvh_program_order": (i32, [P, P, C.c_size_t, C.POINTER(C.c_size_t)]),
        "qsvh_program_destroy": (None, [P]),
        "qsvh_circuit_hea_batch": (i32, [C.c_size_t, u32, u64, P]),
        "qsvh_batch_create": (i32, [P, P, C.c_size_t, i32, PP]),
        "qsvh_batch_destroy": (None, [P]),
        "qsvh_batch_configure": (i32, [P, i32, i32]),
        "qsvh_batch_autotune": (i32, [P, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_double)]),
        "qsvh_batch_run": (i32, [P]),
        "qsvh_batch_passes": (i32, [P, C.c_size_t, C.POINTER(C.c_uint32)]),
        "qsvh_batch_time": (i32, [P, C.POINTER(C.c_double)]),
        "qsvh_batch_info": (i32, [P, C.POINTER(C.c_size_t), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int), C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_uint64)]),
        "qsvh_batch_expectation_z0": (i32, [P, P]),
        "qsvh_batch_download": (i32, [P, C.c_size_t, P, P]),
        "qsvh_batch_probabilities": (i32, [P, C.c_size_t, P]),
        "qsvh_batch_order": (i32, [P, C.c_size_t, P, C.c_size_t, C.POINTER(C.c_size_t)]),
        "qsvh_time_gate": (i32, [P, u32, P, P, P, u32, C.POINTER(C.c_double)]),
        "qsvh_time_copy": (i32, [P, u32, C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


def _ctx_alive(backend) -> bool:
    """False once the backend's context is destroyed: a garbage-collected cycle may finalise
    a backend before the objects that use it, whose handles are then dropped (the process
    is ending) instead of being freed through a destroyed context."""
    return _lib is not None and getattr(backend, "ctx", None) is not None


def _check(rc: int, ctx=None):
    if rc != 0:
        msg = ""
        if ctx is not None:
            msg = (lib().qsvh_last_error(ctx) or b"").decode()
        raise QsvhError(rc, msg)


# ------------------------------------------------------------------------ gates
def gate_matrix(name: str, *params: float) -> np.ndarray:
    """The library gate ``name`` (crates/qsv-core/src/gate.rs:96-180, plus u3/ccx/ccz)."""
    p = np.array(params if params else [0.0], dtype=np.float64)
    re = np.zeros(64)
    im = np.zeros(64)
    m = C.c_uint32(0)
    _check(lib().qsvh_gate_matrix(name.encode(), _ptr(p), len(params), _ptr(re), _ptr(im),
                                  C.byref(m)))
    d = 1 << m.value
    return (re[: d * d] + 1j * im[: d * d]).reshape(d, d)


# ------------------------------------------------------------------------ circuits
class Circuit:
    """``Circuit<R>`` (crates/qsv-core/src/circuit.rs:30-131) held by the C++ engine."""

    def __init__(self, n_qubits: int = 0, _handle=None):
        if _handle is not None:
            self._h = _handle
        else:
            h = C.c_void_p()
            _check(lib().qsvh_circuit_create(n_qubits, C.byref(h)))
            self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _lib is not None:
            _lib.qsvh_circuit_destroy(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @property
    def n_qubits(self) -> int:
        return int(lib().qsvh_circuit_n_qubits(self._h))

    def __len__(self) -> int:
        return int(lib().qsvh_circuit_len(self._h))

    def push(self, matrix, qubits: Sequence[int]) -> "Circuit":
        """Circuit::push (circuit.rs:54-78): validated; raises QsvhError on bad input."""
        mat = np.asarray(matrix, dtype=np.complex128)
        q = np.asarray(qubits, dtype=np.uint32)
        m = len(q)
        if mat.shape != (1 << m, 1 << m):
            raise QsvhError(4, "gate arity must match qubit count")
        re = np.ascontiguousarray(mat.real).ravel()
        im = np.ascontiguousarray(mat.imag).ravel()
        _check(lib().qsvh_circuit_push(self._h, m, _ptr(q), _ptr(re), _ptr(im)))
        return self

    def gate(self, name: str, qubits: Sequence[int], *params: float) -> "Circuit":
        q = np.asarray(qubits, dtype=np.uint32)
        p = np.array(params if params else [0.0], dtype=np.float64)
        _check(lib().qsvh_circuit_push_gate(self._h, name.encode(), _ptr(q), _ptr(p),
                                            len(params)))
        return self

    # fluent builders (circuit.rs:81-130)
    def h(self, q): return self.gate("h", [q])
    def x(self, q): return self.gate("x", [q])
    def y(self, q): return self.gate("y", [q])
    def z(self, q): return self.gate("z", [q])
    def s(self, q): return self.gate("s", [q])
    def t(self, q): return self.gate("t", [q])
    def sx(self, q): return self.gate("sx", [q])
    def rx(self, q, th): return self.gate("rx", [q], th)
    def ry(self, q, th): return self.gate("ry", [q], th)
    def rz(self, q, th): return self.gate("rz", [q], th)
    def phase(self, q, lam): return self.gate("p", [q], lam)
    def u3(self, q, th, ph, lam): return self.gate("u3", [q], th, ph, lam)
    def cx(self, c, t): return self.gate("cx", [c, t])
    def cz(self, a, b): return self.gate("cz", [a, b])
    def cphase(self, c, t, lam): return self.gate("cp", [c, t], lam)
    def swap(self, a, b): return self.gate("swap", [a, b])
    def rzz(self, a, b, th): return self.gate("rzz", [a, b], th)
    def ccx(self, a, b, t): return self.gate("ccx", [a, b, t])
    def ccz(self, a, b, c): return self.gate("ccz", [a, b, c])

    def ops(self):
        """List of (matrix, qubits) -- ``Circuit::ops`` (circuit.rs:48-51)."""
        out = []
        L = lib()
        for i in range(len(self)):
            m = C.c_uint32(0)
            q = np.zeros(6, dtype=np.uint32)
            _check(L.qsvh_circuit_op(self._h, i, C.byref(m), _ptr(q), None, None))
            d = 1 << m.value
            re = np.zeros(d * d)
            im = np.zeros(d * d)
            _check(L.qsvh_circuit_op(self._h, i, None, None, _ptr(re), _ptr(im)))
            out.append(((re + 1j * im).reshape(d, d), tuple(int(x) for x in q[: m.value])))
        return out

    @classmethod
    def from_ops(cls, n_qubits: int, ops: Iterable) -> "Circuit":
        c = cls(n_qubits)
        for mat, q in ops:
            c.push(mat, q)
        return c


def _gen(fn: str, *args) -> Circuit:
    h = C.c_void_p()
    _check(getattr(lib(), fn)(*args, C.byref(h)))
    return Circuit(_handle=h)


def from_qasm(text: str) -> Circuit:
    """An OpenQASM 3 (or 2.0) program as a circuit (qsvh_circuit_from_qasm); raises
    QsvhError(10, "line L: reason") for what the reader does not accept."""
    raw = text.encode()
    h = C.c_void_p()
    err = C.create_string_buffer(512)
    rc = lib().qsvh_circuit_from_qasm(raw, len(raw), C.byref(h), err, len(err))
    if rc != 0:
        raise QsvhError(rc, err.value.decode())
    return Circuit(_handle=h)


def read_qasm(path: str) -> Circuit:
    """from_qasm of a file's contents."""
    with open(path, encoding="utf-8") as f:
        return from_qasm(f.read())


def ghz(n: int) -> Circuit:
    """crates/qsv-core/src/circuits.rs:50-57: H(0); CX(0, q) for q = 1..n."""
    return _gen("qsvh_circuit_ghz", n)


def ghz_chain(n: int) -> Circuit:
    """H(0); CX(i, i+1) for i = 0..n-2 (the GHZ(20) workload)."""
    return _gen("qsvh_circuit_ghz_chain", n)


def qft(n: int) -> Circuit:
    """crates/qsv-core/src/circuits.rs:61-74."""
    return _gen("qsvh_circuit_qft", n)


def random_circuit(n: int, n_gates: int, seed: int) -> Circuit:
    """crates/qsv-core/src/circuits.rs:79-113."""
    return _gen("qsvh_circuit_random", n, n_gates, seed)


def layered(n: int, depth: int, seed: int) -> Circuit:
    """Per layer: U3 on every qubit, RZ on every qubit, CZ on brickwork pairs."""
    return _gen("qsvh_circuit_layered", n, depth, seed)


def hea(n: int, depth: int, seed: int) -> Circuit:
    """Per layer: RY then RZ on every qubit, then CX(i, i+1)."""
    return _gen("qsvh_circuit_hea", n, depth, seed)


def mix(n: int, n_gates: int, seed: int) -> Circuit:
    """50% U3, 30% CX/CZ, 15% RZZ, 5% CCX/CCZ on random qubits."""
    return _gen("qsvh_circuit_mix", n, n_gates, seed)


# ------------------------------------------------------------------------ backend
class State:
    """Device-resident SoA state (``Backend::State``); freed with the object."""

    def __init__(self, backend: "HipBackend", handle, n: int, dtype: int):
        self.backend = backend
        self._h = handle
        self.n_qubits = n
        self.dtype = dtype
        self._programs = weakref.WeakSet()  # programs bound to this state

    @property
    def handle(self):
        return self._h

    @property
    def np_dtype(self):
        return np.float64 if self.dtype == C128 else np.float32

    def free(self):
        """Free the device state; programs bound to it are destroyed first (the C ABI's
        order: a program must not outlive its state)."""
        if self._h is not None:
            for p in list(self._programs):
                p._destroy()
            lib().qsvh_state_free(self._h)
            self._h = None

    def __del__(self):
        if getattr(self, "_h", None) is not None and _ctx_alive(getattr(self, "backend", None)):
            for p in list(getattr(self, "_programs", ())):
                p._destroy()
            _lib.qsvh_state_free(self._h)
        self._h = None


class Program:
    """A circuit bound to a state (matrices resident on the device, optional graph)."""

    def __init__(self, state: State, circuit: Circuit, precision=PREC_EXACT,
                 use_graph=False, tile_qubits=0, subspace=False, mfma_tiles=0, planned=False):
        self.state = state
        self.circuit = circuit
        opts = _ProgramOpts(precision, int(bool(use_graph)), int(tile_qubits),
                            int(bool(subspace)), int(mfma_tiles), int(bool(planned)))
        h = C.c_void_p()
        _check(lib().qsvh_program_create(state.handle, circuit.handle, C.byref(opts),
                                         C.byref(h)), state.backend.ctx)
        self._h = h
        state._programs.add(self)

    def _handle(self):
        if self._h is None:
            raise QsvhError(1, "program destroyed (its state was freed)")
        return self._h

    def _destroy(self):
        state = getattr(self, "state", None)
        if getattr(self, "_h", None) is not None and _ctx_alive(getattr(state, "backend", None)):
            _lib.qsvh_program_destroy(self._h)
        self._h = None

    def launch(self):
        _check(lib().qsvh_program_launch(self._handle()), self.state.backend.ctx)

    def time(self) -> float:
        """Launch once; device milliseconds from HIP events on the engine's stream."""
        ms = C.c_double(0)
        _check(lib().qsvh_program_time(self._handle(), C.byref(ms)), self.state.backend.ctx)
        return ms.value

    def num_kernels(self) -> int:
        n = C.c_uint64(0)
        _check(lib().qsvh_program_num_kernels(self._handle(), C.byref(n)))
        return n.value

    def time_kernels(self):
        """Launch once; device milliseconds of every kernel (HIP events between them)."""
        n = C.c_size_t(0)
        cap = self.num_kernels()
        out = np.zeros(max(cap, 1), dtype=np.float64)
        _check(lib().qsvh_program_time_kernels(self._handle(), _ptr(out), cap, C.byref(n)),
               self.state.backend.ctx)
        return out[: n.value]

    def order(self):
        """The circuit's op indices in the order the program applies them."""
        n = C.c_size_t(0)
        _check(lib().qsvh_program_order(self._handle(), None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.uint64)
        _check(lib().qsvh_program_order(self._handle(), _ptr(out), n.value, C.byref(n)))
        return [int(x) for x in out[: n.value]]

    def __del__(self):
        self._destroy()


TUNE_KEYS = ("max_blocks", "unroll", "nontemporal", "mfma_blocks", "mfma_layout",
             "tile_blocks", "contig", "xcd", "mq_lane", "diag_blocks")


class HipBackend:
    """``Backend<R>`` (crates/qsv-core/src/backend/mod.rs:31-74) on one MI355X."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().qsvh_ctx_create(device, C.byref(h)))
        self.ctx = h

    def close(self):
        """Wait for queued work.  The context itself is destroyed when the last object
        that uses it (states, programs, batches, models -- each keeps a reference to this
        backend) is gone: the C ABI requires every such object to be destroyed first."""
        if self.ctx is not None:
            _check(lib().qsvh_synchronize_ctx(self.ctx), self.ctx)

    def __del__(self):
        if getattr(self, "ctx", None) is not None and _lib is not None:
            _lib.qsvh_ctx_destroy(self.ctx)
            self.ctx = None

    def last_error(self) -> str:
        return (lib().qsvh_last_error(self.ctx) or b"").decode()

    def tune(self, **knobs):
        """Set launch-shape knobs (see TUNE_KEYS); returns all values."""
        for k, v in knobs.items():
            _check(lib().qsvh_tune_set(self.ctx, k.encode(), int(v)), self.ctx)
        out = {}
        for k in TUNE_KEYS:
            v = C.c_int64(0)
            _check(lib().qsvh_tune_get(self.ctx, k.encode(), C.byref(v)), self.ctx)
            out[k] = v.value
        return out

    def device_info(self):
        name = C.create_string_buffer(256)
        cus = C.c_int(0)
        mem = C.c_uint64(0)
        _check(lib().qsvh_device_info(self.ctx, name, 256, C.byref(cus), C.byref(mem)),
               self.ctx)
        return {"name": name.value.decode(), "compute_units": cus.value,
                "hbm_bytes": mem.value}

    # -- trait methods
    def alloc(self, n_qubits: int, dtype: int = C128) -> State:
        h = C.c_void_p()
        _check(lib().qsvh_state_alloc(self.ctx, n_qubits, dtype, C.byref(h)), self.ctx)
        return State(self, h, n_qubits, dtype)

    def init_basis(self, state: State, basis: int = 0):
        _check(lib().qsvh_init_basis(state.handle, basis), self.ctx)

    def apply(self, state: State, matrix, qubits: Sequence[int]):
        mat = np.asarray(matrix, dtype=np.complex128)
        q = np.asarray(qubits, dtype=np.uint32)
        if mat.shape != (1 << len(q), 1 << len(q)):
            raise QsvhError(4, "gate arity must match qubit count")
        re = np.ascontiguousarray(mat.real).ravel()
        im = np.ascontiguousarray(mat.imag).ravel()
        _check(lib().qsvh_apply(state.handle, len(q), _ptr(q), _ptr(re), _ptr(im)), self.ctx)

    def amplitude(self, state: State, index: int) -> complex:
        re = C.c_double(0)
        im = C.c_double(0)
        _check(lib().qsvh_amplitude(state.handle, index, C.byref(re), C.byref(im)), self.ctx)
        return complex(re.value, im.value)

    def probabilities(self, state: State) -> np.ndarray:
        out = np.empty(1 << state.n_qubits, dtype=state.np_dtype)
        _check(lib().qsvh_probabilities(state.handle, _ptr(out)), self.ctx)
        return out

    def norm_sqr(self, state: State) -> float:
        v = C.c_double(0)
        _check(lib().qsvh_norm_sqr(state.handle, C.byref(v)), self.ctx)
        return v.value

    def expectation_z(self, state: State, mask: int) -> float:
        v = C.c_double(0)
        _check(lib().qsvh_expectation_z(state.handle, mask, C.byref(v)), self.ctx)
        return v.value

    def sample(self, state: State, shots: int, seed: int):
        """Backend::sample (crates/qsv-core/src/backend/mod.rs:52-54) on the device by
        inverse CDF (sample.rs:140-153) with SplitMix64 started at `seed`; returns the
        outcomes and the generator state after the `shots` draws."""
        out = np.empty(shots, dtype=np.uint64)
        _check(lib().qsvh_sample(state.handle, shots, seed & (2**64 - 1), _ptr(out)), self.ctx)
        return out, (seed + shots * 0x9E3779B97F4A7C15) & (2**64 - 1)

    def download(self, state: State) -> np.ndarray:
        re = np.empty(1 << state.n_qubits, dtype=state.np_dtype)
        im = np.empty(1 << state.n_qubits, dtype=state.np_dtype)
        _check(lib().qsvh_download(state.handle, _ptr(re), _ptr(im)), self.ctx)
        return re + 1j * im

    def download_parts(self, state: State):
        re = np.empty(1 << state.n_qubits, dtype=state.np_dtype)
        im = np.empty(1 << state.n_qubits, dtype=state.np_dtype)
        _check(lib().qsvh_download(state.handle, _ptr(re), _ptr(im)), self.ctx)
        return re, im

    def upload(self, state: State, values=None, re=None, im=None):
        if values is not None:
            re = np.ascontiguousarray(np.real(values), dtype=state.np_dtype)
            im = np.ascontiguousarray(np.imag(values), dtype=state.np_dtype)
        else:
            re = np.ascontiguousarray(re, dtype=state.np_dtype)
            im = np.ascontiguousarray(im, dtype=state.np_dtype)
        if re.size != (1 << state.n_qubits) or im.size != re.size:
            raise QsvhError(1, "state size mismatch")
        _check(lib().qsvh_upload(state.handle, _ptr(re), _ptr(im)), self.ctx)

    def init_random_normal(self, state: State, seed: int):
        """Random normalised start state generated on the device (see qsvh.h)."""
        _check(lib().qsvh_init_random_normal(state.handle, seed), self.ctx)

    def compare(self, a: State, b: State):
        """(max_i |a_i - b_i|, |<a|b>|^2 / (<a|a><b|b>)) computed on the device in f64."""
        err = C.c_double(0)
        fid = C.c_double(0)
        _check(lib().qsvh_compare(a.handle, b.handle, C.byref(err), C.byref(fid)), self.ctx)
        return err.value, fid.value

    def synchronize(self, state: State):
        _check(lib().qsvh_synchronize(state.handle), self.ctx)

    def execute(self, circuit: Circuit, dtype: int = C128) -> State:
        """Backend::execute (mod.rs:66-73): |0...0> then every op, one submission."""
        h = C.c_void_p()
        _check(lib().qsvh_execute(self.ctx, circuit.handle, dtype, C.byref(h)), self.ctx)
        return State(self, h, circuit.n_qubits, dtype)

    def program(self, state: State, circuit: Circuit, **kw) -> Program:
        return Program(state, circuit, **kw)

    def run(self, state: State, circuit: Circuit, **kw):
        """Apply `circuit` to the state's current contents."""
        p = Program(state, circuit, **kw)
        p.launch()
        self.synchronize(state)

    def fuse(self, circuit: Circuit, max_qubits: int = 4) -> Circuit:
        return fuse(self, circuit, max_qubits)

    # -- measurement helpers
    def time_gate(self, state: State, matrix, qubits, reps: int = 10) -> float:
        mat = np.asarray(matrix, dtype=np.complex128)
        q = np.asarray(qubits, dtype=np.uint32)
        re = np.ascontiguousarray(mat.real).ravel()
        im = np.ascontiguousarray(mat.imag).ravel()
        ms = C.c_double(0)
        _check(lib().qsvh_time_gate(state.handle, len(q), _ptr(q), _ptr(re), _ptr(im), reps,
                                    C.byref(ms)), self.ctx)
        return ms.value

    def time_copy(self, state: State, reps: int = 10) -> float:
        ms = C.c_double(0)
        _check(lib().qsvh_time_copy(state.handle, reps, C.byref(ms)), self.ctx)
        return ms.value


def hea_batch(count: int, depth: int, seed: int):
    """The variational-sweep workload: `count` HEA circuits, n_c = 10 + below(13)."""
    arr = (C.c_void_p * count)()
    _check(lib().qsvh_circuit_hea_batch(count, depth, seed, arr))
    return [Circuit(_handle=C.c_void_p(arr[i])) for i in range(count)]


class Batch:
    """Many circuits run together (persistent tile-pass launches in one HIP graph)."""

    def __init__(self, backend: "HipBackend", circuits, dtype: int = C64):
        self.backend = backend
        self.circuits = list(circuits)
        self.dtype = dtype
        arr = (C.c_void_p * len(self.circuits))(*[c.handle.value for c in self.circuits])
        h = C.c_void_p()
        _check(lib().qsvh_batch_create(backend.ctx, arr, len(self.circuits), dtype,
                                       C.byref(h)), backend.ctx)
        self._h = h

    def configure(self, tile_qubits: int, blocks: int = 0):
        _check(lib().qsvh_batch_configure(self._h, tile_qubits, blocks), self.backend.ctx)

    def autotune(self):
        t, bl, ms = C.c_int(0), C.c_int(0), C.c_double(0)
        _check(lib().qsvh_batch_autotune(self._h, C.byref(t), C.byref(bl), C.byref(ms)),
               self.backend.ctx)
        return t.value, bl.value, ms.value

    def run(self):
        _check(lib().qsvh_batch_run(self._h), self.backend.ctx)

    def time(self) -> float:
        ms = C.c_double(0)
        _check(lib().qsvh_batch_time(self._h, C.byref(ms)), self.backend.ctx)
        return ms.value

    def info(self):
        n, t, bl = C.c_size_t(0), C.c_int(0), C.c_int(0)
        nl, amps = C.c_uint64(0), C.c_uint64(0)
        _check(lib().qsvh_batch_info(self._h, C.byref(n), C.byref(t), C.byref(bl),
                                     C.byref(nl), C.byref(amps)))
        return {"count": n.value, "tile_qubits": t.value, "blocks": bl.value,
                "launches": nl.value, "amplitudes": amps.value}

    def passes(self, i: int) -> int:
        """Passes of circuit i (its first builds |0...0> in LDS without reading HBM)."""
        n = C.c_uint32(0)
        _check(lib().qsvh_batch_passes(self._h, i, C.byref(n)))
        return n.value

    def hbm_bytes(self) -> int:
        """Algorithmic HBM bytes of one run: every pass writes its circuit's state, every
        pass but the first reads it ((2 passes - 1) x 2^n x bytes per amplitude)."""
        amp = 16 if self.dtype == C128 else 8
        return sum((2 * self.passes(i) - 1) * (1 << c.n_qubits) * amp
                   for i, c in enumerate(self.circuits))

    def expectation_z0(self) -> np.ndarray:
        out = np.zeros(len(self.circuits))
        _check(lib().qsvh_batch_expectation_z0(self._h, _ptr(out)), self.backend.ctx)
        return out

    def _np(self):
        return np.float64 if self.dtype == C128 else np.float32

    def download(self, i: int) -> np.ndarray:
        n = self.circuits[i].n_qubits
        re = np.empty(1 << n, dtype=self._np())
        im = np.empty(1 << n, dtype=self._np())
        _check(lib().qsvh_batch_download(self._h, i, _ptr(re), _ptr(im)), self.backend.ctx)
        return re + 1j * im

    def probabilities(self, i: int) -> np.ndarray:
        n = self.circuits[i].n_qubits
        out = np.empty(1 << n, dtype=self._np())
        _check(lib().qsvh_batch_probabilities(self._h, i, _ptr(out)), self.backend.ctx)
        return out

    def order(self, i: int):
        n = C.c_size_t(0)
        lib().qsvh_batch_order(self._h, i, None, 0, C.byref(n))
        out = np.zeros(max(n.value, 1), dtype=np.uint64)
        _check(lib().qsvh_batch_order(self._h, i, _ptr(out), n.value, C.byref(n)))
        return [int(x) for x in out[: n.value]]

    def __del__(self):
        if getattr(self, "_h", None) is not None and _ctx_alive(getattr(self, "backend", None)):
            _lib.qsvh_batch_destroy(self._h)
        self._h = None


def fuse_groups(circuit: Circuit, max_qubits: int = 4):
    """The greedy grouping of fusion::fuse (host only): list of (begin, end, qubits)."""
    n = max(len(circuit), 1)
    b = np.zeros(n, dtype=np.uint64)
    e = np.zeros(n, dtype=np.uint64)
    m = np.zeros(n, dtype=np.uint32)
    q = np.zeros(6 * n, dtype=np.uint32)
    ng = C.c_size_t(0)
    _check(lib().qsvh_fuse_groups(circuit.handle, max_qubits, _ptr(b), _ptr(e), _ptr(m),
                                  _ptr(q), n, C.byref(ng)))
    return [(int(b[g]), int(e[g]), tuple(int(x) for x in q[6 * g: 6 * g + m[g]]))
            for g in range(ng.value)]


def fuse(backend: HipBackend, circuit: Circuit, max_qubits: int = 4) -> Circuit:
    """fusion::fuse (crates/qsv-core/src/fusion.rs:39-77); f64 composition on the GPU."""
    h = C.c_void_p()
    _check(lib().qsvh_fuse(backend.ctx, circuit.handle, max_qubits, C.byref(h)), backend.ctx)
    return Circuit(_handle=h)
