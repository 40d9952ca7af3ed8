"""Learned fusion planning: timing collection, cost-model training and planning.

The greedy ``fusion::fuse`` (nez0b/rustatevec crates/qsv-core/src/fusion.rs:39-77) groups
consecutive gates up to a fixed ``max_qubits``.  Here a small network predicts the
device time of a fused group from its features (include/qsvh.h,
``qsvh_group_features``); it is trained on the device (``qsvh_mlp_train``) on timings
that the engine itself produces for random groups, and ``qsvh_plan_fusion`` picks the
group boundaries and widths that minimise the predicted total -- keeping a greedy plan
when that predicts lower.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import (C64, PREC_MFMA_BF16X3, Circuit, HipBackend, Program, QsvhError, _check,
               _ctx_alive, _ptr, lib)

N_FEATURES = 12
N_TILE_FEATURES = 18
N_MT_FEATURES = 11
HIDDEN = 64


def _bind():
    L = lib()
    P, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    PP = C.POINTER(C.c_void_p)
    sig = {
        "qsvh_mlp_create": (i32, [P, u32, u32, u64, PP]),
        "qsvh_mlp_destroy": (None, [P]),
        "qsvh_mlp_num_params": (i32, [P, C.POINTER(C.c_size_t)]),
        "qsvh_mlp_get_params": (i32, [P, P]),
        "qsvh_mlp_set_params": (i32, [P, P]),
        "qsvh_mlp_forward": (i32, [P, P, C.c_size_t, P]),
        "qsvh_mlp_gradients": (i32, [P, P, P, C.c_size_t, P, C.POINTER(C.c_float)]),
        "qsvh_mlp_train": (i32, [P, P, P, C.c_size_t, u32, C.c_float,
                                 C.POINTER(C.c_float)]),
        "qsvh_mlp_predict": (i32, [P, P, C.c_size_t, P]),
        "qsvh_mlp_target_scale": (i32, [P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "qsvh_group_features": (i32, [u32, u32, P, i32, C.c_double, P]),
        "qsvh_plan_fusion": (i32, [P, P, P, u32, u32, PP, C.POINTER(C.c_double),
                                   C.POINTER(C.c_int)]),
        "qsvh_tile_pass_features": (i32, [P, i32, u32, P, C.c_size_t,
                                          C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
        "qsvh_mt_pass_features": (i32, [P, P, C.c_size_t, C.POINTER(C.c_size_t),
                                        C.POINTER(C.c_size_t)]),
        "qsvh_plan_form": (i32, [P, P, P, P, P, i32, u32, u32, PP, C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), C.POINTER(C.c_double),
                                 C.POINTER(C.c_int)]),
        "qsvh_plan_passes": (i32, [P, P, P, P, i32, PP, C.POINTER(C.c_double)]),
        "qsvh_circuit_plan_len": (i32, [P, C.POINTER(C.c_size_t)]),
        "qsvh_circuit_plan_pass": (i32, [P, C.c_size_t, C.POINTER(C.c_int),
                                         C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_uint64), P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def group_features(n: int, qubits, diagonal: bool, density: float) -> np.ndarray:
    q = np.asarray(sorted(qubits), dtype=np.uint32)
    out = np.zeros(N_FEATURES, dtype=np.float32)
    _check(_bind().qsvh_group_features(n, len(q), _ptr(q), int(bool(diagonal)), density,
                                       _ptr(out)))
    return out


class Mlp:
    """The cost-model network (device parameters, device training)."""

    def __init__(self, backend: HipBackend, n_features=N_FEATURES, hidden=HIDDEN, seed=1):
        self.backend = backend
        self.F, self.H = n_features, hidden
        h = C.c_void_p()
        _check(_bind().qsvh_mlp_create(backend.ctx, n_features, hidden, seed, C.byref(h)),
               backend.ctx)
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) is not None and _ctx_alive(getattr(self, "backend", None)):
            _bind().qsvh_mlp_destroy(self._h)
        self._h = None

    @property
    def handle(self):
        return self._h

    def num_params(self) -> int:
        n = C.c_size_t(0)
        _check(_bind().qsvh_mlp_num_params(self._h, C.byref(n)))
        return n.value

    def params(self) -> np.ndarray:
        out = np.zeros(self.num_params(), dtype=np.float32)
        _check(_bind().qsvh_mlp_get_params(self._h, _ptr(out)), self.backend.ctx)
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params()
        _check(_bind().qsvh_mlp_set_params(self._h, _ptr(p)), self.backend.ctx)

    def forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros(len(X), dtype=np.float32)
        _check(_bind().qsvh_mlp_forward(self._h, _ptr(X), len(X), _ptr(out)), self.backend.ctx)
        return out

    def predict(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros(len(X), dtype=np.float32)
        _check(_bind().qsvh_mlp_predict(self._h, _ptr(X), len(X), _ptr(out)), self.backend.ctx)
        return out

    def gradients(self, X, y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.zeros(self.num_params(), dtype=np.float32)
        loss = C.c_float(0)
        _check(_bind().qsvh_mlp_gradients(self._h, _ptr(X), _ptr(y), len(X), _ptr(g),
                                          C.byref(loss)), self.backend.ctx)
        return float(loss.value), g

    def train(self, X, y, epochs=1500, lr=3e-3) -> float:
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = C.c_float(0)
        _check(_bind().qsvh_mlp_train(self._h, _ptr(X), _ptr(y), len(X), epochs, lr,
                                      C.byref(loss)), self.backend.ctx)
        return float(loss.value)


def _random_group(rng, n, k, kind):
    qs = sorted(int(x) for x in rng.choice(n, size=k, replace=False))
    d = 1 << k
    if kind == "diag":
        U = np.diag(np.exp(1j * rng.uniform(0, 2 * np.pi, size=d)))
    elif kind == "perm":
        U = np.eye(d)[rng.permutation(d)]
        if k == 1:
            U = np.array([[0, 1], [1, 0]], dtype=complex)
    else:
        z = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
        U, _ = np.linalg.qr(z)
    density = float(np.count_nonzero(U)) / (d * d)
    return qs, U, kind == "diag", density


def collect_timings(backend: HipBackend, sizes=(22, 24, 26), per_size=200, seed=7,
                    precision=PREC_MFMA_BF16X3, reps=3):
    """Time random fused groups (k = 1..6; diagonal, permutation or dense) as the engine
    runs them on complex64 states; returns features X and targets y = log2 of the device
    milliseconds scaled to 2^30 amplitudes."""
    rng = np.random.default_rng(seed)
    X, y = [], []
    for n in sizes:
        st = backend.alloc(n, C64)
        backend.init_random_normal(st, seed)
        for _ in range(per_size):
            k = int(rng.integers(1, 7))
            kind = rng.choice(["diag", "perm", "dense"], p=[0.3, 0.2, 0.5])
            qs, U, diag, dens = _random_group(rng, n, k, kind)
            p = Program(st, Circuit(n).push(U, qs), precision=precision)
            p.launch()
            ms = min(p.time() for _ in range(reps))
            X.append(group_features(n, qs, diag, dens))
            y.append(np.log2(ms * 2.0 ** (30 - n)))
            del p
        st.free()
    return np.array(X, dtype=np.float32), np.array(y, dtype=np.float32)


def train_cost_model(backend: HipBackend, X, y, epochs=1500, lr=3e-3, seed=1):
    m = Mlp(backend, seed=seed)
    loss = m.train(X, y, epochs=epochs, lr=lr)
    return m, loss


def plan(backend: HipBackend, model: Mlp, circuit: Circuit, max_qubits=6, window=48):
    """(planned circuit, predicted ms, chosen): chosen 0 = dynamic programme, m > 0 =
    greedy fuse at max_qubits m."""
    h = C.c_void_p()
    ms = C.c_double(0)
    chosen = C.c_int(0)
    _check(_bind().qsvh_plan_fusion(backend.ctx, model.handle, circuit.handle, max_qubits,
                                    window, C.byref(h), C.byref(ms), C.byref(chosen)),
           backend.ctx)
    return Circuit(_handle=h), ms.value, chosen.value


def tile_pass_features(circuit: Circuit, dtype: int, tile_qubits: int):
    """(rows, n_other): the cost-model features of every tile pass the scheduler builds for
    `circuit` (one row per pass) and the number of launches left outside passes."""
    L = _bind()
    n_p, n_o = C.c_size_t(0), C.c_size_t(0)
    _check(L.qsvh_tile_pass_features(circuit.handle, dtype, tile_qubits, None, 0,
                                     C.byref(n_p), C.byref(n_o)))
    out = np.zeros((n_p.value, N_TILE_FEATURES), dtype=np.float32)
    if n_p.value:
        _check(L.qsvh_tile_pass_features(circuit.handle, dtype, tile_qubits, _ptr(out),
                                         n_p.value, C.byref(n_p), C.byref(n_o)))
    return out, n_o.value


_SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=np.complex128)

# gate kinds of the training passes, in the order of the cumulative weights below
_PASS_KINDS = ("u3", "real", "x", "rz", "diag2", "cx", "swap", "ccx")


def _random_pass(rng, n, tb):
    """A random circuit that the tile scheduler runs as one pass of tb tile bits: 1 to 160
    gates with their action qubits in one random tile set (controls and the qubits of
    diagonal gates anywhere), drawn from a gate-kind mix that is itself random per pass
    (Dirichlet), so passes range from the mixed workloads' to layered circuits' (U3, RZ,
    CZ) and single-kind runs."""
    tile = [0, 1, 2, 3] + sorted(int(x) for x in rng.choice(np.arange(4, n), size=tb - 4,
                                                           replace=False))
    w = rng.dirichlet(np.full(len(_PASS_KINDS), 0.7))
    cum = np.cumsum(w)
    c = Circuit(n)
    for _ in range(int(rng.integers(1, 161))):
        kind = _PASS_KINDS[min(int(np.searchsorted(cum, rng.random())), len(_PASS_KINDS) - 1)]
        t = int(rng.choice(tile))
        o = int(rng.choice([q for q in range(n) if q != t]))
        a = lambda: float(rng.uniform(0, 2 * np.pi))  # noqa: E731
        if kind == "u3":
            c.u3(t, a(), a(), a())
        elif kind == "real":
            c.ry(t, a()) if rng.random() < 0.5 else c.h(t)
        elif kind == "x":
            c.x(t)
        elif kind == "rz":
            c.rz(int(rng.integers(n)), a())
        elif kind == "diag2":
            k = rng.integers(3)
            c.cz(t, o) if k == 0 else (c.cphase(o, t, a()) if k == 1 else c.rzz(o, t, a()))
        elif kind == "cx":
            c.cx(o, t)
        elif kind == "swap":
            u = int(rng.choice([q for q in tile if q != t]))
            c.swap(t, u)
        else:
            p = int(rng.choice([q for q in range(n) if q not in (t, o)]))
            c.gate("ccx", [o, p, t])
    return c


def collect_tile_timings(backend: HipBackend, sizes=(26,), per_size=400, seed=7, dtype=C64,
                         reps=3):
    """Time random single tile passes (tile bits 11..13 for complex64, 10..12 for
    complex128) on HBM-resident states; returns pass features X and y = log2 of the
    device milliseconds scaled to 2^30 amplitudes."""
    rng = np.random.default_rng(seed)
    lo = 11 if dtype == C64 else 10
    X, y = [], []
    for n in sizes:
        st = backend.alloc(n, dtype)
        backend.init_random_normal(st, seed)
        for _ in range(per_size):
            tb = int(rng.integers(lo, lo + 3))
            c = _random_pass(rng, n, tb)
            rows, other = tile_pass_features(c, dtype, tb)
            if len(rows) == 1 and other:
                # SWAPs that end the circuit run as a bit-swap launch after the pass: time
                # the pass without them
                ops = c.ops()
                k = len(ops)
                while k and len(ops[k - 1][1]) == 2 and np.array_equal(ops[k - 1][0], _SWAP):
                    k -= 1
                c = Circuit.from_ops(n, ops[:k])
                rows, other = tile_pass_features(c, dtype, tb)
            if len(rows) != 1 or other:
                continue
            p = Program(st, c, tile_qubits=tb)
            p.launch()
            ms = min(p.time() for _ in range(reps))
            X.append(rows[0])
            y.append(np.log2(ms * 2.0 ** (30 - n)))
            del p
        st.free()
    return np.array(X, dtype=np.float32), np.array(y, dtype=np.float32)


def train_tile_model(backend: HipBackend, X, y, epochs=1500, lr=3e-3, seed=2):
    m = Mlp(backend, n_features=N_TILE_FEATURES, seed=seed)
    loss = m.train(X, y, epochs=epochs, lr=lr)
    return m, loss


def mt_pass_features(circuit: Circuit):
    """(rows, n_other): the cost-model features of every MFMA tile pass (2^12-amplitude
    tiles) built for the fused `circuit`, one row per pass, and the launches outside them."""
    L = _bind()
    n_p, n_o = C.c_size_t(0), C.c_size_t(0)
    _check(L.qsvh_mt_pass_features(circuit.handle, None, 0, C.byref(n_p), C.byref(n_o)))
    out = np.zeros((n_p.value, N_MT_FEATURES), dtype=np.float32)
    if n_p.value:
        _check(L.qsvh_mt_pass_features(circuit.handle, _ptr(out), n_p.value, C.byref(n_p),
                                       C.byref(n_o)))
    return out, n_o.value


def _unitary(rng, d):
    q, r = np.linalg.qr(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))
    return q * (np.diag(r) / np.abs(np.diag(r)))


def _random_mt_pass(rng, n):
    """A random fused circuit that runs as one MFMA tile pass: 0 to 10 dense groups of 4..6
    qubits (and a few 1..3-qubit gates) inside one random tile set (qubits 0..3 and 8 more),
    0 to 80 diagonal groups of 1..6 qubits anywhere (phases within the pass's 2048-entry
    table), in random order and with a random mix per pass."""
    tile = [0, 1, 2, 3] + sorted(int(x) for x in rng.choice(np.arange(4, n), size=8,
                                                           replace=False))
    n_dense = int(rng.integers(0, 11))
    n_small = int(rng.integers(0, 4))
    n_diag = int(rng.integers(0 if n_dense + n_small else 1, 81))
    wk = rng.dirichlet(np.ones(3))
    wm = rng.dirichlet(np.ones(6))
    ops, entries = [], 0
    for _ in range(n_dense):
        k = 4 + int(rng.choice(3, p=wk))
        ops.append(("d", k))
    for _ in range(n_small):
        ops.append(("d", int(rng.integers(1, 4))))
    for _ in range(n_diag):
        m = 1 + int(rng.choice(6, p=wm))
        if entries + (1 << m) > 2048:
            break
        entries += 1 << m
        ops.append(("g", m))
    rng.shuffle(ops)
    c = Circuit(n)
    for kind, k in ops:
        if kind == "d":
            qs = [int(x) for x in rng.choice(tile, size=k, replace=False)]
            c.push(_unitary(rng, 1 << k), qs)
        else:
            qs = [int(x) for x in rng.choice(n, size=k, replace=False)]
            c.push(np.diag(np.exp(1j * rng.uniform(0, 2 * np.pi, size=1 << k))), qs)
    return c


def collect_mt_timings(backend: HipBackend, sizes=(26,), per_size=300, seed=11, reps=3):
    """Time random single MFMA tile passes (split-BF16, 2^12-amplitude tiles) on
    HBM-resident complex64 states; returns pass features X and y = log2 of the device
    milliseconds scaled to 2^30 amplitudes."""
    rng = np.random.default_rng(seed)
    X, y = [], []
    for n in sizes:
        st = backend.alloc(n, C64)
        backend.init_random_normal(st, seed)
        for _ in range(per_size):
            c = _random_mt_pass(rng, n)
            rows, other = mt_pass_features(c)
            if len(rows) != 1 or other:
                continue
            p = Program(st, c, precision=PREC_MFMA_BF16X3, mfma_tiles=12)
            p.launch()
            ms = min(p.time() for _ in range(reps))
            X.append(rows[0])
            y.append(np.log2(ms * 2.0 ** (30 - n)))
            del p
        st.free()
    return np.array(X, dtype=np.float32), np.array(y, dtype=np.float32)


def train_mt_model(backend: HipBackend, X, y, epochs=2000, lr=3e-3, seed=3):
    m = Mlp(backend, n_features=N_MT_FEATURES, seed=seed)
    loss = m.train(X, y, epochs=epochs, lr=lr)
    return m, loss


def plan_form(backend: HipBackend, group_model: Mlp, tile_model: Mlp, circuit: Circuit,
              dtype=C64, max_qubits=6, window=48, mt_model: Mlp | None = None):
    """(circuit, tile_qubits, mfma_tiles, predicted ms, chosen): the fused plan (run it with
    an MFMA precision), the circuit as exact tile passes of tile_qubits bits, or (with
    mt_model, complex64) greedy fusion at 4 or 5 qubits as MFMA tile passes (run it
    MFMA_BF16X3 with mfma_tiles), whichever the cost models predict fastest; chosen as in
    ``plan``, -tile_qubits, or 100 + the fusion width."""
    h = C.c_void_p()
    tq, mt, chosen = C.c_int(0), C.c_int(0), C.c_int(0)
    ms = C.c_double(0)
    _check(_bind().qsvh_plan_form(backend.ctx, group_model.handle, tile_model.handle,
                                  mt_model.handle if mt_model is not None else None,
                                  circuit.handle, dtype, max_qubits, window, C.byref(h),
                                  C.byref(tq), C.byref(mt), C.byref(ms), C.byref(chosen)),
           backend.ctx)
    return Circuit(_handle=h), tq.value, mt.value, ms.value, chosen.value


FORM_LEARNED = -1
FORM_EXACT = 0


def plan_passes(backend: HipBackend, tile_model: Mlp | None, mt_model: Mlp | None,
                circuit: Circuit, form=FORM_LEARNED):
    """(planned circuit, predicted ms): the circuit cut into tile passes, each run exactly or
    as an MFMA tile pass of its gates fused up to 3, 4 or 5 qubits -- chosen per pass by the
    two pass cost models (form -1), or one form for every pass (0 exact, k fused at k).  Run
    the result with ``Program(..., precision=PREC_MFMA_BF16X3, planned=True)``."""
    h = C.c_void_p()
    ms = C.c_double(0)
    _check(_bind().qsvh_plan_passes(backend.ctx, tile_model.handle if tile_model else None,
                                    mt_model.handle if mt_model else None, circuit.handle,
                                    int(form), C.byref(h), C.byref(ms)), backend.ctx)
    return Circuit(_handle=h), ms.value


def plan_info(circuit: Circuit, features=False):
    """Passes of a planned circuit: (form, op_begin, n_ops, tile_mask[, feature row])."""
    L = _bind()
    n = C.c_size_t(0)
    _check(L.qsvh_circuit_plan_len(circuit.handle, C.byref(n)))
    out = []
    for i in range(n.value):
        form = C.c_int(0)
        b, k, t = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        f = np.zeros(N_TILE_FEATURES, dtype=np.float32)
        _check(L.qsvh_circuit_plan_pass(circuit.handle, i, C.byref(form), C.byref(b),
                                        C.byref(k), C.byref(t), _ptr(f) if features else None))
        row = (form.value, b.value, k.value, t.value)
        if features:
            row += (f[: N_TILE_FEATURES if form.value == 0 else N_MT_FEATURES].copy(),)
        out.append(row)
    return out


def collect_pass_timings(backend: HipBackend, sizes=(22, 24), seeds=range(1000, 1004),
                         forms=(FORM_EXACT, 3, 4, 5), reps=3):
    """Time every pass of planned random-mix circuits (the planner workload's gate mix, 40 n
    gates) run in each fixed form, one kernel per pass; returns (X_tile, y_tile, X_mt, y_mt):
    the exact passes' and the MFMA passes' feature rows and y = log2 of the device
    milliseconds scaled to 2^30 amplitudes."""
    from . import mix
    Xt, yt, Xm, ym = [], [], [], []
    for n in sizes:
        st = backend.alloc(n, C64)
        for seed in seeds:
            circ = mix(n, 40 * n, seed)
            for form in forms:
                planned, _ = plan_passes(backend, None, None, circ, form)
                info = plan_info(planned, features=True)
                p = Program(st, planned, precision=PREC_MFMA_BF16X3, planned=True)
                backend.init_basis(st, 0)
                p.launch()
                ts = np.min([p.time_kernels() for _ in range(reps)], axis=0)
                del p
                assert len(ts) == len(info)
                for (f, _, _, _, feat), ms in zip(info, ts):
                    y = np.log2(ms * 2.0 ** (30 - n))
                    if f == 0:
                        Xt.append(feat)
                        yt.append(y)
                    else:
                        Xm.append(feat)
                        ym.append(y)
        st.free()
    return (np.array(Xt, dtype=np.float32), np.array(yt, dtype=np.float32),
            np.array(Xm, dtype=np.float32), np.array(ym, dtype=np.float32))


__all__ = ["Mlp", "group_features", "collect_timings", "train_cost_model", "plan",
           "tile_pass_features", "collect_tile_timings", "train_tile_model", "plan_form",
           "N_MT_FEATURES", "mt_pass_features", "collect_mt_timings", "train_mt_model",
           "plan_passes", "plan_info", "collect_pass_timings", "FORM_LEARNED", "FORM_EXACT",
           "QsvhError"]
