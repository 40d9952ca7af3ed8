"""Write tests/golden/golden.npz: gate matrices and final states computed with numpy and
scipy only, independently of the C oracle and of the engine.

* Gate matrices come from the Pauli algebra and ``scipy.linalg.expm``
  (RX = expm(-i t X/2), RZZ = expm(-i t Z(x)Z / 2), U3 = e^{i(p+l)/2} RZ(p) RY(t) RZ(l), ...),
  laid out in the reference's qubit convention (crates/qsv-core/src/gate.rs:8-12:
  internal index bit j <-> qs[j], control first).
* Circuits are gate records ``(name, qubits, params)``.  The random generators are
  re-implemented here from their published description (SplitMix64, Vigna;
  crates/qsv-core/src/circuits.rs:79-113 for ``random_circuit``), so that the engine's and
  the oracle's generators are both compared against a third implementation.
* States are evolved by axis permutation + matrix product on a (2,)*n tensor, and the
  result is cross-checked here against the full 2^n x 2^n Kronecker unitary for n <= 6.

Run: ``python tests/golden/make_golden.py`` (writes golden.npz next to this file).
"""
from __future__ import annotations

import json
import os

import numpy as np
from scipy.linalg import expm

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "golden.npz")

I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
P0 = np.diag([1, 0]).astype(complex)
P1 = np.diag([0, 1]).astype(complex)


def rx(t): return expm(-0.5j * t * X)
def ry(t): return expm(-0.5j * t * Y)
def rz(t): return expm(-0.5j * t * Z)


def u3(t, p, l):
    return np.exp(0.5j * (p + l)) * rz(p) @ ry(t) @ rz(l)


def ctrl(u):
    # internal index = bit_control + 2 * bit_target -> kron(target, control)
    return np.kron(I2, P0) + np.kron(u, P1)


def gate(name: str, params=()):
    p = list(params) + [0.0] * 3
    s = 1 / np.sqrt(2)
    table = {
        "h": lambda: (X + Z) * s,
        "x": lambda: X,
        "y": lambda: Y,
        "z": lambda: Z,
        "s": lambda: np.diag([1, 1j]),
        "t": lambda: np.diag([1, np.exp(0.25j * np.pi)]),
        "sx": lambda: 0.5 * np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]]),
        "rx": lambda: rx(p[0]),
        "ry": lambda: ry(p[0]),
        "rz": lambda: rz(p[0]),
        "p": lambda: np.exp(0.5j * p[0]) * rz(p[0]),
        "u3": lambda: u3(p[0], p[1], p[2]),
        "cx": lambda: ctrl(X),
        "cz": lambda: ctrl(Z),
        "cp": lambda: ctrl(np.exp(0.5j * p[0]) * rz(p[0])),
        "swap": lambda: 0.5 * (np.eye(4) + np.kron(X, X) + np.kron(Y, Y) + np.kron(Z, Z)),
        "rzz": lambda: expm(-0.5j * p[0] * np.kron(Z, Z)),
        # [c0, c1, t]: index = c0 + 2 c1 + 4 t -> kron(t, c1, c0)
        "ccx": lambda: np.kron(I2, np.eye(4) - np.kron(P1, P1)) + np.kron(X, np.kron(P1, P1)),
        "ccz": lambda: np.kron(I2, np.eye(4) - np.kron(P1, P1)) + np.kron(Z, np.kron(P1, P1)),
    }
    return np.asarray(table[name](), dtype=complex)


ARITY = {"h": 1, "x": 1, "y": 1, "z": 1, "s": 1, "t": 1, "sx": 1, "rx": 1, "ry": 1,
         "rz": 1, "p": 1, "u3": 1, "cx": 2, "cz": 2, "cp": 2, "swap": 2, "rzz": 2,
         "ccx": 3, "ccz": 3}


# ------------------------------------------------------------------ state evolution
def apply(psi: np.ndarray, n: int, g: np.ndarray, qs) -> np.ndarray:
    """Apply g (row index bit j <-> qs[j]) by moving qubit axes to the front."""
    m = len(qs)
    t = psi.reshape([2] * n)  # axis k <-> qubit n-1-k
    axes = [n - 1 - qs[j] for j in reversed(range(m))]  # MSB of the internal index first
    t = np.moveaxis(t, axes, list(range(m)))
    shp = t.shape
    t = (g @ t.reshape(1 << m, -1)).reshape(shp)
    t = np.moveaxis(t, list(range(m)), axes)
    return t.reshape(-1)


def full_unitary(n: int, g: np.ndarray, qs) -> np.ndarray:
    """The 2^n x 2^n operator of g on qs, built entry by entry from bit patterns."""
    dim = 1 << n
    U = np.zeros((dim, dim), dtype=complex)
    mask = sum(1 << q for q in qs)
    for c in range(dim):
        gc = sum(((c >> q) & 1) << j for j, q in enumerate(qs))
        for gr in range(1 << len(qs)):
            r = (c & ~mask) | sum(((gr >> j) & 1) << q for j, q in enumerate(qs))
            U[r, c] += g[gr, gc]
    return U


def evolve(n: int, recs, psi0=None) -> np.ndarray:
    psi = np.zeros(1 << n, dtype=complex)
    psi[0] = 1
    if psi0 is not None:
        psi = psi0.astype(complex).copy()
    for name, qs, ps in recs:
        psi = apply(psi, n, gate(name, ps), list(qs))
    return psi


# ------------------------------------------------------------------ generators
class SplitMix64:
    """Vigna's SplitMix64 (https://prng.di.unimi.it/splitmix64.c)."""

    M = (1 << 64) - 1

    def __init__(self, seed):
        self.s = seed & self.M

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & self.M
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.M
        return z ^ (z >> 31)

    def below(self, n):
        return self.next() % n

    def unit(self):
        return (self.next() >> 11) / float(1 << 53)

    def angle(self):
        return self.unit() * 2 * np.pi


def gen_random(n, n_gates, seed):
    rng = SplitMix64(seed)
    fixed = ["h", "x", "y", "z", "s", "t", "sx"]
    rot = ["rx", "ry", "rz", "p"]
    out = []
    for _ in range(n_gates):
        a = rng.below(n)
        c = rng.below(15)
        if c < 7:
            out.append((fixed[c], (a,), ()))
        elif c < 11:
            out.append((rot[c - 7], (a,), (rng.angle(),)))
        else:
            b = rng.below(n)
            while b == a:
                b = rng.below(n)
            w = rng.below(4)
            if w < 3:
                out.append((["cx", "cz", "swap"][w], (a, b), ()))
            else:
                out.append(("rzz", (a, b), (rng.angle(),)))
    return out


def gen_layered(n, depth, seed):
    rng = SplitMix64(seed)
    out = []
    for layer in range(depth):
        for j in range(n):
            t, f, la = rng.angle(), rng.angle(), rng.angle()
            out.append(("u3", (j,), (t, f, la)))
        for j in range(n):
            out.append(("rz", (j,), (rng.angle(),)))
        for j in range(layer & 1, n - 1, 2):
            out.append(("cz", (j, j + 1), ()))
    return out


def gen_hea(n, depth, seed):
    rng = SplitMix64(seed)
    out = []
    for _ in range(depth):
        for j in range(n):
            out.append(("ry", (j,), (rng.angle(),)))
        for j in range(n):
            out.append(("rz", (j,), (rng.angle(),)))
        for j in range(n - 1):
            out.append(("cx", (j, j + 1), ()))
    return out


def gen_mix(n, n_gates, seed):
    rng = SplitMix64(seed)
    out = []

    def other(a, b=None):
        while True:
            x = rng.below(n)
            if x != a and (b is None or x != b):
                return x

    for _ in range(n_gates):
        r = rng.below(100)
        a = rng.below(n)
        if r < 50:
            t, f, la = rng.angle(), rng.angle(), rng.angle()
            out.append(("u3", (a,), (t, f, la)))
        elif r < 80:
            b = other(a)
            out.append(("cx" if rng.below(2) == 0 else "cz", (a, b), ()))
        elif r < 95:
            b = other(a)
            out.append(("rzz", (a, b), (rng.angle(),)))
        else:
            b = other(a)
            c = other(a, b)
            out.append(("ccx" if rng.below(2) == 0 else "ccz", (a, b, c), ()))
    return out


def gen_qft(n):
    out = []
    for j in range(n):
        out.append(("h", (j,), ()))
        for k in range(j + 1, n):
            out.append(("cp", (k, j), (np.pi / (1 << (k - j)),)))
    for j in range(n // 2):
        out.append(("swap", (j, n - 1 - j), ()))
    return out


def gen_ghz_star(n):
    return [("h", (0,), ())] + [("cx", (0, q), ()) for q in range(1, n)]


def gen_ghz_chain(n):
    return [("h", (0,), ())] + [("cx", (i, i + 1), ()) for i in range(n - 1)]


# ------------------------------------------------------------------ cases
def cases():
    cs = []
    cs.append(("bell", 2, [("h", (0,), ()), ("cx", (0, 1), ())]))
    cs.append(("ghz3", 3, gen_ghz_star(3)))
    cs.append(("ghz_star6", 6, gen_ghz_star(6)))
    cs.append(("ghz_chain7", 7, gen_ghz_chain(7)))
    cs.append(("xx4", 4, [("x", (2,), ()), ("x", (2,), ())]))
    cs.append(("rx_pi", 1, [("rx", (0,), (np.pi,))]))
    cs.append(("mixed4", 4, [("h", (0,), ()), ("rx", (1,), (0.4,)), ("ry", (2,), (1.1,)),
                             ("cx", (0, 1), ()), ("cz", (1, 2), ()), ("rzz", (2, 3), (0.7,)),
                             ("swap", (0, 3), ()), ("t", (2,), ())]))
    for n in (1, 2, 3, 5, 6):
        cs.append((f"qft{n}", n, gen_qft(n)))
    for seed in range(12):
        n = 3 + seed % 6
        cs.append((f"random_s{seed}", n, gen_random(n, 20 + seed % 40, seed)))
    cs.append(("random10", 10, gen_random(10, 80, 9000)))
    cs.append(("layered6", 6, gen_layered(6, 3, 0xC0FFEE)))
    cs.append(("layered7", 7, gen_layered(7, 2, 11)))
    cs.append(("hea5", 5, gen_hea(5, 2, 3)))
    cs.append(("mix6", 6, gen_mix(6, 60, 5)))
    cs.append(("mix8", 8, gen_mix(8, 80, 6)))
    cs.append(("ccx_orders", 6, [("h", (q,), ()) for q in range(6)]
               + [("ry", (q,), (0.3 * (q + 1),)) for q in range(6)]
               + [("ccx", (0, 1, 2), ()), ("ccx", (2, 0, 4), ()), ("ccx", (4, 3, 1), ()),
                  ("ccx", (5, 2, 0), ()), ("ccz", (1, 4, 5), ())]))
    return cs


GATE_SAMPLES = [("h", ()), ("x", ()), ("y", ()), ("z", ()), ("s", ()), ("t", ()), ("sx", ()),
                ("rx", (0.7,)), ("ry", (1.3,)), ("rz", (2.1,)), ("p", (0.9,)),
                ("u3", (0.4, 1.7, -2.2)), ("cx", ()), ("cz", ()), ("cp", (0.9,)),
                ("swap", ()), ("rzz", (0.5,)), ("ccx", ()), ("ccz", ())]


def main():
    arrays = {}
    meta = {"cases": [], "gates": []}
    for name, params in GATE_SAMPLES:
        key = f"gate_{name}"
        arrays[key] = gate(name, params)
        meta["gates"].append({"key": key, "name": name, "params": list(params)})
    for cname, n, recs in cases():
        psi = evolve(n, recs)
        if n <= 6:  # independent cross-check: product of full unitaries
            U = np.eye(1 << n, dtype=complex)
            for name, qs, ps in recs:
                U = full_unitary(n, gate(name, ps), list(qs)) @ U
            assert np.abs(U[:, 0] - psi).max() < 1e-12, cname
        arrays[f"state_{cname}"] = psi
        meta["cases"].append({"name": cname, "n": n,
                              "ops": [[g, list(q), list(p)] for g, q, p in recs]})
    # a random normalised start state for upload tests (Box-Muller on SplitMix64)
    rng = SplitMix64(0xC0FFEE)
    n0 = 7
    u1 = np.array([rng.unit() for _ in range(2 << n0)]).reshape(2, -1)
    u1 = np.clip(u1, 1e-300, None)
    v = np.sqrt(-2 * np.log(u1[0])) * np.exp(2j * np.pi * u1[1])
    v /= np.linalg.norm(v)
    arrays["start_random7"] = v
    recs = gen_qft(n0)
    arrays["state_qft7_from_random"] = evolve(n0, recs, v)
    meta["splitmix64_seed0"] = [hex(x) for x in (lambda r: [r.next() for _ in range(4)])(
        SplitMix64(0))]
    arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT}: {len(meta['cases'])} circuits, {len(meta['gates'])} gates")


if __name__ == "__main__":
    main()
