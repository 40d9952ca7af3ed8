"""The reference's known-answer cases (crates/qsv-core/tests/oracle.rs:12-91, TOL 1e-12)
through the GPU engine: per-gate execute on the registers the reference uses, and the same
circuits embedded in a 12-qubit register as exact tile passes (complex128 and complex64)."""
import numpy as np
import pytest

import rustatevec_amd as rs

pytestmark = pytest.mark.gpu

TOL = 1e-12
R = 2 ** -0.5


def bell(n):
    return rs.Circuit(n).h(0).cx(0, 1)


def ghz3(n):
    return rs.Circuit(n).h(0).cx(0, 1).cx(0, 2)


def x_twice(n):
    return rs.Circuit(n).x(2).x(2)


def cx_on_zero(n):
    return rs.Circuit(n).cx(0, 1)


def rx_pi(n):
    return rs.Circuit(n).rx(0, np.pi)


def mixed(n):
    c = rs.Circuit(n).h(0).rx(1, 0.4).ry(2, 1.1).cx(0, 1).cz(1, 2).rzz(2, 3, 0.7)
    return c.swap(0, 3).t(2)


# circuit, reference register size, expected nonzero amplitudes {index: value}
CASES = [
    (bell, 2, {0b00: R, 0b11: R}),
    (ghz3, 3, {0b000: R, 0b111: R}),
    (x_twice, 4, {0: 1.0}),
    (cx_on_zero, 2, {0: 1.0}),
    (rx_pi, 1, {1: -1j}),
]


def _run(backend, circ, dtype, tile):
    if tile:
        st = backend.alloc(circ.n_qubits, dtype)
        backend.init_basis(st, 0)
        p = rs.Program(st, circ, tile_qubits=tile)
        p.launch()
        del p
        return st
    return backend.execute(circ, dtype)


@pytest.mark.parametrize("form", ["plain", "tiles_c128", "tiles_c64"])
def test_known_answers(backend, form):
    for make, n0, want in CASES:
        n, dtype, tile, tol = {"plain": (n0, rs.C128, 0, TOL),
                               "tiles_c128": (12, rs.C128, 12, TOL),
                               "tiles_c64": (12, rs.C64, 12, 1e-7)}[form]
        st = _run(backend, make(n), dtype, tile)
        got = backend.download(st)
        exp = np.zeros(1 << n, dtype=complex)
        for i, v in want.items():
            exp[i] = v
        assert np.abs(got - exp).max() < tol, (make.__name__, form)
        assert abs(backend.norm_sqr(st) - 1.0) < tol
        st.free()


@pytest.mark.parametrize("form", ["plain", "tiles_c128"])
def test_norm_preserved_through_mixed_circuit(backend, form):
    n, tile = (4, 0) if form == "plain" else (12, 11)
    st = _run(backend, mixed(n), rs.C128, tile)
    assert abs(backend.norm_sqr(st) - 1.0) < 1e-10
    # the mixed circuit's state is the same in both forms and on both register sizes:
    # amplitudes of the embedded run are those of the 4-qubit run (qubits 4.. stay |0>)
    small = backend.execute(mixed(4), rs.C128)
    a = backend.download(small)
    b = backend.download(st)[: 16]
    assert np.abs(a - b).max() < TOL
    small.free()
    st.free()
