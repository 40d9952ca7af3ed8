"""Tile passes: runs of gates applied per 2^t-amplitude LDS tile, one HBM round trip per
pass.  A pass applies its gates in the program's recorded order (gates on disjoint
qubits may move past each other), with the per-gate kernels' arithmetic, so:

* the result equals the oracle's CpuBackend on the *reordered* op list bit for bit, and
* it matches the oracle on the original circuit to <= 1e-12 (complex128)."""
import ctypes as C

import numpy as np
import pytest

import oracle
import rustatevec_amd as rs
from tests.helpers import max_abs_diff

pytestmark = pytest.mark.gpu


def order_of(prog):
    n = C.c_size_t(0)
    rs.lib().qsvh_program_order(prog._h, None, 0, C.byref(n))
    out = np.zeros(max(n.value, 1), dtype=np.uint64)
    rs.lib().qsvh_program_order(prog._h, out.ctypes.data_as(C.c_void_p), n.value, C.byref(n))
    return [int(x) for x in out[: n.value]]


def run_tiled(backend, circ, tile, dtype=rs.C128, start=None):
    n = circ.n_qubits
    st = backend.alloc(n, dtype)
    if start is None:
        backend.init_basis(st, 0)
    else:
        backend.upload(st, start)
    p = rs.Program(st, circ, tile_qubits=tile)
    passes = p.num_kernels()
    p.launch()
    out = backend.download(st)
    order = order_of(p)
    del p
    st.free()
    return out, order, passes


def check(backend, circ, tile, dtype=rs.C128, tol=1e-12):
    n = circ.n_qubits
    ops = circ.ops()
    got, order, passes = run_tiled(backend, circ, tile, dtype)
    assert sorted(order) == list(range(len(ops)))
    npdt = np.float64 if dtype == rs.C128 else np.float32
    reordered = oracle.execute(oracle.CPU_PARALLEL, n, [ops[i] for i in order], dtype=npdt)
    assert max_abs_diff(got, reordered) == 0.0
    want = oracle.execute(oracle.CPU_PARALLEL, n, ops)
    assert max_abs_diff(got, want) < tol
    return passes


@pytest.mark.parametrize("tile", [10, 11, 12])
def test_random_circuits(backend, tile):
    for seed in range(12):
        n = tile + seed % 5
        check(backend, rs.random_circuit(n, 150, 300 + seed), tile)


def test_small_states_fall_back_to_plain(backend):
    for n in (3, 6, 9):
        got, order, passes = run_tiled(backend, rs.random_circuit(n, 60, n), 10)
        assert order == list(range(60)) and passes == 60
        want = oracle.execute(oracle.CPU_SERIAL, n, oracle.gen_random(n, 60, n))
        assert max_abs_diff(got, want) == 0.0


def test_workloads_pass_counts(backend):
    passes = check(backend, rs.qft(16), 12)
    assert passes <= 8, passes  # 16 H + 120 CP + 8 SWAP in a handful of passes
    passes = check(backend, rs.layered(16, 6, 0xC0FFEE), 12)
    assert passes < 6 * (16 * 2 + 8) // 4, passes
    check(backend, rs.hea(15, 8, 3), 11)
    check(backend, rs.mix(17, 400, 4), 12)
    check(backend, rs.ghz_chain(18), 10)


def test_controlled_gates_with_outside_controls(backend):
    # CX/CCX controls far above the tile: lowered to ops conditioned on the tile's bits.
    n = 16
    c = rs.Circuit(n)
    for q in range(n):
        c.h(q)
    for a, b in [(15, 0), (14, 3), (13, 12), (0, 15), (9, 2)]:
        c.cx(a, b)
    c.ccx(15, 14, 1).ccx(13, 2, 5).ccz(15, 7, 0).swap(0, 11).cphase(15, 3, 0.3)
    for q in range(n):
        c.ry(q, 0.1 * (q + 1))
    check(backend, c, 10)
    check(backend, c, 12)


def test_complex64_tiles(backend):
    for tile in (11, 12, 13):
        check(backend, rs.random_circuit(15, 200, tile), tile, rs.C64, tol=1e-5)


def test_tiles_with_wide_fused_groups(backend):
    # groups of 4..6 qubits are barriers between tile passes (their own kernels)
    n = 15
    circ = rs.random_circuit(n, 200, 9)
    fused = rs.fuse(backend, circ, 5)
    got, order, _ = run_tiled(backend, fused, 11)
    want = oracle.execute(oracle.CPU_PARALLEL, n, oracle.gen_random(n, 200, 9))
    assert max_abs_diff(got, want) < 1e-12
    ops = fused.ops()
    assert max_abs_diff(got, oracle.execute(oracle.CPU_PARALLEL, n,
                                            [ops[i] for i in order])) == 0.0


def test_tiles_from_random_start_and_graph_replay(backend):
    n = 14
    qft = rs.qft(n)
    ref = backend.alloc(n)
    backend.init_random_normal(ref, 3)
    psi = backend.download(ref)
    rs.Program(ref, qft).launch()
    want = backend.download(ref)
    st = backend.alloc(n)
    for use_graph in (False, True):
        p = rs.Program(st, qft, tile_qubits=11, use_graph=use_graph)
        for rep in range(3):
            backend.upload(st, psi)
            p.launch()
            assert max_abs_diff(backend.download(st), want) < 1e-12, (use_graph, rep)
        del p
    st.free()
    ref.free()


def test_bad_tile_size_is_an_error(backend):
    st = backend.alloc(14)
    with pytest.raises(rs.QsvhError) as e:
        rs.Program(st, rs.qft(14), tile_qubits=13)
    assert e.value.status == "INVALID_ARG"
    st.free()
