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


def _perm_matrix(perm, phases):
    d = len(perm)
    u = np.zeros((d, d), dtype=np.complex128)
    for r, c in enumerate(perm):
        u[r, c] = phases[r]
    return u


def test_lowering_forms(backend):
    """Every form the tile lowering produces, bit for bit against the oracle from a random
    start: real and complex dense 1q ops, anti-diagonal (X, Y) and swap pair variants,
    controls inside and outside the tile as conditions, register blocks of 1q runs,
    permutation blocks with unit and non-unit coefficients (SWAP, CCX, custom phased 2q/3q
    permutations), real dense 2q blocks, and diagonal blocks of non-diagonal gates."""
    rng = np.random.default_rng(5)
    for n, tile, dtype in ((13, 10, rs.C128), (14, 12, rs.C128), (15, 13, rs.C64),
                           (12, 12, rs.C128)):
        c = rs.Circuit(n)
        for rep in range(3):
            for q in range(n):
                c.h(q)               # real dense
                c.ry(q, 0.3 + q)     # real dense
                c.x((q + rep) % n)   # swap variant
                c.y((q + 2) % n)     # anti-diagonal with phases
                c.u3(q, 0.1 * q, 0.2, 0.3 + rep)  # complex dense, 1q runs -> register blocks
            for q in range(0, n - 1, 2):
                c.cx(q, q + 1)       # control inside the tile (condition)
                if n - 1 - q != q:
                    c.cx(n - 1 - q, q)   # control often outside the tile
                c.swap(q, (q + 3) % n)
            c.ccx(0, 1, 2)
            c.ccx(n - 1, n - 2, 3)
            c.ccz(1, 4, n - 1)
            # custom permutations with phases
            ph2 = np.exp(1j * rng.uniform(0, 2 * np.pi, 4))
            c.push(_perm_matrix([2, 0, 3, 1], ph2), [1, 5])
            c.push(_perm_matrix([1, 0, 2, 3], [1, 1, 1, 1j]), [2, 3])
            ph3 = np.exp(1j * rng.uniform(0, 2 * np.pi, 8))
            c.push(_perm_matrix([3, 6, 1, 0, 7, 2, 5, 4], ph3), [0, 4, 6])
            # real dense 2q (not a permutation), and a 2q gate diagonal in one qubit only
            hh = np.kron([[1, 1], [1, -1]], [[1, 1], [1, -1]]) / 2.0
            c.push(hh, [3, 7])
            blk = np.zeros((4, 4), dtype=np.complex128)
            blk[:2, :2] = np.diag([1j, -1])
            blk[2:, 2:] = [[0.6, 0.8], [-0.8, 0.6]]
            c.push(blk, [2, 9])      # bit 0 (qubit 2) acts, bit 1 (qubit 9) controls
        ops = c.ops()
        st = backend.alloc(n, dtype)
        backend.init_random_normal(st, n)
        start = backend.download(st)
        p = rs.Program(st, c, tile_qubits=tile)
        p.launch()
        got = backend.download(st)
        order = order_of(p)
        del p
        st.free()
        assert sorted(order) == list(range(len(ops)))
        npdt = np.float64 if dtype == rs.C128 else np.float32
        want = oracle.execute(oracle.CPU_PARALLEL, n, [ops[i] for i in order], dtype=npdt,
                              start=start)
        assert max_abs_diff(got, want) == 0.0, (n, tile)


def test_controlled_phase_members(backend):
    """Controlled phases get lane-run members of their own -- a phase controlled by a bit
    outside the tile (one test of the tile's base per gate) and stretches of phases on two
    in-tile bits sharing one register bit -- and lane runs choose their register bits from
    the 1-qubit ops and diagonal gates around them.  Ladders of CP / CZ in random partner
    order, diagonals with a phase on another pattern (not lowered that way) and 1-qubit gates
    between them, from a random start: bitwise against the oracle on the recorded order."""
    rng = np.random.default_rng(7)
    for n, tile, dtype in ((14, 10, rs.C128), (15, 12, rs.C128), (15, 11, rs.C64),
                           (16, 13, rs.C64)):
        c = rs.Circuit(n)
        for j in rng.permutation(n)[: n // 2 + 2]:
            j = int(j)
            c.h(j)
            for k in rng.permutation(n):
                k = int(k)
                if k == j:
                    continue
                r = rng.uniform()
                if r < 0.6:
                    c.cphase(k, j, float(rng.uniform(0, 2 * np.pi)))
                elif r < 0.8:
                    c.cz(k, j)
                else:  # phase where bit j is 1 and bit k is 0: another table pattern
                    c.push(np.diag([1, np.exp(1j * rng.uniform(0, 6)), 1, 1]), [j, k])
            c.u3(int(rng.integers(n)), 0.4, 0.5, 0.6)
        ops = c.ops()
        st = backend.alloc(n, dtype)
        backend.init_random_normal(st, n + 1)
        start = backend.download(st)
        p = rs.Program(st, c, tile_qubits=tile)
        p.launch()
        got = backend.download(st)
        order = order_of(p)
        del p
        st.free()
        assert sorted(order) == list(range(len(ops)))
        npdt = np.float64 if dtype == rs.C128 else np.float32
        want = oracle.execute(oracle.CPU_PARALLEL, n, [ops[i] for i in order], dtype=npdt,
                              start=start)
        assert max_abs_diff(got, want) == 0.0, (n, tile)
