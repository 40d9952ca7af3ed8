"""Trailing qubit permutations in tile-pass and fused programs: the SWAP network that ends a
circuit (a QFT's), alone or fused into groups whose matrices only move bits, runs as one or
two bit-swap passes (each exchanges the state bits of disjoint pairs, in place).  A SWAP
moves amplitudes without arithmetic, so the result is the input permuted exactly: a numpy
index permutation agrees with it bit for bit, and the oracle's dense SWAP (products by exact
1 and 0, crates/qsv-core/src/backend/cpu.rs:221-247) in value (the two differ at most in
the sign of a zero amplitude; the random start states have none), in complex128 and complex64, at
state sizes from two qubits to a few tiles' bits, for involutions (one pass) and for
longer cycles of qubits (two passes)."""
import numpy as np
import pytest

import oracle
import rustatevec_amd as rs
from tests.helpers import max_abs_diff
from tests.test_gpu_tile import check, order_of

pytestmark = pytest.mark.gpu


def permuted(psi, dest):
    """psi with the data of state bit b moved to bit dest[b]."""
    n = len(dest)
    i = np.arange(1 << n, dtype=np.int64)
    j = np.zeros_like(i)
    for b in range(n):
        j |= ((i >> b) & 1) << dest[b]
    out = np.empty_like(psi)
    out[j] = psi
    return out


def swap_network(n, rng, length):
    """A circuit of `length` random SWAPs and where each state bit's data ends."""
    c = rs.Circuit(n)
    dest = list(range(n))
    for _ in range(length):
        a, b = rng.choice(n, size=2, replace=False)
        c.swap(int(a), int(b))
        # the data now at bit a goes to b and vice versa
        dest = [int(b) if d == a else int(a) if d == b else d for d in dest]
    return c, dest


def random_state(rng, n, dtype):
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    return psi.astype(np.complex128 if dtype == rs.C128 else np.complex64)


@pytest.mark.parametrize("max_blocks", [0, 64])
@pytest.mark.parametrize("dtype", [rs.C128, rs.C64])
def test_swap_networks_move_amplitudes_exactly(backend, dtype, max_blocks):
    # max_blocks 64: workgroups loop over several pairs of tiles (the next pair's loads
    # issued before this pair's stores)
    rng = np.random.default_rng(11 + max_blocks)
    tile = 10 if dtype == rs.C128 else 11
    backend.tune(max_blocks=max_blocks)
    try:
        swap_networks(backend, dtype, rng, tile)
    finally:
        backend.tune(max_blocks=0)


def swap_networks(backend, dtype, rng, tile):
    for n in (2, 3, 5, 6, 7, 9, 11, 12, 14, 17, 20):
        for length in (1, 2, 3, n, 2 * n):
            circ, dest = swap_network(n, rng, length)
            psi = random_state(rng, n, dtype)
            st = backend.alloc(n, dtype)
            backend.upload(st, psi)
            p = rs.Program(st, circ, tile_qubits=tile)
            kernels = p.num_kernels()
            order = order_of(p)
            p.launch()
            got = backend.download(st)
            del p
            st.free()
            print(f"n={n} swaps={length} kernels={kernels}", flush=True)
            assert order == list(range(length))
            assert kernels <= 2, (n, length, kernels)
            want = permuted(psi, dest)
            assert np.array_equal(got, want), (n, length)
            ops = circ.ops()
            npdt = np.float64 if dtype == rs.C128 else np.float32
            ref = oracle.execute(oracle.CPU_SERIAL, n, ops, dtype=npdt, start=psi)
            assert max_abs_diff(got, ref) == 0.0


def test_bit_reversal_is_one_pass(backend):
    # QFT's swaps (q, n-1-q) form an involution of the state bits: a single pass
    rng = np.random.default_rng(3)
    for n in (12, 15, 20):
        circ = rs.Circuit(n)
        for q in range(n // 2):
            circ.swap(q, n - 1 - q)
        psi = random_state(rng, n, rs.C128)
        st = backend.alloc(n, rs.C128)
        backend.upload(st, psi)
        p = rs.Program(st, circ, tile_qubits=12)
        assert p.num_kernels() == 1
        p.launch()
        got = backend.download(st)
        del p
        st.free()
        assert np.array_equal(got, permuted(psi, [n - 1 - b for b in range(n)]))


@pytest.mark.parametrize("tile", [10, 11, 12])
def test_qft_tile_programs_with_swap_suffix(backend, tile):
    # bitwise on the recorded order (the swaps last, as in the circuit); the tile passes
    # hold only the H and CP gates, and one bit-swap pass reverses the qubits
    want = {10: (3, 3, 4), 11: (3, 3, 3), 12: (3, 3, 3)}[tile]
    for n, kernels in zip((13, 16, 18), want):
        got = check(backend, rs.qft(n), tile)
        print(f"QFT({n}) t={tile}: {got} kernels", flush=True)
        assert got == kernels, (n, got)


def low_then_high_then_reversal(n, seed):
    """H everywhere, gates on qubits 0..7, gates on the top 8 qubits (with controlled
    phases from the low ones), then the qubit reversal."""
    rng = np.random.default_rng(seed)
    c = rs.Circuit(n)
    for q in range(n):
        c.h(q)
    for _ in range(40):
        a, b = rng.choice(8, 2, replace=False)
        c.cx(int(a), int(b)).u3(int(rng.integers(8)), *rng.uniform(0, 6, 3))
    for _ in range(60):
        a, b = rng.choice(np.arange(n - 8, n), 2, replace=False)
        c.cx(int(a), int(b)).u3(int(rng.integers(n - 8, n)), *rng.uniform(0, 6, 3))
        c.cphase(int(rng.integers(0, 6)), int(rng.integers(n - 8, n)), float(rng.uniform(0, 6)))
    for q in range(n // 2):
        c.swap(q, n - 1 - q)
    return c


@pytest.mark.parametrize("n", [16, 20])
def test_mfma_tiles_with_the_swaps_moved_ahead(backend, n):
    # fused at k = 4 as MFMA tile passes, the trailing permutation runs where the passes
    # come out fewest (here before the high part, whose groups then act on low qubits):
    # the launches the planner features count, within the fused forms' error bar
    from rustatevec_amd import planner
    circ = low_then_high_then_reversal(n, 1)
    fused = rs.fuse(backend, circ, 4)
    rows, other = planner.mt_pass_features(fused)
    h = rs.gate_matrix("h")
    kept, _ = planner.mt_pass_features(rs.Circuit.from_ops(n, fused.ops() + [(h, [0])]))
    assert len(rows) < len(kept)
    plain = backend.alloc(n, rs.C128)
    backend.init_random_normal(plain, 2)
    rs.Program(plain, circ).launch()
    st = backend.alloc(n, rs.C64)
    backend.init_random_normal(st, 2)
    p = rs.Program(st, fused, precision=rs.PREC_MFMA_BF16X3, mfma_tiles=12)
    assert p.num_kernels() == len(rows) + other
    assert sorted(p.order()) == list(range(len(fused)))
    p.launch()
    err, fid = backend.compare(plain, st)
    del p
    st.free()
    plain.free()
    assert err < 1e-6 and 1.0 - fid < 1e-8, (err, 1.0 - fid)


def test_circuits_ending_in_swaps_complex64_tiles(backend):
    rng = np.random.default_rng(5)
    for n, tile in ((13, 11), (15, 12), (16, 13)):
        body = rs.random_circuit(n, 120, n)
        net, _ = swap_network(n, rng, n)
        circ = rs.Circuit.from_ops(n, body.ops() + net.ops())
        check(backend, circ, tile, dtype=rs.C64, tol=1e-5)


@pytest.mark.parametrize("mfma_tiles", [0, 12])
def test_fused_swap_groups_are_data_movement(backend, mfma_tiles):
    # fused groups of only SWAPs move amplitudes exactly, where a split-BF16 product by a
    # permutation matrix would round every amplitude to about 2^-17 relative
    rng = np.random.default_rng(9)
    n = 16
    net, dest = swap_network(n, rng, 24)
    fused = rs.fuse(backend, net, 4)
    psi = random_state(rng, n, rs.C64)
    st = backend.alloc(n, rs.C64)
    backend.upload(st, psi)
    p = rs.Program(st, fused, precision=rs.PREC_MFMA_BF16X3, mfma_tiles=mfma_tiles)
    assert p.num_kernels() <= 2
    p.launch()
    got = backend.download(st)
    del p
    st.free()
    assert np.array_equal(got, permuted(psi, dest))


@pytest.mark.parametrize("mfma_tiles", [0, 12])
def test_fused_qft_with_swap_groups(backend, mfma_tiles):
    # the fused QFT: its last group mixes gates and swaps (kept as a matrix), the rest of
    # the swap network runs as bit swaps; within the fused forms' error bar of complex128
    n = 18
    circ = rs.qft(n)
    fused = rs.fuse(backend, circ, 4)
    plain = backend.alloc(n, rs.C128)
    backend.init_random_normal(plain, 4)
    rs.Program(plain, circ).launch()
    st = backend.alloc(n, rs.C64)
    backend.init_random_normal(st, 4)
    p = rs.Program(st, fused, precision=rs.PREC_MFMA_BF16X3, mfma_tiles=mfma_tiles)
    order = p.order()
    p.launch()
    err, fid = backend.compare(plain, st)
    del p
    st.free()
    plain.free()
    assert sorted(order) == list(range(len(fused)))
    assert err < 1e-6 and 1.0 - fid < 1e-8, (err, 1.0 - fid)
