"""MFMA tile passes: runs of fused groups (dense groups as split-BF16x3 MFMA products,
diagonal groups as f32 phases, dense gates of 1..3 qubits widened to 4) applied to one
2^12-amplitude LDS tile per HBM pass, against the complex128 oracle.

Tolerance: each op adds at most the split-BF16x3 product error (the lo*lo term dropped,
~2^-16 relative per product) or one f32 rounding; the per-group form is held to 1e-4
relative for one group, so a circuit of N ops is held to 1e-4 * sqrt(N) relative to max|psi|
and 1 - fidelity < 1e-8 * N (one group's 1e-4 squared, per op); and to within 3x the
per-group form's error on the same ops."""
import numpy as np
import pytest

import oracle
import rustatevec_amd as rs
from tests.helpers import max_abs_diff

pytestmark = pytest.mark.gpu


def random_unitary(d, rng):
    z = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def random_ops(n, count, rng):
    ops = []
    for _ in range(count):
        kind = rng.choice(["dense", "dense", "diag", "small"])
        if kind == "dense":
            k = int(rng.integers(4, 7))
        elif kind == "diag":
            k = int(rng.integers(1, 7))
        else:
            k = int(rng.integers(1, 4))
        qs = [int(x) for x in rng.choice(n, size=k, replace=False)]
        if kind == "diag":
            U = np.diag(np.exp(1j * rng.uniform(0, 2 * np.pi, size=1 << k)))
        else:
            U = random_unitary(1 << k, rng)
        ops.append((U, qs))
    return ops


def run(backend, n, ops, psi, **kw):
    c = rs.Circuit(n)
    for U, qs in ops:
        c.push(U, qs)
    st = backend.alloc(n, rs.C64)
    backend.upload(st, psi)
    p = rs.Program(st, c, precision=rs.PREC_MFMA_BF16X3, **kw)
    p.launch()
    out = backend.download(st)
    nk = p.num_kernels()
    del p
    st.free()
    return out, nk


def exact(n, ops, psi):
    want = psi.astype(np.complex128)
    for U, qs in ops:
        want = oracle.apply(oracle.CPU_SERIAL, n, want, U, qs)
    return want


def errors(got, want):
    # (inner products in f64: complex64 sums carry ~1e-7 of rounding, the size of the bar)
    got = got.astype(np.complex128)
    err = max_abs_diff(got, want) / np.abs(want).max()
    fid = abs(np.vdot(want, got)) ** 2 / (np.vdot(want, want).real * np.vdot(got, got).real)
    return err, 1 - fid


def check(got, want, n_ops):
    err, infid = errors(got, want)
    assert err < 1e-4 * np.sqrt(n_ops), err
    assert infid < 1e-8 * n_ops, infid
    return err


def test_random_group_mixes_match_the_oracle(backend):
    tb = 12
    rng = np.random.default_rng(100 + tb)
    for n in (tb, 15, 17):
        psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
        psi /= np.linalg.norm(psi)
        ops = random_ops(n, 24, rng)
        want = exact(n, ops, psi)
        got, nk = run(backend, n, ops, psi, mfma_tiles=tb)
        err = check(got, want, len(ops))
        assert nk < len(ops), (n, nk)
        per_group, _ = run(backend, n, ops, psi)
        assert err < 3 * errors(per_group, want)[0], (err, errors(per_group, want))


def test_diagonal_groups_outside_the_tile_and_passes_of_only_diagonals(backend):
    # diagonal groups on qubits far outside a pass's tile take their out-of-tile bits from
    # each tile's base; a run of only diagonal groups is one pass
    n = 18
    rng = np.random.default_rng(7)
    psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
    psi /= np.linalg.norm(psi)
    ops = []
    for qs in ([17, 16, 3], [0, 17], [12, 13, 14, 15, 16, 17], [5], [17, 1, 9, 14, 2, 11]):
        ops.append((np.diag(np.exp(1j * rng.uniform(0, 2 * np.pi, size=1 << len(qs)))), qs))
    got, nk = run(backend, n, ops, psi, mfma_tiles=12)
    assert nk == 1
    check(got, exact(n, ops, psi), len(ops))
    ops2 = [(random_unitary(64, rng), [12, 13, 14, 15, 16, 17])] + ops + \
           [(random_unitary(16, rng), [0, 5, 9, 17])]
    got, nk = run(backend, n, ops2, psi, mfma_tiles=12)
    check(got, exact(n, ops2, psi), len(ops2))


def test_commuting_groups_move_into_earlier_passes(backend):
    # {12..17} does not fit the first pass's tile, the diagonal on {17, 0} and the group on
    # {0..3} do not commute with what is left behind before them; the gate on 5 and the
    # diagonal on {6, 18, 19} do and join the first pass: two passes, the oracle's state
    n = 20
    rng = np.random.default_rng(5)
    psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
    psi /= np.linalg.norm(psi)
    ops = []
    for kind, qs in (("dense", [4, 5, 6, 7, 8, 9]), ("dense", [10, 11]),
                     ("dense", [12, 13, 14, 15, 16, 17]), ("diag", [17, 0]),
                     ("dense", [0, 1, 2, 3]), ("dense", [5]), ("diag", [6, 18, 19]),
                     ("dense", [18, 19])):
        d = 1 << len(qs)
        U = (random_unitary(d, rng) if kind == "dense"
             else np.diag(np.exp(1j * rng.uniform(0, 2 * np.pi, size=d))))
        ops.append((U, qs))
    want = exact(n, ops, psi)
    got, nk = run(backend, n, ops, psi, mfma_tiles=12)
    assert nk == 2
    err = check(got, want, len(ops))
    per_group, _ = run(backend, n, ops, psi)
    assert err < 3 * errors(per_group, want)[0], (err, errors(per_group, want))
    # the program reports the order it applies the ops in: the first pass takes ops 0, 1 and
    # the two that moved ahead (5, 6), the second the rest in circuit order
    c = rs.Circuit(n)
    for U, qs in ops:
        c.push(U, qs)
    st = backend.alloc(n, rs.C64)
    p = rs.Program(st, c, precision=rs.PREC_MFMA_BF16X3, mfma_tiles=12)
    assert p.order() == [0, 1, 5, 6, 2, 3, 4, 7]
    del p
    st.free()


def test_phase_table_budget_splits_long_diagonal_runs(backend):
    # a pass's diagonal phases (2048 entries, staged in LDS) bound its run: 40 six-qubit
    # diagonal groups (2560 entries) need two passes, applied in order
    n = 16
    rng = np.random.default_rng(11)
    psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
    psi /= np.linalg.norm(psi)
    ops = []
    for _ in range(40):
        qs = [int(x) for x in rng.choice(n, size=6, replace=False)]
        ops.append((np.diag(np.exp(1j * rng.uniform(0, 2 * np.pi, size=64))), qs))
    got, nk = run(backend, n, ops, psi, mfma_tiles=12)
    assert nk == 2
    check(got, exact(n, ops, psi), len(ops))


def test_fused_qft_in_tile_passes(backend):
    n = 16
    qft = rs.qft(n)
    fused = rs.fuse(backend, qft, 6)
    rng = np.random.default_rng(3)
    psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
    psi /= np.linalg.norm(psi)
    st = backend.alloc(n, rs.C64)
    outs = {}
    for name, kw in (("groups", {}), ("tiles12", {"mfma_tiles": 12})):
        backend.upload(st, psi)
        p = rs.Program(st, fused, precision=rs.PREC_MFMA_BF16X3, **kw)
        p.launch()
        outs[name] = (backend.download(st), p.num_kernels())
        del p
    st.free()
    want = oracle.execute(oracle.CPU_SERIAL, n, qft.ops(), dtype=np.float64,
                          start=psi.astype(np.complex128))
    for name, (got, nk) in outs.items():
        check(got, want, len(fused))
    assert outs["tiles12"][1] < outs["groups"][1] // 3


def test_smaller_than_a_tile_and_argument_errors(backend):
    n = 11
    rng = np.random.default_rng(5)
    psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
    psi /= np.linalg.norm(psi)
    ops = random_ops(n, 8, rng)
    got, _ = run(backend, n, ops, psi, mfma_tiles=12)  # per-group launches
    check(got, exact(n, ops, psi), len(ops))
    st = backend.alloc(14, rs.C64)
    c = rs.Circuit(14).h(0)
    for kw in ({"mfma_tiles": 11}, {"mfma_tiles": 13}, {"mfma_tiles": 12, "tile_qubits": 12},
               {"mfma_tiles": 12, "subspace": True}):
        with pytest.raises(rs.QsvhError):
            rs.Program(st, c, precision=rs.PREC_MFMA_BF16X3, **kw)
    with pytest.raises(rs.QsvhError):
        rs.Program(st, c, precision=rs.PREC_MFMA_F32, mfma_tiles=12)
    st.free()


def test_graph_replays_match_the_eager_run(backend):
    # a program of MFMA tile passes captured in a HIP graph: every replay (with eager work
    # and synchronisations between the replays) gives the eager launch's amplitudes
    n = 16
    rng = np.random.default_rng(21)
    psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
    psi /= np.linalg.norm(psi)
    ops = random_ops(n, 20, rng)
    want, _ = run(backend, n, ops, psi, mfma_tiles=12)
    c = rs.Circuit(n)
    for U, qs in ops:
        c.push(U, qs)
    st = backend.alloc(n, rs.C64)
    other = backend.alloc(n, rs.C64)
    p = rs.Program(st, c, precision=rs.PREC_MFMA_BF16X3, mfma_tiles=12, use_graph=True)
    for _ in range(4):
        backend.upload(st, psi)
        p.launch()
        backend.init_random_normal(other, 3)  # eager work between the replays
        assert abs(backend.norm_sqr(other) - 1) < 1e-5
        assert np.array_equal(backend.download(st), want)
    del p
    st.free()
    other.free()
