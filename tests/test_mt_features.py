"""MFMA tile-pass cost-model features (qsvh_mt_pass_features) on the CPU: fused circuits
whose pass split and op classification are known, so every count is pinned (a miscounted or
misclassified op, a missed merge or a wrong table size moves one of them)."""
import numpy as np
import pytest

import rustatevec_amd as rs
from rustatevec_amd import planner


def dense(rng, k):
    q, _ = np.linalg.qr(rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k)))
    return q


def diag(rng, m):
    return np.diag(np.exp(1j * rng.uniform(0, 2 * np.pi, size=1 << m)))


def test_one_pass_counts_dense_merged_and_table_entries():
    # n = 14: the dense groups' qubits {0..3} and {4..8} with qubits 0..3 fill 9 of the 12
    # tile bits; the tile is qubits 0..11.  The two diagonal groups between them are
    # consecutive with in-tile qubits {5, 4} (qubit 13 is outside): one merged run, a head
    # over 2 tile bits whose base table holds the group on {4} (every qubit in the tile),
    # and one member (the group on {5, 13}).  Phase table: the base table's 4 entries and
    # the member's 4.
    rng = np.random.default_rng(1)
    c = rs.Circuit(14)
    c.push(dense(rng, 4), [0, 1, 2, 3])
    c.push(diag(rng, 2), [5, 13])
    c.push(diag(rng, 1), [4])
    c.push(dense(rng, 5), [4, 5, 6, 7, 8])
    rows, other = planner.mt_pass_features(c)
    assert other == 0 and rows.shape == (1, planner.N_MT_FEATURES)
    want = np.zeros(planner.N_MT_FEATURES, dtype=np.float32)
    want[0] = 1 / 16          # dense 4-qubit
    want[1] = 1 / 16          # dense 5-qubit
    want[4] = 1 / 16          # merged run heads
    want[5] = 1 / 64          # their members
    want[6] = 4 / 2048        # per-tile table of the head (2 tile bits)
    want[7] = 8 / 2048        # the pass's phase table
    want[8] = 1.0             # widest group: 5 qubits
    want[10] = 14 / 32        # the state size
    np.testing.assert_allclose(rows[0], want, rtol=1e-6)


def test_a_run_inside_the_tile_is_one_diagonal_op():
    # the same pass with the diagonal groups on {4, 5} and {6}: every qubit of the run is in
    # the tile, so its product over tile bits {4, 5, 6} is composed once, on the host, into
    # one diagonal op of 8 phases -- no head, no members
    rng = np.random.default_rng(1)
    c = rs.Circuit(14)
    c.push(dense(rng, 4), [0, 1, 2, 3])
    c.push(diag(rng, 2), [4, 5])
    c.push(diag(rng, 1), [6])
    c.push(dense(rng, 5), [4, 5, 6, 7, 8])
    rows, other = planner.mt_pass_features(c)
    assert other == 0 and rows.shape == (1, planner.N_MT_FEATURES)
    want = np.zeros(planner.N_MT_FEATURES, dtype=np.float32)
    want[0] = want[1] = 1 / 16
    want[3] = 1 / 32          # one single diagonal op
    want[7] = 8 / 2048
    want[8] = 1.0
    want[10] = 14 / 32
    np.testing.assert_allclose(rows[0], want, rtol=1e-6)


def test_passes_split_at_the_tile_and_small_gates_widen():
    # n = 20: a dense group on {12..17} cannot join the first pass (tile {0..11} after
    # {4..11}): two passes.  The 6-qubit diagonal after it shares no qubit with it, so it
    # joins the first pass; a lone diagonal between two dense groups stays a single op; a
    # 2-qubit dense gate counts as a widened 4-qubit group.
    rng = np.random.default_rng(2)
    c = rs.Circuit(20)
    c.push(dense(rng, 6), [4, 5, 6, 7, 8, 9])
    c.push(diag(rng, 3), [0, 15, 19])
    c.push(dense(rng, 2), [10, 11])
    c.push(dense(rng, 6), [12, 13, 14, 15, 16, 17])
    c.push(diag(rng, 6), [0, 1, 2, 3, 18, 19])
    rows, other = planner.mt_pass_features(c)
    assert other == 0 and rows.shape[0] == 2
    first, second = rows
    assert first[2] == pytest.approx(1 / 16) and first[0] == pytest.approx(1 / 16)
    assert first[3] == pytest.approx(2 / 32) and first[4] == 0
    assert first[7] == pytest.approx(72 / 2048) and first[9] == 1.0
    want = np.zeros(planner.N_MT_FEATURES, dtype=np.float32)
    want[2] = 1 / 16
    want[9] = 1.0
    want[10] = 20 / 32
    np.testing.assert_allclose(second, want, rtol=1e-6)


def test_ops_that_commute_with_the_ops_left_behind_move_ahead():
    # n = 20.  {12..17} does not fit the first pass's tile; the diagonal on {17, 0} shares
    # qubit 17 with it and the group on {0..3} qubit 0 with that diagonal: both stay behind.
    # The 1-qubit gate on 5 and the diagonal on {6, 18, 19} commute with all three and join
    # the first pass; the gate on {18, 19} commutes too but would widen the tile past 12 bits.
    rng = np.random.default_rng(3)
    c = rs.Circuit(20)
    c.push(dense(rng, 6), [4, 5, 6, 7, 8, 9])
    c.push(dense(rng, 2), [10, 11])
    c.push(dense(rng, 6), [12, 13, 14, 15, 16, 17])
    c.push(diag(rng, 2), [17, 0])
    c.push(dense(rng, 4), [0, 1, 2, 3])
    c.push(dense(rng, 1), [5])
    c.push(diag(rng, 3), [6, 18, 19])
    c.push(dense(rng, 2), [18, 19])
    rows, other = planner.mt_pass_features(c)
    assert other == 0 and rows.shape == (2, planner.N_MT_FEATURES)
    first = np.zeros(planner.N_MT_FEATURES, dtype=np.float32)
    first[0] = 2 / 16      # {10, 11} and {5} widened to 4 qubits
    first[2] = 1 / 16      # {4..9}
    first[3] = 1 / 32      # the diagonal on {6, 18, 19}
    first[7] = 8 / 2048
    first[9] = 1.0
    first[10] = 20 / 32
    second = np.zeros(planner.N_MT_FEATURES, dtype=np.float32)
    second[0] = 2 / 16     # {0..3} and {18, 19} widened
    second[2] = 1 / 16     # {12..17}
    second[3] = 1 / 32     # the diagonal on {17, 0}
    second[7] = 4 / 2048
    second[9] = 1.0
    second[10] = 20 / 32
    np.testing.assert_allclose(rows[0], first, rtol=1e-6)
    np.testing.assert_allclose(rows[1], second, rtol=1e-6)


def test_the_commuting_scan_is_bounded():
    # n = 20: the tile fills with {4..11}; 1100 gates on qubit 13 are left behind one after
    # another, and the scan gives up after 1024 of them, so the gate on qubit 5 at the end
    # (which commutes with all of them) is not pulled into the first pass
    rng = np.random.default_rng(4)
    c = rs.Circuit(20)
    c.push(dense(rng, 6), [4, 5, 6, 7, 8, 9])
    c.push(dense(rng, 2), [10, 11])
    for _ in range(1100):
        c.push(dense(rng, 1), [13])
    c.push(dense(rng, 1), [5])
    rows, other = planner.mt_pass_features(c)
    assert other == 0 and rows.shape[0] == 2
    assert rows[0][2] == pytest.approx(1 / 16) and rows[0][0] == pytest.approx(1 / 16)
    assert rows[1][0] == pytest.approx(1101 / 16)
