"""Tile-pass cost-model features (qsvh_tile_pass_features) on the CPU: one pass whose
lowering is known op by op, so every count is pinned (a miscounted or misclassified op
moves one of them)."""
import numpy as np

import rustatevec_amd as rs
from rustatevec_amd import planner


def test_features_count_every_lowered_op():
    # n = 16, 12 tile bits: the scheduler takes bits 0..3, the action qubits in order and
    # fills up with the lowest bits, so the tile is qubits 0..11.  Tile-local bits 0..3 are
    # qubits 0..3; the other tile qubits take bits 4..11 least active first (activity: 2
    # per non-diagonal op acting on the qubit, 1 per diagonal gate on it): q8, q9 (0), q4,
    # q5, q6, q7, q11 (2), q10 (3) -> bits 4..11.  Lane runs choose 3 register bits each:
    # a 1-qubit op's bit joins them while there is room, an op whose bit is not among them
    # and whose lane position would be 4 or more starts a new run, and a run that closes
    # with room left takes the bits its diagonal gates use most.
    c = rs.Circuit(16)
    c.h(0)                      # run 1 (register bits 0, 1, 11): register, real (cheap)
    c.u3(1, 0.1, 0.2, 0.3)      # run 1: register, general
    c.u3(10, 0.4, 0.5, 0.6)     # run 1: register, general (bit 11)
    c.h(11)                     # bit 10: lane position 8, so run 2: register, cheap
    c.h(7)                      # run 2 (bit 9): register, cheap
    c.rz(3, 0.7)                # run 2: diagonal on a lane bit
    c.rz(10, 0.8)               # run 2: diagonal on a lane bit (bit 11, lane position 8)
    c.rz(14, 0.9)               # run 2: diagonal outside: phase uniform per tile
    c.gate("ccz", [0, 1, 2])    # run 2: diagonal on 3 qubits (run 2 closes with bits 0, 9, 10)
    c.swap(4, 5)                # permutation op (closes run 2)
    c.cx(13, 2)                 # run 3: register, swap variant, outside control
    c.cx(2, 1)                  # run 3: register, swap variant, conditioned on bit 2
    c.h(6)                      # run 3 (bits 1, 2, 8): register, cheap
    rows, other = planner.tile_pass_features(c, rs.C128, 12)
    assert rows.shape == (1, planner.N_TILE_FEATURES) and other == 0
    f = rows[0]
    assert f[0] == 0.0 and np.isclose(f[1], 12 / 13) and np.isclose(f[2], 16 / 32)
    counts = np.rint(f[3:] * 16).astype(int)
    want = {3: 1, 4: 2, 5: 0, 6: 1, 7: 0, 8: 0, 9: 2, 10: 6, 11: 0, 12: 0, 13: 0, 14: 0,
            15: 1, 16: 3, 17: 1}
    assert {i + 3: int(v) for i, v in enumerate(counts)} == want


def test_features_of_the_workloads():
    # QFT(30) complex128 at 12 tile bits: 4 passes holding its 435 CP and 30 H; the 15
    # trailing SWAPs are one bit-swap launch outside the passes
    rows, other = planner.tile_pass_features(rs.qft(30), rs.C128, 12)
    assert rows.shape == (4, planner.N_TILE_FEATURES) and other == 1
    tot = np.rint(rows[:, 3:] * 16).sum(axis=0).astype(int)
    assert tot[0:4].sum() == 435            # diagonal gates, all lowering variants
    assert tot[4:8].sum() + tot[9:11].sum() + tot[11] == 30  # H: run members, block members, LDS
    assert tot[12] == 0                     # no permutation op left in the passes
    # with a gate after them the swaps are no longer a trailing permutation: 6 passes hold
    # the 15 SWAPs as permutation ops
    ops = rs.qft(30).ops()
    c = rs.Circuit.from_ops(30, ops + [(rs.gate_matrix("h"), [0])])
    rows, other = planner.tile_pass_features(c, rs.C128, 12)
    assert rows.shape == (6, planner.N_TILE_FEATURES) and other == 0
    tot = np.rint(rows[:, 3:] * 16).sum(axis=0).astype(int)
    assert tot[12] == 15
    # a gate a tile cannot hold (4-qubit dense) stays outside the passes
    c = rs.Circuit(14)
    q = np.linalg.qr(np.random.default_rng(1).normal(size=(16, 16)))[0]
    c.h(0).push(q, [0, 5, 9, 13]).h(1)
    rows, other = planner.tile_pass_features(c, rs.C64, 11)
    assert other == 1 and len(rows) == 2
