"""Batched circuits: every circuit of a batch against the oracle.  Each circuit's final
state equals the oracle's CpuBackend on the batch's recorded gate order bit for bit, and
its <Z_0> matches sample::expectation_z (crates/qsv-core/src/sample.rs:177-200)."""
import numpy as np
import pytest

import oracle
import rustatevec_amd as rs
from tests.helpers import max_abs_diff

pytestmark = pytest.mark.gpu


def check_batch(batch, circuits, dtype):
    npdt = np.float64 if dtype == rs.C128 else np.float32
    z0 = batch.expectation_z0()
    for i, c in enumerate(circuits):
        n = c.n_qubits
        ops = c.ops()
        order = batch.order(i)
        assert sorted(order) == list(range(len(ops)))
        got = batch.download(i)
        want = oracle.execute(oracle.CPU_PARALLEL, n, [ops[k] for k in order], dtype=npdt)
        assert max_abs_diff(got, want) == 0.0, (i, n)
        ref = oracle.execute(oracle.CPU_PARALLEL, n, ops)
        tol = 1e-12 if dtype == rs.C128 else 2e-5
        assert max_abs_diff(got, ref) < tol
        ez = oracle.expectation_z(oracle.probabilities(ref), 1)
        assert abs(z0[i] - ez) < (1e-12 if dtype == rs.C128 else 1e-5), (i, z0[i], ez)


@pytest.mark.parametrize("dtype", [rs.C64, rs.C128])
def test_hea_batch_against_oracle(backend, dtype):
    rng = np.random.default_rng(3)
    circuits = [rs.hea(int(n), 4, int(s)) for n, s in
                zip(rng.integers(10, 17, size=14), rng.integers(0, 2**62, size=14))]
    circuits += [rs.hea(4, 2, 1), rs.hea(9, 3, 2)]  # smaller than any tile
    b = rs.Batch(backend, circuits, dtype)
    b.run()
    check_batch(b, circuits, dtype)
    # a second replay starts again from |0...0>
    z1 = b.expectation_z0()
    b.run()
    assert np.array_equal(b.expectation_z0(), z1)


def test_tile_sizes_and_mixed_gates(backend):
    circuits = [rs.random_circuit(n, 120, 50 + n) for n in (11, 12, 14, 15, 16)]
    circuits += [rs.mix(13, 150, 7), rs.qft(14), rs.layered(15, 3, 9)]
    for tile in (11, 12, 13):
        b = rs.Batch(backend, circuits, rs.C64)
        b.configure(tile)
        b.run()
        check_batch(b, circuits, rs.C64)
        assert b.info()["tile_qubits"] == tile


def test_autotune_keeps_results(backend):
    circuits = rs.hea_batch(24, 8, 0xC0FFEE)
    b = rs.Batch(backend, circuits, rs.C64)
    b.run()
    z_before = b.expectation_z0()
    tile, blocks, ms = b.autotune()
    assert 11 <= tile <= 13 and blocks > 0 and ms > 0
    b.run()
    z_after = b.expectation_z0()
    assert np.abs(z_after - z_before).max() < 1e-5
    for i in (0, 7, 23):
        n = circuits[i].n_qubits
        ref = oracle.execute(oracle.CPU_PARALLEL, n, circuits[i].ops())
        assert max_abs_diff(b.download(i), ref) < 2e-5
        p = b.probabilities(i)
        assert abs(p.sum() - 1) < 1e-4


def test_wide_gates_are_rejected(backend):
    c = rs.Circuit(12).push(np.eye(16)[::-1], [0, 1, 2, 3])  # acts on all four qubits
    with pytest.raises(rs.QsvhError) as e:
        rs.Batch(backend, [c], rs.C64)
    assert e.value.status == "TOO_WIDE"
