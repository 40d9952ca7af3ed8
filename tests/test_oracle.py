"""Pins for the CPU oracle (oracle/qsv_oracle.c): each check is independent of the
oracle's own formulas -- published generator outputs, the reference's unit-test examples,
scipy/numpy golden matrices and states, brute-force invariants.  CPU only."""
import numpy as np
import pytest
from scipy.linalg import expm

import oracle
from tests.helpers import max_abs_diff
from tests.golden.make_golden import SplitMix64 as PySM, gen_random as py_gen_random
from tests.golden.make_golden import gen_layered, gen_hea, gen_mix, gen_qft


# ------------------------------------------------------------------ SplitMix64
def test_splitmix64_published_outputs():
    # Vigna's splitmix64.c from seed 0: first outputs.
    r = oracle.SplitMix64(0)
    assert [r.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                                 0x06C45D188009454F]


def test_splitmix64_matches_independent_python():
    for seed in (1, 42, 0xC0FFEE, 2**64 - 1):
        a, b = oracle.SplitMix64(seed), PySM(seed)
        for _ in range(200):
            assert a.next_u64() == b.next()
        assert oracle.SplitMix64(seed).unit() == PySM(seed).unit()
        assert oracle.SplitMix64(seed).below(13) == PySM(seed).below(13)


def test_unit_range():  # circuits.rs:126-130
    r = oracle.SplitMix64(1)
    for _ in range(1000):
        assert 0.0 <= r.unit() < 1.0


# ------------------------------------------------------------------ index math
def test_insert_zero_bit_examples():  # layout.rs:136-149
    for i in range(32):
        assert oracle.insert_zero_bit(i, 0) == i << 1
    assert oracle.insert_zero_bit(0b1011, 2) == 0b10011
    for i in range(64):
        for b in range(6):
            assert (oracle.insert_zero_bit(i, b) >> b) & 1 == 0


def test_pairs_differ_in_one_bit():  # layout.rs:151-161
    for bit in range(5):
        for i in range(16):
            a0 = oracle.insert_zero_bit(i, bit)
            a1 = a0 ^ (1 << bit)
            assert a1 == a0 + (1 << bit)


def test_gather_scatter_roundtrip():  # layout.rs:163-174
    qs = [0, 2, 5]
    for sub in range(8):
        assert oracle.gather_bits(oracle.scatter_bits(sub, qs), qs) == sub
    assert oracle.gather_bits(0b100101, qs) == 0b111


def test_insert_zero_bits_tiles_the_space():  # layout.rs:176-193
    for n, targets in [(4, [1, 3]), (6, [0, 2, 5]), (7, [1, 2, 3, 6])]:
        m = len(targets)
        seen = np.zeros(1 << n, dtype=int)
        for o in range(1 << (n - m)):
            base = oracle.insert_zero_bits(o, targets)
            for t in targets:
                assert (base >> t) & 1 == 0
            for s in range(1 << m):
                seen[base | oracle.scatter_bits(s, targets)] += 1
        assert (seen == 1).all()


# ------------------------------------------------------------------ gates
def test_gate_matrices_match_golden(golden):
    z, meta = golden
    for g in meta["gates"]:
        got = oracle.gate_matrix(g["name"], g["params"])
        assert np.abs(got - z[g["key"]]).max() < 1e-14, g["name"]


@pytest.mark.parametrize("t", [0.0, 0.3, 1.7, np.pi, 5.9])
def test_rotations_are_exponentials(t):
    X = np.array([[0, 1], [1, 0]])
    Y = np.array([[0, -1j], [1j, 0]])
    Z = np.diag([1, -1])
    assert np.abs(oracle.gate_matrix("rx", [t]) - expm(-0.5j * t * X)).max() < 1e-14
    assert np.abs(oracle.gate_matrix("ry", [t]) - expm(-0.5j * t * Y)).max() < 1e-14
    assert np.abs(oracle.gate_matrix("rz", [t]) - expm(-0.5j * t * Z)).max() < 1e-14
    assert np.abs(oracle.gate_matrix("rzz", [t]) - expm(-0.5j * t * np.kron(Z, Z))).max() < 1e-14


def test_u3_identities():
    for t in (0.2, 1.1, 2.9):
        assert np.abs(oracle.gate_matrix("u3", [t, -np.pi / 2, np.pi / 2])
                      - oracle.gate_matrix("rx", [t])).max() < 1e-15
        assert np.abs(oracle.gate_matrix("u3", [t, 0, 0]) - oracle.gate_matrix("ry", [t])).max() < 1e-15
        assert np.abs(oracle.gate_matrix("u3", [0, 0, t]) - oracle.gate_matrix("p", [t])).max() < 1e-15
    assert np.abs(oracle.gate_matrix("u3", [np.pi / 2, 0, np.pi])
                  - oracle.gate_matrix("h")).max() < 1e-15


def test_cx_permutes_basis():  # gate.rs:190-202
    g = oracle.gate_matrix("cx")
    for col, row in enumerate([0, 3, 2, 1]):
        e = np.zeros(4)
        e[row] = 1
        assert np.array_equal(g[:, col], e)


def test_is_diagonal_classifies():  # gate.rs:220-236
    for k, p in [("z", []), ("s", []), ("t", []), ("rz", [0.3]), ("p", [1.1]), ("cz", []),
                 ("rzz", [0.5]), ("cp", [0.9]), ("ccz", [])]:
        assert oracle.is_diagonal(oracle.gate_matrix(k, p)), k
    for k, p in [("h", []), ("x", []), ("rx", [0.4]), ("cx", []), ("swap", []), ("ccx", []),
                 ("u3", [0.1, 0.2, 0.3])]:
        assert not oracle.is_diagonal(oracle.gate_matrix(k, p)), k


def test_all_gates_unitary():
    for k in oracle.KINDS:
        g = oracle.gate_matrix(k, [0.37, 1.9, -0.8])
        assert np.abs(g @ g.conj().T - np.eye(len(g))).max() < 1e-14, k


# ------------------------------------------------------------------ generators
def _same_records(a, b):
    assert len(a) == len(b)
    for (ka, qa, pa), (kb, qb, pb) in zip(a, b):
        assert ka == kb and tuple(qa) == tuple(qb)
        assert np.allclose(list(pa)[: len(pb)], list(pb), rtol=0, atol=0) or \
            all(x == y for x, y in zip(pa, pb))


def test_generators_match_independent_python():
    for n, g, s in [(3, 20, 0), (5, 40, 7), (8, 59, 123), (14, 60, 4242)]:
        _same_records([(k, q, p[:len(pp)]) for (k, q, p), (_, _, pp)
                       in zip(oracle.gen_random(n, g, s), py_gen_random(n, g, s))],
                      py_gen_random(n, g, s))
    _same_records([(k, q, p[:3 if k == "u3" else (1 if k in ("rz",) else 0)])
                   for k, q, p in oracle.gen_layered(6, 3, 0xC0FFEE)], gen_layered(6, 3, 0xC0FFEE))
    _same_records([(k, q, p[:1 if k in ("ry", "rz") else 0]) for k, q, p in oracle.gen_hea(5, 2, 3)],
                  gen_hea(5, 2, 3))
    _same_records([(k, q, p[:3 if k == "u3" else (1 if k == "rzz" else 0)])
                   for k, q, p in oracle.gen_mix(8, 80, 6)], gen_mix(8, 80, 6))
    _same_records([(k, q, p[:1 if k == "cp" else 0]) for k, q, p in oracle.gen_qft(6)], gen_qft(6))


def test_generator_shapes():  # circuits.rs:133-138
    assert len(oracle.gen_ghz_star(5)) == 5
    assert len(oracle.gen_random(4, 30, 7)) == 30
    recs = oracle.gen_layered(28, 20, 1)
    kinds = [k for k, _, _ in recs]
    assert kinds.count("u3") == 560 and kinds.count("rz") == 560 and kinds.count("cz") == 270
    assert len(oracle.gen_qft(30)) == 30 + 435 + 15
    recs = oracle.gen_hea(12, 8, 3)
    assert len(recs) == 8 * (3 * 12 - 1)


# ------------------------------------------------------------------ backends vs golden
@pytest.mark.parametrize("backend_id", [oracle.REF, oracle.CPU_SERIAL, oracle.CPU_PARALLEL])
def test_backends_match_golden_states(golden, backend_id):
    z, meta = golden
    for case in meta["cases"]:
        recs = [(k, tuple(q), tuple(p)) for k, q, p in case["ops"]]
        got = oracle.execute(backend_id, case["n"], recs)
        assert max_abs_diff(got, z["state_" + case["name"]]) < 1e-12, case["name"]


def test_ref_and_cpu_agree_bitwise_on_random_circuits():
    # crates/qsv-core/tests/equivalence.rs:28-70 restated (TOL there: 1e-9); the restated
    # Ref and Cpu backends perform the same floating-point operations in the same order.
    for seed in range(200):
        n = 3 + seed % 6
        depth = 20 + seed % 40
        recs = oracle.gen_random(n, depth, seed)
        a = oracle.execute(oracle.REF, n, recs)
        b = oracle.execute(oracle.CPU_SERIAL, n, recs)
        c = oracle.execute(oracle.CPU_PARALLEL, n, recs)
        assert max_abs_diff(a, b) == 0.0 and max_abs_diff(a, c) == 0.0, seed
        assert abs(oracle.norm_sqr(b) - 1) < 1e-12


def test_threaded_paths_at_14_qubits():  # equivalence.rs:83-106
    for seed in range(3):
        recs = oracle.gen_random(14, 60, 4242 + seed)
        a = oracle.execute(oracle.REF, 14, recs)
        c = oracle.execute(oracle.CPU_PARALLEL, 14, recs)
        assert max_abs_diff(a, c) == 0.0


def test_known_answers():  # tests/oracle.rs:12-91
    s = 1 / np.sqrt(2)
    bell = oracle.execute(oracle.REF, 2, [("h", (0,), ()), ("cx", (0, 1), ())])
    assert max_abs_diff(bell, [s, 0, 0, s]) < 1e-12
    assert max_abs_diff(oracle.execute(oracle.REF, 2, [("cx", (0, 1), ())]), [1, 0, 0, 0]) == 0
    assert max_abs_diff(oracle.execute(oracle.REF, 1, [("rx", (0,), (np.pi,))]), [0, -1j]) < 1e-12
    for n in (1, 2, 3, 5, 8):  # equivalence.rs:108-120
        st = oracle.execute(oracle.CPU_SERIAL, n, oracle.gen_qft(n))
        assert np.abs(np.abs(st) ** 2 - 2.0 ** -n).max() < 1e-12


def test_single_gate_against_full_operator():
    # Brute force: one gate on arbitrary qubit orders == the explicit 2^n operator.
    from tests.golden.make_golden import full_unitary
    rng = np.random.default_rng(0)
    n = 5
    psi = rng.normal(size=32) + 1j * rng.normal(size=32)
    for k, qs in [("u3", (3,)), ("cx", (4, 1)), ("swap", (0, 3)), ("ccx", (2, 4, 0)),
                  ("rzz", (1, 3)), ("ccz", (0, 1, 4))]:
        g = oracle.gate_matrix(k, [0.3, 0.2, 0.1])
        want = full_unitary(n, g, list(qs)) @ psi
        for b in (oracle.REF, oracle.CPU_SERIAL):
            assert max_abs_diff(oracle.apply(b, n, psi, g, qs), want) < 1e-13, (k, b)


def test_errors_are_status_codes():
    with pytest.raises(ValueError):
        oracle.apply(oracle.REF, 3, np.ones(8, complex), oracle.gate_matrix("cx"), (0, 0))
    with pytest.raises(ValueError):
        oracle.apply(oracle.CPU_SERIAL, 3, np.ones(8, complex), oracle.gate_matrix("h"), (3,))


# ------------------------------------------------------------------ fusion
def test_fusion_preserves_results():  # equivalence.rs:148-178
    for mq in (2, 3, 4, 5):
        for seed in range(40):
            n = 4 + seed % 5
            recs = oracle.gen_random(n, 50, 70000 + seed)
            flat = oracle.flatten(recs)
            fused = oracle.fuse(flat, mq)
            assert len(fused) <= len(recs)
            want = oracle.execute(oracle.REF, n, recs)
            got = oracle.execute(oracle.CPU_PARALLEL, n, fused)
            assert max_abs_diff(want, got) < 1e-12, (mq, seed)
    for n in (4, 6, 8):
        flat = oracle.flatten(oracle.gen_qft(n))
        got = oracle.execute(oracle.CPU_PARALLEL, n, oracle.fuse(flat, 4))
        assert max_abs_diff(oracle.execute(oracle.REF, n, oracle.gen_qft(n)), got) < 1e-12


def test_fusion_counts():  # fusion.rs:162-186
    flat = oracle.flatten(oracle.gen_qft(6))
    assert len(oracle.fuse(flat, 4)) < 6 + 15 + 3
    flat = oracle.flatten([("h", (0,), ()), ("h", (2,), ())])
    assert len(oracle.fuse(flat, 1)) == 2


def test_fused_groups_are_unitary_and_diagonal_products_stay_diagonal():
    flat = oracle.flatten(oracle.gen_qft(8))
    for b, e, q in oracle.fuse_groups(flat, 5):
        if e - b < 2:
            continue
        U = oracle.build_fused(flat, b, e, q)
        assert np.abs(U @ U.conj().T - np.eye(len(U))).max() < 1e-13
        all_diag = all(oracle.is_diagonal(flat.ops[i][0]) for i in range(b, e))
        if all_diag:
            assert oracle.is_diagonal(U)  # exact zeros survive (fusion.rs:126)


def test_fusion_group_qubits_sorted_and_capped():
    recs = oracle.gen_random(8, 200, 99)
    flat = oracle.flatten(recs)
    for cap in (1, 2, 3, 4, 5, 6):
        for b, e, q in oracle.fuse_groups(flat, cap):
            assert e > b
            if e - b > 1:
                assert list(q) == sorted(q) and len(q) <= cap
            else:
                assert tuple(q) == tuple(flat.ops[b][1])


# ------------------------------------------------------------------ reductions
def test_expectation_z_on_ghz():  # tests/sampling.rs:48-58
    st = oracle.execute(oracle.CPU_PARALLEL, 5, oracle.gen_ghz_star(5))
    p = oracle.probabilities(st)
    assert abs(oracle.expectation_z(p, 0b00001)) < 1e-12
    assert abs(oracle.expectation_z(p, 0b00011) - 1) < 1e-12
    assert abs(oracle.expectation_z(p, 0b01111) - 1) < 1e-12
    assert abs(oracle.expectation_z(p, 0b11111)) < 1e-12


def test_norm_and_probabilities_of_basis():  # state/mod.rs:137-158
    re, im = oracle.basis(3, 5)
    st = re + 1j * im
    assert oracle.norm_sqr(st) == 1.0
    assert (oracle.probabilities(st) > 0).sum() == 1


def test_f32_path_tracks_f64():
    recs = oracle.gen_random(8, 60, 3)
    a = oracle.execute(oracle.CPU_SERIAL, 8, recs)
    b = oracle.execute(oracle.CPU_SERIAL, 8, recs, dtype=np.float32)
    assert max_abs_diff(a, b) < 1e-5


def test_planted_mistake_is_caught(golden):
    # A plausible slip -- conjugating one gate matrix -- must fail the golden check.
    z, meta = golden
    case = next(c for c in meta["cases"] if c["name"] == "mix8")
    ops = oracle.records_to_ops([(k, tuple(q), tuple(p)) for k, q, p in case["ops"]])
    i = next(i for i, (g, _) in enumerate(ops) if not np.isrealobj(g) and np.abs(g.imag).max() > 0)
    bad = list(ops)
    bad[i] = (ops[i][0].conj(), ops[i][1])
    got = oracle.execute(oracle.REF, case["n"], bad)
    assert max_abs_diff(got, z["state_mix8"]) > 1e-3


def test_qft_is_the_discrete_fourier_transform():
    # the oracle's QFT (circuits.rs:61-74) on a random state is numpy's FFT between bit
    # reversals; a wrong phase sign, a dropped swap or a reversed ladder fails this
    from tests.helpers import qft_closed_form
    rng = np.random.default_rng(8)
    for n in (1, 2, 5, 9):
        psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        psi /= np.linalg.norm(psi)
        got = oracle.execute(oracle.CPU_SERIAL, n, oracle.gen_qft(n), start=psi)
        assert np.abs(got - qft_closed_form(psi)).max() < 1e-14, n
        # the plain FFT without the reversals is another map (n >= 2)
        if n >= 2:
            assert np.abs(got - np.fft.ifft(psi) * np.sqrt(psi.size)).max() > 1e-3
