"""The C-ABI library without a device: it loads, exports every symbol include/qsvh.h
declares, and its host-side parts (gate library, circuit validation, generators, fusion
grouping) agree with the oracle.  CPU only -- no compute calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
import rustatevec_amd as rs
from tests.helpers import records_of

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "qsvh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qsvh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(rs.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 40
    for s in syms:
        assert hasattr(L, s), s


def test_python_binding_covers_the_header():
    from rustatevec_amd import planner
    L = planner._bind()
    for s in declared_symbols():
        assert getattr(L, s).restype is not C.c_int or getattr(L, s).argtypes is not None, s


def test_abi_version():
    assert rs.lib().qsvh_abi_version() == 1


def test_gate_library_matches_oracle_bitwise_and_golden(golden):
    z, meta = golden
    for g in meta["gates"]:
        mine = rs.gate_matrix(g["name"], *g["params"])
        assert np.array_equal(mine, oracle.gate_matrix(g["name"], g["params"])), g["name"]
        assert np.abs(mine - z[g["key"]]).max() < 1e-14


def test_unknown_gate_is_an_error():
    with pytest.raises(rs.QsvhError) as e:
        rs.gate_matrix("nope")
    assert e.value.status == "INVALID_ARG"


def test_push_validation_returns_status():  # circuit.rs:54-78
    c = rs.Circuit(3)
    with pytest.raises(rs.QsvhError) as e:
        c.cx(0, 3)
    assert e.value.status == "INVALID_QUBIT"
    with pytest.raises(rs.QsvhError) as e:
        c.cx(1, 1)
    assert e.value.status == "REPEATED_QUBIT"
    with pytest.raises(rs.QsvhError) as e:
        c.push(np.eye(4), [0])
    assert e.value.status == "ARITY"
    big = rs.Circuit(8)
    with pytest.raises(rs.QsvhError) as e:
        big.push(np.eye(128), list(range(7)))
    assert e.value.status == "TOO_WIDE"
    assert len(c) == 0


def _same_ops(circ, recs):
    ops = circ.ops()
    want = oracle.records_to_ops(recs)
    assert len(ops) == len(want)
    for (ga, qa), (gb, qb) in zip(ops, want):
        assert tuple(qa) == tuple(qb)
        assert np.array_equal(ga, gb)


def test_generators_match_oracle():
    _same_ops(rs.ghz(6), oracle.gen_ghz_star(6))
    _same_ops(rs.ghz_chain(20), oracle.gen_ghz_chain(20))
    _same_ops(rs.qft(9), oracle.gen_qft(9))
    for n, g, s in [(3, 20, 0), (8, 59, 123), (14, 60, 4242)]:
        _same_ops(rs.random_circuit(n, g, s), oracle.gen_random(n, g, s))
    _same_ops(rs.layered(10, 3, 0xC0FFEE), oracle.gen_layered(10, 3, 0xC0FFEE))
    _same_ops(rs.hea(7, 2, 9), oracle.gen_hea(7, 2, 9))
    _same_ops(rs.mix(9, 120, 6), oracle.gen_mix(9, 120, 6))


def test_hea_batch_workload_matches_oracle():
    circuits = rs.hea_batch(6, 8, 0xC0FFEE)
    rng = oracle.SplitMix64(0xC0FFEE)
    for c in circuits:
        n = 10 + rng.below(13)
        seed = rng.next_u64()
        assert c.n_qubits == n
        _same_ops(c, oracle.gen_hea(n, 8, seed))


def test_golden_circuits_build_through_the_abi(golden):
    _, meta = golden
    for case in meta["cases"]:
        c = rs.Circuit(case["n"])
        for k, q, p in records_of(case):
            c.gate(k, q, *p[: {"u3": 3}.get(k, 1 if k in ("rx", "ry", "rz", "p", "cp", "rzz") else 0)])
        _same_ops(c, records_of(case))


def test_fusion_grouping_matches_oracle():
    for cap in (1, 2, 3, 4, 5, 6):
        for seed in range(12):
            n = 4 + seed % 6
            recs = oracle.gen_random(n, 80, 500 + seed)
            mine = rs.fuse_groups(rs.random_circuit(n, 80, 500 + seed), cap)
            want = oracle.fuse_groups(oracle.flatten(recs), cap)
            assert mine == want, (cap, seed)
    qft = rs.qft(12)
    assert rs.fuse_groups(qft, 5) == oracle.fuse_groups(oracle.flatten(oracle.gen_qft(12)), 5)


def test_fuse_rejects_groups_wider_than_six():
    with pytest.raises(rs.QsvhError) as e:
        rs.fuse_groups(rs.qft(8), 7)
    assert e.value.status == "TOO_WIDE"


def test_context_needs_a_device():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a device is present; covered by the gpu tests")
    h = C.c_void_p()
    assert rs.lib().qsvh_ctx_create(0, C.byref(h)) == 9  # QSVH_ERR_NO_DEVICE
