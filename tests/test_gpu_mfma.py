"""Fused k-qubit groups on the matrix cores (complex64 states): FP32 MFMA and split-BF16x3
MFMA against the exact complex128 engine and the oracle.

Tolerances (written here, per the precision of each form): a single dense group on a
normalised state -- FP32: 2e-6 relative to max|psi|; BF16x3 (hi/lo split, lo*lo dropped,
~2^-16 per product): 1e-4 relative.  Whole fused circuits are judged by max amplitude error
and 1 - fidelity against the complex128 plain run."""
import numpy as np
import pytest

import oracle
import rustatevec_amd as rs
from tests.helpers import max_abs_diff

pytestmark = pytest.mark.gpu

PRECISIONS = [(rs.PREC_MFMA_F32, 2e-6), (rs.PREC_MFMA_BF16X3, 1e-4)]


def random_unitary(d, seed):
    rng = np.random.default_rng(seed)
    z = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def run_c64(backend, n, circ, psi, precision):
    st = backend.alloc(n, rs.C64)
    backend.upload(st, psi)
    p = rs.Program(st, circ, precision=precision)
    p.launch()
    out = backend.download(st)
    del p
    st.free()
    return out


@pytest.mark.parametrize("prec,tol", PRECISIONS)
def test_single_group_all_placements(backend, prec, tol):
    n = 13
    rng = np.random.default_rng(5)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    placements = [(0, 1, 2, 3), (3, 2, 1, 0), (9, 10, 11, 12), (0, 5, 7, 12), (4, 6, 1, 11),
                  (0, 1, 2, 3, 4), (12, 2, 7, 3, 9), (0, 1, 2, 3, 4, 5), (7, 8, 9, 10, 11, 12),
                  (1, 3, 5, 7, 9, 11), (12, 0, 6, 2, 10, 4)]
    for i, qs in enumerate(placements):
        U = random_unitary(1 << len(qs), i)
        circ = rs.Circuit(n).push(U, qs)
        want = oracle.apply(oracle.CPU_SERIAL, n, psi, U, qs)
        got = run_c64(backend, n, circ, psi, prec)
        err = max_abs_diff(got, want) / np.abs(want).max()
        assert err < tol, (qs, err)


@pytest.mark.parametrize("prec,tol", PRECISIONS)
def test_group_on_larger_state_matches_exact_c64(backend, prec, tol):
    n = 20
    rng = np.random.default_rng(11)
    psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
    psi /= np.linalg.norm(psi)
    for i, qs in enumerate([(0, 1, 2, 3), (15, 16, 17, 18, 19), (2, 9, 13, 14, 17, 19)]):
        U = random_unitary(1 << len(qs), 100 + i)
        circ = rs.Circuit(n).push(U, qs)
        exact = run_c64(backend, n, circ, psi, rs.PREC_EXACT)
        got = run_c64(backend, n, circ, psi, prec)
        err = max_abs_diff(got, exact) / np.abs(exact).max()
        assert err < tol, (qs, err)


@pytest.mark.parametrize("cap", [4, 5, 6])
def test_fused_random_circuit_against_c128(backend, cap):
    n = 16
    circ = rs.random_circuit(n, 300, 40 + cap)
    fused = rs.fuse(backend, circ, cap)
    ref = backend.execute(circ, rs.C128)
    for prec, _ in PRECISIONS:
        st = backend.alloc(n, rs.C64)
        backend.init_basis(st, 0)
        p = rs.Program(st, fused, precision=prec)
        p.launch()
        err, fid = backend.compare(ref, st)
        print(f"cap={cap} prec={prec}: max|err|={err:.3e} 1-F={1 - fid:.3e}")
        assert err < 1e-5 and 1 - fid < 1e-6
        del p
        st.free()
    ref.free()


def test_fused_qft_from_random_state(backend):
    n = 18
    qft = rs.qft(n)
    ref = backend.alloc(n, rs.C128)
    backend.init_random_normal(ref, 7)
    psi = backend.download(ref)
    rs.Program(ref, qft).launch()
    for cap in (4, 5, 6):
        fused = rs.fuse(backend, qft, cap)
        for prec, _ in PRECISIONS:
            st = backend.alloc(n, rs.C64)
            backend.upload(st, psi)
            p = rs.Program(st, fused, precision=prec)
            p.launch()
            err, fid = backend.compare(ref, st)
            assert err < 1e-6 and 1 - fid < 1e-7, (cap, prec, err, fid)
            del p
            st.free()
    ref.free()


def test_random_normal_start_state(backend):
    # qsvh_init_random_normal: Box-Muller on SplitMix64 outputs 2i, 2i+1, normalised.
    from tests.golden.make_golden import SplitMix64
    n, seed = 10, 12345
    st = backend.alloc(n, rs.C128)
    backend.init_random_normal(st, seed)
    got = backend.download(st)
    r = SplitMix64(seed)
    u = np.array([(r.next() >> 11) / 2.0 ** 53 for _ in range(2 << n)]).reshape(-1, 2)
    want = np.sqrt(-2 * np.log(1 - u[:, 0])) * np.exp(2j * np.pi * u[:, 1])
    want /= np.linalg.norm(want)
    assert max_abs_diff(got, want) < 1e-13
    assert abs(backend.norm_sqr(st) - 1) < 1e-13
    st.free()


def test_mfma_precision_needs_c64(backend):
    st = backend.alloc(12, rs.C128)
    with pytest.raises(rs.QsvhError) as e:
        rs.Program(st, rs.qft(12), precision=rs.PREC_MFMA_F32)
    assert e.value.status == "DTYPE"
    st.free()


def test_fp32_lds_layouts_and_grid_sizes_agree_bitwise(backend):
    # The FP32 group kernel stages a tile row-major or column-major in LDS (chosen from the
    # group's qubits, or forced by the "mfma_layout" knob) and runs on any persistent grid:
    # each output is the same chain of MFMA products either way, so every layout and grid
    # gives bit-identical amplitudes, and the automatic choice matches the forced ones.
    n = 16
    rng = np.random.default_rng(17)
    psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)).astype(np.complex64)
    psi /= np.linalg.norm(psi)
    old = {k: v for k, v in backend.tune().items() if k in ("mfma_layout", "mfma_blocks")}
    try:
        for i, qs in enumerate([(0, 1, 2, 3, 4, 5), (1, 2, 3, 4), (3, 9, 14, 15, 10, 12),
                                (0, 2, 4, 13, 11)]):
            U = random_unitary(1 << len(qs), 40 + i)
            circ = rs.Circuit(n).push(U, qs)
            want = oracle.apply(oracle.CPU_SERIAL, n, psi.astype(np.complex128), U, qs)
            outs = []
            for lay, blocks in ((0, 0), (1, 0), (2, 0), (1, 3), (2, 512), (0, 2048)):
                backend.tune(mfma_layout=lay, mfma_blocks=blocks)
                outs.append(run_c64(backend, n, circ, psi, rs.PREC_MFMA_F32))
            for o in outs[1:]:
                assert np.array_equal(o, outs[0]), qs
            assert max_abs_diff(outs[0], want) / np.abs(want).max() < 2e-6, qs
    finally:
        backend.tune(**old)
