"""Parity at the workloads' full sizes, through properties that do not need the oracle to
finish: QFT(30) from the random-normal start and layered(28) depth 20, plain complex128 per
gate, keep the norm to 1e-12; QFT(30) in tile passes and with subspace launches equals the
plain run bit for bit (compared on the device); layered(28) in tile passes is within the
1e-12 complex128 bar of it (its passes move gates past gates on disjoint qubits, which
changes the rounding; bitwise parity on the recorded order is checked against the oracle at
smaller sizes); and the fused form the bench reports (fusion at 4 qubits, split-BF16 MFMA tile
passes, complex64) stays within the 1e-6 max amplitude error bar of the plain complex128 run,
with 1 - fidelity below 1e-8.  The batched sweep runs at its full size, 512 circuits, and
the 28-qubit random mix in the forms the learned planner chooses among.  QFT(30)'s fused
per-group forms (FP32 and split-BF16x3 matrix-core groups at k = 4, 5, 6) meet the same bars,
and QFT(20) is pinned to its closed form: numpy's FFT of the start state, in the qubit order
of the reference's generator (circuits.rs:61-74: H on qubit 0 first and the final swaps make
QFT|x> = 2^(-n/2) sum_y exp(2 pi i rev(x) rev(y) / 2^n) |y>, rev the n-bit reversal)."""
import numpy as np
import pytest

import rustatevec_amd as rs
from tests.helpers import qft_closed_form

pytestmark = pytest.mark.gpu


def run(backend, circ, n, dtype, seed, **kw):
    st = backend.alloc(n, dtype)
    if seed is None:
        backend.init_basis(st, 0)
    else:
        backend.init_random_normal(st, seed)
    p = rs.Program(st, circ, **kw)
    p.launch()
    del p
    return st


@pytest.mark.parametrize("workload", ["qft30", "layered28"])
def test_full_size_forms_agree(backend, workload):
    if workload == "qft30":
        n, circ, seed = 30, rs.qft(30), 1
    else:
        n, circ, seed = 28, rs.layered(28, 20, 0xC0FFEE), None
    print(f"{workload}: plain complex128", flush=True)
    plain = run(backend, circ, n, rs.C128, seed)
    assert abs(backend.norm_sqr(plain) - 1.0) < 1e-12
    forms = [("tiles12", {"tile_qubits": 12})]
    if workload == "qft30":
        forms.append(("subspace", {"subspace": True}))
    for name, kw in forms:
        print(f"{workload}: {name}", flush=True)
        st = run(backend, circ, n, rs.C128, seed, **kw)
        err, fid = backend.compare(plain, st)
        st.free()
        if workload == "qft30":
            assert err == 0.0, (name, err)
        else:
            assert err < 1e-12 and 1.0 - fid < 1e-12, (name, err, 1.0 - fid)
    print(f"{workload}: fused k=4, split-BF16 MFMA tile passes", flush=True)
    fused = rs.fuse(backend, circ, 4)
    st = run(backend, fused, n, rs.C64, seed, precision=rs.PREC_MFMA_BF16X3, mfma_tiles=12)
    err, fid = backend.compare(plain, st)
    st.free()
    plain.free()
    assert err < 1e-6 and 1.0 - fid < 1e-8, (err, 1.0 - fid)


def test_full_size_hea_batch(backend):
    # the batched sweep at its full size: 512 HEA circuits of 10..22 qubits, depth 8,
    # complex64, one batch; circuits of the smallest, largest and some middle sizes against
    # their own plain per-gate runs (the complex64 bar, 2e-5), every <Z_0> within [-1, 1]
    circuits = rs.hea_batch(512, 8, 0xC0FFEE)
    sizes = [c.n_qubits for c in circuits]
    assert len(circuits) == 512 and min(sizes) == 10 and max(sizes) == 22
    b = rs.Batch(backend, circuits, rs.C64)
    b.run()
    z0 = b.expectation_z0()
    assert z0.shape == (512,) and abs(z0).max() <= 1 + 1e-6
    picks = {sizes.index(10), sizes.index(22), len(sizes) - 1 - sizes[::-1].index(22), 0, 255}
    for i in sorted(picks):
        c = circuits[i]
        st = run(backend, c, c.n_qubits, rs.C64, None)
        want = backend.download(st)
        ez = backend.expectation_z(st, 1)
        st.free()
        got = b.download(i)
        assert abs(got - want).max() < 2e-5, i
        assert abs(z0[i] - ez) < 1e-5, (i, z0[i], ez)
        p = b.probabilities(i)
        assert abs(p.sum() - 1.0) < 1e-4, i


def test_full_size_random_mix(backend):
    # the learned planner's 28-qubit random mix (1120 gates) in the forms it chooses among:
    # 13-bit exact complex64 tiles and fusion at 4 qubits in split-BF16 MFMA tile passes,
    # against the plain complex128 run (the 1e-6 bar)
    n = 28
    circ = rs.mix(n, 40 * n, 0xC0FFEE)
    plain = run(backend, circ, n, rs.C128, None)
    for name, c, kw in (("tiles13", circ, {"tile_qubits": 13}),
                        ("mfma_tiles", rs.fuse(backend, circ, 4),
                         {"precision": rs.PREC_MFMA_BF16X3, "mfma_tiles": 12})):
        st = run(backend, c, n, rs.C64, None, **kw)
        err, fid = backend.compare(plain, st)
        st.free()
        assert err < 1e-6 and 1.0 - fid < 1e-8, (name, err, 1.0 - fid)
    plain.free()


@pytest.mark.parametrize("k", [4, 5, 6])
def test_full_size_qft30_per_group_mfma(backend, k):
    # QFT(30) fused at k, one launch per group, on the matrix cores in FP32 and split-BF16x3,
    # from the random-normal start: the 1e-6 / 1e-8 bars against plain complex128
    n, seed = 30, 1
    circ = rs.qft(n)
    plain = run(backend, circ, n, rs.C128, seed)
    fused = rs.fuse(backend, circ, k)
    for prec in (rs.PREC_MFMA_F32, rs.PREC_MFMA_BF16X3):
        st = run(backend, fused, n, rs.C64, seed, precision=prec)
        err, fid = backend.compare(plain, st)
        st.free()
        print(f"QFT(30) fused k={k} precision {prec}: max err {err:.3e}, 1-F {1 - fid:.3e}",
              flush=True)
        assert err < 1e-6 and 1.0 - fid < 1e-8, (k, prec, err, 1.0 - fid)
    plain.free()


def test_qft20_equals_the_discrete_fourier_transform(backend):
    # every form on the closed form: plain per gate, tile passes and subspace launches in
    # complex128 (the 1e-12 bar), fused MFMA tile passes in complex64 (1e-6)
    n, seed = 20, 5
    st = backend.alloc(n, rs.C128)
    backend.init_random_normal(st, seed)
    psi = backend.download(st)
    st.free()
    want = qft_closed_form(psi)
    circ = rs.qft(n)
    for name, dtype, c, kw, tol in (
            ("plain", rs.C128, circ, {}, 1e-12),
            ("tiles12", rs.C128, circ, {"tile_qubits": 12}, 1e-12),
            ("subspace", rs.C128, circ, {"subspace": True}, 1e-12),
            ("mfma_tiles_k4", rs.C64, rs.fuse(backend, circ, 4),
             {"precision": rs.PREC_MFMA_BF16X3, "mfma_tiles": 12}, 1e-6)):
        st = backend.alloc(n, dtype)
        backend.upload(st, psi.astype(np.complex128 if dtype == rs.C128 else np.complex64))
        p = rs.Program(st, c, **kw)
        p.launch()
        del p
        got = backend.download(st)
        st.free()
        err = float(np.abs(got - want).max())
        print(f"QFT(20) {name}: max |got - DFT| = {err:.3e}", flush=True)
        assert err < tol, (name, err)
