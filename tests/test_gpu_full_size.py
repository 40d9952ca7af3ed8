"""Parity at the workloads' full sizes, through properties that do not need the oracle to
finish: QFT(30) from the random-normal start and layered(28) depth 20, plain complex128 per
gate, keep the norm to 1e-12; QFT(30) in tile passes and with subspace launches equals the
plain run bit for bit (compared on the device); layered(28) in tile passes is within the
1e-12 complex128 bar of it (its passes move gates past gates on disjoint qubits, which
changes the rounding; bitwise parity on the recorded order is checked against the oracle at
smaller sizes); and the fused form the bench reports (fusion at 4 qubits, split-BF16 MFMA tile
passes, complex64) stays within the 1e-6 max amplitude error bar of the plain complex128 run,
with 1 - fidelity below 1e-8.  The batched sweep runs at its full size, 512 circuits, and
the 28-qubit random mix in the forms the learned planner chooses among."""
import pytest

import rustatevec_amd as rs

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
