"""GPU parity: the HIP engine through the C ABI against the CPU oracle and the golden
fixtures, on the reference's own differential recipes
(crates/qsv-core/tests/equivalence.rs, crates/qsv-cuda/tests/equivalence.rs).

Tolerances: the plain complex128 path must match to <= 1e-12 max amplitude error; it is
in fact bit-identical to the CpuBackend restatement (same operations, same order, no
contraction), which is asserted as well.  complex64 is compared with the oracle's f32
CpuBackend (also bit-identical) and with the f64 result to 1e-5."""
import numpy as np
import pytest

import oracle
import rustatevec_amd as rs
from tests.helpers import max_abs_diff, records_of

pytestmark = pytest.mark.gpu

TOL = 1e-12


def gpu_state(backend, circ, dtype=rs.C128):
    st = backend.execute(circ, dtype)
    out = backend.download(st)
    st.free()
    return out


def cpu_state(n, recs_or_ops, dtype=np.float64, backend=oracle.CPU_SERIAL):
    return oracle.execute(backend, n, recs_or_ops, dtype=dtype)


def test_device_is_present(backend):
    info = backend.device_info()
    print(info)
    assert info["compute_units"] > 0


def test_golden_states(backend, golden):
    z, meta = golden
    for case in meta["cases"]:
        recs = records_of(case)
        c = rs.Circuit.from_ops(case["n"], oracle.records_to_ops(recs))
        got = gpu_state(backend, c)
        assert max_abs_diff(got, z["state_" + case["name"]]) < TOL, case["name"]


@pytest.mark.parametrize("chunk", range(4))
def test_random_circuits_bitwise(backend, chunk):
    # equivalence.rs:28-70: 200 random circuits, 3..8 qubits, depth 20..59.
    for seed in range(50 * chunk, 50 * (chunk + 1)):
        n = 3 + seed % 6
        depth = 20 + seed % 40
        recs = oracle.gen_random(n, depth, seed)
        want_ref = cpu_state(n, recs, backend=oracle.REF)
        want_cpu = cpu_state(n, recs)
        got = gpu_state(backend, rs.random_circuit(n, depth, seed))
        assert max_abs_diff(got, want_ref) < TOL, seed
        assert max_abs_diff(got, want_cpu) == 0.0, seed


def test_larger_registers(backend):
    # equivalence.rs:72-106 (10q depth 80, 14q depth 60), plus sizes where every kernel
    # takes its vectorised, multi-block path.
    for n, depth, seed in [(10, 80, 9000), (14, 60, 4242), (16, 200, 7), (18, 120, 8)]:
        recs = oracle.gen_random(n, depth, seed)
        want = cpu_state(n, recs, backend=oracle.CPU_PARALLEL)
        got = gpu_state(backend, rs.random_circuit(n, depth, seed))
        assert max_abs_diff(got, want) == 0.0, (n, seed)


def test_qft_and_ghz(backend):  # qsv-cuda/tests/equivalence.rs:45-62
    for n in (1, 2, 3, 5, 8, 12):
        got = gpu_state(backend, rs.qft(n))
        assert max_abs_diff(got, cpu_state(n, oracle.gen_qft(n), backend=oracle.REF)) < TOL
        assert np.abs(np.abs(got) ** 2 - 2.0 ** -n).max() < 1e-12
    for n in (3, 6, 10):
        got = gpu_state(backend, rs.ghz(n))
        assert max_abs_diff(got, cpu_state(n, oracle.gen_ghz_star(n), backend=oracle.REF)) < TOL


def test_ghz20_chain_probabilities(backend):
    # GHZ(20) workload: probabilities nonzero only at 0 and 2^20-1 (0.5 each).
    st = backend.execute(rs.ghz_chain(20))
    p = backend.probabilities(st)
    assert p.dtype == np.float64 and p.shape == (1 << 20,)
    assert abs(p[0] - 0.5) < 1e-15 and abs(p[-1] - 0.5) < 1e-15
    assert np.count_nonzero(p) == 2
    st.free()


def test_three_qubit_gates(backend):  # equivalence.rs:195-225
    ccx = rs.gate_matrix("ccx")
    for seed, qs in [(1, [0, 1, 2]), (2, [2, 0, 4]), (3, [4, 3, 1]), (4, [5, 2, 0])]:
        recs = oracle.gen_random(6, 40, seed)
        ops = oracle.records_to_ops(recs) + [(ccx, qs), (oracle.gate_matrix("h"), (0,)),
                                              (oracle.gate_matrix("rx", [0.9]), (3,)),
                                              (ccx, (1, 4, 5))]
        c = rs.Circuit.from_ops(6, ops)
        got = gpu_state(backend, c)
        assert max_abs_diff(got, cpu_state(6, ops, backend=oracle.REF)) < TOL
        assert max_abs_diff(got, cpu_state(6, ops)) == 0.0


def test_mix_and_layered_workloads(backend):
    for n, g, s in [(9, 200, 1), (12, 300, 2), (15, 200, 3)]:
        got = gpu_state(backend, rs.mix(n, g, s))
        assert max_abs_diff(got, cpu_state(n, oracle.gen_mix(n, g, s))) == 0.0
    for n, d, s in [(8, 4, 0xC0FFEE), (13, 3, 5)]:
        got = gpu_state(backend, rs.layered(n, d, s))
        assert max_abs_diff(got, cpu_state(n, oracle.gen_layered(n, d, s))) == 0.0


def test_apply_every_target_and_size(backend):
    # Every 1q target, both vector paths (q below/above the 16-byte vector width) and the
    # tiny-state scalar path; dense 2q/3q on low and high qubits.
    rng = np.random.default_rng(1)
    for n in (1, 2, 3, 5, 9):
        psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        st = backend.alloc(n)
        backend.upload(st, psi)
        want = psi.copy()
        for q in range(n):
            g = oracle.gate_matrix("u3", [0.3 + q, 1.1, -0.4])
            backend.apply(st, g, [q])
            want = oracle.apply(oracle.CPU_SERIAL, n, want, g, [q])
        for qs in ([0, 1], [1, 0], [n - 1, 0]) if n >= 2 else ():
            g = oracle.gate_matrix("u3", [0.2, 0.3, 0.4])
            g2 = np.kron(g, oracle.gate_matrix("rx", [0.7])) @ oracle.gate_matrix("cx")
            backend.apply(st, g2, qs)
            want = oracle.apply(oracle.CPU_SERIAL, n, want, g2, qs)
        assert max_abs_diff(backend.download(st), want) == 0.0, n
        st.free()


@pytest.mark.parametrize("dtype", [rs.C128, rs.C64])
def test_1q_lane_and_pair_kernels_agree_on_every_target(backend, dtype):
    """The default 1q kernel pairs lanes 0..31 with 32..63 of a wave (one chunk per lane);
    "unroll" 2 selects the pair-per-thread kernel.  Both must give the oracle's result,
    bitwise, for every target, on states from one wave of chunks upward."""
    npt = np.float64 if dtype == rs.C128 else np.float32
    rng = np.random.default_rng(3)
    old = backend.tune()["unroll"]
    try:
        for n in (6, 7, 8, 13, 18):
            psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)) * 0.01
            ops = [(oracle.gate_matrix("u3", [0.3 + q, 1.1, -0.4 * q]), [q]) for q in range(n)]
            want = oracle.execute(oracle.CPU_SERIAL, n, ops, dtype=npt, start=psi)
            for unroll in (1, 2):
                backend.tune(unroll=unroll)
                st = backend.alloc(n, dtype)
                backend.upload(st, psi)
                for g, qs in ops:
                    backend.apply(st, g, qs)
                got = backend.download(st)
                st.free()
                assert np.array_equal(got, want), (n, unroll)
    finally:
        backend.tune(unroll=old)


def random_unitary(d, rng):
    q, r = np.linalg.qr(rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d)))
    return q * (np.diag(r) / np.abs(np.diag(r)))


@pytest.mark.parametrize("dtype", [rs.C128, rs.C64])
def test_mq_lane_and_block_kernels_agree_on_every_placement(backend, dtype):
    """Dense 2- and 3-qubit gates run with one chunk per lane by default (complex64 2-qubit
    gates: members gathered across lanes 8, 16, 32 apart; complex128 2-qubit gates on a
    qubit below 6 and one at 6 or above: runs of 64 amplitudes per wave, members gathered by
    lane permutes) and one block per thread with "mq_lane" 0; both must give the oracle's
    result bitwise for every ordered placement of the gate's qubits (low, high, scattered,
    reversed), on states from below one wave of blocks upward."""
    npt = np.float64 if dtype == rs.C128 else np.float32
    rng = np.random.default_rng(12)
    old = backend.tune()["mq_lane"]
    try:
        for n in (3, 5, 8, 9, 14):
            psi = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)) * 0.01
            ops = []
            for a in range(n):
                b = (a * 5 + 3) % n
                if b != a:
                    ops.append((random_unitary(4, rng), [a, b]))
                    ops.append((oracle.gate_matrix("swap", []), [b, a]))
                    ops.append((oracle.gate_matrix("cx", []), [a, b]))
                c = (a * 3 + 1) % n
                if len({a, b, c}) == 3:
                    ops.append((random_unitary(8, rng), [c, a, b]))
                    ops.append((oracle.gate_matrix("ccx", []), [a, b, c]))
            want = oracle.execute(oracle.CPU_SERIAL, n, ops, dtype=npt, start=psi)
            for lane in (1, 0):
                backend.tune(mq_lane=lane)
                st = backend.alloc(n, dtype)
                backend.upload(st, psi)
                for g, qs in ops:
                    backend.apply(st, g, qs)
                got = backend.download(st)
                st.free()
                assert np.array_equal(got, want), (n, lane)
    finally:
        backend.tune(mq_lane=old)


def test_upload_download_roundtrip_and_qft_from_random_start(backend, golden):
    z, _ = golden
    v = z["start_random7"]
    st = backend.alloc(7)
    backend.upload(st, v)
    assert max_abs_diff(backend.download(st), v) == 0.0
    p = rs.Program(st, rs.qft(7))
    p.launch()
    assert max_abs_diff(backend.download(st), z["state_qft7_from_random"]) < TOL
    st.free()


@pytest.mark.parametrize("dtype", [rs.C128, rs.C64])
def test_upload_download_roundtrip_over_several_staging_chunks(backend, dtype):
    # 2^26 amplitudes cross the ABI in four 2^24-amplitude device-staged chunks; every
    # amplitude lands back bit-identical, and the split parts match the joined download
    n = 26
    npt = np.float64 if dtype == rs.C128 else np.float32
    rng = np.random.default_rng(26)
    re = rng.normal(size=1 << n).astype(npt)
    im = rng.normal(size=1 << n).astype(npt)
    st = backend.alloc(n, dtype)
    backend.upload(st, re=re, im=im)
    gre, gim = backend.download_parts(st)
    assert np.array_equal(gre, re) and np.array_equal(gim, im)
    got = backend.download(st)
    assert np.array_equal(got.real, re) and np.array_equal(got.imag, im)
    st.free()


def test_fusion_composes_like_build_fused(backend):
    # fusion.rs:102-139: group matrices composed in f64 on the device, bit-identical.
    for cap in (2, 3, 4, 5, 6):
        for seed in range(6):
            n = 5 + seed % 4
            recs = oracle.gen_random(n, 60, 70000 + seed)
            flat = oracle.flatten(recs)
            want = oracle.fuse(flat, cap)
            got = rs.fuse(backend, rs.random_circuit(n, 60, 70000 + seed), cap).ops()
            assert len(got) == len(want)
            for (ga, qa), (gb, qb) in zip(got, want):
                assert tuple(qa) == tuple(qb)
                assert np.array_equal(ga, gb), (cap, seed)


def test_fusion_preserves_results(backend):  # equivalence.rs:148-178
    for cap in (2, 3, 4, 5, 6):
        for seed in range(10):
            n = 4 + seed % 5
            recs = oracle.gen_random(n, 50, 70000 + seed)
            want = cpu_state(n, recs, backend=oracle.REF)
            fused = rs.fuse(backend, rs.random_circuit(n, 50, 70000 + seed), cap)
            assert len(fused) <= 50
            got = gpu_state(backend, fused)
            assert max_abs_diff(got, want) < TOL, (cap, seed)
            # the fused circuit run on the oracle's CpuBackend: identical bits
            assert max_abs_diff(got, cpu_state(n, fused.ops())) == 0.0
    for n in (4, 6, 8, 12):
        fused = rs.fuse(backend, rs.qft(n), 4)
        got = gpu_state(backend, fused)
        assert max_abs_diff(got, cpu_state(n, oracle.gen_qft(n), backend=oracle.REF)) < TOL


def test_wide_fused_groups_on_large_states(backend):
    # k = 4..6 groups over the LDS-tiled kernel at sizes with many tiles.
    for n, cap, seed in [(14, 4, 1), (15, 5, 2), (16, 6, 3)]:
        fused = rs.fuse(backend, rs.random_circuit(n, 150, seed), cap)
        got = gpu_state(backend, fused)
        assert max_abs_diff(got, cpu_state(n, fused.ops(), backend=oracle.CPU_PARALLEL)) == 0.0


def test_complex64(backend):
    for n, g, s in [(6, 50, 1), (11, 120, 2), (15, 200, 3)]:
        got = gpu_state(backend, rs.random_circuit(n, g, s), rs.C64)
        want32 = cpu_state(n, oracle.gen_random(n, g, s), dtype=np.float32)
        want64 = cpu_state(n, oracle.gen_random(n, g, s))
        assert max_abs_diff(got, want32) == 0.0
        assert max_abs_diff(got, want64) < 1e-5
    fused = rs.fuse(backend, rs.random_circuit(12, 100, 9), 5)
    got = gpu_state(backend, fused, rs.C64)
    assert max_abs_diff(got, cpu_state(12, fused.ops(), dtype=np.float32)) == 0.0


def test_reductions(backend):
    st = backend.execute(rs.ghz(5))
    # sampling.rs:48-58
    assert abs(backend.expectation_z(st, 0b00001)) < 1e-12
    assert abs(backend.expectation_z(st, 0b00011) - 1) < 1e-12
    assert abs(backend.expectation_z(st, 0b01111) - 1) < 1e-12
    assert abs(backend.expectation_z(st, 0b11111)) < 1e-12
    st.free()
    # probabilities sum to one (qsv-cuda/tests/equivalence.rs:64-71)
    st = backend.execute(rs.qft(10))
    assert abs(backend.probabilities(st).sum() - 1) < 1e-9
    st.free()
    for n, dtype in [(16, rs.C128), (18, rs.C64), (3, rs.C128), (1, rs.C64)]:
        circ = rs.random_circuit(max(n, 2), 80, n) if n >= 2 else rs.Circuit(1).h(0)
        st = backend.execute(circ, dtype)
        psi = backend.download(st)
        p = oracle.probabilities(psi.astype(np.complex128))
        for mask in (0, 1, 0b101, (1 << n) - 1):
            want = oracle.expectation_z(p, mask)
            got = backend.expectation_z(st, mask) if mask else backend.norm_sqr(st)
            assert abs(got - want) < 1e-12, (n, mask)
        probs = backend.probabilities(st)
        want_p = psi.real ** 2 + psi.imag ** 2
        assert np.array_equal(probs, want_p.astype(probs.dtype)) or \
            np.abs(probs - want_p).max() < 1e-7
        for i in (0, (1 << n) - 1, (1 << n) // 3):
            assert backend.amplitude(st, i) == complex(psi[i])
        st.free()


def test_errors_are_status_codes(backend):
    st = backend.alloc(3)
    backend.init_basis(st, 0)
    with pytest.raises(rs.QsvhError) as e:
        backend.apply(st, rs.gate_matrix("cx"), [0, 3])
    assert e.value.status == "INVALID_QUBIT"
    with pytest.raises(rs.QsvhError) as e:
        backend.apply(st, rs.gate_matrix("cx"), [2, 2])
    assert e.value.status == "REPEATED_QUBIT"
    with pytest.raises(rs.QsvhError) as e:
        backend.apply(st, np.eye(2), [0, 1])
    assert e.value.status == "ARITY"
    with pytest.raises(rs.QsvhError) as e:
        backend.amplitude(st, 8)
    assert e.value.status == "INVALID_ARG"
    with pytest.raises(rs.QsvhError) as e:
        backend.init_basis(st, 9)
    assert e.value.status == "INVALID_ARG"
    with pytest.raises(rs.QsvhError) as e:
        rs.Program(st, rs.qft(4))
    assert e.value.status == "INVALID_ARG"
    with pytest.raises(rs.QsvhError) as e:
        backend.alloc(3, dtype=7)
    assert e.value.status == "DTYPE"
    # the state is untouched by rejected calls
    assert backend.amplitude(st, 0) == 1
    st.free()


def test_graph_replays_match_eager(backend):
    # Replay a captured program several times with eager work in between; every replay
    # must equal the eager application of the same gates.
    n = 12
    circ = rs.random_circuit(n, 100, 77)
    fused = rs.fuse(backend, circ, 5)
    a = backend.alloc(n)
    b = backend.alloc(n)
    backend.init_basis(a, 0)
    backend.init_basis(b, 0)
    pg = rs.Program(a, fused, use_graph=True)
    pe = rs.Program(b, fused)
    for rep in range(4):
        pg.launch()
        pe.launch()
        backend.synchronize(a)
        assert backend.norm_sqr(b) > 0.99  # eager work on the same stream in between
        assert max_abs_diff(backend.download(a), backend.download(b)) == 0.0, rep
    del pg, pe
    a.free()
    b.free()


def test_32_qubit_state_launch_limits(backend):
    """A complex128 state of 32 qubits has 2^32 one-chunk work items, one more than a
    dispatch can hold: the streaming kernels must cap their grid and stride over the rest
    (regression: init_basis failed with an invalid launch configuration)."""
    n = 32
    st = backend.alloc(n, rs.C128)
    backend.init_basis(st, 0)
    for q in (0, n - 1):
        backend.apply(st, rs.gate_matrix("h"), [q])
    backend.apply(st, rs.gate_matrix("cp", 0.5), [0, n - 1])
    h = 0.5
    for i, want in ((0, h), (1, h), (1 << (n - 1), h), ((1 << (n - 1)) + 1, h * np.exp(0.5j)),
                    (2, 0.0), ((1 << n) - 1, 0.0)):
        a = backend.amplitude(st, i)
        assert abs(a - want) < 1e-15, (i, a)
    assert abs(backend.norm_sqr(st) - 1.0) < 1e-14
    out, _ = backend.sample(st, 4000, 3)
    assert set(np.unique(out).tolist()) <= {0, 1, 1 << (n - 1), (1 << (n - 1)) + 1}
    st.free()


@pytest.mark.parametrize("dtype,n", [(rs.C128, 25), (rs.C64, 26)])
def test_nontemporal_knob_is_bitwise_neutral(backend, dtype, n):
    """States beyond the 256 MiB infinity cache take non-temporal loads and stores in the
    streaming kernels by default ("nontemporal" 2); the hints change no arithmetic, so the
    1q (pair halves sharing a cache line or not), diagonal (1, 2 and 3 qubits) and cx
    results are bitwise those of the plain accesses (0) and of hints everywhere (1)."""
    gates = [("h", [0]), ("h", [2]), ("h", [3]), ("h", [4]), ("h", [n - 1]),
             ("rz", [3], 0.4), ("cp", [1, n - 2], 0.7), ("ccz", [0, 9, n - 1]),
             ("cx", [n - 1, 5]), ("h", [12])]
    outs = []
    old = backend.tune()["nontemporal"]
    try:
        for nt in (0, 1, 2):
            backend.tune(nontemporal=nt)
            st = backend.alloc(n, dtype)
            backend.init_random_normal(st, 5)
            for g in gates:
                backend.apply(st, rs.gate_matrix(g[0], *g[2:]), g[1])
            outs.append(backend.download(st))
            st.free()
    finally:
        backend.tune(nontemporal=old)
    assert np.all(np.isfinite(outs[0]))
    for got in outs[1:]:
        assert np.array_equal(got, outs[0])


def _run_program(backend, circ, dtype, **opts):
    st = backend.alloc(circ.n_qubits, dtype)
    backend.init_basis(st, 0)
    p = rs.Program(st, circ, **opts)
    p.launch()
    out = backend.download(st)
    k = p.num_kernels()
    del p
    st.free()
    return out, k


@pytest.mark.parametrize("dtype", [rs.C128, rs.C64])
def test_subspace_launches_match_the_full_passes(backend, dtype):
    """Program option `subspace`: controlled 1-qubit gates on their control subspace,
    diagonal gates with one non-unit phase on that phase's amplitudes and SWAP on its
    |01>, |10> amplitudes give the same values
    as the full passes (bit for bit against the oracle's CpuBackend), with the same number
    of launches."""
    npdt = np.float64 if dtype == rs.C128 else np.float32
    circs = [rs.qft(11), rs.ghz_chain(12), rs.random_circuit(12, 300, 3),
             rs.mix(12, 400, 4), rs.layered(10, 4, 5)]
    c = rs.Circuit(9)
    for q in range(9):
        c.h(q)
    u = rs.gate_matrix("u3", 0.3, 0.5, 0.7)
    cu0 = np.eye(8, dtype=np.complex128)  # U on qubit 2 when qubit 0 = 1 and qubit 5 = 0
    for r in range(2):
        for cc in range(2):
            cu0[(r << 2) | 1, (cc << 2) | 1] = u[r, cc]
    c.push(cu0, [0, 5, 2])
    c.push(np.diag([1, 1j, 1, 1]), [3, 7])  # one non-unit entry at sub-index 1
    c.ccx(8, 1, 0).ccz(2, 6, 7).cx(4, 0).cz(1, 3).cphase(0, 8, 0.4)
    c.swap(0, 8).swap(4, 2).swap(7, 6).swap(1, 0)  # SWAP on its |01>, |10> amplitudes
    circs.append(c)
    for circ in circs:
        n = circ.n_qubits
        got, k_sub = _run_program(backend, circ, dtype, subspace=True)
        _, k_full = _run_program(backend, circ, dtype)
        want = oracle.execute(oracle.CPU_PARALLEL, n, circ.ops(), dtype=npdt)
        assert max_abs_diff(got, want) == 0.0, n
        assert k_sub == k_full


def test_freeing_a_state_destroys_its_programs(backend):
    # a program must not outlive its state: freeing the state destroys the programs bound
    # to it first, and a destroyed program refuses to launch instead of touching freed memory
    st = backend.alloc(12)
    p = rs.Program(st, rs.qft(12), tile_qubits=12)
    backend.init_basis(st, 0)
    p.launch()
    st.free()
    with pytest.raises(rs.QsvhError):
        p.launch()
    del p


@pytest.mark.parametrize("kw", [{}, {"tile_qubits": 12}, {"subspace": True}, {"use_graph": True},
                                {"precision": rs.PREC_MFMA_F32},
                                {"precision": rs.PREC_MFMA_BF16X3},
                                {"precision": rs.PREC_MFMA_BF16X3, "mfma_tiles": 12}],
                         ids=["plain", "tiles", "subspace", "graph", "mfma_f32", "mfma_bf16x3",
                              "mfma_tiles"])
def test_empty_circuit_leaves_the_state_unchanged(backend, kw):
    # Backend::execute over no ops is the initial state (backend/mod.rs:66-73); every form
    # of an empty program runs no kernel and leaves an uploaded state bit-identical
    n = 14
    dtype = rs.C64 if "precision" in kw else rs.C128
    rng = np.random.default_rng(14)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi = psi.astype(np.complex64 if dtype == rs.C64 else np.complex128)
    st = backend.alloc(n, dtype)
    backend.upload(st, psi)
    p = rs.Program(st, rs.Circuit(n), **kw)
    assert p.num_kernels() == 0
    p.launch()
    p.launch()
    assert np.array_equal(backend.download(st), psi)
    del p
    st.free()
    st = backend.execute(rs.Circuit(3))
    want = np.zeros(8)
    want[0] = 1
    assert np.array_equal(backend.download(st), want)
    st.free()
