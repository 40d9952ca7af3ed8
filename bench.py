"""Benchmark of the statevector hot path on one MI355X.

Workload (one "step"): QFT(30) -- H + controlled-phase ladder + final swaps, 480 gates --
applied in complex128 to a random-normal 30-qubit state already resident in HBM
(re, im ~ N(0,1) per amplitude, normalised; generated on the device).  The plain engine
applies one gate per HBM pass, so a step moves 480 x 2^30 x 32 B = 16.49 TB of
algorithmic traffic; ``value`` is that traffic divided by the measured step time (the
reference's own convention: GB/s = amplitude-updates/s x 32 B).

The same line reports, for the same start state:
* ``fused``: fusion::fuse with k in {4, 5, 6}, run on a complex64 copy with the wide
  groups on the matrix cores in FP32 and in split-BF16x3 -- circuit seconds, speed-up over
  plain, max amplitude error and 1 - fidelity against the plain complex128 result;
* ``tiled``: the exact complex128 engine with LDS tile passes (many gates per pass);
* the dominant plain kernel's roofline (HIP-event timing on the engine's stream), a copy
  kernel's bandwidth on the same state, GHZ(20) latency, the random layered 28-qubit
  circuit, and the CPU baseline (the oracle's CpuBackend restatement on the host cores, on
  a bounded sample of the same workload).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--quick]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E
FP32_MFMA_PEAK_TFLOPS = 157.3
BF16_MFMA_PEAK_TFLOPS = 2500.0
METRIC = "circuit sec & HBM GB/s (% roofline) at 30q; fused-MFMA vs plain at err<=1e-6"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic():
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE scaled by the copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_k_diag_small_30q_c128.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def dense_group_6():
    """A Haar-like dense 64 x 64 unitary (QR of a complex normal matrix)."""
    return np.linalg.qr(np.random.default_rng(1).normal(size=(64, 64))
                        + 1j * np.random.default_rng(2).normal(size=(64, 64)))[0]


def timed_steps(backend, state, prog, steps, warmup, reset=None):
    """Mean wall seconds per launch; `reset` (untimed) restores the start state."""
    total = 0.0
    for i in range(warmup + steps):
        if reset:
            reset()
        backend.synchronize(state)
        t0 = time.perf_counter()
        prog.launch()
        backend.synchronize(state)
        if i >= warmup:
            total += time.perf_counter() - t0
    return total / steps


def cpu_baseline(n_sample_gates: int, n: int = 30):
    """The oracle's CpuBackend::parallel restatement on the first gates of QFT(n)."""
    import oracle

    recs = oracle.gen_qft(n)[:n_sample_gates]
    re, im = oracle.basis(n, 0)
    oracle.run(oracle.CPU_PARALLEL, n, oracle.flatten(recs[:1]), re, im)  # touch pages
    t0 = time.perf_counter()
    oracle.run(oracle.CPU_PARALLEL, n, oracle.flatten(recs), re, im)
    dt = time.perf_counter() - t0
    del re, im
    return n_sample_gates * (1 << n) * 32 / dt / 1e9, dt, oracle.num_threads()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="smaller sizes (debugging)")
    ap.add_argument("--cpu-gates", type=int, default=160)
    ap.add_argument("--tile", type=int, default=12)
    ap.add_argument("--kernels-only", action="store_true",
                    help="only 5 launches each of the dominant kernel (controlled-phase "
                         "diagonal, 30q c128) and of the copy kernel, for counter runs")
    ap.add_argument("--mfma-only", action="store_true",
                    help="only 5 launches each of one dense 6-qubit group on a 30q c64 "
                         "state in FP32 and in split-BF16x3 MFMA, for counter runs")
    ap.add_argument("--mfma4-only", action="store_true",
                    help="only 5 launches each of one dense 4-qubit group on a 30q c64 state "
                         "in split-BF16x3 MFMA, on qubits {3, 9, 20, 29} and {10, 15, 22, 27}, "
                         "with the default grid (one workgroup per tile) and then with the "
                         "persistent grid of 2 workgroups per CU, for profiler runs")
    ap.add_argument("--mt-only", action="store_true",
                    help="only 3 launches of QFT(30) fused at k = 4 as split-BF16 MFMA tile "
                         "passes (2^12-amplitude tiles), c64, for profiler runs")
    ap.add_argument("--tile-only", action="store_true",
                    help="only 3 launches of the random mix circuit (28 qubits, 1120 gates) "
                         "as exact 2^13-amplitude tile passes, c64, for profiler runs")
    ap.add_argument("--qft-tile-only", action="store_true",
                    help="only 2 launches of QFT(30) complex128 as exact 2^12-amplitude tile "
                         "passes from the random-normal state, for profiler runs")
    ap.add_argument("--batch-only", action="store_true",
                    help="only 3 runs of the batched sweep (512 HEA circuits, 10..22 qubits, "
                         "complex64) at the autotuned shape, for profiler runs")
    ap.add_argument("--passes-only", action="store_true",
                    help="only 3 launches of the random mix circuit (28 qubits, 1120 gates) as "
                         "pass-planned MFMA tile passes fused at k = 4, c64, for profiler runs")
    ap.add_argument("--layered-mt-only", action="store_true",
                    help="only 3 launches of layered(28) depth 20 fused at k = 4 as split-BF16 "
                         "MFMA tile passes (2^12-amplitude tiles), c64, from |0>, for profiler runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only in this build).", file=sys.stderr)
        sys.exit(2)

    import rustatevec_amd as rs

    if args.kernels_only:
        b = rs.HipBackend(0)
        st = b.alloc(30, rs.C128)
        b.init_random_normal(st, 1)
        ms_cp = b.time_gate(st, rs.gate_matrix("cp", np.pi / 8), [29, 15], 4)
        ms_copy = b.time_copy(st, 4)
        print(json.dumps({"diag_cp_ms": ms_cp, "copy_ms": ms_copy,
                          "bytes_per_launch": (1 << 30) * 32}), flush=True)
        return

    if args.mt_only:
        b = rs.HipBackend(0)
        st = b.alloc(30, rs.C64)
        b.init_random_normal(st, 1)
        p = rs.Program(st, rs.fuse(b, rs.qft(30), 4), precision=rs.PREC_MFMA_BF16X3,
                       mfma_tiles=12)
        ms = [round(p.time(), 3) for _ in range(3)]
        print(json.dumps({"qft30_fused_k4_mfma_tiles12_ms": ms, "passes": p.num_kernels()}),
              flush=True)
        return

    if args.batch_only:
        b = rs.HipBackend(0)
        batch = rs.Batch(b, rs.hea_batch(512, 8, 0xC0FFEE), rs.C64)
        tb_best, bl_best, _ = batch.autotune()
        batch.run()
        ms = [round(batch.time(), 3) for _ in range(3)]
        print(json.dumps({"batch_ms": ms, "tile_qubits": tb_best, "blocks": bl_best,
                          "launches": batch.info()["launches"],
                          "hbm_bytes_per_run": batch.hbm_bytes()}), flush=True)
        return

    if args.passes_only:
        from rustatevec_amd import planner
        b = rs.HipBackend(0)
        st = b.alloc(28, rs.C64)
        pc, _ = planner.plan_passes(b, None, None, rs.mix(28, 40 * 28, 0xC0FFEE), 4)
        p = rs.Program(st, pc, precision=rs.PREC_MFMA_BF16X3, planned=True)
        ms = []
        for _ in range(3):
            b.init_basis(st, 0)
            ms.append(round(p.time(), 3))
        print(json.dumps({"mix28_passes_k4_ms": ms, "passes": p.num_kernels()}), flush=True)
        return

    if args.layered_mt_only:
        b = rs.HipBackend(0)
        st = b.alloc(28, rs.C64)
        p = rs.Program(st, rs.fuse(b, rs.layered(28, 20, 0xC0FFEE), 4),
                       precision=rs.PREC_MFMA_BF16X3, mfma_tiles=12)
        ms = []
        for _ in range(3):
            b.init_basis(st, 0)
            ms.append(round(p.time(), 3))
        print(json.dumps({"layered28_fused_k4_mfma_tiles12_ms": ms, "passes": p.num_kernels()}),
              flush=True)
        return

    if args.qft_tile_only:
        b = rs.HipBackend(0)
        st = b.alloc(30, rs.C128)
        b.init_random_normal(st, 1)
        p = rs.Program(st, rs.qft(30), tile_qubits=12)
        ms = [round(p.time(), 3) for _ in range(2)]
        print(json.dumps({"qft30_tiles12_c128_ms": ms, "passes": p.num_kernels()}), flush=True)
        return

    if args.tile_only:
        b = rs.HipBackend(0)
        st = b.alloc(28, rs.C64)
        p = rs.Program(st, rs.mix(28, 40 * 28, 0xC0FFEE), tile_qubits=13)
        ms = []
        for _ in range(3):
            b.init_basis(st, 0)
            ms.append(round(p.time(), 3))
        print(json.dumps({"mix28_tiles13_ms": ms, "passes": p.num_kernels()}), flush=True)
        return

    if args.mfma_only:
        b = rs.HipBackend(0)
        st = b.alloc(30, rs.C64)
        b.init_random_normal(st, 1)
        g6 = rs.Circuit(30).push(dense_group_6(), [3, 9, 14, 20, 25, 29])
        rec = {}
        for prec, name in ((rs.PREC_MFMA_F32, "f32"), (rs.PREC_MFMA_BF16X3, "bf16x3")):
            p = rs.Program(st, g6, precision=prec)
            rec[name + "_ms"] = [round(p.time(), 4) for _ in range(5)]
            del p
        print(json.dumps(rec), flush=True)
        st.free()
        b.close()
        return

    if args.mfma4_only:
        b = rs.HipBackend(0)
        st = b.alloc(30, rs.C64)
        b.init_random_normal(st, 1)
        u4 = np.linalg.qr(np.random.default_rng(1).normal(size=(16, 16))
                          + 1j * np.random.default_rng(2).normal(size=(16, 16)))[0]
        cus = b.device_info().get("compute_units", 256)
        rec = {}
        for blocks, name in ((0, "one_per_tile"), (2 * cus, "two_per_cu")):
            b.tune(mfma_blocks=blocks)
            for qs in ([3, 9, 20, 29], [10, 15, 22, 27]):
                p = rs.Program(st, rs.Circuit(30).push(u4, qs), precision=rs.PREC_MFMA_BF16X3)
                rec[f"{name}_q{qs[0]}_{qs[-1]}_ms"] = [round(p.time(), 4) for _ in range(5)]
                del p
        b.tune(mfma_blocks=0)
        print(json.dumps(rec), flush=True)
        st.free()
        b.close()
        return

    n = 24 if args.quick else 30
    n_layered = 22 if args.quick else 28
    seed = 0xC0FFEE
    b = rs.HipBackend(0)
    log(f"device {b.device_info()}; tuning {b.tune()}")

    # ------------------------------------------------------------ QFT(30) plain c128
    qft = rs.qft(n)
    st = b.alloc(n, rs.C128)
    reset = lambda: b.init_random_normal(st, seed)  # noqa: E731
    prog = rs.Program(st, qft)
    passes = prog.num_kernels()
    step_s = timed_steps(b, st, prog, args.steps, args.warmup, reset)
    bytes_step = passes * (1 << n) * 32
    value = bytes_step / step_s / 1e9
    log(f"QFT({n}) plain c128: {passes} passes, {step_s:.4f} s/step, {value:.1f} GB/s")
    # the state across the ABI (SoA host arrays <-> interleaved HBM): download and upload
    # rates, and the circuit rate with both transfers included (never the reported value)
    t0 = time.perf_counter()
    host_re, host_im = b.download_parts(st)
    dl_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    b.upload(st, re=host_re, im=host_im)
    b.synchronize(st)
    ul_s = time.perf_counter() - t0
    del host_re, host_im
    sbytes = (1 << n) * 16
    host_rec = {"state_bytes": sbytes, "download_s": round(dl_s, 3), "upload_s": round(ul_s, 3),
                "download_gbs": round(sbytes / dl_s / 1e9, 2),
                "upload_gbs": round(sbytes / ul_s / 1e9, 2),
                "pcie_inclusive_gbs": round(bytes_step / (step_s + dl_s + ul_s) / 1e9, 1)}
    log(f"host transfers of the {n}q state: {host_rec}")
    del prog
    reset()
    rs.Program(st, qft).launch()  # the complex128 reference result stays resident

    # dominant kernel: the controlled-phase diagonal pass (435 of the 480 launches)
    per_launch = (1 << n) * 32
    ms_cp = b.time_gate(st, rs.gate_matrix("cp", np.pi / 8), [n - 1, n // 2], 10)
    ms_h = {q: b.time_gate(st, rs.gate_matrix("h"), [q], 10) for q in (0, 1, n // 2, n - 1)}
    ms_copy = b.time_copy(st, 10)
    achieved = per_launch / (ms_cp * 1e-3) / 1e9
    copy_gbs = per_launch / (ms_copy * 1e-3) / 1e9
    log(f"diag CP kernel {ms_cp:.3f} ms -> {achieved:.1f} GB/s; copy {copy_gbs:.1f} GB/s")
    # undo the timing gates' effect on the reference: recompute it
    reset()
    rs.Program(st, qft).launch()

    # the same per-gate program with subspace launches: CP touches only the amplitudes
    # whose phase is not 1 (a quarter), the rest as before; must equal the full passes
    st_s = b.alloc(n, rs.C128)
    reset_s = lambda: b.init_random_normal(st_s, seed)  # noqa: E731
    prog_s = rs.Program(st_s, qft, subspace=True)
    sub_s = timed_steps(b, st_s, prog_s, max(1, args.steps), 1, reset_s)
    reset_s()
    prog_s.launch()
    err_s, _ = b.compare(st, st_s)
    subspace = {"passes": prog_s.num_kernels(), "circuit_s": round(sub_s, 5),
                "speedup_vs_plain": round(step_s / sub_s, 2), "max_abs_err": err_s}
    del prog_s
    st_s.free()
    log(f"QFT({n}) per-gate with subspace launches: {subspace}")

    # on-device sampling of the QFT(30) output state (inverse CDF, 1e6 shots)
    shots = 1_000_000
    b.sample(st, shots, seed)
    t0 = time.perf_counter()
    outcomes, _ = b.sample(st, shots, seed)
    sample_s = time.perf_counter() - t0
    sampling = {"qubits": n, "shots": shots, "wall_ms": round(sample_s * 1e3, 3),
                "state_bytes_read": (1 << n) * 16,
                "distinct_outcomes": int(len(np.unique(outcomes)))}
    log(f"sampling {shots} shots of QFT({n}): {sampling}")

    # ------------------------------------------------------------ tiled exact c128
    tiled = {}
    st2 = b.alloc(n, rs.C128)
    reset2 = lambda: b.init_random_normal(st2, seed)  # noqa: E731
    p2 = rs.Program(st2, qft, tile_qubits=args.tile)
    t_s = timed_steps(b, st2, p2, max(1, args.steps), 1, reset2)
    reset2()
    p2.launch()
    err, fid = b.compare(st, st2)
    tiled = {"tile_qubits": args.tile, "passes": p2.num_kernels(), "circuit_s": round(t_s, 5),
             "speedup_vs_plain": round(step_s / t_s, 2), "max_abs_err": err,
             "one_minus_fidelity": 1 - fid}
    log(f"QFT({n}) tiled c128 t={args.tile}: {p2.num_kernels()} passes, {t_s:.4f} s "
        f"({step_s / t_s:.1f}x), err {err:.2e}")
    del p2
    st2.free()

    # ------------------------------------------------------------ fused c64 on MFMA
    fused_out = []
    mfma_roof = None
    st64 = b.alloc(n, rs.C64)
    reset64 = lambda: b.init_random_normal(st64, seed)  # noqa: E731
    # per-group launches in FP32 and split-BF16, and split-BF16 runs of groups in MFMA tile
    # passes (one HBM round trip per 2^12-amplitude tile run)
    forms = ((rs.PREC_MFMA_F32, "mfma_f32", {}), (rs.PREC_MFMA_BF16X3, "mfma_bf16x3", {}),
             (rs.PREC_MFMA_BF16X3, "mfma_bf16x3_tiles12", {"mfma_tiles": 12}))
    for k in (4, 5, 6):
        fq = rs.fuse(b, qft, k)
        for prec, name, kw in forms:
            p = rs.Program(st64, fq, precision=prec, **kw)
            f_s = timed_steps(b, st64, p, max(1, args.steps), 1, reset64)
            reset64()
            p.launch()
            err, fid = b.compare(st, st64)
            rec = {"k": k, "precision": name, "passes": p.num_kernels(),
                   "circuit_s": round(f_s, 5), "speedup_vs_plain": round(step_s / f_s, 2),
                   "max_abs_err": err, "one_minus_fidelity": 1 - fid}
            fused_out.append(rec)
            log(f"QFT({n}) fused k={k} {name}: {rec}")
            del p
    # matrix-core roofline: one dense 6-qubit group, FP32 MFMA, on the 30q c64 state
    g6 = rs.Circuit(n).push(dense_group_6(), [3, 9, 14, 20, 25, 29])
    for prec, peak, name in ((rs.PREC_MFMA_F32, FP32_MFMA_PEAK_TFLOPS, "f32"),
                             (rs.PREC_MFMA_BF16X3, BF16_MFMA_PEAK_TFLOPS, "bf16x3")):
        p = rs.Program(st64, g6, precision=prec)
        p.launch()
        ms = min(p.time() for _ in range(5))
        flops = 8 * 64 * (1 << n) * (3 if name == "bf16x3" else 1)
        gbs = (1 << n) * 16 / (ms * 1e-3) / 1e9
        tf = flops / (ms * 1e-3) / 1e12
        log(f"k=6 group {name}: {ms:.3f} ms, {tf:.1f} TFLOP/s of {peak}, {gbs:.0f} GB/s")
        if name == "f32":
            mfma_roof = {"kernel": "k_fused_f32<6> (one 6-qubit group, 30q c64)",
                         "ms_per_launch": round(ms, 4), "achieved_tflops": round(tf, 2),
                         "peak_tflops": peak, "frac_mfma": round(tf / peak, 4),
                         "hbm_gbs": round(gbs, 1), "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}
        else:
            mfma_roof["bf16x3"] = {"ms_per_launch": round(ms, 4), "achieved_tflops": round(tf, 2),
                                   "peak_tflops": peak, "hbm_gbs": round(gbs, 1),
                                   "frac_hbm": round(gbs / HBM_PEAK_GBS, 4)}
        del p
    st64.free()
    st.free()

    # ------------------------------------------------------------ layered 28q
    lay = rs.layered(n_layered, 20, seed)
    st = b.alloc(n_layered, rs.C128)
    resetl = lambda: b.init_basis(st, 0)  # noqa: E731
    prog = rs.Program(st, lay)
    lay_s = timed_steps(b, st, prog, 1, 1, resetl)
    lay_passes = prog.num_kernels()
    del prog
    pt = rs.Program(st, lay, tile_qubits=args.tile)
    lay_t = timed_steps(b, st, pt, 1, 1, resetl)
    lay_tp = pt.num_kernels()
    del pt
    log(f"layered({n_layered}, 20) c128: plain {lay_s:.4f} s ({lay_passes} passes), "
        f"tiled {lay_t:.4f} s ({lay_tp} passes)")
    st.free()

    # ------------------------------------------------------------ batched HEA sweep
    count = 64 if args.quick else 512
    circuits = rs.hea_batch(count, 8, seed)
    batch = rs.Batch(b, circuits, rs.C64)
    tb_best, bl_best, _ = batch.autotune()
    batch.run()
    b_ms = min(batch.time() for _ in range(3))
    z0 = batch.expectation_z0()
    rng = np.random.default_rng(seed)
    sampled = sorted(int(i) for i in rng.choice(count, size=8, replace=False))
    psums = [float(batch.probabilities(i).sum()) for i in sampled]
    ns = [c.n_qubits for c in circuits]
    equiv = sum(8 * (3 * q - 1) * (1 << q) * 16 for q in ns)  # plain-pass bytes, c64
    batch_rec = {"circuits": count, "qubits_min": min(ns), "qubits_max": max(ns),
                 "gates": sum(len(c) for c in circuits), "tile_qubits": tb_best,
                 "blocks": bl_best, "launches": batch.info()["launches"],
                 "batch_ms": round(b_ms, 3),
                 "plain_pass_equiv_gbs": round(equiv / (b_ms * 1e-3) / 1e9, 1),
                 # HBM: each pass writes its circuit's state, each but the first reads it
                 # (circuits that fit one tile are built and run in LDS, written once)
                 "roofline": {"bound": "hbm", "achieved": round(batch.hbm_bytes()
                                                                / (b_ms * 1e-3) / 1e9, 1),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(batch.hbm_bytes() / (b_ms * 1e-3) / 1e9
                                            / HBM_PEAK_GBS, 4),
                              "traffic": batch.hbm_bytes(),
                              "algorithmic_bytes_per_run": batch.hbm_bytes()},
                 "z0_mean": float(z0.mean()), "z0_first": [round(float(x), 6) for x in z0[:4]],
                 "sampled_prob_sums": [round(x, 6) for x in psums]}
    log(f"batched HEA x{count}: {batch_rec}")
    del batch, circuits

    # ------------------------------------------------------------ learned planner
    from rustatevec_amd import planner

    t0 = time.perf_counter()
    X, y = planner.collect_timings(b, sizes=(22, 24) if args.quick else (24, 26),
                                   per_size=150, seed=seed)
    model, train_loss = planner.train_cost_model(b, X, y, epochs=2000)
    t_model = time.perf_counter() - t0
    # the tile-pass cost model of the form planner (fused groups vs exact tile passes)
    t0 = time.perf_counter()
    XT, yT = planner.collect_tile_timings(b, sizes=(24,) if args.quick else (26,),
                                          per_size=300, seed=seed)
    tile_model, tile_loss = planner.train_tile_model(b, XT, yT, epochs=2000)
    t_tile_model = time.perf_counter() - t0
    # and its MFMA-tile-pass model (greedy fusion run as split-BF16 MFMA tile passes)
    t0 = time.perf_counter()
    XM, yM = planner.collect_mt_timings(b, sizes=(24,) if args.quick else (26,),
                                        per_size=300, seed=seed)
    mt_model, mt_loss = planner.train_mt_model(b, XM, yM, epochs=2000)
    t_mt_model = time.perf_counter() - t0

    # the pass planner's models: exact / MFMA pass timings of mix circuits (training seeds
    # 1000..1003, none of the measured circuits), HBM-resident sizes
    t0 = time.perf_counter()
    Xpt, ypt, Xpm, ypm = planner.collect_pass_timings(
        b, sizes=(22, 24) if args.quick else (24, 26, 28), seeds=range(1000, 1004))
    pt_model, pt_loss = planner.train_tile_model(b, Xpt, ypt, epochs=3000)
    pm_model, pm_loss = planner.train_mt_model(b, Xpm, ypm, epochs=3000)
    t_pass_models = time.perf_counter() - t0

    def pass_legs(nq_, circ, ref_, st_, plain_s_, reset_, fixed_forms):
        """The learned pass plan against fixed forms (each timed and judged against the
        plain complex128 result ref_) and the uniform forms over the same passes."""
        t0_ = time.perf_counter()
        pc_, pred_ = planner.plan_passes(b, pt_model, pm_model, circ, planner.FORM_LEARNED)
        plan_s_ = time.perf_counter() - t0_
        pforms = [f for f, _, _, _ in planner.plan_info(pc_)]
        runs = [("learned_passes", pc_, {"precision": rs.PREC_MFMA_BF16X3, "planned": True})]
        for f in (0, 3, 4, 5):
            pcf, _ = planner.plan_passes(b, None, None, circ, f)
            runs.append((f"passes_uniform_{f}", pcf,
                         {"precision": rs.PREC_MFMA_BF16X3, "planned": True}))
        runs += fixed_forms
        # every program built and judged first, then timed in rounds over all of them (each
        # program's best round): clock and cache drift over the run reach every form alike
        progs_ = []
        for name, c_, kw in runs:
            p_ = rs.Program(st_, c_, **kw)
            reset_()
            p_.launch()
            err_, fid_ = b.compare(ref_, st_)
            progs_.append((name, p_, err_, fid_))
        best_s = {name: float("inf") for name, _, _, _ in progs_}
        for r_ in range(3 if nq_ <= 28 else 2):
            for name, p_, _, _ in progs_:
                best_s[name] = min(best_s[name],
                                   timed_steps(b, st_, p_, 1, 1 if r_ == 0 else 0, reset_))
        legs_ = {}
        for name, p_, err_, fid_ in progs_:
            s_ = best_s[name]
            legs_[name] = {"passes": p_.num_kernels(), "circuit_s": round(s_, 5),
                           "speedup_vs_plain_c128": round(plain_s_ / s_, 2),
                           "max_abs_err": err_, "one_minus_fidelity": 1 - fid_}
            log(f"  mix({nq_}) {name}: {s_:.4f} s, err {err_:.1e}")
        del progs_
        lp = legs_["learned_passes"]
        fixed = {k: v["circuit_s"] for k, v in legs_.items()
                 if not k.startswith("passes_") and k != "learned_passes"}
        unif = {k: v["circuit_s"] for k, v in legs_.items() if k.startswith("passes_uniform")}
        best_f = min(fixed, key=fixed.get)
        best_u = min(unif, key=unif.get)
        # every pass reads and writes the state once (rocprof: 1.00x the algorithmic bytes)
        hbm = lp["passes"] * 2 * (1 << nq_) * 8
        return {"qubits": nq_, "gates": len(circ), "plan_s": round(plan_s_, 3),
                "predicted_ms": round(pred_, 3),
                "pass_forms": {str(k): pforms.count(k) for k in sorted(set(pforms))},
                "legs": legs_,
                "gain_vs_greedy_k4": round(fixed["greedy_k4"] / lp["circuit_s"], 3),
                "best_fixed_form": best_f,
                "gain_vs_best_fixed_form": round(fixed[best_f] / lp["circuit_s"], 3),
                "best_uniform_pass_form": best_u,
                "gain_vs_best_uniform_pass_form": round(unif[best_u] / lp["circuit_s"], 3),
                "roofline": {"bound": "lds/valu (tile passes)",
                             "achieved": round(hbm / lp["circuit_s"] / 1e9, 1),
                             "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(hbm / lp["circuit_s"] / 1e9 / HBM_PEAK_GBS, 4),
                             "traffic": hbm}}

    def fixed_forms_of(circ, full):
        """The fixed forms the pass planner is judged against: greedy fusion at 4/5/6
        (per-group matrix-core launches), exact tile passes at 11..13 bits, greedy fusion at
        4/5 in MFMA tile passes (full=False: greedy 4, tiles 13 and MFMA tiles 4 only)."""
        mf = rs.PREC_MFMA_BF16X3
        out_ = [("greedy_k%d" % k, rs.fuse(b, circ, k), {"precision": mf})
                for k in ((4, 5, 6) if full else (4,))]
        out_ += [("tiles_t%d" % t, circ, {"tile_qubits": t})
                 for t in ((11, 12, 13) if full else (13,))]
        out_ += [("mfma_tiles_k%d" % k, rs.fuse(b, circ, k), {"precision": mf, "mfma_tiles": 12})
                 for k in ((4, 5) if full else (4,))]
        return out_

    def form_name(tq, chosen):
        if chosen >= 100:
            return f"mfma_tiles_k{chosen - 100}"
        return f"tiles_t{tq}" if tq else ("dp" if chosen == 0 else f"greedy_k{chosen}")

    def form_program(st_, circ, tq, mt):
        """A program in the form plan_form chose: exact tiles, MFMA tiles or fused groups."""
        return rs.Program(st_, circ, precision=rs.PREC_EXACT if tq else rs.PREC_MFMA_BF16X3,
                          tile_qubits=tq, mfma_tiles=mt)
    nq = 22 if args.quick else 28
    mixc = rs.mix(nq, 40 * nq, seed)
    ref = b.alloc(nq, rs.C128)
    resetr = lambda: b.init_basis(ref, 0)  # noqa: E731
    p = rs.Program(ref, mixc)
    mix_plain_s = timed_steps(b, ref, p, 1, 1, resetr)
    resetr()
    p.launch()
    del p
    st64 = b.alloc(nq, rs.C64)
    reset64 = lambda: b.init_basis(st64, 0)  # noqa: E731
    legs = {}
    t0 = time.perf_counter()
    planned, pred_ms, chosen = planner.plan(b, model, mixc, 6)
    t_plan = time.perf_counter() - t0
    formc, form_tq, form_mt, form_pred, form_chosen = planner.plan_form(
        b, model, tile_model, mixc, rs.C64, mt_model=mt_model)
    cands = [("greedy_k%d" % k, rs.fuse(b, mixc, k), 0, 0) for k in (4, 5, 6)]
    cands += [("learned", planned, 0, 0), ("learned_form", formc, form_tq, form_mt)]
    cands += [("greedy_k%d_mfma_tiles12" % k, rs.fuse(b, mixc, k), 0, 12) for k in (4, 5)]
    for name, circ, tq, mt in cands:
        p = rs.Program(st64, circ, precision=rs.PREC_EXACT if tq else rs.PREC_MFMA_BF16X3,
                       tile_qubits=tq, mfma_tiles=mt)
        s = timed_steps(b, st64, p, 2, 1, reset64)
        reset64()
        p.launch()
        err, fid = b.compare(ref, st64)
        legs[name] = {"passes": p.num_kernels(), "circuit_s": round(s, 5),
                      "speedup_vs_plain_c128": round(mix_plain_s / s, 2),
                      "max_abs_err": err, "one_minus_fidelity": 1 - fid}
        del p
    learned_rec = {"qubits": nq, "gates": len(mixc), "timings": int(len(X)),
                   "train_mse_std": round(train_loss, 4), "model_s": round(t_model, 2),
                   "plan_s": round(t_plan, 3), "predicted_ms": round(pred_ms, 3),
                   "chosen": "dp" if chosen == 0 else f"greedy_k{chosen}",
                   "plain_c128_s": round(mix_plain_s, 4), "legs": legs,
                   "gain_vs_greedy_k4": round(legs["greedy_k4"]["circuit_s"]
                                              / legs["learned"]["circuit_s"], 3),
                   "form_planner": {"tile_timings": int(len(XT)),
                                    "tile_train_mse_std": round(tile_loss, 4),
                                    "tile_model_s": round(t_tile_model, 2),
                                    "mt_timings": int(len(XM)),
                                    "mt_train_mse_std": round(mt_loss, 4),
                                    "mt_model_s": round(t_mt_model, 2),
                                    "chosen": form_name(form_tq, form_chosen),
                                    "predicted_ms": round(form_pred, 3),
                                    "gain_vs_greedy_k4": round(
                                        legs["greedy_k4"]["circuit_s"]
                                        / legs["learned_form"]["circuit_s"], 3),
                                    "gain_vs_greedy_k6": round(
                                        legs["greedy_k6"]["circuit_s"]
                                        / legs["learned_form"]["circuit_s"], 3)}}
    # the learned pass planner on the same circuit
    pass_recs = [pass_legs(nq, mixc, ref, st64, mix_plain_s, reset64,
                           fixed_forms_of(mixc, True))]
    learned_rec["pass_planner"] = {"exact_timings": int(len(Xpt)), "mfma_timings": int(len(Xpm)),
                                   "tile_train_mse_std": round(pt_loss, 4),
                                   "mt_train_mse_std": round(pm_loss, 4),
                                   "model_s": round(t_pass_models, 2), "sizes": pass_recs}
    # the learned form's HBM rate (each pass reads and writes the state once)
    lf = legs["learned_form"]
    lf_bytes = lf["passes"] * 2 * (1 << nq) * 8
    lf["roofline"] = {"bound": "lds/valu (tile passes)",
                      "achieved": round(lf_bytes / lf["circuit_s"] / 1e9, 1),
                      "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": round(lf_bytes / lf["circuit_s"] / 1e9 / HBM_PEAK_GBS, 4),
                      "traffic": lf_bytes}
    log(f"learned planner: {learned_rec}")
    st64.free()
    ref.free()

    # the same model across the C5 size range (24..32 qubits, 40 n gates): the reference's
    # default greedy k = 4, greedy k = 6 and the learned plan, each against plain c128
    sweep = []
    for nq2 in (() if args.quick else (24, 32)):
        mc = rs.mix(nq2, 40 * nq2, seed + nq2)
        ref = b.alloc(nq2, rs.C128)
        resetr = lambda: b.init_basis(ref, 0)  # noqa: E731
        p = rs.Program(ref, mc)
        plain_s = timed_steps(b, ref, p, 1, 0, resetr)
        log(f"  mix({nq2}) plain complex128: {plain_s:.3f} s")
        resetr()
        p.launch()
        del p
        st64 = b.alloc(nq2, rs.C64)
        reset64 = lambda: b.init_basis(st64, 0)  # noqa: E731
        planned, _, chosen = planner.plan(b, model, mc, 6)
        formc, form_tq, form_mt, _, form_chosen = planner.plan_form(b, model, tile_model, mc,
                                                                    rs.C64, mt_model=mt_model)
        rec = {"qubits": nq2, "gates": len(mc), "plain_c128_s": round(plain_s, 4),
               "chosen": "dp" if chosen == 0 else f"greedy_k{chosen}",
               "form_chosen": form_name(form_tq, form_chosen)}
        for name, circ, tq, mt in (("greedy_k4", rs.fuse(b, mc, 4), 0, 0),
                                   ("greedy_k6", rs.fuse(b, mc, 6), 0, 0),
                                   ("learned", planned, 0, 0),
                                   ("learned_form", formc, form_tq, form_mt)):
            p = form_program(st64, circ, tq, mt)
            s_ = timed_steps(b, st64, p, 1, 1, reset64)
            reset64()
            p.launch()
            err, fid = b.compare(ref, st64)
            log(f"  mix({nq2}) {name}: {s_:.4f} s, err {err:.1e}")
            rec[name] = {"circuit_s": round(s_, 5), "speedup_vs_plain_c128": round(plain_s / s_, 2),
                         "max_abs_err": err, "one_minus_fidelity": 1 - fid}
            del p
        pr = pass_legs(nq2, mc, ref, st64, plain_s, reset64, fixed_forms_of(mc, nq2 <= 28))
        log(f"learned pass planner at {nq2} qubits: {pr}")
        learned_rec["pass_planner"]["sizes"].append(pr)
        log(f"learned planner at {nq2} qubits: {rec}")
        sweep.append(rec)
        st64.free()
        ref.free()
    learned_rec["size_sweep"] = sweep

    # layered(28) through the same forms: greedy / learned fusion on the matrix cores and
    # exact complex64 tile passes, each judged against the plain complex128 run
    ref = b.alloc(n_layered, rs.C128)
    b.init_basis(ref, 0)
    p = rs.Program(ref, lay)
    p.launch()
    del p
    st64 = b.alloc(n_layered, rs.C64)
    reset64 = lambda: b.init_basis(st64, 0)  # noqa: E731
    planned, _, chosen = planner.plan(b, model, lay, 6)
    formc, form_tq, form_mt, _, form_chosen = planner.plan_form(b, model, tile_model, lay,
                                                                rs.C64, mt_model=mt_model)
    lay_legs = {}
    cands = [("greedy_k%d" % k, rs.fuse(b, lay, k), rs.PREC_MFMA_BF16X3, 0, 0)
             for k in (4, 5, 6)]
    cands += [("learned", planned, rs.PREC_MFMA_BF16X3, 0, 0),
              ("learned_form", formc, rs.PREC_EXACT if form_tq else rs.PREC_MFMA_BF16X3,
               form_tq, form_mt),
              ("tiled_c64_exact", lay, rs.PREC_EXACT, 13, 0)]
    cands += [("greedy_k%d_mfma_tiles12" % k, rs.fuse(b, lay, k), rs.PREC_MFMA_BF16X3, 0, 12)
              for k in (4, 5)]
    for name, circ, prec, tq, mt in cands:
        p = rs.Program(st64, circ, precision=prec, tile_qubits=tq, mfma_tiles=mt)
        s = timed_steps(b, st64, p, 2, 1, reset64)
        reset64()
        p.launch()
        err, fid = b.compare(ref, st64)
        lay_legs[name] = {"passes": p.num_kernels(), "circuit_s": round(s, 5),
                          "speedup_vs_plain_c128": round(lay_s / s, 2),
                          "max_abs_err": err, "one_minus_fidelity": 1 - fid}
        del p
    lay_legs["learned"]["chosen"] = "dp" if chosen == 0 else f"greedy_k{chosen}"
    lay_legs["learned_form"]["chosen"] = form_name(form_tq, form_chosen)
    log(f"layered({n_layered}) fused / learned / tiled c64: {lay_legs}")
    st64.free()
    ref.free()

    # ------------------------------------------------------------ GHZ(20)
    st = b.alloc(20, rs.C128)
    b.init_basis(st, 0)
    prog = rs.Program(st, rs.ghz_chain(20), use_graph=True)
    prog.launch()
    b.synchronize(st)
    ghz_ms = min(prog.time() for _ in range(20))
    del prog
    # the same chain as tile passes (a few launches instead of 20), checked on its output
    ghz_tiled = {}
    for tq in (10, 12):
        b.init_basis(st, 0)
        prog = rs.Program(st, rs.ghz_chain(20), use_graph=True, tile_qubits=tq)
        prog.launch()
        b.synchronize(st)
        h = 2 ** -0.5
        ok = (abs(b.amplitude(st, 0) - h) < 1e-15 and abs(b.amplitude(st, (1 << 20) - 1) - h) < 1e-15
              and abs(b.norm_sqr(st) - 1.0) < 1e-14)
        us = min(prog.time() for _ in range(20)) * 1e3
        if ok and ("us" not in ghz_tiled or us < ghz_tiled["us"]):
            ghz_tiled = {"us": round(us, 2), "tile_qubits": tq, "passes": prog.num_kernels()}
        del prog
    st.free()
    log(f"GHZ(20) graph replay: {ghz_ms * 1e3:.1f} us per-gate, tiled {ghz_tiled}")

    # ------------------------------------------------------------ CPU baseline
    cpu = None
    if not args.no_cpu:
        gbs, dt, cores = cpu_baseline(args.cpu_gates, n)
        cpu = {"value": round(gbs, 3), "unit": "GB/s", "cores": cores, "kind": "port",
               "sample": f"first {args.cpu_gates} of {passes} QFT({n}) gates, complex128, "
                         f"CpuBackend::parallel restatement, {dt:.1f} s",
               "circuit_s_extrapolated": round(dt * passes / args.cpu_gates, 2)}
        log(f"CPU baseline: {gbs:.2f} GB/s on {cores} threads ({dt:.1f} s)")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"QFT({n}) plain complex128 from a random-normal state",
                   "qubits": n, "gates": passes, "bytes_per_step": bytes_step,
                   "circuit_s": round(step_s, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic() if n == 30 else None,
                     "kernel": f"k_diag_small (controlled phase, {n}q c128)",
                     "bytes_per_launch": per_launch, "ms_per_launch": round(ms_cp, 4)},
        "copy_kernel_gbs": round(copy_gbs, 1),
        "h_gate_gbs": {str(q): round(per_launch / (t * 1e-3) / 1e9, 1) for q, t in ms_h.items()},
        "tiled": tiled,
        "host_transfer": host_rec,
        "plain_subspace": subspace,
        "sampling": sampling,
        "fused": fused_out,
        # the headline's second half: the fastest fused form of QFT(30) within the error bar
        "fused_best": min((r for r in fused_out if r["max_abs_err"] <= 1e-6),
                          key=lambda r: r["circuit_s"], default=None),
        "mfma_roofline": mfma_roof,
        "layered28": {"plain_s": round(lay_s, 4), "plain_passes": lay_passes,
                      "tiled_s": round(lay_t, 4), "tiled_passes": lay_tp,
                      "c64_forms": lay_legs},
        "batched_hea": batch_rec,
        "learned_planner": learned_rec,
        "ghz20_us": round(ghz_ms * 1e3, 2),
        "ghz20_tiled": ghz_tiled,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    b.close()


if __name__ == "__main__":
    main()
