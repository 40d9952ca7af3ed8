"""Benchmark of the statevector hot path on one MI355X.

Workload (one "step"): QFT(30) -- H + controlled-phase ladder + final swaps, 480 gates --
applied in complex128 to a random-normal 30-qubit state already resident in HBM
(re, im ~ N(0,1) per amplitude, normalised; generated on the device).  The plain engine
applies one gate per HBM pass, so a step moves 480 x 2^30 x 32 B = 16.49 TB of
algorithmic traffic; ``value`` is that traffic divided by the measured step time.

Also reported on the same line: the dominant kernel's roofline (HIP-event timing of the
controlled-phase diagonal kernel on the engine's stream), a copy kernel's bandwidth on
the same state, GHZ(20) latency, the random layered 28-qubit circuit, and the CPU
baseline (the oracle's CpuBackend restatement on the host cores, on a bounded sample of
the same workload).

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

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth
METRIC = ("circuit sec & HBM GB/s (% roofline) at 30q; fused-MFMA vs plain at err<=1e-6")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed_steps(backend, state, prog, steps, warmup):
    for _ in range(warmup):
        prog.launch()
    backend.synchronize(state)
    t0 = time.perf_counter()
    for _ in range(steps):
        prog.launch()
    backend.synchronize(state)
    return (time.perf_counter() - t0) / steps


def cpu_baseline(n_sample_gates: int, n: int = 30):
    """The oracle's CpuBackend::parallel restatement on the first gates of QFT(n)."""
    import oracle

    recs = oracle.gen_qft(n)[:n_sample_gates]
    flat = oracle.flatten(recs)
    re, im = oracle.basis(n, 0)
    oracle.run(oracle.CPU_PARALLEL, n, oracle.flatten(recs[:1]), re, im)  # warm-up touch
    t0 = time.perf_counter()
    oracle.run(oracle.CPU_PARALLEL, n, flat, re, im)
    dt = time.perf_counter() - t0
    del re, im
    gbs = n_sample_gates * (1 << n) * 32 / dt / 1e9
    return gbs, dt, oracle.num_threads()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick", action="store_true", help="smaller sizes (debugging)")
    ap.add_argument("--cpu-gates", type=int, default=40)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only in this build).", file=sys.stderr)
        sys.exit(2)

    import rustatevec_amd as rs

    n = 24 if args.quick else 30
    n_layered = 22 if args.quick else 28
    b = rs.HipBackend(0)
    info = b.device_info()
    log(f"device {info}")

    # ---------------------------------------------------------------- QFT(30) plain
    qft = rs.qft(n)
    st = b.alloc(n, rs.C128)
    b.init_random_normal(st, 0xC0FFEE)
    log(f"QFT({n}) c128 state ready; norm^2 = {b.norm_sqr(st):.15f}")
    prog = rs.Program(st, qft)
    passes = prog.num_kernels()
    step_s = timed_steps(b, st, prog, args.steps, args.warmup)
    bytes_step = passes * (1 << n) * 32
    value = bytes_step / step_s / 1e9
    log(f"QFT({n}) plain c128: {passes} passes, {step_s:.4f} s/step, {value:.1f} GB/s")

    # dominant kernel: the controlled-phase diagonal pass (435 of the 480 launches)
    cp = rs.gate_matrix("cp", np.pi / 8)
    reps = 10
    ms_cp = b.time_gate(st, cp, [n - 1, n // 2], reps)
    h = rs.gate_matrix("h")
    ms_h = {q: b.time_gate(st, h, [q], reps) for q in (0, 1, n // 2, n - 1)}
    ms_copy = b.time_copy(st, reps)
    per_launch = (1 << n) * 32
    achieved = per_launch / (ms_cp * 1e-3) / 1e9
    copy_gbs = per_launch / (ms_copy * 1e-3) / 1e9
    log(f"diag CP kernel {ms_cp:.3f} ms -> {achieved:.1f} GB/s; copy {copy_gbs:.1f} GB/s; "
        + ", ".join(f"H(q={q}) {per_launch / (t * 1e-3) / 1e9:.0f} GB/s" for q, t in ms_h.items()))
    del prog
    st.free()

    # ---------------------------------------------------------------- layered 28q plain
    lay = rs.layered(n_layered, 20, 0xC0FFEE)
    st = b.alloc(n_layered, rs.C128)
    b.init_basis(st, 0)
    prog = rs.Program(st, lay)
    lay_s = timed_steps(b, st, prog, 1, 1)
    lay_bytes = prog.num_kernels() * (1 << n_layered) * 32
    log(f"layered({n_layered}, depth 20) plain c128: {lay_s:.4f} s "
        f"({lay_bytes / lay_s / 1e9:.1f} GB/s)")
    del prog
    st.free()

    # ---------------------------------------------------------------- GHZ(20)
    st = b.alloc(20, rs.C128)
    ghz = rs.ghz_chain(20)
    b.init_basis(st, 0)
    prog = rs.Program(st, ghz, use_graph=True)
    prog.launch()
    b.synchronize(st)
    ghz_ms = min(prog.time() for _ in range(20))
    log(f"GHZ(20) graph replay: {ghz_ms * 1e3:.1f} us")
    del prog
    st.free()

    # ---------------------------------------------------------------- CPU baseline
    cpu = None
    if not args.no_cpu:
        gbs, dt, cores = cpu_baseline(args.cpu_gates, n)
        cpu = {"value": round(gbs, 3), "unit": "GB/s", "cores": cores, "kind": "port",
               "sample": f"first {args.cpu_gates} of {passes} QFT({n}) gates, c128, "
                         f"CpuBackend::parallel restatement ({dt:.1f} s)",
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
                     "traffic": None, "kernel": "k_diag (controlled phase, 30q c128)",
                     "bytes_per_launch": per_launch, "ms_per_launch": round(ms_cp, 4)},
        "copy_kernel_gbs": round(copy_gbs, 1),
        "h_gate_gbs": {str(q): round(per_launch / (t * 1e-3) / 1e9, 1) for q, t in ms_h.items()},
        "layered28_plain_s": round(lay_s, 4),
        "ghz20_us": round(ghz_ms * 1e3, 2),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    b.close()


if __name__ == "__main__":
    main()
