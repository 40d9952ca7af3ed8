"""Runner: ``python -m rustatevec_amd [PROGRAM.qasm | N] [--dtype c128|c64] [--shots K]``.

With no program (or a qubit count N, default 4) it is the reference's smoke runner
(crates/qsv-cli/src/main.rs:6-31): GHZ(N) = H(0), CX(0, q) for q = 1..N-1, executed, then the
outcome probabilities above 1e-12 as ``|bits> : p`` (qubit N-1 leftmost) and <psi|psi>.
With an OpenQASM 3 / 2.0 file (the input that runner names for a later milestone,
main.rs:1-2) it executes that program instead.  ``--shots K`` draws K outcomes on the
device (Backend::sample) and prints their counts instead of the probabilities.
"""
import argparse
import os
import sys

import numpy as np

import rustatevec_amd as rs


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m rustatevec_amd", description=__doc__.split("\n")[0])
    ap.add_argument("program", nargs="?", default="4",
                    help="OpenQASM file, or a qubit count for the GHZ smoke run (default 4)")
    ap.add_argument("--dtype", choices=("c128", "c64"), default="c128")
    ap.add_argument("--shots", type=int, default=0, help="sample this many outcomes")
    ap.add_argument("--seed", type=int, default=0xC0FFEE)
    ap.add_argument("--top", type=int, default=64,
                    help="print at most this many outcomes (largest first)")
    args = ap.parse_args(argv)

    if args.program.isdigit() and not os.path.exists(args.program):
        n = int(args.program)
        circ, title = rs.ghz(n), f"GHZ({n})"
    else:
        try:
            circ = rs.read_qasm(args.program)
        except OSError as e:
            print(f"cannot read {args.program}: {e}", file=sys.stderr)
            return 2
        except rs.QsvhError as e:
            print(f"{args.program}: {e}", file=sys.stderr)
            return 2
        n, title = circ.n_qubits, os.path.basename(args.program)

    backend = rs.HipBackend(0)
    st = backend.execute(circ, rs.C128 if args.dtype == "c128" else rs.C64)
    try:
        if args.shots > 0:
            out, _ = backend.sample(st, args.shots, args.seed)
            vals, counts = np.unique(out, return_counts=True)
            order = np.argsort(-counts, kind="stable")[: args.top]
            print(f"{title} ({n} qubits, {len(circ)} gates) -- {args.shots} shots, "
                  f"{len(vals)} distinct outcomes:")
            for i in order:
                print(f"  |{int(vals[i]):0{n}b}⟩ : {int(counts[i])}")
        else:
            p = backend.probabilities(st)
            idx = np.nonzero(p > 1e-12)[0]
            if len(idx) > args.top:
                idx = idx[np.argsort(-p[idx], kind="stable")[: args.top]]
                note = f" (largest {args.top})"
            else:
                note = ""
            print(f"{title} ({n} qubits, {len(circ)} gates) — non-negligible outcome "
                  f"probabilities{note}:")
            for i in idx:
                print(f"  |{int(i):0{n}b}⟩ : {p[i]:.6f}")
            print(f"⟨ψ|ψ⟩ = {backend.norm_sqr(st):.6f}")
    finally:
        st.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
