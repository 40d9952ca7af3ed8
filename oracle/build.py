"""Build recipe for the CPU oracle library (test infrastructure only).

The oracle is the C restatement in ``qsv_oracle.c``; it is compiled with the flags that
mirror the reference's arithmetic: no floating-point contraction (Rust does not fuse
``a*b + c`` in ``Cplx`` arithmetic, crates/qsv-core/src/complex.rs:95-105) and OpenMP in
place of rayon.  ``-march=x86-64-v3`` stands in for the reference's
``target-cpu=native`` while staying portable between this build host and the machine
the library later runs on.

The reference itself is Rust (no rustc/cargo in this image), so there is no
``oracle/_ref`` build: the restatement is pinned by the checks in tests/test_oracle.py.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

CFLAGS = [
    "-O3", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
    "-march=x86-64-v3", "-Wall", "-Wextra", "-Wno-unused-parameter",
]


def build(verbose: bool = False) -> str:
    src = os.path.join(HERE, "qsv_oracle.c")
    tmp = LIB + ".tmp"
    cmd = ["gcc", *CFLAGS, src, "-o", tmp, "-lm"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
