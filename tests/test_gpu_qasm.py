"""OpenQASM programs through the engine on the GPU: the parsed circuit executes to the
oracle's state (CpuBackend restatement, bitwise for complex128 and complex64), and the
runner (the reference's CLI smoke run, crates/qsv-cli/src/main.rs:6-31) prints GHZ's two
outcomes at 0.5 and a QASM program's sampled counts."""
import numpy as np
import pytest

import oracle
import rustatevec_amd as rs
from rustatevec_amd.__main__ import main as run_main

pytestmark = pytest.mark.gpu

PROGRAM = """OPENQASM 3.0;
include "stdgates.inc";
qubit[4] a;
qubit[6] b;
gate layer(t) x, y { ry(t) x; rzz(t / 3) x, y; crx(2 * t) y, x; }
h a;
ry(0.3) b;
layer(0.7) a[0], b[5];
cswap a[1], b[0], b[3];
ctrl(2) @ rx(pi / 5) a[2], a[3], b[1];
inv @ layer(-1.1) b[2], a[0];
negctrl @ u3(0.4, 0.5, 0.6) b[4], a[1];
rxx(0.8) a[3], b[2];
cu(0.2, 0.3, 0.4, 0.5) b[0], b[1];
cx a, b[0:3];
gphase(0.25);
sdg b; tdg a[0]; sxdg b[5];
barrier a, b;
"""


@pytest.mark.parametrize("dtype", [rs.C128, rs.C64])
def test_qasm_program_executes_to_the_oracle_state(backend, dtype):
    circ = rs.from_qasm(PROGRAM)
    assert circ.n_qubits == 10
    npt = np.float64 if dtype == rs.C128 else np.float32
    want = oracle.execute(oracle.CPU_SERIAL, 10, circ.ops(), dtype=npt)
    st = backend.execute(circ, dtype)
    got = backend.download(st)
    st.free()
    assert np.array_equal(got, want)
    assert abs(np.vdot(want, want).real - 1) < (1e-12 if dtype == rs.C128 else 1e-5)


def test_runner_ghz_smoke(capsys):
    assert run_main(["5"]) == 0
    out = capsys.readouterr().out.splitlines()
    assert out[0].startswith("GHZ(5)")
    assert out[1:3] == ["  |00000⟩ : 0.500000", "  |11111⟩ : 0.500000"]
    assert out[3] == "⟨ψ|ψ⟩ = 1.000000"


def test_runner_qasm_file_with_shots(tmp_path, capsys):
    f = tmp_path / "bell.qasm"
    f.write_text('OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[2];\ncreg c[2];\n'
                 "h q[0];\ncx q[0],q[1];\nmeasure q -> c;\n")
    assert run_main([str(f), "--shots", "4000", "--dtype", "c64"]) == 0
    out = capsys.readouterr().out.splitlines()
    assert "4000 shots, 2 distinct outcomes" in out[0]
    counts = {line.split()[0]: int(line.split()[-1]) for line in out[1:]}
    assert set(counts) == {"|00⟩", "|11⟩"} and sum(counts.values()) == 4000
    assert abs(counts["|00⟩"] - 2000) < 200
    bad = tmp_path / "bad.qasm"
    bad.write_text("OPENQASM 3;\nqubit q;\nfoo q;\n")
    assert run_main([str(bad)]) == 2
    assert "line 3" in capsys.readouterr().err
