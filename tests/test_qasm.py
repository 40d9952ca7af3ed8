"""The OpenQASM reader (qsvh_circuit_from_qasm) on the CPU: the front end the reference's CLI
names for a later milestone (crates/qsv-cli/src/main.rs:1-2, README.md:63).

Pins: circuits written in QASM reproduce the engine's own generators op for op and bit for
bit (QFT, circuits.rs:61-74, in OpenQASM 3 and 2.0 syntax); every standard gate equals an
independent numpy construction (controlled gates as sums of Kronecker products with
projectors on the control, bit 0 = first operand per gate.rs:8-12; RXX/RYY through
scipy's expm); modifiers, definitions, broadcast, registers and the rejected forms."""
import numpy as np
import pytest
from scipy.linalg import expm

import rustatevec_amd as rs

I2 = np.eye(2)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]])
Z = np.diag([1.0, -1.0]).astype(complex)
P0 = np.diag([1.0, 0.0])
P1 = np.diag([0.0, 1.0])


def ctrl(u, k=1, neg=False):
    """k controls on the low index bits, u on the high bits."""
    d = 1 << k
    on = np.zeros((d, d))
    on[0 if neg else d - 1, 0 if neg else d - 1] = 1.0
    return np.kron(np.eye(u.shape[0]), np.eye(d) - on) + np.kron(u, on)


def ops(text):
    return rs.from_qasm(text).ops()


def one(text):
    o = ops(text)
    assert len(o) == 1, o
    return o[0]


HDR = 'OPENQASM 3.0;\ninclude "stdgates.inc";\n'


def qft_qasm3(n):
    lines = [HDR, f"qubit[{n}] q;"]
    for j in range(n):
        lines.append(f"h q[{j}];")
        for k in range(j + 1, n):
            lines.append(f"cp(pi / 2**{k - j}) q[{k}], q[{j}];")
    for j in range(n // 2):
        lines.append(f"swap q[{j}], q[{n - 1 - j}];")
    return "\n".join(lines)


def qft_qasm2(n):
    lines = ['OPENQASM 2.0;', 'include "qelib1.inc";', f"qreg q[{n}];", f"creg c[{n}];"]
    for j in range(n):
        lines.append(f"h q[{j}];")
        for k in range(j + 1, n):
            lines.append(f"cu1(pi/2^{k - j}) q[{k}],q[{j}];")
    for j in range(n // 2):
        lines.append(f"swap q[{j}],q[{n - 1 - j}];")
    lines.append("barrier q;")
    lines.append("measure q -> c;")
    return "\n".join(lines)


def assert_same_ops(got, want):
    assert len(got) == len(want)
    for (gm, gq), (wm, wq) in zip(got, want):
        assert gq == wq
        assert np.array_equal(gm, wm)


@pytest.mark.parametrize("src", [qft_qasm3, qft_qasm2])
def test_qft_text_reproduces_the_generator_bitwise(src):
    for n in (1, 2, 5, 9):
        c = rs.from_qasm(src(n))
        assert c.n_qubits == n
        assert_same_ops(c.ops(), rs.qft(n).ops())


def test_library_gates_are_the_gate_library():
    for name, params, qs in (("h", (), "q[0]"), ("x", (), "q[0]"), ("y", (), "q[0]"),
                             ("z", (), "q[0]"), ("s", (), "q[0]"), ("t", (), "q[0]"),
                             ("sx", (), "q[0]"), ("rx", (0.3,), "q[0]"),
                             ("ry", (-1.2,), "q[0]"), ("rz", (2.5,), "q[0]"),
                             ("p", (0.7,), "q[0]"), ("u3", (0.1, 0.2, 0.3), "q[0]"),
                             ("cx", (), "q[0], q[1]"), ("cz", (), "q[0], q[1]"),
                             ("cp", (0.4,), "q[0], q[1]"), ("swap", (), "q[0], q[1]"),
                             ("rzz", (0.9,), "q[0], q[1]"), ("ccx", (), "q[0], q[1], q[2]"),
                             ("ccz", (), "q[0], q[1], q[2]")):
        ps = "(" + ", ".join(repr(p) for p in params) + ")" if params else ""
        m, _ = one(HDR + f"qubit[3] q;\n{name}{ps} {qs};")
        assert np.array_equal(m, rs.gate_matrix(name, *params)), name
    # aliases
    assert np.array_equal(one(HDR + "qubit[1] q; U(0.1, 0.2, 0.3) q[0];")[0],
                          rs.gate_matrix("u3", 0.1, 0.2, 0.3))
    assert np.array_equal(one(HDR + "qubit[2] q; CX q[0], q[1];")[0], rs.gate_matrix("cx"))
    assert np.array_equal(one(HDR + "qubit[1] q; u1(0.25) q[0];")[0], rs.gate_matrix("p", 0.25))
    assert np.array_equal(one(HDR + "qubit[1] q; u2(0.5, -0.4) q[0];")[0],
                          rs.gate_matrix("u3", np.pi / 2, 0.5, -0.4))
    assert np.array_equal(one(HDR + "qubit[2] q; cphase(0.3) q[0], q[1];")[0],
                          rs.gate_matrix("cp", 0.3))


def test_other_standard_gates_match_independent_constructions():
    th, ph, la, ga = 0.37, -1.1, 2.3, 0.6
    u = rs.gate_matrix("u3", th, ph, la)
    cases = {
        "sdg q[0];": np.diag([1, -1j]),
        "tdg q[0];": np.diag([1, np.exp(-1j * np.pi / 4)]),
        "sxdg q[0];": rs.gate_matrix("sx").conj().T,
        f"cy q[0], q[1];": ctrl(Y),
        f"ch q[0], q[1];": ctrl(rs.gate_matrix("h")),
        f"csx q[0], q[1];": ctrl(rs.gate_matrix("sx")),
        f"crx({th}) q[0], q[1];": ctrl(expm(-0.5j * th * X)),
        f"cry({th}) q[0], q[1];": ctrl(expm(-0.5j * th * Y)),
        f"crz({th}) q[0], q[1];": ctrl(expm(-0.5j * th * Z)),
        f"cu3({th}, {ph}, {la}) q[0], q[1];": ctrl(u),
        f"cu({th}, {ph}, {la}, {ga}) q[0], q[1];": ctrl(np.exp(1j * ga) * u),
        f"rxx({th}) q[0], q[1];": expm(-0.5j * th * np.kron(X, X)),
        f"ryy({th}) q[0], q[1];": expm(-0.5j * th * np.kron(Y, Y)),
        "cswap q[0], q[1], q[2];": ctrl(rs.gate_matrix("swap")),
        "c3x q[0], q[1], q[2], q[3];": ctrl(X, 3),
        "c4x q[0], q[1], q[2], q[3], q[4];": ctrl(X, 4),
    }
    for stmt, want in cases.items():
        m, _ = one(HDR + "qubit[5] q;\n" + stmt)
        assert np.allclose(m, want, atol=1e-15, rtol=0), stmt


def test_operand_order_and_registers():
    c = rs.from_qasm(HDR + "qubit[2] a;\nqubit[3] b;\ncx b[2], a[1];\nccx a[0], b[0], b[1];")
    assert c.n_qubits == 5
    (m0, q0), (m1, q1) = c.ops()
    assert q0 == (4, 1) and np.array_equal(m0, rs.gate_matrix("cx"))
    assert q1 == (0, 2, 3)
    # a single qubit declaration is a register of one
    assert one(HDR + "qubit a; qubit b; cx a, b;")[1] == (0, 1)


def test_broadcast_and_ranges():
    o = ops(HDR + "qubit[4] q; h q;")
    assert [q for _, q in o] == [(0,), (1,), (2,), (3,)]
    o = ops(HDR + "qubit[3] a; qubit[3] b; cx a, b;")
    assert [q for _, q in o] == [(0, 3), (1, 4), (2, 5)]
    o = ops(HDR + "qubit[3] a; qubit[1] t; cx a, t[0];")
    assert [q for _, q in o] == [(0, 3), (1, 3), (2, 3)]
    o = ops(HDR + "qubit[6] q; x q[1:3]; z q[0:2:4];")
    assert [q for _, q in o] == [(1,), (2,), (3,), (0,), (2,), (4,)]


def test_modifiers():
    base = HDR + "qubit[4] q;\n"
    # inv of a rotation is the rotation by -theta, bit for bit (sin is odd)
    assert np.array_equal(one(base + "inv @ rx(0.3) q[0];")[0], rs.gate_matrix("rx", -0.3))
    assert np.array_equal(one(base + "inv @ s q[0];")[0], np.diag([1, -1j]))
    assert np.array_equal(one(base + "pow(-1) @ t q[0];")[0], one(base + "tdg q[0];")[0])
    assert len(ops(base + "pow(3) @ t q[0];")) == 3
    assert len(ops(base + "pow(0) @ h q[0];")) == 0
    assert np.array_equal(one(base + "ctrl @ x q[0], q[1];")[0], rs.gate_matrix("cx"))
    m, q = one(base + "ctrl(2) @ z q[2], q[0], q[1];")
    assert q == (2, 0, 1) and np.array_equal(m, rs.gate_matrix("ccz"))
    m, q = one(base + "ctrl @ ctrl @ x q[0], q[1], q[2];")
    assert np.array_equal(m, rs.gate_matrix("ccx"))
    assert np.allclose(one(base + "negctrl @ x q[0], q[1];")[0], ctrl(X, neg=True), atol=0)
    m, q = one(base + "ctrl @ negctrl @ h q[0], q[1], q[2];")
    # control 0 must be 1 and control 1 must be 0: index bits (b0, b1) = (1, 0) -> 1
    sel = np.zeros((4, 4))
    sel[1, 1] = 1
    want = np.kron(np.eye(2), np.eye(4) - sel) + np.kron(rs.gate_matrix("h"), sel)
    assert q == (0, 1, 2) and np.allclose(m, want, atol=1e-16)
    m, q = one(base + "inv @ ctrl @ rz(0.5) q[3], q[1];")
    assert q == (3, 1) and np.allclose(m, ctrl(rs.gate_matrix("rz", -0.5)), atol=0)


def test_gate_definitions_and_gphase():
    src = HDR + """
    qubit[3] q;
    gate bell a, b { h a; cx a, b; }
    gate rot(t, s) a { rz(t / 2) a; rx(s * t) a; }
    gate twice(t) a, b { rot(t, 2) a; bell b, a; }
    bell q[0], q[2];
    twice(0.8) q[1], q[0];
    inv @ bell q[0], q[1];
    ctrl @ bell q[2], q[0], q[1];
    """
    o = ops(src)
    got = [(q, m) for m, q in o]
    assert [q for q, _ in got] == [(0,), (0, 2), (1,), (1,), (0,), (0, 1), (0, 1), (0,),
                                   (2, 0), (2, 0, 1)]
    assert np.array_equal(got[2][1], rs.gate_matrix("rz", 0.4))
    assert np.array_equal(got[3][1], rs.gate_matrix("rx", 1.6))
    assert np.allclose(got[6][1], rs.gate_matrix("cx").conj().T, atol=0)  # inv: reversed
    assert np.allclose(got[7][1], rs.gate_matrix("h").conj().T, atol=0)
    assert np.allclose(got[8][1], ctrl(rs.gate_matrix("h")), atol=0)
    assert np.allclose(got[9][1], ctrl(rs.gate_matrix("cx")), atol=0)
    # gphase: a global phase on the whole state; under ctrl, a phase on the control
    m, q = one(HDR + "qubit[2] q; gphase(0.5);")
    assert q == (0,) and np.allclose(m, np.exp(0.5j) * np.eye(2), atol=1e-16)
    m, q = one(HDR + "qubit[2] q; ctrl @ gphase(0.5) q[1];")
    assert q == (1,) and np.allclose(m, np.diag([1, np.exp(0.5j)]), atol=1e-16)


def test_expressions_comments_and_ignored_statements():
    src = """// header comment
    OPENQASM 3;
    include "stdgates.inc";
    /* block
       comment */
    qubit[2] q; bit[2] c;
    rz(-pi/4 + 2*tau - tau*2) q[0];   // -pi/4
    ry(sqrt(2) * cos(0) - 1.5e-1 + arctan(1) * 0) q[1];
    p(π / 2) q[0];
    rx(2 ** 3 ** 0 * euler - ln(euler) * 0) q[1];
    barrier q;
    c[0] = measure q[0];
    c = measure q;
    """
    o = ops(src)
    assert len(o) == 4
    assert np.array_equal(o[0][0], rs.gate_matrix("rz", -np.pi / 4 + 2 * (2 * np.pi) - (2 * np.pi) * 2))
    assert np.allclose(o[1][0], rs.gate_matrix("ry", np.sqrt(2) - 0.15), atol=1e-16)
    assert np.array_equal(o[2][0], rs.gate_matrix("p", np.pi / 2))
    assert np.allclose(o[3][0], rs.gate_matrix("rx", 2 * np.e), atol=1e-16)


@pytest.mark.parametrize("body,line,what", [
    ("qubit[2] q;\nfoo q[0];", 4, "unknown gate"),
    ("qubit[2] q;\nrx q[0];", 4, "parameters"),
    ("qubit[2] q;\ncx q[0];", 4, "qubits"),
    ("qubit[2] q;\nh q[2];", 4, "out of range"),
    ("qubit[2] q;\ncx q[1], q[1];", 4, "twice"),
    ("qubit[2] q;\nreset q[0];", 4, "reset"),
    ("qubit[2] q;\nbit c;\nif (c) x q[0];", 5, "unsupported"),
    ("qubit[2] q;\nbit c;\nc = measure q[0];\nx q[0];", 6, "measured"),
    ('include "mygates.inc";', 3, "include"),
    ("qubit[2] q;\n/* open", 4, "unterminated"),
    ("qubit[2] q;\npow(0.5) @ x q[0];", 4, "integer"),
    ("qubit[7] q;\nctrl(6) @ x q[0], q[1], q[2], q[3], q[4], q[5], q[6];", 4, "6 qubits"),
    ("qubit[2] q;\ngate g a { g a; }\ng q[0];", 4, "nested"),  # the recursive call
    ("qubit[2] a; qubit[3] b;\ncx a, b;", 4, "differ"),
])
def test_rejected_programs_name_their_line(body, line, what):
    with pytest.raises(rs.QsvhError) as e:
        rs.from_qasm(HDR + body)
    assert e.value.status == "PARSE"
    assert f"line {line}:" in str(e.value) and what in str(e.value), str(e.value)
