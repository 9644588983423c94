"""OpenQASM 2.0 ingress on the CPU: expressions, registers, broadcasting, gate definitions, the
standard-library decompositions (checked by dense simulation against the gates they implement), the
classical statements, and the errors of quafu/qfasm (tests/quafu/qasm/parser_test.py)."""

import math

import numpy as np
import pytest

import pyquafu_amd as pq
from oracle import dense
from pyquafu_amd.qfasm import LexerError, ParserError, qasm_to_quafu
from pyquafu_amd.simulators.program import compile_circuit

# ------------------------------------------------------------------ expressions


def rz_param(text):
    qc = qasm_to_quafu(f"qreg q[1]; U(0, 0, {text}) q[0];")
    assert qc.gates[0].name == "RZ"
    return qc.gates[0].paras[0]


@pytest.mark.parametrize("text,want", [
    ("0.25 + 1 + 1", 2.25), ("3 - 0.75", 2.25), ("2 * 0.4", 0.8), ("0.9 / 2", 0.45), ("2^3", 8),
    ("2.0-1.0-1.0", 0.0), ("2.0/2.0/2.0", 0.5), ("2^3^2", 2 ** 9), ("-2^2", 4), ("2*(3-1)", 4),
    ("0.5 + 1.5 * -0.3 ^ 2 - 0.2 / 0.5", 0.5 + 1.5 * (-0.3) ** 2 - 0.2 / 0.5), ("pi/4", math.pi / 4),
    ("1e-3 * 2", 2e-3), (".5", 0.5)])
def test_expressions(text, want):
    assert math.isclose(rz_param(text), want, abs_tol=1e-15)


@pytest.mark.parametrize("fn,f", [("sin", math.sin), ("cos", math.cos), ("tan", math.tan), ("exp", math.exp),
                                  ("ln", math.log), ("sqrt", math.sqrt), ("asin", math.asin),
                                  ("acos", math.acos), ("atan", math.atan)])
def test_functions(fn, f):
    assert math.isclose(rz_param(f"{fn}(0.5)"), f(0.5))


def test_u_expands_to_rz_ry_rz():
    qc = qasm_to_quafu("qreg q[1]; U(0.1, 0.2, 0.3) q[0];")
    assert [(g.name, g.paras[0]) for g in qc.gates] == [("RZ", 0.3), ("RY", 0.1), ("RZ", 0.2)]


def test_symbolic_gate_parameters():
    qasm = """
    gate test(a,b,c) q { U(a+b, -(a-c), a+c-b) q; }
    qreg q[1]; test(0.5, 1.0, 2.0) q[0];
    """
    qc = qasm_to_quafu(qasm)
    assert [g.paras[0] for g in qc.gates] == [0.5 + 2.0 - 1.0, 0.5 + 1.0, -(0.5 - 2.0)]


# ------------------------------------------------------------------ registers and statements


def test_registers_flatten_in_declaration_order():
    qc = qasm_to_quafu("qreg a[2]; creg c[2]; qreg b[3]; creg d[1]; x b[1]; measure b[2] -> d[0]; measure a -> c;")
    assert qc.num == 5 and qc.cbits_num == 3
    assert qc.gates[0].pos == [3]
    assert qc.measures == {4: 2, 0: 0, 1: 1}


def test_broadcasting():
    qc = qasm_to_quafu("qreg a[3]; qreg b[3]; qreg c[1]; h a; cx a, b; cx c[0], b;")
    names = [(g.name, g.pos) for g in qc.gates]
    assert names[:3] == [("H", [0]), ("H", [1]), ("H", [2])]
    assert names[3:6] == [("CX", [0, 3]), ("CX", [1, 4]), ("CX", [2, 5])]
    assert names[6:] == [("CX", [6, 3]), ("CX", [6, 4]), ("CX", [6, 5])]


def test_classical_statements():
    qasm = """OPENQASM 2.0;
    include "qelib1.inc";
    qreg q[2]; creg c[2];
    h q[0]; barrier q; measure q[0] -> c[0];
    if (c == 1) x q[1];
    reset q;
    """
    qc = qasm_to_quafu(qasm)
    kinds = [type(op).__name__ for op in qc.instructions]
    assert kinds == ["HGate", "Barrier", "Measure", "Cif", "Reset"]
    cif = qc.instructions[3]
    assert cif.cbits == [0, 1] and cif.condition == 1 and cif.instructions[0].name == "X"
    assert not qc.executable_on_backend


def test_native_gates_and_multi_control():
    qasm = """qreg q[4];
    rxx(0.3) q[0], q[1]; cp(0.2) q[1], q[2]; ccx q[0], q[1], q[2]; mcx q[0], q[1], q[2], q[3];
    cswap q[0], q[1], q[2]; sw q[3]; delay(100) q[2];"""
    qc = qasm_to_quafu(qasm)
    got = [(g.name, g.pos) for g in qc.gates]
    assert got[0] == ("RXX", [0, 1]) and got[1] == ("CP", [1, 2])
    assert got[3] == ("MCX", [0, 1, 2, 3])
    assert qc.gates[3].ctrls == [0, 1, 2]


def test_roundtrip_through_to_openqasm():
    qc = pq.QuantumCircuit(3)
    qc.h(0)
    qc.cx(0, 1)
    qc.rx(2, 0.5)
    qc.cp(1, 2, 0.25)
    qc.measure([0, 1, 2])
    back = pq.QuantumCircuit(0).from_openqasm(qc.to_openqasm())
    # the text has one measure line per qubit, so compare the gates and the measure map
    a, b = compile_circuit(qc), compile_circuit(back)
    ga, gb = a.ops[a.ops["code"] != dense.MEASURE], b.ops[b.ops["code"] != dense.MEASURE]
    assert np.array_equal(ga[["code", "n_ctrl", "n_qubits", "param"]], gb[["code", "n_ctrl", "n_qubits", "param"]])
    assert [list(a.idx[o["q_off"]: o["q_off"] + o["n_qubits"]]) for o in ga] == \
        [list(b.idx[o["q_off"]: o["q_off"] + o["n_qubits"]]) for o in gb]
    assert back.measures == qc.measures and back.num == qc.num


# ------------------------------------------------------------------ library decompositions


def unitary_of(qasm_body: str, n: int) -> np.ndarray:
    """Dense unitary of a QASM program over n qubits (column k = the circuit applied to |k>)."""
    prog = compile_circuit(qasm_to_quafu(f"qreg q[{n}]; " + qasm_body))
    cols = []
    for k in range(1 << n):
        psi = np.zeros(1 << n, dtype=complex)
        psi[k] = 1
        cols.append(dense.run_dense(prog, n, psi=psi)[0])
    return np.array(cols).T


def controlled(u: np.ndarray, n: int, ctrls, targ) -> np.ndarray:
    m = np.eye(1 << n, dtype=complex)
    for k in range(1 << n):
        if all((k >> c) & 1 for c in ctrls):
            b = (k >> targ) & 1
            k0 = k & ~(1 << targ)
            m[k0 | (0 << targ), k] = u[0, b]
            m[k0 | (1 << targ), k] = u[1, b]
            if b == 0:
                m[k0 | (1 << targ), k] = u[1, 0]
                m[k0, k] = u[0, 0]
    return m


def same_up_to_phase(a, b, tol=1e-13):
    k = np.unravel_index(np.argmax(np.abs(b)), b.shape)
    ph = a[k] / b[k]
    return abs(abs(ph) - 1) < tol and np.allclose(a, ph * b, atol=tol)


def test_u3_is_the_u3_matrix_times_its_phase():
    t, p, lam = 0.7, -0.4, 1.3
    got = unitary_of(f"u3({t}, {p}, {lam}) q[0];", 1)
    assert np.allclose(got, np.exp(-0.5j * (p + lam)) * dense.u3(t, p, lam), atol=1e-14)


@pytest.mark.parametrize("name,u", [
    ("crz(0.9)", dense.rz(0.9)), ("cry(0.9)", dense.ry(0.9)), ("cu1(0.9)", dense.pm(0.9)),
    ("ch", dense.HM), ("csx", np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]]) / 2)])
def test_controlled_library_gates(name, u):
    """The decompositions implement the controlled gates up to a global phase (u1 reaches the
    engine as an rz, so gates built from it carry one)."""
    got = unitary_of(f"{name} q[0], q[1];", 2)
    assert same_up_to_phase(got, controlled(u, 2, [0], 1))


def test_cu1_global_phase_is_kept():
    """cu1(l) = u1(l/2) a; cx; u1(-l/2) b; cx; u1(l/2) b with u1 = rz: e^{-i l/4} CP(l) exactly."""
    lam = 0.9
    got = unitary_of(f"cu1({lam}) q[0], q[1];", 2)
    assert np.allclose(got, np.exp(-0.25j * lam) * controlled(dense.pm(lam), 2, [0], 1), atol=1e-14)


def test_crx_up_to_its_phase():
    got = unitary_of("crx(0.9) q[0], q[1];", 2)
    assert same_up_to_phase(got, controlled(dense.rx(0.9), 2, [0], 1))


def test_cu3_controlled_u3():
    t, p, lam = 0.7, -0.4, 1.3
    got = unitary_of(f"cu3({t}, {p}, {lam}) q[0], q[1];", 2)
    assert same_up_to_phase(got, controlled(dense.u3(t, p, lam), 2, [0], 1))


@pytest.mark.parametrize("name,n", [("c3x", 4), ("c4x", 5)])
def test_multi_controlled_x(name, n):
    args = ", ".join(f"q[{k}]" for k in range(n))
    got = unitary_of(f"{name} {args};", n)
    assert same_up_to_phase(got, controlled(dense.XM, n, list(range(n - 1)), n - 1), tol=1e-12)


def test_user_gate_nesting():
    qasm = """
    gate bell a, b { h a; cx a, b; }
    gate twice(t) a, b { bell a, b; rz(t) b; bell a, b; }
    twice(0.3) q[0], q[1];
    """
    got = unitary_of(qasm, 2)
    bell = unitary_of("h q[0]; cx q[0], q[1];", 2)
    rz = np.kron(dense.rz(0.3), np.eye(2))
    assert np.allclose(got, bell @ rz @ bell, atol=1e-14)


# ------------------------------------------------------------------ errors


@pytest.mark.parametrize("text,err,match", [
    ("qreg Qcav[1];", LexerError, "Illegal character"),
    ('include "qe\nlib1.inc";', LexerError, "Invalid include"),
    ('include "qelib1.inc" qreg q[1];', LexerError, 'Expecting ";" for INCLUDE'),
    ("include qelib1.inc;", LexerError, "need a quoted string"),
    ('include "qe.inc";', LexerError, "cannot be found"),
    ("include 'qe.inc';", LexerError, "Illegal character"),
    ("!", LexerError, "Illegal character"),
    ("if (a=2) U(0,0,0)q[0];", ParserError, "Illegal IF statement"),
    ("qreg q[1]; foo q[0];", ParserError, "undefined"),
    ("qreg q[1]; h q[1];", ParserError, "out of bounds"),
    ("qreg q[2]; cx q[0], q[0];", ParserError, "different argument"),
    ("qreg q[2]; qreg q[1];", ParserError, "Duplicate"),
    ("qreg q[2]; creg c[1]; measure q -> c;", ParserError, "doesn't match"),
    ("qreg q[1]; rx q[0];", ParserError, "parameters"),
    ("gate g a { g a; } qreg q[1];", ParserError, "calls itself"),
    ("qreg a[2]; qreg b[3]; cx a, b;", ParserError, "inconsistent"),
    ("opaque g a; qreg q[1]; g q[0];", ParserError, "opaque"),
    ("qreg q[1]; U(0, 0, theta) q[0];", ParserError, "undefined"),
])
def test_errors(text, err, match):
    with pytest.raises(err, match=match):
        qasm_to_quafu(text)


def test_named_parameter_declaration():
    """`name = value;` declares a Parameter; expressions of it stay symbolic on the gate and
    evaluate when the circuit is compiled."""
    qc = qasm_to_quafu("theta = 0.5; qreg q[1]; rx(theta) q[0]; rz(theta * 2) q[0];")
    assert qc.gates[0].paras[0].get_value() == 0.5
    assert list(compile_circuit(qc).ops["param"]) == [0.5, 1.0]
