"""OpenQASM text through simulate() on the GPU (simulators/__init__.py:52-57): the same results as
the parsed circuit, and the oracle's state."""

import numpy as np
import pytest

import pyquafu_amd as pq
from oracle import dense
from pyquafu_amd.qfasm import qasm_to_quafu
from pyquafu_amd.simulators.program import compile_circuit

pytestmark = pytest.mark.gpu

PROGRAM = """OPENQASM 2.0;
include "qelib1.inc";
qreg a[3];
qreg b[4];
creg c[3];
creg d[4];
gate layer(t) x, y { u3(t, t/2, -t) x; crz(2*t) x, y; cu1(t^2) y, x; }
h a; h b[0];
layer(0.3) a[0], b[1]; layer(-1.1) a[2], b[3];
ccx a[0], a[1], b[2]; c3x a[0], a[1], a[2], b[0]; cswap b[0], b[1], b[2];
u2(0.4, 0.9) b; rxx(0.2) a[1], b[2]; cu3(0.5, 0.1, -0.7) b[3], a[1];
measure a -> c;
measure b -> d;
"""


def test_qasm_string_matches_circuit_and_oracle():
    res_text = pq.simulate(PROGRAM, shots=2000, seed=7)
    qc = qasm_to_quafu(PROGRAM)
    res_qc = pq.simulate(qc, shots=2000, seed=7)
    assert res_text["counts"] == res_qc["counts"]
    psi = res_text["statevector"]
    assert np.abs(psi - res_qc["statevector"]).max() == 0.0
    prog = compile_circuit(qc)
    want, _ = dense.run_dense(prog, qc.num, skip_measure=True)
    assert np.abs(psi - want).max() < 1e-12
    assert sum(res_text["counts"].values()) == 2000 and all(len(k) == 7 for k in res_text["counts"])


def test_qasm_dynamic_program():
    """measure, if and reset from text: the per-shot path (q1 copies q0's outcome, then q0 is
    reset). The register value of `if` reads its cbits big-endian, first cbit most significant
    (check_cif, statevector.hpp:186-193), so c[0] = 1 alone is c == 2."""
    text = """qreg q[2]; creg c[2];
    h q[0]; measure q[0] -> c[0];
    if (c == 2) x q[1];
    reset q[0];
    measure q[1] -> c[1];
    """
    res = pq.simulate(text, shots=400, seed=3)
    assert set(res["counts"]) <= {"00", "11"}
    assert 120 < res["counts"].get("11", 0) < 280
