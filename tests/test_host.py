"""Host logic on the CPU: circuit -> program compilation (the from_pyops restatement), gate
matrices, parameters and their gradients, count keys, SimuResult and Hamiltonian matrices.
The compiled programs are executed by the oracle here; the GPU runs them in test_gpu_*.py."""

import json
import math
import os

import numpy as np
import pytest

import pyquafu_amd as pq
from oracle import dense
from oracle.cpu import CpuState
from pyquafu_amd.algorithms.hamiltonian import Hamiltonian, PauliMats, PauliOp
from pyquafu_amd.elements import element_gates as g
from pyquafu_amd.elements.parameters import Parameter
from pyquafu_amd.results.results import SimuResult
from pyquafu_amd.simulators import program as P
from pyquafu_amd.simulators.program import compile_circuit, compile_ops
from pyquafu_amd.simulators.qfvm import counts_from_indices

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.npz")


def run_oracle(qc, psi=None):
    prog = compile_circuit(qc)
    s = CpuState(qc.num)
    if psi is not None:
        s.load(psi)
    s.run(prog, [], qc.cbits_num, skip_measure=True)
    return s.store()


def test_named_dispatch_and_positions():
    qc = pq.QuantumCircuit(4)
    qc.h(0)
    qc.cx(1, 2)
    qc.cnot(2, 3)
    qc.toffoli(0, 1, 3)
    qc.rz(2, 0.3)
    qc.rzz(0, 3, 0.4)
    qc.cp(3, 0, 0.5)
    qc.sdg(1)
    prog = compile_circuit(qc)
    assert prog.ops["code"].tolist() == [P.H, P.CX, P.CX, P.CCX, P.RZ, P.RZZ, P.CP, P.SDG]
    qs = [prog.idx[r["q_off"]: r["q_off"] + r["n_qubits"]].tolist() for r in prog.ops]
    assert qs == [[0], [1, 2], [2, 3], [0, 1, 3], [2], [0, 3], [3, 0], [1]]
    assert prog.ops["param"][4] == 0.3 and prog.ops["m_off"][0] == -1


def test_general_and_multi_matrices():
    qc = pq.QuantumCircuit(4)
    qc << g.U3Gate(1, 0.1, 0.2, 0.3)
    qc << g.CRZGate(3, 0, 0.7)
    qc << g.CRYGate(0, 2, 0.9)
    qc.swap(0, 2)
    qc << g.RXXGate(3, 1, 0.4)
    qc.fredkin(1, 0, 3)
    qc.sx(2)
    prog = compile_circuit(qc)
    assert prog.ops["code"].tolist() == [P.GENERAL, P.GENERAL, P.GENERAL, P.MULTI, P.MULTI, P.MULTI, P.GENERAL]
    assert prog.ops["n_ctrl"].tolist() == [0, 1, 1, 0, 0, 1, 0]

    def mat(k):
        r = prog.ops[k]
        d = 1 << (r["n_qubits"] - r["n_ctrl"])
        a = prog.mat[r["m_off"]: r["m_off"] + 2 * d * d]
        return (a[0::2] + 1j * a[1::2]).reshape(d, d)

    assert np.array_equal(mat(0), dense.u3(0.1, 0.2, 0.3)) or np.abs(mat(0) - dense.u3(0.1, 0.2, 0.3)).max() < 1e-16
    assert np.abs(mat(1) - dense.rz(0.7)).max() < 1e-16
    assert np.abs(mat(2) - dense.ry(0.9)).max() < 1e-16
    # CSWAP: the target matrix is handed over with reverse_order=True (LSB = first target); SWAP is
    # symmetric so it is unchanged
    assert np.array_equal(mat(5), pq.elements.matrices.SwapMatrix)


def test_skips_wrappers_and_final_measure_flag():
    inner = pq.QuantumCircuit(2)
    inner.h(0)
    inner.cx(0, 1)
    qc = pq.QuantumCircuit(3)
    qc.barrier()
    qc.id(0)
    qc.delay(1, 10)
    qc << inner.wrap([1, 2])
    qc.measure([0, 1])
    prog = compile_circuit(qc)
    assert prog.ops["code"].tolist() == [P.H, P.CX, P.MEASURE]
    assert prog.idx[prog.ops[1]["q_off"]: prog.ops[1]["q_off"] + 2].tolist() == [1, 2]
    assert prog.final_measure and prog.measure_vec == [(0, 0), (1, 1)]
    qc.x(2)  # a gate after a measure: per-shot mode
    assert not compile_circuit(qc).final_measure


def test_cif_body_compiled():
    qc = pq.QuantumCircuit(3, 3)
    qc.x(0)
    qc.measure([0], [0])
    with qc.cif([0], 1):
        qc.x(1)
        qc.h(2)
    qc.measure([1], [1])
    prog = compile_circuit(qc)
    codes = prog.ops["code"].tolist()
    assert codes == [P.X, P.MEASURE, P.CIF, P.X, P.H, P.MEASURE]
    assert prog.ops[2]["body"] == 2 and prog.ops[2]["cond"] == 1
    assert not prog.final_measure and prog.n_random == 2


def test_compiled_circuits_simulate_on_oracle():
    # Bell pair and a controlled rotation, executed by the oracle through the compiled program
    qc = pq.QuantumCircuit(2)
    qc.h(0)
    qc.cx(0, 1)
    assert np.allclose(run_oracle(qc), [2**-0.5, 0, 0, 2**-0.5], atol=1e-15)
    qc = pq.QuantumCircuit(2)
    qc.x(0)
    qc << g.CRYGate(0, 1, np.pi)
    assert np.allclose(run_oracle(qc), [0, 0, 0, 1], atol=1e-15)
    # CS carries the X matrix in the reference, so it acts as CX
    qc = pq.QuantumCircuit(2)
    qc.x(0)
    qc.cs(0, 1)
    assert np.allclose(run_oracle(qc), [0, 0, 0, 1], atol=1e-15)


def test_circuit_identities_on_oracle(rng):
    """building_circuit_test.py:47-172: power, dagger, wrap, control-wrap equivalences (|<a|b>| = 1)."""
    n = 4
    qc = pq.QuantumCircuit(n)
    qc << g.U3Gate(0, 0.3, 0.5, 0.7)
    qc << g.RYYGate(1, 2, 0.8)
    qc << g.RXXGate(2, 3, 0.6)
    qc << g.CRYGate(0, 3, 1.1)
    qc << g.CRZGate(2, 1, 0.4)
    psi = np.zeros(1 << n, complex)
    psi[0] = 1
    a = run_oracle(qc)
    back = run_oracle(qc.dagger(), a.copy())
    assert abs(abs(np.vdot(psi, back)) - 1) < 1e-12
    w = pq.QuantumCircuit(n)
    w << qc.wrap()
    assert abs(abs(np.vdot(a, run_oracle(w))) - 1) < 1e-12
    sq = run_oracle(qc.power(2))
    assert abs(abs(np.vdot(run_oracle(qc, a.copy()), sq)) - 1) < 1e-12


def test_parameter_autograd_reference_values():
    # varational_test.py:16-23
    a = Parameter("v1", 1.0)
    b = Parameter("v2", 2.0)
    c = -a * b + 2.0 * b.log()
    assert math.isclose(c.value, -0.6137056388801094)
    assert math.isclose(c.get_value(), -0.6137056388801094)
    assert math.isclose(c.grad()[0], -2.0)
    assert math.isclose(c.grad()[1], 0.0, abs_tol=1e-15)
    e = (a * 3.0).sin() / b + a**2
    x, y = 1.0, 2.0
    assert math.isclose(e.get_value(), math.sin(3 * x) / y + x * x)
    gr = e.grad()
    assert math.isclose(gr[0], 3 * math.cos(3 * x) / y + 2 * x)
    assert math.isclose(gr[1], -math.sin(3 * x) / y**2)


def test_parameter_grads_of_circuit():
    pq_ = pq.QuantumCircuit(4)
    theta = [Parameter(f"theta_{i}", i + 1) for i in range(4)]
    for i in range(4):
        pq_ << g.RXGate(i, theta[i])
    pq_ << g.RYGate(2, theta[0] * theta[1] - 3.0 * theta[0])
    grads = pq_.get_parameter_grads()
    assert pq_.variables[0] is theta[0]
    assert grads[theta[0]][1][1] == pytest.approx(theta[1].value - 3.0)
    assert grads[theta[1]][1][1] == pytest.approx(theta[0].value)
    pq_._update_params([0.1, 0.2, 0.3, 0.4])
    assert pq_.gates[4]._paras[0] == pytest.approx(0.1 * 0.2 - 0.3)


def test_count_keys_big_endian_by_cbit():
    # any_cbit_measure (basis_test.py:98-104): measures (1->1), (2->0), (3->2), (0->3) on |0011>
    idx = np.array([0b0011] * 5, dtype=np.uint64)
    keys = counts_from_indices(idx, [(1, 1), (2, 0), (3, 2), (0, 3)])
    assert keys == {0b0101: 5}
    assert SimuResult({"counts": keys, "measures": {1: 1, 2: 0, 3: 2, 0: 3}, "qbitnum": 4}).counts == {"0101": 5}
    with pytest.raises(IndexError):
        counts_from_indices(idx, [(0, 5)])
    # the 2^n sentinel has no bits on qubits < n
    assert counts_from_indices(np.array([16], np.uint64), [(0, 0), (3, 1)]) == {0: 1}


def test_simuresult_probabilities_bell():
    psi = np.array([2**-0.5, 0, 0, 2**-0.5], complex)
    r = SimuResult({"counts": {}, "measures": {}, "qbitnum": 2, "statevector": psi})
    assert np.allclose(r.probabilities, [0.5, 0, 0, 0.5])
    # measured subset in cbit order: qubit 1 -> cbit 0 only
    psi = np.zeros(8, complex)
    psi[0b110] = 1.0
    r = SimuResult({"counts": {}, "measures": {1: 0}, "qbitnum": 3, "statevector": psi})
    assert np.allclose(r.probabilities, [0, 1])


def test_hamiltonian_matrices_against_reference_fixtures():
    d = np.load(GOLDEN, allow_pickle=False)
    meta = json.loads(str(d["meta"]))
    for key, info in meta.items():
        h = Hamiltonian.from_pauli_list([tuple(x) for x in info["paulis"]])
        assert np.abs(h.get_matrix(info["n"]).toarray() - d[key]).max() < 1e-15, key
    # varational_test.py:54-66
    import scipy.sparse as sp

    op1 = PauliOp("X2 Z0")
    m2 = sp.kron(sp.kron(sp.kron(PauliMats["Z"], PauliMats["I"]), PauliMats["X"]), PauliMats["I"])
    assert np.linalg.norm(op1.get_matrix(4, True).toarray() - m2.toarray()) == 0.0


def test_compile_ops_for_applyop():
    prog = compile_ops([g.RXGate(1, 0.5), g.CZGate(0, 1)], 2)
    assert prog.ops["code"].tolist() == [P.RX, P.CZ]


def test_parametric_program_bind_equals_recompile():
    """Rebinding writes the same op stream and matrix pool as compiling the circuit again with the
    new values: named rotations of one variable through the vectorised slot map, expression angles
    and matrix gates (U3, controlled rotations) re-derived from their gates."""
    from pyquafu_amd.circuits import QuantumCircuit
    from pyquafu_amd.elements import CRZGate, CZGate, RXGate, RYGate, U3Gate
    from pyquafu_amd.elements.parameters import Parameter
    from pyquafu_amd.simulators.binding import ParametricProgram
    from pyquafu_amd.simulators.program import compile_gates

    a, b, c = Parameter("a", 0.1), Parameter("b", 0.2), Parameter("c", -0.3)
    qc = QuantumCircuit(4)
    qc << U3Gate(0, a, b, 0.3)
    qc << CRZGate(0, 1, a * 2)
    qc << RYGate(2, b)
    qc << RXGate(3, c)
    qc << CZGate(1, 2)
    qc << RYGate(1, c)
    pp = ParametricProgram(qc)
    assert [v.name for v in pp.variables] == ["a", "b", "c"]
    assert len(pp.fast_rows) == 3 and len(pp.slow) == 2
    rng = np.random.default_rng(8)
    for _ in range(3):
        vals = rng.uniform(-np.pi, np.pi, size=3)
        pp.bind(vals)
        want, _ = compile_gates(qc.gates, 4)
        assert np.array_equal(pp.prog.ops, want.ops)
        assert np.array_equal(pp.prog.mat, want.mat)
        assert [v.value for v in (a, b, c)] == list(vals)
    with pytest.raises(ValueError):
        pp.bind([0.0, 1.0])




def test_bound_adjoint_program_rebinds_and_recompiles():
    """grad_adjoint's compiled-program cache: the same gate list with new variable values rebinds
    the cached program (same arrays, new values); a changed gate list compiles afresh."""
    from pyquafu_amd import workloads
    from pyquafu_amd.algorithms.gradients.gradients import adjoint_program, bound_adjoint_program
    from pyquafu_amd.elements import RYGate
    from pyquafu_amd.elements.parameters import Parameter

    qc, _ = workloads.hea_ansatz(6, 2, parametrised=True)
    p1, ops1, gens1, diff1 = bound_adjoint_program(qc)
    qc._update_params(list(np.linspace(-1, 1, len(qc.variables))))  # pylint: disable=protected-access
    p2, ops2, gens2, diff2 = bound_adjoint_program(qc)
    assert p2 is p1 and (ops2, gens2, diff2) == (ops1, gens1, diff1)
    want = adjoint_program(qc)
    assert np.array_equal(p2.ops, want[0].ops) and (ops2, gens2, diff2) == tuple(want[1:])
    qc << RYGate(0, Parameter("extra", 0.25))
    p3, ops3, _, _ = bound_adjoint_program(qc)
    assert p3 is not p1 and len(ops3) == len(ops1) + 1
    assert np.array_equal(p3.ops, adjoint_program(qc)[0].ops)
