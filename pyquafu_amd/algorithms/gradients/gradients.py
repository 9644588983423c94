"""Gradients of <H> for parameterised circuits (quafu/algorithms/gradients/gradiant.py:24-173).

grad_adjoint runs the whole reverse-mode sweep on the device in one engine call (qfvm_adjoint):
the forward state, lambda = H psi and the reverse sweep over both states with the gradient markers
fused into the same LDS-tiled passes. The reference instead loops over gates in Python with one
engine round trip, a state copy and a host inner product per gate (gradiant.py:160-172); the
values are the same (d<H>/d theta = 2 Re <lambda| dU |phi>, grad_gate's dU = -i/2 P U on the
control-satisfied subspace, gradiant.py:96-141), and the chain rule through ParameterExpression
is assembled on the host exactly as assemble_grads does (gradiant.py:24-35).

grad_para_shift and grad_finit_diff keep the reference's algorithms (gradiant.py:38-93): one full
simulation per shifted value through `backend.run`.
"""

from __future__ import annotations

from typing import List

import numpy as np

from ... import _native
from ...elements.parameters import Parameter, ParameterExpression
from ...elements.quantum_gate import ControlledGate
from ...exceptions import CircuitError
from ...simulators.program import compile_gates
from ...simulators.qfvm import _call
from ...simulators.simulator import SVSimulator

_GENERATORS = {"RX": "X", "RY": "Y", "RZ": "Z"}


def assemble_grads(para_grads, gate_grads) -> List[float]:
    """Chain rule: d/d var = sum over (gate, parameter slot) uses of d<H>/d slot * d slot/d var."""
    grads = []
    for var in para_grads:
        full = 0.0
        for (i, j), gp in para_grads[var]:
            full += gate_grads[i][j] * gp
        grads.append(full)
    return grads


def _is_symbolic(op) -> bool:
    return len(op.paras) > 0 and isinstance(op.paras[0], (Parameter, ParameterExpression))


def generator_of(op) -> str:
    """Pauli generator of a differentiable rotation (grad_gate supports RX, RY, RZ and their
    controlled forms, gradiant.py:96-141; anything else raises NotImplementedError)."""
    name = op._targ_name if isinstance(op, ControlledGate) else op.name  # pylint: disable=protected-access
    gen = _GENERATORS.get(str(name).upper())
    if gen is None or (isinstance(op, ControlledGate) and len(op.targs) != 1):
        raise NotImplementedError(f"no adjoint gradient for gate {op.name}")
    return gen


def hamiltonian_terms(hamiltonian):
    """([(paulistr, pos), ...], coefficients) of a Hamiltonian's PauliOps."""
    paulis = hamiltonian.paulis
    return [(p.paulistr, list(p.pos)) for p in paulis], [complex(p.coeff) for p in paulis]


def adjoint_program(qc):
    """(program, op rows of the differentiated gates, generator letters, their gate indices)."""
    gates = qc.gates
    prog, rows = compile_gates(gates, qc.num)
    diff = [i for i, op in enumerate(gates) if _is_symbolic(op)]
    return prog, [rows[i] for i in diff], "".join(generator_of(gates[i]) for i in diff), diff


def _gate_stamp(gates):
    """Identity of a gate list for the compiled-program cache: the gate objects and their numeric
    parameters (symbolic ones are rebound on every call)."""
    return tuple((id(g), tuple(p if isinstance(p, (int, float)) else id(p) for p in g.paras)) for g in gates)


def bound_adjoint_program(qc):
    """adjoint_program with the circuit compiled once: the ParametricProgram is kept on the circuit
    while its gate list (gate objects, numeric parameters) stays the same, and each call only
    rebinds the current variable values into it (training loops re-running one circuit with new
    parameters). Returns (program, op rows, generators, gate indices)."""
    from ...simulators.binding import ParametricProgram  # pylint: disable=import-outside-toplevel

    stamp = _gate_stamp(qc.gates)
    cached = getattr(qc, "_pq_bound", None)
    if cached is None or cached[0] != stamp:
        pp = ParametricProgram(qc)
        diff = [i for i, op in enumerate(pp.gates) if _is_symbolic(op)]
        cached = (stamp, pp, [pp.rows[i] for i in diff], "".join(generator_of(pp.gates[i]) for i in diff), diff)
        qc._pq_bound = cached  # pylint: disable=protected-access
    else:
        pp = cached[1]
        pp.bind([v.value for v in pp.variables])
    return cached[1].prog, cached[2], cached[3], cached[4]


def grad_adjoint(qc, hamiltonian, psi_in=np.array([], dtype=complex), *, precision: str = "c128",
                 return_expectation: bool = False, checkpoints: int = 0):
    """Reverse-mode gradient (arXiv:2009.02823) of <H> with respect to qc.variables' order of
    `_calc_parameter_grads`. With return_expectation, returns (grads, <H>). `checkpoints` > 0 keeps
    that many copies of the forward state on the device (fewer when device memory is short) and
    restores them during the reverse sweep instead of uncomputing through them."""
    if not qc._parameter_grads:  # pylint: disable=protected-access
        qc.get_parameter_grads()
    para_grads = qc._calc_parameter_grads()  # pylint: disable=protected-access
    prog, ops, gens, diff = bound_adjoint_program(qc)
    terms, coeffs = hamiltonian_terms(hamiltonian)
    dev = _call(_native.pool.get, qc.num, precision)
    _call(dev.set_option, "checkpoints", int(checkpoints))
    psi_in = np.asarray(psi_in)
    if psi_in.size:
        _call(dev.load, psi_in)
    else:
        _call(dev.init_zero)
    values, expectation = _call(dev.adjoint, prog, ops, gens, terms, coeffs)
    gate_grads = [[] for _ in qc.gates]
    for i, v in zip(diff, values):
        gate_grads[i].append(float(v))
    grads = assemble_grads(para_grads, gate_grads)
    return (grads, expectation) if return_expectation else grads


def grad_para_shift(qc, hamiltonian, backend=None, psi_in=np.array([], dtype=complex)):
    """Parameter-shift gradients (gradiant.py:38-73): each symbolic gate must be RX, RY or RZ."""
    backend = SVSimulator() if backend is None else backend
    if not qc._parameter_grads:  # pylint: disable=protected-access
        qc.get_parameter_grads()
    para_grads = qc._calc_parameter_grads()  # pylint: disable=protected-access
    gate_grads = [[] for _ in qc.gates]
    for i, op in enumerate(qc.gates):
        if _is_symbolic(op):
            if op.name not in ["RX", "RY", "RZ"]:
                raise CircuitError(
                    "It seems the circuit can not apply parameter-shift rule to calculate gradient."
                    " You may need compile the circuit first"
                )
            op.paras[0] = op.paras[0] + np.pi / 2
            res1 = sum(backend.run(qc, hamiltonian=hamiltonian, psi=psi_in)["pauli_expects"])
            op.paras[0] = op.paras[0] - np.pi
            res2 = sum(backend.run(qc, hamiltonian=hamiltonian, psi=psi_in)["pauli_expects"])
            op.paras[0]._undo(2)  # pylint: disable=protected-access
            gate_grads[i].append((res1 - res2) / 2.0)
    return assemble_grads(para_grads, gate_grads)


def grad_para_shift_device(qc, hamiltonian, psi_in=np.array([], dtype=complex), *, precision: str = "c128"):
    """Parameter shift with every shifted circuit on the engine: the circuit is compiled once, the
    rotation angle of one op row is shifted by +-pi/2 in place, and <H> comes from one
    qfvm_expect_hamiltonian pass. Same values as grad_para_shift (RX / RY / RZ only), without the
    per-run host work; the comparator for grad_adjoint."""
    if not qc._parameter_grads:  # pylint: disable=protected-access
        qc.get_parameter_grads()
    para_grads = qc._calc_parameter_grads()  # pylint: disable=protected-access
    prog, ops, _, diff = adjoint_program(qc)
    for i in diff:
        if qc.gates[i].name not in ("RX", "RY", "RZ"):
            raise CircuitError("parameter shift needs uncontrolled RX, RY or RZ gates")
    terms, coeffs = hamiltonian_terms(hamiltonian)
    dev = _call(_native.pool.get, qc.num, precision)
    psi_in = np.asarray(psi_in)
    gate_grads = [[] for _ in qc.gates]
    for i, row in zip(diff, ops):
        vals = []
        theta = float(prog.ops["param"][row])
        for shift in (np.pi / 2, -np.pi / 2):
            prog.ops["param"][row] = theta + shift
            if psi_in.size:
                _call(dev.load, psi_in)
            else:
                _call(dev.init_zero)
            _call(dev.apply, prog)
            vals.append(_call(dev.expect_hamiltonian, terms, coeffs))
        prog.ops["param"][row] = theta
        gate_grads[i].append((vals[0] - vals[1]) / 2.0)
    return assemble_grads(para_grads, gate_grads)


def grad_finit_diff(qc, hamiltonian, backend=None, psi_in=np.array([], dtype=complex), h: float = 1e-10):
    """Central finite differences over qc.variables (gradiant.py:76-93; the step is a keyword here)."""
    backend = SVSimulator() if backend is None else backend
    grads = []
    for v in qc.variables:
        v.value += h
        res1 = sum(backend.run(qc, hamiltonian=hamiltonian, psi=psi_in)["pauli_expects"])
        v.value -= 2 * h
        res2 = sum(backend.run(qc, hamiltonian=hamiltonian, psi=psi_in)["pauli_expects"])
        v.value += h
        grads.append((res1 - res2) / (2 * h))
    return grads
