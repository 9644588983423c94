"""SVSimulator (quafu/simulators/simulator.py:41-145) on the MI355X engine.

`use_gpu` and `use_custatevec` are accepted for signature compatibility; every run uses the HIP
engine. Unlike the reference's GPU branch, the state is always stored in the result
(simulator.py:120-128 drops it) and classical operations (reset, cif) are supported on the GPU.
"""

from __future__ import annotations

import numpy as np

from .. import _native
from ..elements.noise import KrausChannel, UnitaryChannel
from ..elements.quantum_gate import CircuitWrapper, QuantumGate
from ..exceptions import QuafuError
from ..results.results import SimuResult
from .program import KRAUS, UCHAN, compile_gates
from .qfvm import _call, applyop_statevec, bitstring_counts, simulate_circuit


MAX_DEVICE_MARGINAL = 12  # measured qubits reduced on the device while the final state is resident


def measured_marginals(qc, precision: str):
    """Probabilities over the measured qubits in SimuResult.calc_probabilities' layout
    (results.py:148-176): index bit M-1-j <-> the qubit with the j-th smallest cbit. Taken from the
    final state still on the device after simulate_circuit (the engine pool keeps one state per
    qubit count and dtype); None when nothing is measured or too many qubits are."""
    meas = qc.measures
    if not meas or len(meas) > MAX_DEVICE_MARGINAL:
        return None
    qubits, cbits = list(meas.keys()), list(meas.values())
    by_cbit = [qubits[k] for k in np.argsort(cbits)]
    dev = _call(_native.pool.get, qc.num, precision)
    return _call(dev.marginals, by_cbit[::-1])


class Simulator:
    def run(self, *args, **kwargs):
        raise NotImplementedError


class SVSimulator(Simulator):
    def __init__(self, use_gpu: bool = False, use_custatevec: bool = False, precision: str = "c128", seed=None):
        self.use_gpu = use_gpu
        self.use_custatevec = use_custatevec
        self.precision = precision
        self.seed = seed

    def config(self, **kwargs):
        for attr, value in kwargs.items():
            if hasattr(self, attr):
                setattr(self, attr, value)
            else:
                raise ValueError("No such attribute")

    def _apply_op(self, op, psi):
        if isinstance(op, CircuitWrapper):
            return self.run(op.circuit, psi)["statevector"]
        if isinstance(op, QuantumGate):
            return applyop_statevec(op, psi, dtype=self.precision)
        raise NotImplementedError

    def _apply_hamil(self, hamil, psi):
        """lambda = H psi, one Pauli string at a time (simulator.py:78-88)."""
        psi_out = np.zeros(len(psi), dtype=complex)
        for pauli in hamil.paulis:
            psi1 = np.copy(psi)
            for name, pos in zip(pauli.paulistr, pauli.pos):
                op = QuantumGate.gate_classes[name.lower()](pos)
                psi1 = self._apply_op(op, psi1)
            psi_out += psi1 * pauli.coeff
        return psi_out

    def run(self, qc, psi: np.ndarray = np.array([]), shots: int = 0, hamiltonian=None):
        if qc.noised:
            raise QuafuError("Can not run noisy circuits with statevector simulator, please use the noisy version.")
        res_info = {}
        count_dict, psi = simulate_circuit(qc, psi, shots, seed=self.seed, dtype=self.precision)
        res_info["statevector"] = psi
        res_info["counts"] = count_dict
        marginals = measured_marginals(qc, self.precision)
        if marginals is not None:
            res_info["probabilities"] = marginals
        if hamiltonian:
            # expect_statevec on the final state while it is still on the device (the engine pool
            # keeps it), all terms in few passes, instead of uploading the returned copy again
            paulis = hamiltonian.paulis
            dev = _call(_native.pool.get, qc.num, self.precision)
            res = [float(v) for v in _call(dev.expect, [(p.paulistr, list(p.pos)) for p in paulis])]
            for i, j in enumerate(paulis):
                res[i] *= j.coeff
            res_info["pauli_expects"] = res
        else:
            res_info["pauli_expects"] = []
        res_info["qbitnum"] = qc.num
        res_info["measures"] = qc.measures
        res_info["simulator"] = "statevector"
        return SimuResult(res_info)


class NoiseSVSimulator(Simulator):
    """Noisy statevector trajectories (quafu/simulators/simulator.py:148-223).

    Each shot runs one trajectory of the circuit's gates and noise channels on the engine (channels
    are program ops, include/qfvm.h QFVM_OP_KRAUS / QFVM_OP_UCHAN, one uniform each), then draws
    one sample over `qc.measures` and, with a Hamiltonian, the coefficient-weighted Pauli
    expectations, averaged over the shots. As in the reference's gen_circuit, only gates and
    channels run: measure, reset and cif instructions are left out of the trajectory.
    `seed` (an extension) seeds the trajectory uniforms."""

    def __init__(self, use_gpu: bool = False, use_custatevec: bool = False, precision: str = "c128", seed=None):
        self.backend = SVSimulator(use_gpu=use_gpu, use_custatevec=use_custatevec, precision=precision, seed=seed)
        self.precision = precision
        self.seed = seed

    @staticmethod
    def trajectory_program(qc):
        """The program one trajectory runs: gates and channels in circuit order."""
        if qc._has_wrap:  # pylint: disable=protected-access
            qc.unwrap()
        ops = [op for op in qc.instructions if isinstance(op, (QuantumGate, KrausChannel))]
        prog, _ = compile_gates(ops, qc.num)
        prog.n_random = int(np.isin(prog.ops["code"], [KRAUS, UCHAN]).sum()) if prog.n_ops else 0
        return prog

    @staticmethod
    def gen_circuit(qc, rng=None):
        """One sampled circuit: unitary channels replaced by a drawn gate (identities dropped),
        Kraus channels kept between wrapped gate blocks (simulator.py:199-223)."""
        from ..circuits import QuantumCircuit  # pylint: disable=import-outside-toplevel

        num = qc.num
        new_qc, temp_qc = QuantumCircuit(num), QuantumCircuit(num)
        if qc._has_wrap:  # pylint: disable=protected-access
            qc.unwrap()
        for op in qc.instructions:
            if isinstance(op, QuantumGate):
                temp_qc << op  # pylint: disable=pointless-statement
            elif isinstance(op, UnitaryChannel):
                g = op.gen_gate(rng)
                if g.name != "ID":
                    temp_qc << g  # pylint: disable=pointless-statement
            elif isinstance(op, KrausChannel):
                new_qc << temp_qc.wrap()  # pylint: disable=pointless-statement
                new_qc << op  # pylint: disable=pointless-statement
                temp_qc = QuantumCircuit(num)
        new_qc << temp_qc.wrap()  # pylint: disable=pointless-statement
        return new_qc

    def run_once(self, qc, psi, hamiltonian=None, rng=None):
        """One trajectory from `psi` (an empty array: |0...0>): returns (the sampled bitstring of
        qc.measures, the coefficient-weighted Pauli expectations or None), as the reference's
        run_once (simulator.py:152-164). `rng` (an extension) supplies the uniforms."""
        prog = self.trajectory_program(qc)
        n = qc.num
        has_input = psi is not None and np.asarray(psi).size != 0
        if has_input and np.asarray(psi).size != (1 << n):
            raise ValueError(f"input state length {np.asarray(psi).size} does not match {n} qubits")
        rng = np.random.default_rng(self.seed) if rng is None else rng
        dev = _call(_native.pool.get, n, self.precision)
        if has_input:
            _call(dev.load, np.asarray(psi, dtype=np.complex128))
        else:
            _call(dev.init_zero)
        u = rng.random(prog.n_random + 1)
        _call(dev.run, prog, u[:-1], 0, True)
        sample = _call(dev.sample, u[-1:])
        meas = [(int(q), int(c)) for q, c in qc.measures.items()]
        key = next(iter(bitstring_counts(np.asarray(sample, dtype=np.uint64), meas)))
        if not hamiltonian:
            return key, None
        paulis = hamiltonian.paulis
        res = _call(dev.expect, [(p.paulistr, [int(x) for x in p.pos]) for p in paulis])
        return key, [float(v) * p.coeff for v, p in zip(res, paulis)]

    def run(self, qc, psi: np.ndarray = np.array([]), shots: int = 0, hamiltonian=None):  # pylint: disable=arguments-differ
        prog = self.trajectory_program(qc)
        n = qc.num
        has_input = psi is not None and np.asarray(psi).size != 0
        if has_input and np.asarray(psi).size != (1 << n):
            raise ValueError(f"input state length {np.asarray(psi).size} does not match {n} qubits")
        shots = int(shots)
        meas = [(int(q), int(c)) for q, c in qc.measures.items()]
        paulis = hamiltonian.paulis if hamiltonian else []
        terms = [(p.paulistr, [int(x) for x in p.pos]) for p in paulis]
        coeffs = np.array([p.coeff for p in paulis])
        if shots == 0:  # the reference divides its float accumulator by the shot count
            raise ZeroDivisionError("float division by zero")
        rng = np.random.default_rng(self.seed)
        dev = _call(_native.pool.get, n, self.precision)
        samples = np.zeros(max(shots, 0), dtype=np.uint64)
        acc = np.zeros(len(paulis))
        for shot in range(shots):
            if has_input:
                _call(dev.load, np.asarray(psi, dtype=np.complex128))
            else:
                _call(dev.init_zero)
            u = rng.random(prog.n_random + 1)
            _call(dev.run, prog, u[:-1], 0, True)
            samples[shot] = _call(dev.sample, u[-1:])[0]
            if paulis:
                acc += _call(dev.expect, terms)
        counts = bitstring_counts(samples, meas) if shots > 0 else {}
        pauli_expects = list(acc * coeffs / shots) if paulis else []
        res_info = {"counts": counts, "pauli_expects": pauli_expects}
        res_info["qbitnum"] = qc.num
        res_info["measures"] = qc.measures
        res_info["simulator"] = "noisy statevector"
        return SimuResult(res_info)
