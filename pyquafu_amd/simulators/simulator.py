"""SVSimulator (quafu/simulators/simulator.py:41-145) on the MI355X engine.

`use_gpu` and `use_custatevec` are accepted for signature compatibility; every run uses the HIP
engine. Unlike the reference's GPU branch, the state is always stored in the result
(simulator.py:120-128 drops it) and classical operations (reset, cif) are supported on the GPU.
"""

from __future__ import annotations

import numpy as np

from ..elements.quantum_gate import CircuitWrapper, QuantumGate
from ..exceptions import QuafuError
from ..results.results import SimuResult
from .qfvm import applyop_statevec, expect_statevec, simulate_circuit


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
        if hamiltonian:
            paulis = hamiltonian.paulis
            res = expect_statevec(res_info["statevector"], paulis)
            for i, j in enumerate(paulis):
                res[i] *= j.coeff
            res_info["pauli_expects"] = res
        else:
            res_info["pauli_expects"] = []
        res_info["qbitnum"] = qc.num
        res_info["measures"] = qc.measures
        res_info["simulator"] = "statevector"
        return SimuResult(res_info)
