"""Expectation estimator (quafu/algorithms/estimator.py:31-230), simulator backend.

Estimator.run binds a parameter list to the circuit's variables and returns the sum of the
coefficient-weighted Pauli expectations from simulate() (execute_circuit, estimator.py:31-35).
Real-machine tasks (the reference's backend != "sim" branch, which submits to the cloud service)
are not part of this package and raise QuafuError.
"""

from __future__ import annotations

from typing import List, Optional

from ..exceptions import QuafuError
from ..simulators import simulate


def execute_circuit(circ, observables):
    """Sum of the Pauli expectations of `observables` after `circ` on the simulator."""
    sim_res = simulate(circ, hamiltonian=observables)
    return sum(sim_res["pauli_expects"])


class Estimator:
    """Estimate expectations of observables for a parameterised circuit."""

    def __init__(self, circ, backend: str = "sim", task=None, **task_options) -> None:
        if backend != "sim":
            raise QuafuError("only the simulator backend ('sim') is available in this package")
        self._circ = circ
        self._circ.get_parameter_grads()  # parameter shift relies on this initialisation
        self._backend = backend
        self._task = task
        self._task_options = task_options
        self._exp_cache = {}

    def clear_cache(self):
        self._exp_cache.clear()

    def run(self, observables, params: Optional[List[float]], cache_key: Optional[str] = None):
        if params is not None:
            self._circ._update_params(params)  # pylint: disable=protected-access
        return execute_circuit(self._circ, observables)
