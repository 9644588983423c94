"""Parameter shift through an Estimator, Jacobians and vector-Jacobian products
(quafu/algorithms/gradients/param_shift.py:23-78, vjp.py:35-140).

`jacobian(..., method="adjoint")` is an addition: one engine adjoint sweep per output observable
(qfvm_adjoint) instead of two full simulations per parameter and output.
"""

from __future__ import annotations

from typing import List, Optional

import numpy as np

from ..estimator import Estimator
from ..hamiltonian import Hamiltonian


class ParamShift:
    """Parameter-shift rule through an estimator: E(p + pi/2 e_k) - E(p - pi/2 e_k), halved."""

    def __init__(self, estimator: Estimator) -> None:
        self._est = estimator

    def __call__(self, obs: Hamiltonian, params: List[float], cache_key: Optional[str] = None):
        return self.grad(obs, params, cache_key=cache_key)

    @staticmethod
    def _gen_param_shift_vals(params):
        params = np.array(params, dtype=float)
        offsets = np.identity(params.size)
        return (params + offsets * np.pi / 2).tolist() + (params - offsets * np.pi / 2).tolist()

    def grad(self, obs: Hamiltonian, params: List[float], cache_key: Optional[str] = None):
        shifted = self._gen_param_shift_vals(params)
        res = np.zeros(len(shifted))
        for i, p in enumerate(shifted):
            key = None if cache_key is None else cache_key + str(i)
            res[i] = self._est.run(obs, p, cache_key=key)
        half = len(res) // 2
        return (res[:half] - res[half:]) / 2


def _generate_expval_z(num_qubits: int):
    return [Hamiltonian.from_pauli_list([("Z" + str(i), 1)]) for i in range(num_qubits)]


def run_circ(circ, params: Optional[List[float]] = None, backend: str = "sim",
             estimator: Optional[Estimator] = None):
    """<Z_i> for every qubit after binding `params` (None keeps the current variable values)."""
    obs_list = _generate_expval_z(circ.num)
    if estimator is None:
        estimator = Estimator(circ, backend=backend)
    if params is None:
        params = [v.value for v in circ.variables]
    output = [estimator.run(obs, params, cache_key="00") for obs in obs_list]
    estimator.clear_cache()
    return np.array(output)


def jacobian(circ, params_input: np.ndarray, estimator: Optional[Estimator] = None, method: str = "param_shift"):
    """d<Z_i>/d p_k for every row of params_input: shape [batch, num_qubits, num_params]."""
    params_input = np.asarray(params_input, dtype=float)
    batch_size, num_params = params_input.shape
    obs_list = _generate_expval_z(circ.num)
    output = np.zeros((batch_size, len(obs_list), num_params))
    if method == "adjoint":
        from .gradients import grad_adjoint  # pylint: disable=import-outside-toplevel

        for i in range(batch_size):
            circ._update_params(params_input[i].tolist())  # pylint: disable=protected-access
            for j, obs in enumerate(obs_list):
                output[i, j, :] = grad_adjoint(circ, obs)
        return output
    if estimator is None:
        estimator = Estimator(circ)
    calc_grad = ParamShift(estimator)
    for i in range(batch_size):
        output[i] = np.stack([np.array(calc_grad(obs, params_input[i].tolist(), cache_key=str(i)))
                              for obs in obs_list])
    estimator.clear_cache()
    return output


def compute_vjp(jac: np.ndarray, dy: np.ndarray) -> np.ndarray:
    """sum_j jac[i, j, k] * dy[i, j] -> [batch, num_params]."""
    batch_size, num_outputs, _ = jac.shape
    assert dy.shape[0] == batch_size and dy.shape[1] == num_outputs
    return np.einsum("ijk,ij->ik", jac, dy)
