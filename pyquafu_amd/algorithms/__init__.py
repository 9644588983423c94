"""Algorithms on the statevector path: Hamiltonians, gradients, estimator, torch layer."""

from .hamiltonian import Hamiltonian, PauliMats, PauliOp

__all__ = ["Hamiltonian", "PauliOp", "PauliMats"]
