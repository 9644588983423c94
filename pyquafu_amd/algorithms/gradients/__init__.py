"""Gradients of expectation values (quafu/algorithms/gradients/__init__.py)."""

from .gradients import assemble_grads, grad_adjoint, grad_finit_diff, grad_para_shift, grad_para_shift_device
from .vjp import ParamShift, compute_vjp, jacobian, run_circ

__all__ = [
    "assemble_grads",
    "grad_adjoint",
    "grad_finit_diff",
    "grad_para_shift",
    "grad_para_shift_device",
    "ParamShift",
    "compute_vjp",
    "jacobian",
    "run_circ",
]
