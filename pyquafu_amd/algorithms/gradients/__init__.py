"""Gradients of expectation values (quafu/algorithms/gradients/__init__.py)."""

from .gradients import assemble_grads, grad_adjoint, grad_finit_diff, grad_para_shift

__all__ = ["assemble_grads", "grad_adjoint", "grad_finit_diff", "grad_para_shift"]
