"""PyTorch layer (quafu/algorithms/interface/torch.py:29-89) on the MI355X engine.

ExecuteCircuits keeps the reference's contract: parameters [batch, num_params] (float64, CPU)
-> outputs [batch, num_qubits] = <Z_i>, with the backward pass a vector-Jacobian product. The
Jacobian comes from parameter shift (the reference's only method) or, with method="adjoint", from
one engine adjoint sweep per output.

ExpectationFunction is the adjoint-differentiated loss <H>(params): its forward pass makes one
qfvm_adjoint call, which returns <H> and every d<H>/d param together, and its backward pass scales
the stored gradient by the incoming one.
"""

from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from ..gradients import compute_vjp, jacobian, run_circ
from ..gradients.gradients import grad_adjoint


class TorchTransformer:
    @staticmethod
    def init_weights(shape):
        """A float64 weight tensor that requires grad."""
        return torch.randn(*shape, requires_grad=True, dtype=torch.double)

    @staticmethod
    def execute(circ, parameters: torch.Tensor, run_fn=run_circ, grad_fn=None, method: str = "internal",
                estimator: Optional[object] = None):
        kwargs = {"circ": circ, "run_fn": run_fn, "grad_fn": grad_fn, "estimator": estimator, "method": method}
        return ExecuteCircuits.apply(parameters, kwargs)


class ExecuteCircuits(torch.autograd.Function):
    """Parameters are inputs from previous layers."""

    @staticmethod
    def forward(ctx, parameters, kwargs):  # pylint: disable=arguments-differ
        ctx.run_fn = kwargs["run_fn"]
        ctx.circ = kwargs["circ"]
        ctx.estimator = kwargs["estimator"]
        ctx.method = kwargs.get("method", "internal")
        ctx.save_for_backward(parameters)
        outputs = [ctx.run_fn(ctx.circ, para, estimator=ctx.estimator)
                   for para in parameters.detach().cpu().numpy().tolist()]
        return torch.from_numpy(np.stack(outputs))

    @staticmethod
    def backward(ctx, grad_out):  # pylint: disable=arguments-differ
        (parameters,) = ctx.saved_tensors
        method = "adjoint" if ctx.method == "adjoint" else "param_shift"
        jac = jacobian(ctx.circ, parameters.detach().cpu().numpy(), estimator=ctx.estimator, method=method)
        return torch.from_numpy(compute_vjp(jac, grad_out.detach().cpu().numpy())), None


class ExpectationFunction(torch.autograd.Function):
    """<H> of a parameterised circuit; gradients by the engine's adjoint sweep."""

    @staticmethod
    def forward(ctx, params, circ, hamiltonian, precision, checkpoints=0):  # pylint: disable=arguments-differ
        circ._update_params(params.detach().cpu().numpy().tolist())  # pylint: disable=protected-access
        grads, e = grad_adjoint(circ, hamiltonian, precision=precision, return_expectation=True,
                                checkpoints=checkpoints)
        ctx.save_for_backward(torch.tensor(grads, dtype=torch.float64))
        return torch.tensor(e, dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_out):  # pylint: disable=arguments-differ
        (grads,) = ctx.saved_tensors
        return grad_out * grads, None, None, None, None


def expectation(circ, hamiltonian, params: torch.Tensor, precision: str = "c128", checkpoints: int = 0) -> torch.Tensor:
    """Differentiable <H>(params); params follow circ.variables' order. `checkpoints` > 0 keeps that
    many forward-state copies for the reverse sweep (see grad_adjoint)."""
    return ExpectationFunction.apply(params, circ, hamiltonian, precision, checkpoints)
