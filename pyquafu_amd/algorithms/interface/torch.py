"""PyTorch layer (quafu/algorithms/interface/torch.py:29-89) on the MI355X engine.

ExecuteCircuits keeps the reference's contract: parameters [batch, num_params] (float64, CPU)
-> outputs [batch, num_qubits] = <Z_i>, with the backward pass a vector-Jacobian product. The
Jacobian comes from parameter shift (the reference's only method) or, with method="adjoint", from
one engine adjoint sweep per output.

ExpectationFunction is the adjoint-differentiated loss <H>(params): its forward pass makes one
qfvm_adjoint call, which returns <H> and every d<H>/d param together, and its backward pass scales
the stored gradient by the incoming one.

QuantumNeuralNetwork and ModuleWrapper follow quafu/algorithms/ansatz.py:149-260 and
interface/torch.py:92-130 for the gate-list ("legacy") form of the QNN: the layers are gates whose
parameters are the circuit's variables, the tunable ones become a (1, D) weight tensor, and the
module concatenates its inputs with the expanded weights into each row's parameter vector.
Embedding layers (the reference's BaseEmebdding subclasses) are not part of this package.
"""

from __future__ import annotations

from typing import Optional

import numpy as np
import torch
from torch import nn

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


class QuantumNeuralNetwork:
    """A parameterised circuit used as a network layer: outputs <Z_i> per row of parameters
    (ansatz.py:149-260, gate-list layers)."""

    def __init__(self, num_qubits: int, layers, interface: str = "torch", backend: str = "sim"):
        from ...circuits import QuantumCircuit  # pylint: disable=import-outside-toplevel
        from ...elements.quantum_gate import QuantumGate  # pylint: disable=import-outside-toplevel

        if interface != "torch":
            raise ValueError(f"unknown interface {interface!r} (this package provides 'torch')")
        if not layers or not isinstance(layers[0], QuantumGate):
            raise TypeError("layers must be a list of gates (embedding layers are not part of this package)")
        self.circ = QuantumCircuit(num_qubits)
        for g in layers:
            self.circ << g  # pylint: disable=pointless-statement
        self.circ.get_parameter_grads()
        self._layers = layers
        self._backend = backend
        self._num_tunable_params = 0
        self._weights = TorchTransformer.init_weights((1, self.num_tunable_parameters))

    @property
    def num(self) -> int:
        return self.circ.num

    @property
    def weights(self):
        return self._weights

    @property
    def num_tunable_parameters(self) -> int:
        """Tunable parameters of the gates (counted once, as the reference does)."""
        if not self._num_tunable_params:
            for g in self.circ.gates:
                for p in g.paras:
                    if getattr(p, "tunable", False):
                        self._num_tunable_params += 1
        return self._num_tunable_params

    def __call__(self, inputs: torch.Tensor) -> torch.Tensor:
        """<Z_i> for each row of `inputs` (one full parameter vector per row)."""
        from ..estimator import Estimator  # pylint: disable=import-outside-toplevel

        estimator = Estimator(self.circ, backend=self._backend)
        return TorchTransformer.execute(self.circ, inputs, estimator=estimator)


class ModuleWrapper(nn.Module):
    """A QuantumNeuralNetwork as a torch module (interface/torch.py:92-130): its weights are a
    trainable (1, D) parameter appended to every input row."""

    def __init__(self, qnn: QuantumNeuralNetwork):
        super().__init__()
        self._qnn = qnn
        self.weights = nn.parameter.Parameter(qnn.weights) if qnn.weights is not None else None

    def forward(self, inputs: torch.Tensor) -> torch.Tensor:  # pylint: disable=arguments-differ
        bsz, _ = inputs.shape
        if self.weights is not None:
            weights_expanded = self.weights.expand(bsz, self.weights.size(-1))
            inputs_to_circ = torch.cat((inputs, weights_expanded), dim=1)
        else:
            inputs_to_circ = inputs
        return self._qnn(inputs_to_circ)
