"""Circuit elements: gates, parameters and classical instructions (quafu.elements)."""

from . import matrices
from .element_gates import *  # noqa: F401,F403
from .element_gates import __all__ as _gate_names
from .instruction import Barrier, Cif, Delay, Instruction, Measure, Reset
from .noise import AmplitudeDamping, BitFlip, Decoherence, Dephasing, Depolarizing, KrausChannel, UnitaryChannel
from .parameters import Parameter, ParameterExpression, ParameterType
from .quantum_gate import CircuitWrapper, ControlledCircuitWrapper, ControlledGate, ControlledU, QuantumGate

__all__ = list(_gate_names) + [
    "matrices",
    "Barrier",
    "Cif",
    "Delay",
    "Instruction",
    "Measure",
    "Reset",
    "Parameter",
    "ParameterExpression",
    "ParameterType",
    "CircuitWrapper",
    "ControlledCircuitWrapper",
    "ControlledGate",
    "ControlledU",
    "QuantumGate",
    "KrausChannel",
    "UnitaryChannel",
    "BitFlip",
    "Dephasing",
    "Depolarizing",
    "AmplitudeDamping",
    "Decoherence",
]
