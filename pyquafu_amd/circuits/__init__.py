from .quantum_circuit import QuantumCircuit

__all__ = ["QuantumCircuit"]
