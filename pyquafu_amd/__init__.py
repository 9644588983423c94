"""MI355X statevector engine with pyquafu's host API for the qfvm path.

    from pyquafu_amd import QuantumCircuit, simulate
    qc = QuantumCircuit(2); qc.h(0); qc.cx(0, 1); qc.measure([0, 1])
    simulate(qc, shots=1000).counts

Simulation runs on the GPU through libqfvm_hip.so (built by pyquafu_amd.build); there is no CPU
fallback.
"""

from .circuits import QuantumCircuit
from .elements import *  # noqa: F401,F403
from .exceptions import CircuitError, QuafuError
from .results import SimuResult
from .simulators import simulate

__version__ = "0.5.0"

__all__ = ["QuantumCircuit", "SimuResult", "simulate", "QuafuError", "CircuitError"]
