"""simulate() (quafu/simulators/__init__.py:25-78) on the MI355X engine."""

from __future__ import annotations

import numpy as np

from .program import Program, compile_circuit, compile_ops


def simulate(
    qc,
    psi: np.ndarray = np.array([]),
    simulator: str = "statevector",
    shots: int = 0,
    hamiltonian=None,
    use_gpu: bool = False,
    use_custatevec: bool = False,
    precision: str = "c128",
    seed=None,
):
    """Simulate a circuit; returns a SimuResult.

    "statevector" and "noisy statevetor" (trajectories, the reference's spelling) run on the
    engine; the reference's "clifford" simulator is another engine and is not provided. `precision` ("c128" default, or "c64") and `seed` are extensions."""
    if isinstance(qc, str):  # OpenQASM 2.0 text (simulators/__init__.py:52-57)
        from ..qfasm import parse_qasm  # pylint: disable=import-outside-toplevel

        qc = parse_qasm(qc)
    if simulator == "statevector":
        from .simulator import SVSimulator  # pylint: disable=import-outside-toplevel

        backend = SVSimulator(use_gpu, use_custatevec, precision=precision, seed=seed)
        return backend.run(qc, psi, shots, hamiltonian)
    if simulator == "noisy statevetor":  # the reference's spelling (simulators/__init__.py:66)
        from .simulator import NoiseSVSimulator  # pylint: disable=import-outside-toplevel

        backend = NoiseSVSimulator(use_gpu, use_custatevec, precision=precision, seed=seed)
        return backend.run(qc, psi, shots, hamiltonian)
    if simulator == "clifford":
        raise NotImplementedError(f"simulator {simulator!r} is not part of this engine")
    raise ValueError("invalid simulator name")


__all__ = ["simulate", "Program", "compile_circuit", "compile_ops"]
