"""Batched execution of many parameterised circuits of uneven sizes on one GPU.

Each circuit is evaluated as <H> plus its full adjoint gradient. A whole batch is one engine call
(qfvm_batch_run, include/qfvm.h): worker threads inside the engine take the circuits largest first
onto the least-loaded worker, each worker owns one handle (one HIP stream) per qubit count, and the
handles stay alive between calls, so a circuit evaluated again -- a training step with new
parameters -- meets its cached fusion plans (option "plan_cache") and, with graphs on, its cached
executable hipGraph. Parameters are rebound in place (ParametricProgram.bind): nothing is
recompiled or re-converted between steps. Tile shapes and fusion width come from `Autotuner`, which
times a probe circuit per (qubit count, dtype) once.

Not in the reference: pyquafu evaluates one circuit at a time (Estimator.run,
quafu/algorithms/estimator.py:186-229; the per-row loop of quafu/algorithms/interface/torch.py:76-79).
"""

from __future__ import annotations

import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native
from .algorithms.gradients.gradients import _is_symbolic, generator_of, hamiltonian_terms
from .simulators.binding import ParametricProgram


@dataclass
class BatchItem:
    """One compiled circuit: its parametric program, the differentiated op rows and their
    generators, the observable, and the engine call's arguments (built once)."""

    n: int
    pp: ParametricProgram
    ops: List[int]
    gens: str
    terms: list
    coeffs: list
    call: Optional[_native.AdjointCall] = None

    def __post_init__(self):
        if self.call is None:
            self.call = _native.AdjointCall(self.pp.prog, self.ops, self.gens, self.terms, self.coeffs)

    @property
    def prog(self):
        return self.pp.prog

    def bind(self, values) -> None:
        """New parameter values (circuit.variables order), written into the program in place."""
        self.pp.bind(values)


@dataclass
class BatchResult:
    expectation: float
    grads: np.ndarray
    latency_s: float


def compile_item(qc, hamiltonian) -> BatchItem:
    pp = ParametricProgram(qc)
    gates = pp.gates
    diff = [i for i, op in enumerate(gates) if _is_symbolic(op)]
    terms, coeffs = hamiltonian_terms(hamiltonian)
    return BatchItem(qc.num, pp, [pp.rows[i] for i in diff], "".join(generator_of(gates[i]) for i in diff), terms,
                     coeffs)


class Autotuner:
    """Picks (tile_bits, low_bits, max_groups) per (qubit count, dtype) by timing a probe adjoint
    run; 0 keeps the engine default."""

    CANDIDATES = [(0, 0, 0), (10, 5, 0), (11, 6, 0), (12, 6, 0), (12, 4, 0), (11, 4, 0), (0, 0, 6), (0, 0, 24),
                  (11, 6, 6), (12, 6, 24)]

    def __init__(self, probe_layers: int = 4):
        self.probe_layers = probe_layers
        self.choice: Dict[Tuple[int, str], Tuple[int, int, int]] = {}
        self._lock = threading.Lock()

    def tune(self, n: int, dtype: str) -> Tuple[int, int, int]:
        key = (n, dtype)
        with self._lock:
            if key in self.choice:
                return self.choice[key]
        from . import workloads  # pylint: disable=import-outside-toplevel
        from .algorithms.hamiltonian import Hamiltonian, PauliOp  # pylint: disable=import-outside-toplevel

        qc, _ = workloads.hea_ansatz(n, self.probe_layers, parametrised=True)
        item = compile_item(qc, Hamiltonian([PauliOp(f"Z{i}") for i in range(n)]))
        dev = _native.DeviceState(n, dtype)
        best, best_t = (0, 0, 0), float("inf")
        try:
            for tb, lb, mg in self.CANDIDATES:
                if tb and (tb > n or (1 << tb) * (8 if dtype == "c64" else 16) * 2 > 160 * 1024):
                    continue
                dev.set_option("tile_bits", tb)
                dev.set_option("low_bits", lb)
                dev.set_option("max_groups", mg or 12)
                dev.init_zero()
                dev.adjoint_prepared(item.call)
                t0 = time.perf_counter()
                for _ in range(3):
                    dev.init_zero()
                    dev.adjoint_prepared(item.call)
                dt = time.perf_counter() - t0
                if dt < best_t:
                    best, best_t = (tb, lb, mg), dt
        finally:
            dev.close()
        with self._lock:
            self.choice[key] = best
        return best


@dataclass
class BatchRunner:
    """Evaluates lists of BatchItems through one engine batch context. `graphs=True` (option
    "graphs") captures each circuit's engine work into a hipGraph that later calls replay;
    circuits of at least `pin_qubits` qubits run one after another on one worker (qfvm_batch_set_option)."""

    dtype: str = "c64"
    workers: int = 4
    tuner: Optional[Autotuner] = None
    graphs: bool = True
    pin_qubits: int = 20
    _batch: Optional[_native.NativeBatch] = field(default=None, repr=False)
    _tuned: set = field(default_factory=set, repr=False)
    _prepared: Optional[Tuple[tuple, _native.PreparedBatch]] = field(default=None, repr=False)

    def __post_init__(self):
        self._batch = _native.NativeBatch(self.dtype, self.workers)
        self._batch.set_option("pin_qubits", self.pin_qubits)
        if self.graphs:
            self._batch.set_option("graphs", 1)

    def warm(self, ns: Sequence[int]) -> None:
        """Tunes the tile shapes of these qubit counts (outside any timing)."""
        if self.tuner is None:
            return
        for n in sorted(set(int(n) for n in ns)):
            if n in self._tuned:
                continue
            tb, lb, mg = self.tuner.tune(n, self.dtype)
            self._batch.set_option("tile_bits", tb, n)
            self._batch.set_option("low_bits", lb, n)
            self._batch.set_option("max_groups", mg or 12, n)
            self._tuned.add(n)

    def prepare(self, items: Sequence[BatchItem]) -> _native.PreparedBatch:
        key = tuple(id(it) for it in items)
        if self._prepared is None or self._prepared[0] != key:
            self._prepared = (key, _native.PreparedBatch([(it.n, it.call) for it in items]))
        return self._prepared[1]

    def run(self, items: Sequence[BatchItem], values: Optional[Sequence] = None) -> Tuple[List[BatchResult], float]:
        """Evaluates every item (after binding values[k] into item k, if given); returns the results
        in input order and the wall time of the engine call."""
        self.warm([it.n for it in items])
        if values is not None:
            for it, v in zip(items, values):
                it.bind(v)
        pb = self.prepare(items)
        t0 = time.perf_counter()
        self._batch.run(pb)
        wall = time.perf_counter() - t0
        results = [BatchResult(float(pb.expectations[k]), pb.grads_of(k).copy(), float(pb.seconds[k]))
                   for k in range(len(items))]
        return results, wall

    def close(self) -> None:
        if self._batch is not None:
            self._batch.close()
            self._batch = None
