"""Batched execution of many parameterised circuits of uneven sizes on one GPU.

Each circuit is evaluated as <H> plus its full adjoint gradient (one qfvm_adjoint call). Circuits
run concurrently: `workers` host threads each own one engine handle per qubit count, every handle
has its own HIP stream, and ctypes releases the GIL during engine calls, so small circuits (which
cannot fill the GPU alone) overlap on the device. Tile shapes come from `Autotuner`, which times a
probe circuit per (qubit count, dtype) once and keeps the fastest (tile_bits, low_bits).

Not in the reference: pyquafu evaluates one circuit at a time (Estimator / grad_adjoint).
"""

from __future__ import annotations

import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _native
from .algorithms.gradients.gradients import adjoint_program, hamiltonian_terms


@dataclass
class BatchItem:
    """One compiled circuit: program, differentiated op rows, generators, observable, and the
    engine call's arguments."""

    n: int
    prog: object
    ops: List[int]
    gens: str
    terms: list
    coeffs: list
    call: Optional[_native.AdjointCall] = None

    def __post_init__(self):
        if self.call is None:
            self.call = _native.AdjointCall(self.prog, self.ops, self.gens, self.terms, self.coeffs)


@dataclass
class BatchResult:
    expectation: float
    grads: np.ndarray
    latency_s: float
    start_s: float


def compile_item(qc, hamiltonian) -> BatchItem:
    prog, ops, gens, _ = adjoint_program(qc)
    terms, coeffs = hamiltonian_terms(hamiltonian)
    return BatchItem(qc.num, prog, ops, gens, terms, coeffs)


class Autotuner:
    """Picks (tile_bits, low_bits) per (qubit count, dtype) by timing a probe adjoint run."""

    CANDIDATES = [(0, 0), (10, 5), (11, 6), (12, 6), (12, 4), (11, 4)]

    def __init__(self, probe_layers: int = 4):
        self.probe_layers = probe_layers
        self.choice: Dict[Tuple[int, str], Tuple[int, int]] = {}
        self._lock = threading.Lock()

    def tune(self, n: int, dtype: str) -> Tuple[int, int]:
        key = (n, dtype)
        with self._lock:
            if key in self.choice:
                return self.choice[key]
        from . import workloads  # pylint: disable=import-outside-toplevel
        from .algorithms.hamiltonian import Hamiltonian, PauliOp  # pylint: disable=import-outside-toplevel

        qc, _ = workloads.hea_ansatz(n, self.probe_layers, parametrised=True)
        item = compile_item(qc, Hamiltonian([PauliOp(f"Z{i}") for i in range(n)]))
        dev = _native.DeviceState(n, dtype)
        best, best_t = (0, 0), float("inf")
        try:
            for tb, lb in self.CANDIDATES:
                if tb and (tb > n or (1 << tb) * (8 if dtype == "c64" else 16) * 2 > 160 * 1024):
                    continue
                dev.set_option("tile_bits", tb)
                dev.set_option("low_bits", lb)
                dev.init_zero()
                dev.adjoint(item.prog, item.ops, item.gens, item.terms, item.coeffs)
                t0 = time.perf_counter()
                for _ in range(3):
                    dev.init_zero()
                    dev.adjoint(item.prog, item.ops, item.gens, item.terms, item.coeffs)
                dt = time.perf_counter() - t0
                if dt < best_t:
                    best, best_t = (tb, lb), dt
        finally:
            dev.close()
        with self._lock:
            self.choice[key] = best
        return best


@dataclass
class BatchRunner:
    """`graphs=True` runs every engine call as a captured hipGraph (option "graphs"): each worker
    slot owns its handles and a fixed share of the items (longest-first onto the least-loaded
    slot), so a circuit evaluated again (a training loop with new parameters) lands on the handle
    that already holds its executable graph and replays it."""

    dtype: str = "c64"
    workers: int = 8
    tuner: Optional[Autotuner] = None
    graphs: bool = False
    _local: threading.local = field(default_factory=threading.local, repr=False)
    _slots: dict = field(default_factory=dict, repr=False)
    _handles: list = field(default_factory=list, repr=False)
    _lock: threading.Lock = field(default_factory=threading.Lock, repr=False)
    _pool: Optional[ThreadPoolExecutor] = field(default=None, repr=False)

    def __post_init__(self):
        self._pool = ThreadPoolExecutor(self.workers)

    def _device(self, n: int, slot: Optional[int] = None) -> _native.DeviceState:
        if slot is None:
            states = getattr(self._local, "states", None)
            if states is None:
                states = self._local.states = {}
        else:
            with self._lock:
                states = self._slots.setdefault(slot, {})
        dev = states.get(n)
        if dev is None:
            dev = _native.DeviceState(n, self.dtype)
            if self.tuner is not None:
                tb, lb = self.tuner.tune(n, self.dtype)
                dev.set_option("tile_bits", tb)
                dev.set_option("low_bits", lb)
            if self.graphs:
                dev.set_option("graphs", 1)
            states[n] = dev
            with self._lock:
                self._handles.append(dev)
        return dev

    def _run_one(self, item: BatchItem, t0: float, slot: Optional[int] = None) -> BatchResult:
        start = time.perf_counter()
        dev = self._device(item.n, slot)
        dev.init_zero()
        grads, e = dev.adjoint_prepared(item.call)
        end = time.perf_counter()
        return BatchResult(e, grads, end - start, start - t0)

    def _partition(self, items: Sequence[BatchItem]) -> List[List[int]]:
        """Item indices per worker slot: longest first (cost ~ 2^n x ops) onto the least-loaded
        slot; the same items give the same partition."""
        cost = [(1 << it.n) * (int(it.call.cp.n_ops) + 8) for it in items]
        load = [0] * self.workers
        parts: List[List[int]] = [[] for _ in range(self.workers)]
        for k in sorted(range(len(items)), key=lambda k: (-cost[k], k)):
            w = min(range(self.workers), key=lambda w: (load[w], w))
            parts[w].append(k)
            load[w] += cost[k]
        return parts

    def warm(self, ns: Sequence[int]) -> None:
        """Tune, then create every worker's handles for these qubit counts (outside any timing)."""
        ns = sorted(set(int(n) for n in ns))
        if self.tuner is not None:
            for n in ns:
                self.tuner.tune(n, self.dtype)
        barrier = threading.Barrier(self.workers)

        def per_worker(w):
            barrier.wait()  # one task per worker thread
            for n in ns:
                self._device(n, w if self.graphs else None)

        list(self._pool.map(per_worker, range(self.workers)))

    def run(self, items: Sequence[BatchItem]) -> Tuple[List[BatchResult], float]:
        """Evaluates every item; returns the results in input order and the wall time."""
        if self.graphs:
            parts = self._partition(items)
            results: List[Optional[BatchResult]] = [None] * len(items)
            t0 = time.perf_counter()

            def per_slot(w):
                for k in parts[w]:
                    results[k] = self._run_one(items[k], t0, w)

            list(self._pool.map(per_slot, range(self.workers)))
            return results, time.perf_counter() - t0
        t0 = time.perf_counter()
        results = list(self._pool.map(lambda it: self._run_one(it, t0), items))
        return results, time.perf_counter() - t0

    def close(self) -> None:
        self._pool.shutdown(wait=True)
        with self._lock:
            for dev in self._handles:
                dev.close()
            self._handles.clear()
