"""Drop-in replacement of the reference's pybind module `qfvm` (src/qfvm/qfvm.cpp:263-288).

Same function names, argument meaning, return shapes and in-place behaviour; the work runs on the
MI355X engine through the C-ABI (include/qfvm.h). Randomness: the reference draws from
std::random_device and cannot be seeded (src/qfvm/util.h:35-46); here the uniforms come from
numpy's PCG64, seedable through the extra keyword `seed`.
"""

from __future__ import annotations

from collections import Counter
from typing import Dict, List, Sequence, Tuple

import numpy as np

from .. import _native
from .program import Program, compile_circuit, compile_ops

_ERRORS = {
    _native.ERR_VALUE: ValueError,
    _native.ERR_INDEX: IndexError,
    _native.ERR_MEMORY: MemoryError,
    _native.ERR_DEVICE: RuntimeError,
    _native.ERR_RUNTIME: RuntimeError,
}


def _call(fn, *args, **kw):
    try:
        return fn(*args, **kw)
    except _native.EngineError as e:
        raise _ERRORS.get(e.code, RuntimeError)(str(e)) from None


def _uniforms(seed, n: int) -> np.ndarray:
    return np.random.default_rng(seed).random(n)


def _n_from_state(psi: np.ndarray, n_circuit: int) -> int:
    size = psi.size
    n = size.bit_length() - 1
    if size != (1 << n):
        raise ValueError(f"input state length {size} is not a power of two")
    if n != n_circuit:
        raise ValueError(f"input state has {n} qubits, circuit has {n_circuit}")
    return n


def counts_from_indices(indices: np.ndarray, measure_vec: Sequence[Tuple[int, int]]) -> Dict[int, int]:
    """Integer count keys exactly as simulate_circuit builds them (qfvm.cpp:90-97 over
    measure_samples' bitstring, statevector.hpp:1397-1401): a string of len(measure_vec) chars, char
    `cbit` set from bit `qubit` of the sampled index (later pairs overwrite earlier ones), read as a
    big-endian binary number. The index 2^n (no index reached) has no bits set on qubits < n."""
    m = len(measure_vec)
    keys = np.zeros(indices.shape, dtype=np.uint64)
    for q, c in measure_vec:
        if c < 0 or c >= m:
            raise IndexError("basic_string::replace: cbit position exceeds the number of measured bits")
        w = np.uint64(1) << np.uint64(m - 1 - c)
        bit = (indices >> np.uint64(q)) & np.uint64(1)
        keys = (keys & ~w) | (bit * w)
    cnt = Counter(keys.tolist())
    return {int(k): int(v) for k, v in sorted(cnt.items())}


def simulate_program(
    prog: Program,
    inputstate: np.ndarray,
    shots: int,
    seed=None,
    dtype: str = "c128",
):
    """Core of simulate_circuit (qfvm.cpp:68-134) on a compiled program.

    Returns (counts: dict[int, int], state: ndarray complex128)."""
    if prog.n_cbits <= 0:
        # StateVector::set_creg (statevector.hpp:108-115)
        raise RuntimeError("The number of cbit must be positive.")
    has_input = inputstate is not None and np.asarray(inputstate).size != 0
    n = prog.n_qubits
    if has_input:
        n = _n_from_state(np.asarray(inputstate), prog.n_qubits)
    dev = _call(_native.pool.get, n, dtype)
    rng = np.random.default_rng(seed)
    shots = int(shots)

    def fresh():
        if has_input:
            _call(dev.load, inputstate)
        else:
            _call(dev.init_zero)

    if prog.final_measure:
        fresh()
        _call(dev.run, prog, rng.random(prog.n_random), prog.n_cbits, True)
        counts = {}
        if prog.measure_vec:
            idx = _call(dev.sample, rng.random(max(shots, 0)))
            counts = counts_from_indices(idx, prog.measure_vec)
        if has_input and isinstance(inputstate, np.ndarray) and inputstate.dtype == np.complex128 \
                and inputstate.flags.c_contiguous and inputstate.flags.writeable:
            _call(dev.store, inputstate)
            return counts, inputstate
        return counts, _call(dev.store)

    # per-shot path: a full re-simulation per shot, creg read out over the measured cbits
    measured = sorted({c for _, c in prog.measure_vec})
    outcount: Counter = Counter()
    for _ in range(max(shots, 1)):
        fresh()
        creg = _call(dev.run, prog, rng.random(prog.n_random), prog.n_cbits, False)
        if shots <= 0:
            break
        outcome = 0
        for j in measured:
            if j < prog.n_cbits:
                outcome = outcome * 2 + int(creg[j])
        outcount[outcome] += 1
    counts = {int(k): int(v) for k, v in sorted(outcount.items())}
    return counts, _call(dev.store)


def simulate_circuit(circuit, inputstate: np.ndarray = np.array([], dtype=complex), shots: int = 0, *, seed=None,
                     dtype: str = "c128"):
    """qfvm.simulate_circuit(circuit, inputstate, shots) -> (dict[int, int], ndarray)."""
    prog = compile_circuit(circuit)
    return simulate_program(prog, inputstate, shots, seed=seed, dtype=dtype)


def applyop_statevec(operation, inputstate: np.ndarray, *, dtype: str = "c128"):
    """Apply one gate to `inputstate` in place and return it (qfvm.cpp:36-50)."""
    psi = np.asarray(inputstate)
    if psi.size == 0:
        return inputstate
    n = psi.size.bit_length() - 1
    prog = compile_ops([operation], n)
    dev = _call(_native.pool.get, n, dtype)
    _call(dev.load, psi)
    _call(dev.apply, prog)
    if isinstance(inputstate, np.ndarray) and inputstate.dtype == np.complex128 and inputstate.flags.writeable \
            and inputstate.flags.c_contiguous:
        _call(dev.store, inputstate)
        return inputstate
    return _call(dev.store)


def expect_statevec(inputstate: np.ndarray, paulis) -> List[float]:
    """<psi|P|psi> for each PauliOp (`.pos`, `.paulistr`), unscaled (qfvm.cpp:246-261)."""
    psi = np.asarray(inputstate)
    n = psi.size.bit_length() - 1
    dev = _call(_native.pool.get, n, "c128")
    _call(dev.load, psi)
    return [float(v) for v in _call(dev.expect, [(p.paulistr, list(p.pos)) for p in paulis])]


def sampling_statevec(measures: dict, inputstate: np.ndarray, shots: int, *, seed=None) -> Dict[str, int]:
    """String-keyed counts of `shots` samples of `inputstate` (qfvm.cpp:52-66)."""
    psi = np.asarray(inputstate)
    n = psi.size.bit_length() - 1
    dev = _call(_native.pool.get, n, "c128")
    _call(dev.load, psi)
    idx = _call(dev.sample, _uniforms(seed, int(shots)))
    meas = [(int(q), int(c)) for q, c in measures.items()]
    m = len(meas)
    return bitstring_counts(idx, meas)


def bitstring_counts(indices: np.ndarray, measure_vec: Sequence[Tuple[int, int]]) -> Dict[str, int]:
    """counts_from_indices with the keys as measure_samples' strings (statevector.hpp:1381-1403):
    len(measure_vec) characters, "" when nothing is measured."""
    m = len(measure_vec)
    keys = counts_from_indices(np.asarray(indices, dtype=np.uint64), measure_vec)
    return {(bin(k)[2:].zfill(m) if m else ""): v for k, v in keys.items()}
