"""Seeded benchmark and parity workloads (synthetic circuits; numpy PCG64 seeds).

Each builder fixes the order in which it draws random numbers, so the same seed gives the same
circuit on every machine:

* `random_u3_cx(n, depth, crz_per_layer, seed)`: per layer, angles (theta, phi, lambda) ~ U[0, 2pi)
  drawn as one (n, 3) array (qubit-major), U3 on every qubit, CX(2i, 2i+1) for all i, then
  CX(2i+1, 2i+2) for all i; then `crz_per_layer` CRZ gates, each drawn as an ordered pair of distinct
  qubits (rng.choice(n, 2, replace=False)) followed by theta ~ U[0, 2pi). 20 qubits, depth 20, no CRZ
  is the CPU-parity workload; 30 qubits, depth 40, 15 CRZ per layer the HBM workload.
* `hea_ansatz(n, layers, seed)`: per layer RY on every qubit, RZ on every qubit, CZ(i, i+1 mod n);
  2*n*layers parameters ~ U[-pi, pi] drawn as one array, in gate order.
* `two_target_dense(n, layers, seed)`: per layer U3 on every qubit (three angles per qubit as in
  random_u3_cx, qubit by qubit), then n // 2 gates on random pairs cycling SWAP, iSWAP, RXX, RYY
  (RXX / RYY draw their angle after the pair), then one controlled SWAP on three random qubits.
* `deep_blocks(n, block, layers, seed)`: per layer, angles drawn as in random_u3_cx and U3 on every
  qubit, then within each run of `block` consecutive qubits (0 .. block-1, block .. 2 block-1, ...;
  the last run may be shorter) CX(q, q+1) down the run: deep runs of gates on few qubits, whose
  gates never leave their run.
* `u3_runs(n, layers, cx_every, seed)`: per layer, angles drawn as in random_u3_cx and U3 on every
  qubit; after every `cx_every`-th layer CX(2i, 2i+1) for all i: runs of one-qubit gates.
* `tfim_hamiltonian(n, h)`: sum_i Z_i Z_{i+1 mod n} + h sum_i X_i.
* `qft(n, x)`: X on the set bits of x, then the swap-free QFT in textbook order: for target j from
  n-1 down to 0, H(j) then CP(k -> j, 2 pi / 2^(j-k+1)) for k = j-1 .. 0. Output amplitude of |y> is
  2^(-n/2) exp(2 pi i x rev(y) / 2^n), rev = bit reversal over n bits.
"""

from __future__ import annotations

import numpy as np

from .algorithms.hamiltonian import Hamiltonian, PauliOp
from .circuits import QuantumCircuit
from .elements import (CPGate, CRZGate, CSwapGate, CZGate, ISwapGate, Parameter, RXXGate, RYGate, RYYGate, RZGate,
                       SwapGate, U3Gate)


def random_u3_cx(n: int, depth: int, crz_per_layer: int = 0, seed: int = 20250001, measure: bool = True):
    rng = np.random.Generator(np.random.PCG64(seed))
    qc = QuantumCircuit(n)
    for _ in range(depth):
        ang = rng.uniform(0.0, 2 * np.pi, size=(n, 3))
        for q in range(n):
            qc << U3Gate(q, float(ang[q, 0]), float(ang[q, 1]), float(ang[q, 2]))
        for i in range(n // 2):
            qc.cx(2 * i, 2 * i + 1)
        for i in range((n - 1) // 2):
            qc.cx(2 * i + 1, 2 * i + 2)
        for _ in range(crz_per_layer):
            c, t = rng.choice(n, 2, replace=False)
            qc << CRZGate(int(c), int(t), float(rng.uniform(0.0, 2 * np.pi)))
    if measure:
        qc.measure(list(range(n)))
    return qc


def two_target_dense(n: int, layers: int, seed: int = 20250007):
    """Per layer: U3 on every qubit (angles uniform in [0, 2 pi)), then n // 2 gates on random qubit
    pairs cycling SWAP, iSWAP, RXX(theta), RYY(theta), and one controlled SWAP: gates that lower to
    dense 2-target matrices."""
    rng = np.random.Generator(np.random.PCG64(seed))
    qc = QuantumCircuit(n)
    for _ in range(layers):
        for q in range(n):
            qc << U3Gate(q, *map(float, rng.uniform(0, 2 * np.pi, 3)))
        for k in range(n // 2):
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            kind = k % 4
            if kind == 0:
                qc << SwapGate(a, b)
            elif kind == 1:
                qc << ISwapGate(a, b)
            elif kind == 2:
                qc << RXXGate(a, b, float(rng.uniform(0, 2 * np.pi)))
            else:
                qc << RYYGate(a, b, float(rng.uniform(0, 2 * np.pi)))
        c, a, b = (int(x) for x in rng.choice(n, 3, replace=False))
        qc << CSwapGate(c, a, b)
    return qc


def deep_blocks(n: int, block: int, layers: int, seed: int = 20250008):
    rng = np.random.Generator(np.random.PCG64(seed))
    qc = QuantumCircuit(n)
    for _ in range(layers):
        ang = rng.uniform(0.0, 2 * np.pi, size=(n, 3))
        for q in range(n):
            qc << U3Gate(q, float(ang[q, 0]), float(ang[q, 1]), float(ang[q, 2]))
        for b in range(0, n, block):
            for q in range(b, min(b + block, n) - 1):
                qc.cx(q, q + 1)
    return qc


def u3_runs(n: int, layers: int, cx_every: int = 2, seed: int = 20250009):
    rng = np.random.Generator(np.random.PCG64(seed))
    qc = QuantumCircuit(n)
    for layer in range(layers):
        ang = rng.uniform(0.0, 2 * np.pi, size=(n, 3))
        for q in range(n):
            qc << U3Gate(q, float(ang[q, 0]), float(ang[q, 1]), float(ang[q, 2]))
        if (layer + 1) % cx_every == 0:
            for i in range(n // 2):
                qc.cx(2 * i, 2 * i + 1)
    return qc


def hea_ansatz(n: int, layers: int, seed: int = 20250003, parametrised: bool = True):
    """Returns (circuit, values); gate parameters are `Parameter`s when `parametrised`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    values = rng.uniform(-np.pi, np.pi, size=2 * n * layers)
    qc = QuantumCircuit(n)
    k = 0
    for _ in range(layers):
        for q in range(n):
            v = Parameter(f"t{k}", values[k]) if parametrised else float(values[k])
            qc << RYGate(q, v)
            k += 1
        for q in range(n):
            v = Parameter(f"t{k}", values[k]) if parametrised else float(values[k])
            qc << RZGate(q, v)
            k += 1
        for i in range(n):
            qc << CZGate(i, (i + 1) % n)
    return qc, values


def batch_circuits(count: int = 256, seed: int = 20250004, n_range=(10, 22), depth_range=(4, 16)):
    """Uneven batch: per circuit, n ~ U{n_range} and depth ~ U{depth_range} drawn first (in that
    order), then 2 n d parameters ~ U[-pi, pi]; layers of RY and RZ on every qubit and a CZ ring
    (hea_ansatz's layout). Returns [(circuit, values)]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        d = int(rng.integers(depth_range[0], depth_range[1] + 1))
        values = rng.uniform(-np.pi, np.pi, size=2 * n * d)
        qc = QuantumCircuit(n)
        k = 0
        for _ in range(d):
            for q in range(n):
                qc << RYGate(q, Parameter(f"t{k}", values[k]))
                k += 1
            for q in range(n):
                qc << RZGate(q, Parameter(f"t{k}", values[k]))
                k += 1
            for i in range(n):
                qc << CZGate(i, (i + 1) % n)
        out.append((qc, values))
    return out


def z_sum(n: int) -> Hamiltonian:
    """sum_i Z_i."""
    return Hamiltonian([PauliOp(f"Z{i}") for i in range(n)])


def tfim_hamiltonian(n: int, h: float = 0.7) -> Hamiltonian:
    terms = [PauliOp(f"Z{i} Z{(i + 1) % n}") for i in range(n)]
    terms += [PauliOp(f"X{i}", h) for i in range(n)]
    return Hamiltonian(terms)


def qft(n: int, x: int, measure: bool = False):
    qc = QuantumCircuit(n)
    for q in range(n):
        if (x >> q) & 1:
            qc.x(q)
    for j in range(n - 1, -1, -1):
        qc.h(j)
        for k in range(j - 1, -1, -1):
            qc << CPGate(k, j, 2 * np.pi / 2 ** (j - k + 1))
    if measure:
        qc.measure(list(range(n)))
    return qc


def qft_amplitudes(n: int, x: int, indices=None) -> np.ndarray:
    """Closed form of qft(n, x)'s output (complex128) at `indices` (default: every basis state),
    with the phase x * rev(y) taken exactly as an integer mod 2^n."""
    y = np.arange(1 << n, dtype=np.int64) if indices is None else np.asarray(indices, dtype=np.int64)
    rev = np.zeros_like(y)
    for b in range(n):
        rev |= ((y >> b) & 1) << (n - 1 - b)
    # products wrap mod 2^64, a multiple of 2^n, so the masked result is exact for every n <= 34
    ph = (np.uint64(x) * rev.astype(np.uint64)) & np.uint64((1 << n) - 1)
    return np.exp(2j * np.pi * (ph.astype(np.float64) / (1 << n))) / np.sqrt(1 << n)
