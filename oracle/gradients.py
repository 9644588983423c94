"""ORACLE -- test infrastructure, not product code.

CPU restatement of the reference's gradient routines (quafu/algorithms/gradients/gradiant.py) on
programs (include/qfvm.h op records), in dense NumPy (oracle/dense.py), for small n:

* hamiltonian_apply: _apply_hamil (quafu/simulators/simulator.py:78-88): for every Pauli term, a
  copy of psi with the term's X / Y / Z gates applied one by one, times the coefficient, summed.
* expectation: sum_t c_t <psi| P_t |psi> with dense little-endian Pauli matrices.
* adjoint: grad_adjoint (gradiant.py:144-173) step for step: forward run; lam = H psi; for every op
  from the last down to the first differentiated one: phi <- U^dagger phi; for a differentiated op,
  mu = projCtrl (dU phi) with dU = -1/2 i P U on the target (grad_gate, gradiant.py:96-141) and the
  gradient Re(2 <lam, mu>); then lam <- U^dagger lam.
* param_shift: (E(theta + pi/2) - E(theta - pi/2)) / 2 per parameter (gradiant.py:38-73).
* finite_diff: central differences with step h (gradiant.py:76-93).

`make(thetas) -> program` rebuilds the program for a parameter vector (param_shift, finite_diff).
"""

from __future__ import annotations

import numpy as np

from . import dense

PAULI = {"X": dense.XM, "Y": dense.YM, "Z": dense.ZM}


def hamiltonian_apply(psi: np.ndarray, n: int, terms, coeffs) -> np.ndarray:
    out = np.zeros_like(psi, dtype=complex)
    for (s, pos), c in zip(terms, coeffs):
        v = np.array(psi, dtype=complex)
        for letter, q in zip(s, pos):
            if letter != "I":
                v = dense.apply_unitary(v, n, [], [int(q)], PAULI[letter])
        out += v * c
    return out


def expectation(psi: np.ndarray, n: int, terms, coeffs) -> float:
    tot = 0.0
    for (s, pos), c in zip(terms, coeffs):
        m = dense.pauli_matrix_le(s, pos, n)
        tot += (c * np.vdot(psi, m @ psi)).real
    return float(tot)


def _ctrl_projector(psi: np.ndarray, n: int, ctrls) -> np.ndarray:
    idx = np.arange(1 << n)
    keep = np.ones(1 << n, dtype=bool)
    for c in ctrls:
        keep &= ((idx >> c) & 1) == 1
    return np.where(keep, psi, 0.0)


def adjoint(prog, n: int, grad_ops, gens: str, terms, coeffs, psi0=None):
    """Returns (grads in grad_ops order, <H>)."""
    r = dense.DenseRunner(prog, n, psi0)
    r.exec(0, len(prog.ops))
    phi = r.psi.copy()
    lam = hamiltonian_apply(phi, n, terms, coeffs)
    e = float(np.vdot(phi, lam).real)
    slot = {int(k): g for g, k in enumerate(grad_ops)}
    begin = min(slot) if slot else len(prog.ops)
    grads = np.zeros(len(grad_ops))
    for k in range(len(prog.ops) - 1, begin - 1, -1):
        ctrls, targs, u = r.gate_parts(prog.ops[k])
        ud = u.conj().T
        phi = dense.apply_unitary(phi, n, ctrls, targs, ud)
        if k in slot:
            g = slot[k]
            deriv = -0.5j * PAULI[gens[g]] @ u
            mu = dense.apply_unitary(phi, n, [], targs, deriv)
            mu = _ctrl_projector(mu, n, ctrls)
            grads[g] = np.real(2.0 * np.inner(lam.conj(), mu))
        lam = dense.apply_unitary(lam, n, ctrls, targs, ud)
    return grads, e


def energy(make, thetas, n: int, terms, coeffs, psi0=None) -> float:
    psi, _ = dense.run_dense(make(np.asarray(thetas, dtype=float)), n, psi0)
    return expectation(psi, n, terms, coeffs)


def param_shift(make, thetas, n: int, terms, coeffs, psi0=None) -> np.ndarray:
    thetas = np.asarray(thetas, dtype=float)
    out = np.zeros(thetas.size)
    for k in range(thetas.size):
        tp, tm = thetas.copy(), thetas.copy()
        tp[k] += np.pi / 2
        tm[k] -= np.pi / 2
        out[k] = (energy(make, tp, n, terms, coeffs, psi0) - energy(make, tm, n, terms, coeffs, psi0)) / 2.0
    return out


def finite_diff(make, thetas, n: int, terms, coeffs, h: float = 1e-6, psi0=None) -> np.ndarray:
    thetas = np.asarray(thetas, dtype=float)
    out = np.zeros(thetas.size)
    for k in range(thetas.size):
        tp, tm = thetas.copy(), thetas.copy()
        tp[k] += h
        tm[k] -= h
        out[k] = (energy(make, tp, n, terms, coeffs, psi0) - energy(make, tm, n, terms, coeffs, psi0)) / (2 * h)
    return out
