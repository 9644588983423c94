"""PauliOp and Hamiltonian (quafu/algorithms/hamiltonian.py:33-140)."""

from __future__ import annotations

from typing import Iterable

import numpy as np
import scipy.sparse as sp

from ..exceptions import QuafuError

IMat = sp.coo_matrix(np.array([[1.0, 0.0], [0.0, 1.0]], dtype=complex))
XMat = sp.coo_matrix(np.array([[0.0, 1.0], [1.0, 0.0]], dtype=complex))
YMat = sp.coo_matrix(np.array([[0.0, -1.0j], [1.0j, 0.0]], dtype=complex))
ZMat = sp.coo_matrix(np.array([[1.0, 0.0], [0.0, -1.0]], dtype=complex))
PauliMats = {"X": XMat, "Y": YMat, "Z": ZMat, "I": IMat}


class PauliOp:
    def __init__(self, paulis: str, coeff: complex = 1.0):
        self.paulistr = ""
        self.pos = []
        for p in paulis.split(" "):
            assert p[0] in "XYZ"
            self.paulistr += p[0]
            self.pos.append(int(p[1:]))
        self.coeff = coeff

    def __repr__(self):
        s = "" if self.coeff == 1.0 else str(self.coeff) + "*"
        for i, j in enumerate(self.pos):
            s += f"{self.paulistr[i]}{j}*"
        return s[:-1]

    __str__ = __repr__

    def get_matrix(self, qnum, big_endian=False):
        if qnum - 1 < max(self.pos):
            raise ValueError("The support of the paulis exceed the total qubit number")
        pos = np.array(self.pos)
        if not big_endian:
            pos = qnum - 1 - pos
        inds = np.argsort(pos)
        iq = ip = 0
        mat = 1.0
        while iq < qnum:
            if ip < len(pos) and iq == pos[inds[ip]]:
                mat = sp.kron(mat, PauliMats[self.paulistr[inds[ip]]])
                ip += 1
            else:
                mat = sp.kron(mat, PauliMats["I"])
            iq += 1
        return self.coeff * mat


class Hamiltonian:
    def __init__(self, paulis: list):
        self.paulis = paulis

    @staticmethod
    def from_pauli_list(pauli_list: Iterable):
        pauli_list = list(pauli_list)
        if len(pauli_list) == 0:
            raise QuafuError("Pauli list cannot be empty.")
        return Hamiltonian([PauliOp(p[0], coeff=p[1]) for p in pauli_list])

    def __repr__(self):
        return "+".join(str(p) for p in self.paulis)

    __str__ = __repr__

    def get_matrix(self, qnum, big_endian=False):
        mat = 0.0
        for pauli in self.paulis:
            mat += pauli.get_matrix(qnum, big_endian)
        return mat

    def to_pauli_list(self):
        return [[p.paulistr, p.pos] for p in self.paulis]
