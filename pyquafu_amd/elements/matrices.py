"""Gate matrices handed to the engine.

The formulas are those of quafu/elements/matrices/mat_lib.py:18-224 (numpy complex128,
row-major, most significant qubit first for multi-qubit matrices). The engine receives them
exactly as the reference's pybind layer receives `QuantumGate.matrix` / `_get_targ_matrix`
(src/qfvm/instructions.hpp:250-255), so the same numpy trig calls give the same bits.
"""

from typing import List

import numpy as np

IdMatrix = np.eye(2, dtype=complex)
XMatrix = np.array([[0.0, 1.0], [1.0, 0.0]], dtype=complex)
YMatrix = np.array([[0.0, -1.0j], [1.0j, 0.0]], dtype=complex)
ZMatrix = np.array([[1.0, 0.0], [0.0, -1.0]], dtype=complex)
SMatrix = np.array([[1.0, 0.0], [0.0, 1.0j]], dtype=complex)
SXMatrix = np.array([[1.0, 1.0j], [1.0j, 1.0]], dtype=complex) / np.sqrt(2)
SYMatrix = np.array([[1.0, -1.0], [1.0, 1.0]], dtype=complex) / np.sqrt(2)
TMatrix = np.array([[1.0, 0.0], [0.0, np.exp(1.0j * np.pi / 4)]], dtype=complex)
WMatrix = (XMatrix + YMatrix) / np.sqrt(2)
SWMatrix = np.array([[0.5 + 0.5j, -np.sqrt(0.5) * 1j], [np.sqrt(0.5), 0.5 + 0.5j]], dtype=complex)
HMatrix = (XMatrix + ZMatrix) / np.sqrt(2)


def _perm4(a, b, c, d, e=1.0):
    m = np.zeros((4, 4), dtype=complex)
    m[0, a] = 1.0
    m[1, b] = e
    m[2, c] = e
    m[3, d] = 1.0
    return m


SwapMatrix = _perm4(0, 2, 1, 3)
ISwapMatrix = _perm4(0, 2, 1, 3, 1.0j)
CXMatrix = np.eye(4, dtype=complex)[[0, 1, 3, 2]]
CYMatrix = np.eye(4, dtype=complex)
CYMatrix[2:, 2:] = YMatrix
CZMatrix = np.diag([1.0, 1.0, 1.0, -1.0]).astype(complex)
CTMatrix = np.diag([1.0, 1.0, 1.0, np.exp(1.0j * np.pi / 4)]).astype(complex)
ToffoliMatrix = np.eye(8, dtype=complex)[[0, 1, 2, 3, 4, 5, 7, 6]]
FredkinMatrix = np.eye(8, dtype=complex)[[0, 1, 2, 3, 4, 6, 5, 7]]


def rx_mat(theta):
    return np.array(
        [
            [np.cos(0.5 * theta), -1.0j * np.sin(0.5 * theta)],
            [-1.0j * np.sin(0.5 * theta), np.cos(0.5 * theta)],
        ],
        dtype=complex,
    )


def ry_mat(theta):
    return np.array(
        [[np.cos(0.5 * theta), -np.sin(0.5 * theta)], [np.sin(0.5 * theta), np.cos(0.5 * theta)]],
        dtype=complex,
    )


def rz_mat(theta):
    return np.array([[np.exp(-0.5j * theta), 0.0], [0.0, np.exp(0.5j * theta)]], dtype=complex)


def pmatrix(labda):
    return np.array([[1, 0], [0, np.exp(1j * labda)]], dtype=complex)


def rxx_mat(theta):
    c, s = np.cos(theta / 2), -1j * np.sin(theta / 2)
    return np.array([[c, 0, 0, s], [0, c, s, 0], [0, s, c, 0], [s, 0, 0, c]])


def ryy_mat(theta):
    c = np.cos(theta / 2)
    s = 1j * np.sin(theta / 2)
    return np.array([[c, 0, 0, +s], [0, c, -s, 0], [0, -s, c, 0], [+s, 0, 0, c]])


def rzz_mat(theta):
    a, b = np.exp(-1j * theta / 2), np.exp(1j * theta / 2)
    return np.diag([a, b, b, a])


def u2matrix(_phi=0.0, _lambda=0.0):
    return np.array(
        [[1.0, np.exp(-1.0j * _lambda)], [np.exp(1.0j * _phi), np.exp((_phi + _lambda) * 1.0j)]],
        dtype=complex,
    )


def u3matrix(_theta=0.0, _phi=0.0, _lambda=0.0):
    """OpenQASM 3.0 U gate (mat_lib.py:186-197)."""
    return np.array(
        [
            [np.cos(0.5 * _theta), -np.exp(_lambda * 1.0j) * np.sin(0.5 * _theta)],
            [
                np.exp(_phi * 1.0j) * np.sin(0.5 * _theta),
                np.exp((_phi + _lambda) * 1.0j) * np.cos(0.5 * _theta),
            ],
        ],
        dtype=complex,
    )


mat_dict = {
    "id": IdMatrix,
    "x": XMatrix,
    "y": YMatrix,
    "z": ZMatrix,
    "h": HMatrix,
    "w": WMatrix,
    "s": SMatrix,
    "t": TMatrix,
    "cx": CXMatrix,
    "cnot": CXMatrix,
    "cy": CYMatrix,
    "cz": CZMatrix,
    "swap": SwapMatrix,
    "iswap": ISwapMatrix,
    "rx": rx_mat,
    "ry": ry_mat,
    "rz": rz_mat,
    "p": pmatrix,
    "rxx": rxx_mat,
    "ryy": ryy_mat,
    "rzz": rzz_mat,
    "u2": u2matrix,
    "u3": u3matrix,
}


def reorder_matrix(matrix: np.ndarray, pos: List):
    """Reorder a matrix written for sorted qubits into the order of `pos` (mat_utils.py:25-32)."""
    qnum = len(pos)
    dim = 2**qnum
    inds = np.argsort(pos)
    inds = np.concatenate([inds, inds + qnum])
    tensorm = np.reshape(matrix, [2] * 2 * qnum)
    return np.transpose(tensorm, inds).reshape([dim, dim])
