"""SimuResult (quafu/results/results.py:101-246), statevector subset."""

from __future__ import annotations

import copy
from typing import Iterable, List, Union

import numpy as np


def permutebits(mat: np.ndarray, order: Iterable) -> np.ndarray:
    """Permute the qubit axes of a state or operator (default_simulator.py:50-60)."""
    order = np.array(list(order))
    num = len(order)
    r = len(mat.shape)
    mat = np.reshape(mat, [2] * r * num)
    order = np.concatenate([order + num * i for i in range(r)])
    mat = np.transpose(mat, order)
    return np.reshape(mat, [2**num] * r)


def ptrace(psi: np.ndarray, ind_a: List, diag: bool = True) -> np.ndarray:
    """Reduced probabilities (diag) or density matrix of qubits ind_a (default_simulator.py:63-76)."""
    num = int(np.log2(psi.shape[0]))
    order = copy.deepcopy(list(ind_a))
    order.extend([p for p in range(num) if p not in ind_a])
    psi = permutebits(psi, order)
    if diag:
        psi = np.abs(psi) ** 2
        psi = np.reshape(psi, [2 ** len(ind_a), 2 ** (num - len(ind_a))])
        return np.sum(psi, axis=1)
    psi = np.reshape(psi, [2 ** len(ind_a), 2 ** (num - len(ind_a))])
    return psi @ np.conj(np.transpose(psi))


class Result:
    """Base class of results."""


class SimuResult(Result):
    def __init__(self, res_info: dict):
        self._meta_data = res_info
        counts = res_info["counts"]
        if counts and isinstance(list(counts.keys())[0], (int, np.integer)):
            cbits_num = len(res_info["measures"])
            self._meta_data["counts"] = {bin(k)[2:].zfill(cbits_num): v for k, v in counts.items()}
        self._probabilities = []

    def __getitem__(self, key: str):
        return self._meta_data[key]

    def get_statevector(self):
        try:
            return self["statevector"]
        except KeyError as exc:
            raise KeyError(f"no statevector saved from {self['simulator']} simulator") from exc

    @property
    def probabilities(self):
        if len(self._probabilities) == 0:
            self.calc_probabilities()
        return self._probabilities

    @property
    def counts(self):
        return self["counts"]

    def calc_probabilities(self, from_counts=False):
        num = self["qbitnum"]
        if from_counts and self._meta_data["counts"]:
            counts = self._meta_data["counts"]
            total = sum(counts.values())
            self._probabilities = np.zeros(2**num)
            for key, v in counts.items():
                self._probabilities[int(key, 2)] = v / total
        elif self._meta_data.get("probabilities") is not None:
            # measured-qubit marginals reduced on the device by the simulator, same layout
            self._probabilities = np.asarray(self._meta_data["probabilities"])
        elif "statevector" in self._meta_data:
            psi = self.get_statevector()
            measures = list(self["measures"].keys())
            values = np.argsort(list(self["measures"].values()))
            psi = permutebits(psi, range(num)[::-1])
            if measures:
                self._probabilities = permutebits(ptrace(psi, measures), values)
            else:
                self._probabilities = np.abs(psi) ** 2
        else:
            raise ValueError("No data saved for probs")

    def calc_density_matrix(self):
        psi = self.get_statevector()
        num = self["qbitnum"]
        measures = list(self["measures"].keys())
        values: Union[list, np.ndarray] = np.argsort(list(self["measures"].values()))
        if len(measures) == 0:
            measures = list(range(num))
            values = list(range(num))
        psi = permutebits(psi, range(num)[::-1])
        rho = ptrace(psi, measures, diag=False)
        return permutebits(rho, values)
