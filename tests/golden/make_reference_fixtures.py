"""Writes tests/golden/reference_fixtures.npz from data the reference's own tests hold.

Run in a container that has the reference checkout at /root/reference (read as text / data only,
nothing is imported or executed from it):

    python tests/golden/make_reference_fixtures.py

Contents (all inputs and expected outputs, no code):
  * M_0, M_1, M_2: dense 3-qubit Hamiltonian matrices written out literally in
    tests/quafu/algorithms/hamiltonian_test.py:18-271, expected for the Pauli lists of
    test_to_matrix (hamiltonian_test.py:278-292); parsed from the file with ast.literal_eval.
  * QAOA_H, VQE_H: tests/quafu/algorithms/data/{qaoa,vqe}_hamiltonian.npy (numpy, no pickle),
    expected for the Pauli lists at integration_test.py:111-117 and 144-149.
"""

import ast
import json
import os

import numpy as np

REF = "/root/reference/tests/quafu/algorithms"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_fixtures.npz")

PAULI_LISTS = {
    "M_0": (3, [("Z0 Z1", 1), ("Z0 Z2", 1), ("Z1 Z2", 1)]),
    "M_1": (3, [("X0 Z1", 1), ("X0 Z2", 1), ("X1 Z2", 1)]),
    "M_2": (3, [("Z1 Y2", 1), ("Y0 Z2", 1), ("X1 Z2", 1)]),
    "QAOA_H": (5, [("Z0 Z1", 1), ("Z0 Z2", 1), ("Z0 Z3", 1), ("Z0 Z4", 1)]),
    "VQE_H": (2, [("Z0 Y1", 0.3980), ("Z1", -0.3980), ("Z0 Z1", -0.0113), ("X0 X1", 0.1810)]),
}


def _literal_matrices(path):
    tree = ast.parse(open(path, encoding="utf-8").read())
    out = {}
    for node in tree.body:
        if isinstance(node, ast.Assign) and len(node.targets) == 1 and isinstance(node.targets[0], ast.Name):
            name = node.targets[0].id
            if name.startswith("M_") and isinstance(node.value, ast.Call):
                out[name] = np.array(ast.literal_eval(node.value.args[0]), dtype=complex)
    return out


def main():
    arrays = _literal_matrices(os.path.join(REF, "hamiltonian_test.py"))
    arrays["QAOA_H"] = np.load(os.path.join(REF, "data", "qaoa_hamiltonian.npy"), allow_pickle=False)
    arrays["VQE_H"] = np.load(os.path.join(REF, "data", "vqe_hamiltonian.npy"), allow_pickle=False)
    meta = {k: {"n": n, "paulis": [[s, c] for s, c in lst]} for k, (n, lst) in PAULI_LISTS.items()}
    np.savez_compressed(OUT, meta=np.array(json.dumps(meta)), **arrays)
    print("wrote", OUT, sorted(arrays))


if __name__ == "__main__":
    main()
