"""Pins the oracle's gradient routines (oracle/gradients.py) on the CPU: closed forms, the exact
parameter-shift rule for uncontrolled rotations, and central finite differences."""

import numpy as np
import pytest

from oracle import gradients as og
from pyquafu_amd.simulators.program import GENERAL, RX, RY, RZ, H, X, ProgramBuilder
from tests.progs import _rot, random_pauli_terms, rotation_program


def one_rotation(code, pre=()):
    def make(th):
        b = ProgramBuilder()
        for c in pre:
            b.add(c, [0])
        b.add(code, [0], param=float(th[0]))
        return b.program(n_qubits=1)

    return make


@pytest.mark.parametrize("theta", [-2.3, -0.4, 0.0, 0.9, 2.7])
def test_closed_forms(theta):
    cases = [
        (RY, (), "Z", np.cos(theta), -np.sin(theta), "Y"),
        (RX, (), "Y", -np.sin(theta), -np.cos(theta), "X"),
        (RZ, (H,), "X", np.cos(theta), -np.sin(theta), "Z"),
    ]
    for code, pre, obs, e_want, g_want, gen in cases:
        make = one_rotation(code, pre)
        prog = make([theta])
        g, e = og.adjoint(prog, 1, [len(pre)], gen, [(obs, [0])], [1.0])
        assert abs(e - e_want) < 1e-14
        assert abs(g[0] - g_want) < 1e-14


@pytest.mark.parametrize("control_set", [True, False])
def test_controlled_rotation_closed_form(control_set):
    theta = 0.77

    def make(th):
        b = ProgramBuilder()
        if control_set:
            b.add(X, [0])
        b.add(GENERAL, [0, 1], n_ctrl=1, matrix=_rot("Y", float(th[0])))
        return b.program(n_qubits=2)

    k = 1 if control_set else 0
    g, e = og.adjoint(make([theta]), 2, [k], "Y", [("Z", [1])], [1.0])
    if control_set:
        assert abs(e - np.cos(theta)) < 1e-14 and abs(g[0] + np.sin(theta)) < 1e-14
    else:
        assert abs(e - 1.0) < 1e-14 and abs(g[0]) < 1e-14


@pytest.mark.parametrize("seed", range(4))
def test_adjoint_equals_parameter_shift(seed):
    # uncontrolled rotations only: the two-term shift rule is exact for them
    rng = np.random.Generator(np.random.PCG64(900 + seed))
    n = 4
    make, thetas, ops, gens = rotation_program(n, 30, rng, max_ctrl=0)
    terms, coeffs = random_pauli_terms(n, 5, rng)
    g, _ = og.adjoint(make(thetas), n, ops, gens, terms, coeffs)
    ps = og.param_shift(make, thetas, n, terms, coeffs)
    assert len(g) == len(ps) == len(thetas) > 0
    assert np.abs(g - ps).max() < 1e-12


@pytest.mark.parametrize("seed", range(3))
def test_adjoint_equals_finite_differences_with_controls(seed):
    rng = np.random.Generator(np.random.PCG64(950 + seed))
    n = 5
    make, thetas, ops, gens = rotation_program(n, 30, rng, max_ctrl=2)
    terms, coeffs = random_pauli_terms(n, 4, rng)
    g, e = og.adjoint(make(thetas), n, ops, gens, terms, coeffs)
    fd = og.finite_diff(make, thetas, n, terms, coeffs, h=1e-6)
    assert abs(e - og.energy(make, thetas, n, terms, coeffs)) < 1e-12
    assert np.abs(g - fd).max() < 1e-7


def test_hamiltonian_apply_matches_dense_matrices():
    rng = np.random.Generator(np.random.PCG64(5))
    n = 4
    terms, coeffs = random_pauli_terms(n, 6, rng)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    from oracle import dense

    want = sum(c * (dense.pauli_matrix_le(s, p, n) @ psi) for (s, p), c in zip(terms, coeffs))
    assert np.abs(og.hamiltonian_apply(psi, n, terms, coeffs) - want).max() < 1e-12


# ------------------------------------------------------------------ the C++ oracle's grad_adjoint
@pytest.mark.parametrize("theta", [-2.3, 0.0, 0.9])
def test_cpp_adjoint_closed_forms(theta):
    """qfvm_adjoint of the C++ oracle (oracle/qfvm_cpu.cpp) on the one-rotation closed forms."""
    from oracle.cpu import CpuState

    cases = [
        (RY, (), "Z", np.cos(theta), -np.sin(theta), "Y"),
        (RX, (), "Y", -np.sin(theta), -np.cos(theta), "X"),
        (RZ, (H,), "X", np.cos(theta), -np.sin(theta), "Z"),
    ]
    for code, pre, obs, e_want, g_want, gen in cases:
        prog = one_rotation(code, pre)([theta])
        s = CpuState(1)
        g, e = s.adjoint(prog, [len(pre)], gen, [(obs, [0])], [1.0])
        assert abs(e - e_want) < 1e-14
        assert abs(g[0] - g_want) < 1e-14


@pytest.mark.parametrize("seed,max_ctrl", [(0, 0), (1, 1), (2, 2), (3, 2)])
def test_cpp_adjoint_matches_dense_oracle(seed, max_ctrl):
    """The C++ restatement against the dense NumPy one (itself pinned above by closed forms, the
    exact shift rule and finite differences) on random rotation programs with controls, complex
    coefficients and a random input state; the state left behind is the one before the first
    differentiated op."""
    from oracle.cpu import CpuState
    from tests.conftest import rand_state

    rng = np.random.Generator(np.random.PCG64(1200 + seed))
    n = 6
    make, thetas, ops, gens = rotation_program(n, 40, rng, max_ctrl=max_ctrl)
    terms, coeffs = random_pauli_terms(n, 5, rng)
    coeffs = [complex(c) for c in coeffs]
    psi0 = rand_state(n, rng)
    prog = make(thetas)
    g_want, e_want = og.adjoint(prog, n, ops, gens, terms, coeffs, psi0=psi0)
    s = CpuState(n)
    s.load(psi0)
    g, e = s.adjoint(prog, ops, gens, terms, coeffs)
    assert abs(e - e_want) < 1e-12
    assert np.abs(g - g_want).max() < 1e-12
    from oracle import dense

    r = dense.DenseRunner(prog, n, psi0)
    r.exec(0, min(ops))
    assert np.abs(s.store() - r.psi).max() < 1e-12
