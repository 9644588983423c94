"""Pins the oracle before anything is checked against it (CPU only).

* the dense NumPy runner (oracle/dense.py) against closed forms: Bell / GHZ states, gate identities,
  the swap-free QFT against numpy's FFT;
* the C++ restatement (oracle/qfvm_cpu.cpp) against the dense runner on random programs that use
  every op code, controls anywhere and dense k-target matrices;
* expect_pauli against dense Pauli matrices, which are themselves pinned by the reference's
  Hamiltonian fixtures (tests/golden/reference_fixtures.npz);
* the sampler (one sequential pass + binary search) against the reference's per-shot scan,
  exact dyadic cases and deterministic states;
* measure / reset / cif against the dense runner with the same uniforms.
"""

import json
import os

import numpy as np
import pytest

from oracle import dense
from oracle.cpu import OP_DTYPE as ORACLE_OP_DTYPE
from oracle.cpu import CpuState
from pyquafu_amd.simulators.program import OP_DTYPE, ProgramBuilder
from tests.conftest import rand_state
from tests.progs import random_classical_program, random_gate_program

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.npz")


def prog_of(*ops, n):
    b = ProgramBuilder()
    for o in ops:
        b.add(*o[:2], **(o[2] if len(o) > 2 else {}))
    return b.program(n_qubits=n, n_cbits=n)


def test_op_layout_matches_header():
    assert ORACLE_OP_DTYPE == OP_DTYPE
    assert OP_DTYPE.itemsize == 64
    assert [OP_DTYPE.fields[k][1] for k in OP_DTYPE.names] == [0, 4, 8, 12, 16, 24, 32, 40, 48, 56]


# ------------------------------------------------------------------ dense runner vs closed forms
def test_dense_bell_and_ghz():
    p = prog_of((dense.H, [0]), (dense.CX, [0, 1], {"n_ctrl": 1}), n=2)
    psi, _ = dense.run_dense(p, 2)
    assert np.allclose(psi, [2**-0.5, 0, 0, 2**-0.5], atol=1e-15)
    p = prog_of((dense.H, [0]), (dense.CX, [0, 1], {"n_ctrl": 1}), (dense.CX, [1, 2], {"n_ctrl": 1}), n=3)
    psi, _ = dense.run_dense(p, 3)
    want = np.zeros(8, complex)
    want[0] = want[7] = 2**-0.5
    assert np.allclose(psi, want, atol=1e-15)


def test_dense_identities(rng):
    n = 3
    psi0 = rand_state(n, rng)
    # H Z H = X on qubit 1
    a, _ = dense.run_dense(prog_of((dense.H, [1]), (dense.Z, [1]), (dense.H, [1]), n=n), n, psi0)
    b, _ = dense.run_dense(prog_of((dense.X, [1]), n=n), n, psi0)
    assert np.allclose(a, b, atol=1e-14)
    # CX is an involution and permutes basis states |c t> -> |c, t^c>
    c, _ = dense.run_dense(prog_of((dense.CX, [2, 0], {"n_ctrl": 1}), (dense.CX, [2, 0], {"n_ctrl": 1}), n=n), n, psi0)
    assert np.allclose(c, psi0, atol=1e-15)
    e, _ = dense.run_dense(prog_of((dense.CX, [2, 0], {"n_ctrl": 1}), n=n), n, np.eye(8)[0b100])
    assert np.allclose(e, np.eye(8)[0b101])
    # U3(t, -pi/2, pi/2) = RX(t); U3(t, 0, 0) = RY(t)
    t = 0.731
    assert np.allclose(dense.u3(t, -np.pi / 2, np.pi / 2), dense.rx(t), atol=1e-15)
    assert np.allclose(dense.u3(t, 0, 0), dense.ry(t), atol=1e-15)
    # RZ(t) = exp(-i t/2) P(t)
    assert np.allclose(dense.rz(t), np.exp(-0.5j * t) * dense.pm(t), atol=1e-15)
    # target bit j of the matrix belongs to targs[j]: a 2-target matrix that is X (x) I in LSB-first
    # order acts as X on targs[0]
    m = np.kron(np.eye(2), dense.XM)  # row bit 0 <-> targs[0]
    f, _ = dense.run_dense(prog_of((dense.MULTI, [2, 0], {"matrix": m}), n=n), n, np.eye(8)[0])
    assert np.allclose(f, np.eye(8)[0b100])


def _full_matrix(n, ctrls, targs, U):
    """The whole 2^n x 2^n operator of a controlled k-target gate, built entry by entry from basis
    states: column i maps to rows with the target bits replaced (U[r', r], bit j of r <-> targs[j])
    where every control bit of i is set, and to itself elsewhere."""
    N = 1 << n
    M = np.zeros((N, N), dtype=complex)
    for i in range(N):
        if not all((i >> c) & 1 for c in ctrls):
            M[i, i] = 1.0
            continue
        r = sum(((i >> t) & 1) << j for j, t in enumerate(targs))
        base = i & ~sum(1 << t for t in targs)
        for r2 in range(1 << len(targs)):
            row = base | sum(((r2 >> j) & 1) << t for j, t in enumerate(targs))
            M[row, i] += U[r2, r]
    return M


@pytest.mark.parametrize("seed", range(6))
def test_multi_target_dense_against_full_matrix(seed):
    """Brute force on 5 qubits: the oracle's controlled k-target gate (k = 1..4, 0..2 controls, any
    target order) equals the explicit 2^n x 2^n operator applied to the state, and a Kronecker
    product gate equals its factors applied one target at a time."""
    rng = np.random.Generator(np.random.PCG64(4200 + seed))
    n = 5
    k = int(rng.integers(1, 5))
    nc = int(rng.integers(0, min(2, n - k) + 1))
    qs = [int(x) for x in rng.permutation(n)]
    targs, ctrls = qs[:k], qs[k:k + nc]
    a = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
    U, _ = np.linalg.qr(a)
    psi = rand_state(n, rng)
    got = dense.apply_unitary(psi, n, ctrls, targs, U)
    assert np.allclose(got, _full_matrix(n, ctrls, targs, U) @ psi, atol=1e-13)
    if k >= 2:
        A, _ = np.linalg.qr(rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)))
        B, _ = np.linalg.qr(rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2)))
        # row bit 0 <-> targs[0]: kron(A, B) is B on targs[0] and A on targs[1]
        both = dense.apply_unitary(psi, n, ctrls, targs[:2], np.kron(A, B))
        step = dense.apply_unitary(dense.apply_unitary(psi, n, ctrls, [targs[0]], B), n, ctrls, [targs[1]], A)
        assert np.allclose(both, step, atol=1e-13)


@pytest.mark.parametrize("n", [1, 3, 6, 9])
def test_qft_closed_form_against_fft(n):
    x = (0b1011011 * 7919) % (1 << n)
    amp = dense.qft_closed_form(n, x)
    # swap-free QFT output is the DFT read at bit-reversed indices
    rev = np.array([int(format(y, f"0{n}b")[::-1], 2) for y in range(1 << n)])
    e = np.zeros(1 << n)
    e[x] = 1.0
    dft = np.fft.ifft(e) * np.sqrt(1 << n)
    assert np.allclose(amp[rev], dft, atol=1e-12)


@pytest.mark.parametrize("n", [2, 5, 8])
def test_qft_circuit_matches_closed_form(n):
    from pyquafu_amd import workloads
    from pyquafu_amd.simulators.program import compile_circuit

    x = (12345 * n) % (1 << n)
    prog = compile_circuit(workloads.qft(n, x))
    psi, _ = dense.run_dense(prog, n)
    assert np.abs(psi - dense.qft_closed_form(n, x)).max() < 1e-13
    assert np.abs(workloads.qft_amplitudes(n, x) - dense.qft_closed_form(n, x)).max() < 1e-15


# ------------------------------------------------------------------ C++ oracle vs dense runner
@pytest.mark.parametrize("n,seed", [(1, 1), (2, 2), (3, 3), (5, 4), (7, 5), (9, 6)])
def test_cpp_oracle_matches_dense(n, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    prog = random_gate_program(n, 60, rng)
    psi0 = rand_state(n, rng)
    want, _ = dense.run_dense(prog, n, psi0)
    s = CpuState(n)
    s.load(psi0)
    s.apply(prog)
    got = s.store()
    assert np.abs(got - want).max() < 1e-12
    assert abs(s.norm2() - 1.0) < 1e-12


def test_cpp_oracle_target0_control0_and_unroll_paths():
    # exercise the three loop shapes of the reference kernels: target 0, control 0, unroll-by-2
    rng = np.random.Generator(np.random.PCG64(99))
    n = 6
    for ctrl, targ in [(1, 0), (0, 1), (0, 5), (3, 2), (5, 4), (2, 5)]:
        for code in (dense.CX, dense.CZ, dense.CP):
            prog = prog_of((code, [ctrl, targ], {"n_ctrl": 1, "param": 0.37}), n=n)
            psi0 = rand_state(n, rng)
            want, _ = dense.run_dense(prog, n, psi0)
            s = CpuState(n)
            s.load(psi0)
            s.apply(prog)
            assert np.abs(s.store() - want).max() < 1e-13


def test_cpp_oracle_errors():
    s = CpuState(3)
    with pytest.raises(Exception) as e:
        s.apply(prog_of((dense.X, [3]), n=3))
    assert e.value.code == 2  # QFVM_ERR_INDEX
    with pytest.raises(Exception) as e:
        s.apply(prog_of((dense.CX, [1, 1], {"n_ctrl": 1}), n=3))
    assert e.value.code == 1


# ------------------------------------------------------------------ expectations
def test_pauli_matrix_against_reference_fixtures():
    d = np.load(GOLDEN, allow_pickle=False)
    meta = json.loads(str(d["meta"]))
    for key, info in meta.items():
        n = info["n"]
        tot = np.zeros((1 << n, 1 << n), complex)
        for s, c in info["paulis"]:
            parts = s.split(" ")
            tot += c * dense.pauli_matrix_le("".join(p[0] for p in parts), [int(p[1:]) for p in parts], n)
        assert np.array_equal(tot, d[key]) or np.abs(tot - d[key]).max() < 1e-15, key


def test_expect_pauli_against_dense(rng):
    n = 6
    psi = rand_state(n, rng)
    s = CpuState(n)
    s.load(psi)
    terms = [("Z", [0]), ("ZZ", [1, 4]), ("X", [5]), ("XY", [0, 3]), ("YZX", [2, 0, 5]), ("Y", [1]),
             ("XXYY", [0, 1, 2, 3]), ("IZ", [2, 3])]
    got = s.expect(terms)
    for (ps, pos), g in zip(terms, got):
        m = dense.pauli_matrix_le(ps, pos, n)
        want = np.real(np.vdot(psi, m @ psi))
        assert abs(g - want) < 1e-13, (ps, pos)


def test_expect_against_reference_hamiltonian_fixtures(rng):
    d = np.load(GOLDEN, allow_pickle=False)
    meta = json.loads(str(d["meta"]))
    for key, info in meta.items():
        n = info["n"]
        psi = rand_state(n, rng)
        s = CpuState(n)
        s.load(psi)
        terms, coeffs = [], []
        for ps, c in info["paulis"]:
            parts = ps.split(" ")
            terms.append(("".join(p[0] for p in parts), [int(p[1:]) for p in parts]))
            coeffs.append(c)
        got = float(np.dot(s.expect(terms), coeffs))
        want = float(np.real(np.vdot(psi, d[key] @ psi)))
        assert abs(got - want) < 1e-13, key


# ------------------------------------------------------------------ sampler
def sampler_states(rng):
    """Named states for the sampler checks: random, dyadic with zeros, tie-heavy (|amp|^2 exactly
    half an ulp of the running sum), unnormalised, leading / trailing zeros, tiny probabilities."""
    n = 10
    out = {"random": rand_state(n, rng)}
    psi = np.array([0, 0.5, 0, 0.5j, 0.25 + 0.25j, 0, 0.25 - 0.25j, -0.5], complex)
    out["dyadic"] = psi
    t = np.full(1 << n, 2.0**-27, complex)  # p = 2^-54: a tie once the sum reaches [0.5, 1)
    t[0] = np.sqrt(0.5)
    t[rng.random(1 << n) < 0.3] *= np.sqrt(3.0)
    out["ties"] = t
    out["unnormalised"] = rand_state(n, rng) * 37.5
    z = rand_state(n, rng)
    z[:700] = 0.0
    z[1000:] = 0.0
    out["zeros"] = z
    out["tiny"] = rand_state(n, rng) * 1e-150
    return out


@pytest.mark.parametrize("name", ["random", "dyadic", "ties", "unnormalised", "zeros", "tiny"])
def test_sampler_is_the_sequential_rule(name):
    """qfvm_sample (one sequential pass + binary search) against qfvm_sample_sequential, which is
    the reference's per-shot scan compiled with the reference's flags: identical for every u,
    including u on a boundary, 0, >= the total, and values within an ulp of boundaries."""
    rng = np.random.Generator(np.random.PCG64(sum(map(ord, name))))
    psi = sampler_states(rng)[name]
    n = int(np.log2(psi.size))
    s = CpuState(n)
    s.load(psi)
    p = np.abs(psi) ** 2
    cdf = np.cumsum(p)
    u = np.concatenate([rng.random(4000) * cdf[-1] * 1.01, [0.0, cdf[-1], 2.0 * cdf[-1] + 1.0],
                        cdf[::7], np.nextafter(cdf[::5], 0.0), np.nextafter(cdf[::11], np.inf)])
    assert np.array_equal(s.sample(u), s.sample_sequential(u))


def test_sampler_exact_dyadic():
    # probabilities that are exact binary fractions: every partial sum is exact, so the rule is
    # exactly "first index with u < CDF", skipping zero-probability states
    psi = np.array([0, 0.5, 0, 0.5j, 0.25 + 0.25j, 0, 0.25 - 0.25j, -0.5], complex)
    p = np.abs(psi) ** 2
    assert p.sum() == 1.0
    s = CpuState(3)
    s.load(psi)
    u = np.arange(0, 64) / 64.0
    got = s.sample(u)
    cdf = np.cumsum(p)
    want = np.array([np.argmax(x < cdf) for x in u])
    assert np.array_equal(got, want)
    assert set(got.tolist()) == {1, 3, 4, 6, 7}


def test_sampler_special_states():
    s = CpuState(4)
    s.init_basis(11)
    u = np.array([0.0, 0.3, 0.999999, 1.0, 1.5])
    assert s.sample(u).tolist() == [11, 11, 11, 16, 16]  # u >= total -> 2^n (bits all zero)
    h = np.full(16, 0.25, complex)
    s.load(h)
    u = np.random.Generator(np.random.PCG64(5)).random(1000)
    assert np.array_equal(s.sample(u), np.floor(u * 16).astype(np.uint64))


# ------------------------------------------------------------------ mid-circuit semantics
@pytest.mark.parametrize("seed", [11, 12, 13, 14])
def test_classical_ops_match_dense(seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 4
    prog = random_classical_program(n, rng)
    u = rng.random(prog.n_random)
    psi0 = rand_state(n, rng)
    want, wcreg = dense.run_dense(prog, n, psi0, u, n_creg=n)
    s = CpuState(n)
    s.load(psi0)
    creg = s.run(prog, u, n)
    assert np.array_equal(creg, wcreg)
    assert np.abs(s.store() - want).max() < 1e-12
