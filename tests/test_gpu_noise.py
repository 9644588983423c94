"""Noisy statevector trajectories on the GPU (QFVM_OP_KRAUS / QFVM_OP_UCHAN through qfvm_run, and
NoiseSVSimulator, quafu/simulators/simulator.py:148-223) against the dense oracle run with the same
uniforms, plus the statistics of the reference's own noise test circuits
(tests/quafu/simulator/noise_test.py) against closed forms."""

import numpy as np
import pytest

import pyquafu_amd as pq
from oracle import dense
from oracle.cpu import sample_on
from pyquafu_amd import _native
from pyquafu_amd.algorithms.hamiltonian import Hamiltonian, PauliOp
from pyquafu_amd.elements import (AmplitudeDamping, BitFlip, Decoherence, Dephasing, Depolarizing,
                                  KrausChannel, QuantumGate)
from pyquafu_amd.simulators.simulator import NoiseSVSimulator
from tests.conftest import rand_state
from tests.progs import random_unitary

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}


def random_kraus(q: int, m: int, rng) -> KrausChannel:
    """m operators K_k = blocks of the first two columns of a random 2m x 2m unitary: sum K^+K = I."""
    u = random_unitary(2 * m, rng)[:, :2]
    gates = [QuantumGate(f"K{k}", [q], [], u[2 * k: 2 * k + 2, :].copy()) for k in range(m)]
    return KrausChannel("rand", q, gates)


def noisy_circuit(n: int, depth: int, rng):
    qc = pq.QuantumCircuit(n)
    for _ in range(depth):
        q = int(rng.integers(n))
        kind = int(rng.integers(9))
        if kind == 0:
            qc.h(q)
        elif kind == 1:
            qc.rx(q, float(rng.uniform(-3, 3)))
        elif kind == 2:
            qc.ry(q, float(rng.uniform(-3, 3)))
        elif kind == 3 and n > 1:
            t = int((q + 1 + rng.integers(n - 1)) % n)
            qc.cx(q, t)
        elif kind == 4:
            qc << [BitFlip, Dephasing, Depolarizing][int(rng.integers(3))](q, float(rng.uniform(0, 0.5)))  # noqa
        elif kind == 5:
            qc << AmplitudeDamping(q, float(rng.uniform(0.05, 0.9)))  # noqa
        elif kind == 6:
            qc << Decoherence(q, float(rng.uniform(0.1, 2)), 3.0, 2.5)  # noqa
        elif kind == 7:
            qc << random_kraus(q, int(rng.integers(1, 4)), rng)  # noqa
        else:
            qc.rz(q, float(rng.uniform(-3, 3)))
    return qc


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n,seed", [(1, 1), (3, 2), (6, 3), (9, 4), (13, 5)])
def test_trajectories_match_oracle(n, seed, dtype):
    """Each trajectory: same uniforms into the engine and the oracle; the final states agree, so
    every channel took the same branch."""
    rng = np.random.Generator(np.random.PCG64(7100 + seed))
    qc = noisy_circuit(n, 40 + 4 * n, rng)
    prog = NoiseSVSimulator.trajectory_program(qc)
    assert prog.n_random > 0
    dev = _native.DeviceState(n, dtype)
    try:
        for t in range(6):
            u = rng.random(prog.n_random)
            if t == 1:
                u[:] = 0.999999  # the last branches, and no branch at all where the ratios sum below 1
            if t == 2:
                u[:] = 0.0
            psi0 = rand_state(n, rng)
            dev.load(psi0)
            dev.run(prog, u, 0, True)
            got = dev.store()
            want, _ = dense.run_dense(prog, n, psi=psi0, uniforms=u)
            print(f"n={n} seed={seed} dtype={dtype} trajectory={t}", flush=True)
            assert np.abs(got - want).max() < TOL[dtype] * 10
    finally:
        dev.close()


def test_kraus_branch_boundaries_on_device():
    """AmplitudeDamping on |1>: u just below sqrt(1-p) keeps |1>, just above it decays (the norm
    ratio rule); a Kraus set whose ratios sum to 1/2 leaves the state for u >= 1/2."""
    p = 0.3
    b0 = np.sqrt(1 - p)
    dev = _native.DeviceState(2, "c128")
    try:
        qc = pq.QuantumCircuit(2)
        qc << AmplitudeDamping(1, p)  # noqa
        prog = NoiseSVSimulator.trajectory_program(qc)
        for u, want_idx in [(b0 - 1e-9, 2), (b0 + 1e-9, 0)]:
            dev.init_basis(2)
            dev.run(prog, [u], 0, True)
            got = dev.store()
            assert abs(abs(got[want_idx]) - 1.0) < 1e-12, (u, got)
        half = QuantumGate("HALF", [0], [], 0.5 * np.eye(2, dtype=complex))
        qc = pq.QuantumCircuit(2)
        qc << KrausChannel("k", 0, [half])  # noqa
        prog = NoiseSVSimulator.trajectory_program(qc)
        psi0 = rand_state(2, np.random.default_rng(1))
        for u in (0.2, 0.7):
            dev.load(psi0)
            dev.run(prog, [u], 0, True)
            assert np.abs(dev.store() - psi0).max() < 1e-14
    finally:
        dev.close()


def oracle_noisy_run(qc, shots, seed, hamiltonian=None, psi=None):
    """The trajectory loop of NoiseSVSimulator.run restated on the oracle: per shot the uniforms
    for the channels, then one for the sample (the sequential-rule sampler on the oracle's state)."""
    prog = NoiseSVSimulator.trajectory_program(qc)
    n = qc.num
    rng = np.random.default_rng(seed)
    meas = sorted(qc.measures.items(), key=lambda kv: kv[1])
    counts = {}
    acc = np.zeros(len(hamiltonian.paulis) if hamiltonian else 0)
    for _ in range(shots):
        u = rng.random(prog.n_random + 1)
        out, _ = dense.run_dense(prog, n, psi=psi, uniforms=u[:-1])
        idx = int(sample_on(out, u[-1:])[0])
        bits = ["0"] * len(qc.measures)
        for q, c in qc.measures.items():
            bits[c] = str((idx >> q) & 1) if idx < (1 << n) else "0"
        key = "".join(bits)
        counts[key] = counts.get(key, 0) + 1
        if hamiltonian:
            for i, pl in enumerate(hamiltonian.paulis):
                m = dense.pauli_matrix_le(pl.paulistr, pl.pos, n)
                acc[i] += float(np.real(np.vdot(out, m @ out)))
    del meas
    expects = [acc[i] * pl.coeff / shots for i, pl in enumerate(hamiltonian.paulis)] if hamiltonian else []
    return counts, expects


@pytest.mark.parametrize("seed", [11, 12])
def test_noise_simulator_matches_oracle_loop(seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 5
    qc = noisy_circuit(n, 30, rng)
    qc.measure([0, 2, 3], [0, 1, 2])
    ham = Hamiltonian([PauliOp("Z0", 0.5), PauliOp("X1 Y3"), PauliOp("Z2 Z4", -1.5)])
    res = NoiseSVSimulator(seed=seed).run(qc, shots=150, hamiltonian=ham)
    want_counts, want_exp = oracle_noisy_run(qc, 150, seed, ham)
    assert res["counts"] == want_counts
    assert np.abs(np.array(res["pauli_expects"]) - np.array(want_exp)).max() < 1e-11
    assert res["simulator"] == "noisy statevector"


def test_noise_simulator_input_state():
    rng = np.random.Generator(np.random.PCG64(77))
    n = 4
    qc = noisy_circuit(n, 20, rng)
    qc.measure([1, 3], [0, 1])
    psi0 = rand_state(n, rng)
    res = pq.simulate(qc, psi=psi0, simulator="noisy statevetor", shots=100, seed=5)
    want_counts, _ = oracle_noisy_run(qc, 100, 5, psi=psi0)
    assert res["counts"] == want_counts


def test_reference_noise_simu_statistics():
    """noise_test.py test_noise_simu: X0, Depolarizing(0, 0.1), X1, AmplitudeDamping(1, 0.2),
    measure q0, <Z0>, <Z1> over 3000 shots. Closed forms: q0 is flipped back by X or Y (2p/3), so
    P(0) = 0.0667 and <Z0> = -(1 - 4p/3); q1 decays when u >= sqrt(1 - 0.2) (the norm-ratio rule),
    so <Z1> = -(2 sqrt(0.8) - 1)."""
    q = pq.QuantumCircuit(2)
    q.x(0)
    q << Depolarizing(0, 0.1)  # noqa
    q.x(1)
    q << AmplitudeDamping(1, 0.2)  # noqa
    q.measure([0])
    ham = Hamiltonian([PauliOp("Z0"), PauliOp("Z1")])
    shots = 3000
    res = NoiseSVSimulator(seed=2024).run(q, shots=shots, hamiltonian=ham)
    p0 = res["counts"].get("0", 0) / shots
    assert sum(res["counts"].values()) == shots and set(res["counts"]) <= {"0", "1"}
    sd = np.sqrt(0.0667 * 0.9333 / shots)
    assert abs(p0 - 0.2 / 3) < 5 * sd
    z0, z1 = (float(np.real(v)) for v in res["pauli_expects"])
    assert abs(z0 - (-(1 - 0.4 / 3))) < 5 / np.sqrt(shots)
    assert abs(z1 - (-(2 * np.sqrt(0.8) - 1))) < 5 / np.sqrt(shots)
    # per shot, <Z0> is exactly -1 or +1 and agrees with the sampled bit of q0
    assert abs(z0 - (2 * p0 - 1)) < 1e-12


def test_reference_noise_circuit_ghz():
    """noise_test.py test_noise_circuit: 5-qubit GHZ with depolarizing noise after the cx gates on
    qubits 0, 2, 3; measure q0 and q4. Without noise only 00 and 11 occur; with p = 0.02 the
    correlated outcomes dominate."""
    q = pq.QuantumCircuit(5)
    q.h(0)
    for k in range(4):
        q.cx(k, k + 1)
    q.add_noise("depolarizing", channel_args=(0.02,), gates=["cx"], qubits=[0, 2, 3])
    q.measure([0, 4])
    res = pq.simulate(q, simulator="noisy statevetor", shots=3000, seed=9)
    counts = res["counts"]
    assert sum(counts.values()) == 3000 and set(counts) <= {"00", "01", "10", "11"}
    assert counts.get("00", 0) + counts.get("11", 0) > 0.9 * 3000
    assert abs(counts.get("00", 0) - counts.get("11", 0)) < 0.1 * 3000


def test_noise_run_once_matches_run():
    """run_once (simulator.py:152-164): one trajectory's sample and weighted expectations; with the
    same uniforms it equals a one-shot run."""
    rng = np.random.Generator(np.random.PCG64(91))
    n = 4
    qc = noisy_circuit(n, 20, rng)
    qc.measure([0, 2], [0, 1])
    ham = Hamiltonian([PauliOp("Z0", 0.5), PauliOp("X1 Y3", -1.0)])
    sim = NoiseSVSimulator(seed=21)
    key, res = sim.run_once(qc, np.array([]), ham, rng=np.random.default_rng(21))
    one = NoiseSVSimulator(seed=21).run(qc, shots=1, hamiltonian=ham)
    assert one["counts"] == {key: 1}
    assert np.abs(np.array(res) - np.array(one["pauli_expects"])).max() < 1e-12
    key2, none = sim.run_once(qc, np.array([]), None, rng=np.random.default_rng(21))
    assert key2 == key and none is None
