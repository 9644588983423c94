"""GPU parity: the HIP engine, called through its C-ABI, against the CPU oracle on the same seeded
inputs. complex128 amplitudes to 1e-12 max-abs, complex64 to 1e-5; sampled indices bit-exact."""

import numpy as np
import pytest

import pyquafu_amd as pq
from oracle import dense
from oracle.cpu import CpuState, sample_on
from pyquafu_amd import _native, workloads
from pyquafu_amd.simulators.program import compile_circuit
from tests.conftest import rand_state
from tests.progs import random_classical_program, random_gate_program

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 1e-5}


def gpu_state(n, dtype="c128"):
    return _native.DeviceState(n, dtype)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 11, 14])
def test_random_programs(n, dtype):
    rng = np.random.Generator(np.random.PCG64(100 + n))
    prog = random_gate_program(n, 80, rng, max_multi=min(5, n))
    psi0 = rand_state(n, rng)
    cpu = CpuState(n)
    cpu.load(psi0)
    cpu.apply(prog)
    want = cpu.store()
    dev = gpu_state(n, dtype)
    dev.load(psi0)
    dev.apply(prog)
    got = dev.store()
    assert np.abs(got - want).max() < TOL[dtype]


def test_large_dense_multi_target():
    rng = np.random.Generator(np.random.PCG64(7))
    from tests.progs import random_unitary
    from pyquafu_amd.simulators.program import MULTI, ProgramBuilder

    n = 12
    b = ProgramBuilder()
    b.add(MULTI, [3, 9, 0, 11, 5, 7, 1], n_ctrl=1, matrix=random_unitary(64, rng))
    b.add(MULTI, [2, 4, 6, 8, 10, 1, 0, 11], n_ctrl=0, matrix=random_unitary(256, rng))
    prog = b.program(n_qubits=n, n_cbits=n)
    psi0 = rand_state(n, rng)
    cpu = CpuState(n)
    cpu.load(psi0)
    cpu.apply(prog)
    dev = gpu_state(n)
    dev.load(psi0)
    dev.apply(prog)
    assert np.abs(dev.store() - cpu.store()).max() < 1e-12


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_expectations(dtype):
    rng = np.random.Generator(np.random.PCG64(5))
    n = 13
    psi = rand_state(n, rng)
    terms = [("Z", [0]), ("ZZ", [1, 12]), ("X", [12]), ("XY", [0, 3]), ("YZX", [2, 0, 11]), ("Y", [7]),
             ("XXYY", [0, 1, 2, 3]), ("IZ", [2, 3]), ("ZZZZZZ", [0, 2, 4, 6, 8, 10])]
    cpu = CpuState(n)
    cpu.load(psi)
    dev = gpu_state(n, dtype)
    dev.load(psi)
    assert np.abs(dev.expect(terms) - cpu.expect(terms)).max() < (1e-12 if dtype == "c128" else 1e-5)
    assert abs(dev.norm2() - 1.0) < (1e-12 if dtype == "c128" else 1e-5)
    assert np.abs(dev.probabilities() - cpu.probabilities()).max() < (1e-14 if dtype == "c128" else 1e-6)


def tfim_terms(n):
    return [("ZZ", [i, (i + 1) % n]) for i in range(n)] + [("X", [i]) for i in range(n)]


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [1, 3, 12, 20])
def test_many_term_expectations(n, dtype):
    """expect_statevec over many terms at once (LDS-tiled passes that each evaluate every term
    whose X / Y qubits fit the tile) against the oracle's per-term expect_pauli: the periodic TFIM,
    random k-local strings and strings too wide for one tile. A term that names a qubit twice is
    refused (the reference adds its bit masks, which then carry past the last qubit)."""
    rng = np.random.Generator(np.random.PCG64(700 + n))
    psi = rand_state(n, rng)
    terms = tfim_terms(n) if n > 1 else [("Z", [0]), ("X", [0]), ("Y", [0])]
    for _ in range(40):
        k = int(rng.integers(1, n + 1))
        qs = [int(x) for x in rng.choice(n, k, replace=False)]
        terms.append(("".join(rng.choice(list("IXYZ"), k)), qs))
    if n >= 12:
        terms.append(("X" * n, list(range(n))))
        terms.append(("Y" * (n // 2) + "X" * (n - n // 2), list(range(n))[::-1]))
    cpu = CpuState(n)
    cpu.load(psi)
    want = cpu.expect(terms)
    dev = gpu_state(n, dtype)
    dev.load(psi)
    dev.profile(reset=True)
    got = dev.expect(terms)
    prof = dev.profile(reset=True)
    err = float(np.abs(got - want).max())
    print(f"[expect n={n} {dtype}] {len(terms)} terms, max-abs vs oracle {err:.2e}, launches "
          f"{ {k: v['launches'] for k, v in prof.items()} }", flush=True)
    assert err < (1e-12 if dtype == "c128" else 1e-5)
    if n == 20:
        # the TFIM alone: the X terms on qubits 5..19 fit 6 chosen tile qubits per pass (three
        # passes); the Z strings and the X terms on the 5 low qubits fill those passes
        dev.profile(reset=True)
        got = dev.expect(tfim_terms(n))
        prof = dev.profile(reset=True)
        assert np.abs(got - cpu.expect(tfim_terms(n))).max() < (1e-12 if dtype == "c128" else 1e-5)
        assert prof["expect_tile"]["launches"] == 3 and "pauli_expect" not in prof
    if n >= 3:
        with pytest.raises(_native.EngineError) as e:
            dev.expect([("XZX", [2, 0, 2])])
        assert e.value.code == _native.ERR_VALUE and "repeated qubit" in str(e.value)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [0, 1, 4, 5, 6, 12, 17])
def test_sampler_bit_exact(n, dtype):
    rng = np.random.Generator(np.random.PCG64(40 + n))
    psi = rand_state(n, rng)
    psi[rng.random(psi.size) < 0.3] = 0.0  # zero-probability states must never be drawn
    if not psi.any():
        psi[0] = 1.0
    psi /= np.linalg.norm(psi)
    dev = gpu_state(n, dtype)
    dev.load(psi)
    u = rng.random(100000)
    u[:4] = [0.0, 1.0 - 2**-53, 1.0, 2.0]
    got = dev.sample(u)
    on_gpu_state = dev.store()  # c64 states come back widened exactly
    want = sample_on(on_gpu_state, u)
    assert np.array_equal(got, want)
    nz = np.abs(on_gpu_state) > 0
    drawn = got[got < (1 << n)].astype(np.int64)
    assert nz[drawn].all()


def crafted_state(kind: str, n: int, rng) -> np.ndarray:
    """States that stress the exact evaluation of the sequential sum: ties (|amp|^2 exactly half an
    ulp of the running sum), many binade crossings spread through the array, long zero runs,
    unnormalised and tiny magnitudes."""
    size = 1 << n
    if kind == "random":
        return rand_state(n, rng)
    if kind == "ties":
        t = np.full(size, 2.0**-27, complex)  # p = 2^-54 while the sum is in [0.5, 1)
        t[0] = np.sqrt(0.5)
        t[rng.random(size) < 0.3] *= np.sqrt(3.0)
        t[rng.random(size) < 0.1] = 0.0
        return t
    if kind == "growing":
        # |amp|^2 grows by 2^40 over the array: the sum crosses ~40 binades at scattered places
        mag = 2.0 ** (-60 + 40 * np.arange(size) / size) * (0.5 + rng.random(size))
        return np.sqrt(mag) * np.exp(2j * np.pi * rng.random(size))
    if kind == "zeros":
        z = rand_state(n, rng)
        z[: size // 3] = 0.0
        z[rng.random(size) < 0.5] = 0.0
        z[-size // 5:] = 0.0
        return z
    if kind == "unnormalised":
        return rand_state(n, rng) * 37.5
    if kind == "tiny":
        return rand_state(n, rng) * 1e-150
    raise KeyError(kind)


@pytest.mark.parametrize("kind", ["random", "ties", "growing", "zeros", "unnormalised", "tiny"])
@pytest.mark.parametrize("n", [9, 16, 22])
def test_sampler_sequential_rule_states(kind, n):
    """The GPU sampler against measure_samples' rule (the oracle's qfvm_sample, itself pinned
    against the reference's per-shot scan) on crafted states, with uniforms spread over the range
    and placed exactly on, just below and just above cumulative boundaries."""
    rng = np.random.Generator(np.random.PCG64(500 + n + sum(map(ord, kind))))
    psi = crafted_state(kind, n, rng)
    dev = gpu_state(n, "c128")
    dev.load(psi)
    cdf = np.cumsum(np.abs(psi) ** 2)
    total = float(cdf[-1])
    pick = rng.integers(0, psi.size, size=3000)
    u = np.concatenate([rng.random(20000) * total * 1.001, [0.0, total, 2.0 * total], cdf[pick],
                        np.nextafter(cdf[pick], 0.0), np.nextafter(cdf[pick], np.inf)])
    got = dev.sample(u)
    want = sample_on(psi, u)
    print(f"[sampler {kind} n={n}] {int((got != want).sum())} mismatches of {u.size}", flush=True)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("seed", [21, 22, 23, 24, 25])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_classical_programs(seed, dtype):
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 6
    prog = random_classical_program(n, rng)
    u = rng.random(prog.n_random)
    psi0 = rand_state(n, rng)
    cpu = CpuState(n)
    cpu.load(psi0)
    want_creg = cpu.run(prog, u, n)
    dev = gpu_state(n, dtype)
    dev.load(psi0)
    got_creg = dev.run(prog, u, n)
    assert np.array_equal(got_creg, want_creg)
    assert np.abs(dev.store() - cpu.store()).max() < (1e-12 if dtype == "c128" else 1e-5)


@pytest.mark.parametrize("seed", [61, 62])
def test_layout_with_mid_circuit_operations(seed):
    """Option "layout" (with 2^10-amplitude tiles on 13 qubits) on gate runs of 150 gates between
    measurements, resets and classically controlled gates: every gate run ends in the standard qubit
    order, so the collapses, the classical register and the final state match the oracle."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 13
    prog = random_classical_program(n, rng, n_blocks=4, block_gates=150)
    u = rng.random(prog.n_random)
    psi0 = rand_state(n, rng)
    cpu = CpuState(n)
    cpu.load(psi0)
    want_creg = cpu.run(prog, u, n)
    dev = gpu_state(n)
    dev.set_option("tile_bits", 10)
    dev.set_option("layout", 1)
    dev.load(psi0)
    got_creg = dev.run(prog, u, n)
    assert np.array_equal(got_creg, want_creg)
    assert np.abs(dev.store() - cpu.store()).max() < 1e-12


def test_parity_workload_20q():
    """20 qubits, complex128, 20 layers of U3 on every qubit + CX brickwork, 10^5 shots."""
    n = 20
    qc = workloads.random_u3_cx(n, 20, 0, seed=20250001)
    prog = compile_circuit(qc)
    cpu = CpuState(n)
    cpu.run(prog, [], n, skip_measure=True)
    want = cpu.store()
    dev = gpu_state(n)
    dev.init_zero()
    dev.run(prog, [], n, True)
    got = dev.store()
    err = np.abs(got - want).max()
    print(f"20q max-abs vs oracle: {err:.3e}")
    assert err < 1e-12
    u = np.random.Generator(np.random.PCG64(20250001)).random(100000)
    idx = dev.sample(u)
    # measure_samples' rule on the GPU's state: all 10^5 shots through the oracle's sequential
    # pass + search, and the first 2,000 through the reference's own per-shot scan
    held = CpuState(n)
    held.load(got)
    assert np.array_equal(idx, held.sample(u))
    assert np.array_equal(idx[:2000], held.sample_sequential(u[:2000]))


# ------------------------------------------------------------------ reference known answers
def sim(qc, **kw):
    return pq.simulate(qc, seed=7, **kw)


def test_reference_basis_known_answers():
    """tests/quafu/simulator/basis_test.py:128-266 through simulate()."""
    qc = pq.QuantumCircuit(2)
    qc.h(0)
    qc.cx(0, 1)
    r = sim(qc)
    assert np.allclose(r.probabilities, [0.5, 0, 0, 0.5]) and r.counts == {}
    qc = pq.QuantumCircuit(2, 2)
    qc.x(0)
    qc.x(1)
    qc.measure([0, 1])
    r = sim(qc, shots=1)
    assert np.allclose(r.probabilities, [0, 0, 0, 1]) and r.counts == {"11": 1}
    qc = pq.QuantumCircuit(2)
    qc.x(0)
    qc.measure([0], [0])
    qc.x(1)
    qc.measure([1], [1])
    r = sim(qc, shots=10)
    assert np.allclose(r.probabilities, [0, 0, 0, 1]) and r.counts == {"11": 10}
    qc = pq.QuantumCircuit(2)
    qc.x(0)
    qc.measure([0], [0])
    qc.cx(0, 1)
    qc.measure([1], [1])
    r = sim(qc, shots=10)
    assert r.counts == {"11": 10}
    qc = pq.QuantumCircuit(4, 4)
    qc.x(0)
    qc.x(1)
    qc.measure([1, 2], [1, 0])
    qc.measure([3, 0], [2, 3])
    r = sim(qc, shots=10)
    assert abs(r.probabilities[5] - 1) < 1e-12 and r.counts == {"0101": 10}
    qc = pq.QuantumCircuit(3, 2)  # bell_measure_normal with shots=1: collapsed to 00 or 11
    qc.h(0)
    qc.cx(0, 1)
    qc.measure([0, 1])
    qc.h(2)
    p = sim(qc, shots=1).probabilities
    d00 = np.linalg.norm(np.array([1, 0, 0, 0]) - p) ** 2
    d11 = np.linalg.norm(np.array([0, 0, 0, 1]) - p) ** 2
    assert np.allclose([d00, d11], [0, 2]) or np.allclose([d00, d11], [2, 0])
    qc = pq.QuantumCircuit(2, 22)  # after_measure: measure then reset
    qc.h(0)
    qc.cx(0, 1)
    qc.measure([0], [0])
    qc.measure([1], [1])
    qc.reset([0, 1])
    p = sim(qc, shots=10).probabilities
    d00 = np.linalg.norm(np.array([1, 0, 0, 0]) - p) ** 2
    d11 = np.linalg.norm(np.array([0, 0, 0, 1]) - p) ** 2
    assert np.allclose([d00, d11], [0, 2]) or np.allclose([d00, d11], [2, 0])


def test_reference_classic_known_answers():
    """tests/quafu/simulator/classic_test.py:94-185 through simulate()."""
    qc = pq.QuantumCircuit(2, 2)
    qc.x(0)
    qc.measure([0], [0])
    with qc.cif([0], 1):
        qc.x(1)
    qc.measure([1], [1])
    assert sim(qc, shots=10).counts == {"11": 10}
    qc = pq.QuantumCircuit(2)
    qc.measure([0], [0])
    with qc.cif([0], 1):
        qc.x(1)
    qc.measure([1], [1])
    assert sim(qc, shots=10).counts == {"00": 10}
    qc = pq.QuantumCircuit(3, 3)
    qc.x(0)
    qc.cx(0, 1)
    qc.measure([0, 1], [0, 1])
    with qc.cif([0, 1], 3):
        qc.x(2)
    qc.measure([2], [2])
    r = sim(qc, shots=10)
    assert abs(r.probabilities[7] - 1) < 1e-12 and r.counts == {"111": 10}
    qc = pq.QuantumCircuit(3)
    qc.x(0)
    qc.cx(0, 1)
    qc.measure([0, 1], [0, 1])
    for cond in (0, 1, 2, 4):
        with qc.cif([0, 1], cond):
            qc.x(2)
    qc.measure([2], [2])
    r = sim(qc, shots=10)
    assert abs(r.probabilities[6] - 1) < 1e-12 and r.counts == {"110": 10}
    qc = pq.QuantumCircuit(2)
    qc.x(0)
    qc.measure([0], [0])
    qc.reset([0])
    qc.measure([0], [1])
    r = sim(qc, shots=10)
    assert abs(r.probabilities[0] - 1) < 1e-12 and r.counts == {"10": 10}
    psi = np.zeros(4, dtype=np.complex128)
    psi[0] = 1.0
    r = sim(qc, shots=10, psi=psi)
    assert abs(r.probabilities[0] - 1) < 1e-12 and r.counts == {"10": 10}
    qc = pq.QuantumCircuit(4)
    qc.x(0)
    qc.cx(0, 1)
    qc.measure([0, 1], [0, 1])
    qc.reset([0, 1])
    qc.measure([0, 1], [2, 3])
    r = sim(qc, shots=10)
    assert abs(r.probabilities[0] - 1) < 1e-12 and r.counts == {"1100": 10}


def test_estimator_reference_value():
    """estimator_test.py:47-62 + varational_test.py:194-198: <Z0Z1+Z3Z4+Z2Z3+Z0Z4> = 1.0."""
    from pyquafu_amd.algorithms.hamiltonian import Hamiltonian, PauliOp

    circ = pq.QuantumCircuit(5)
    for i in range(5):
        if i % 2 == 0:
            circ.h(i)
    circ.cnot(0, 1)
    circ.measure(list(range(5)))
    ham = Hamiltonian([PauliOp("Z0 Z1"), PauliOp("Z3 Z4"), PauliOp("Z2 Z3"), PauliOp("Z0 Z4")])
    assert abs(sum(pq.simulate(circ, hamiltonian=ham)["pauli_expects"]) - 1.0) < 1e-12


def test_input_state_adopted_in_place():
    qc = pq.QuantumCircuit(2)
    qc.x(1)
    psi = np.zeros(4, dtype=np.complex128)
    psi[0] = 1.0
    out = pq.simulate(qc, psi=psi)["statevector"]
    assert out is psi and psi[2] == 1.0


def test_errors_are_reported():
    qc = pq.QuantumCircuit(3, 0)
    qc.x(0)
    with pytest.raises(RuntimeError):
        pq.simulate(qc)
    dev = gpu_state(3)
    from pyquafu_amd.simulators.program import X, ProgramBuilder

    b = ProgramBuilder()
    b.add(X, [5])
    with pytest.raises(_native.EngineError) as e:
        dev.apply(b.program(n_qubits=3))
    assert e.value.code == _native.ERR_INDEX


# ------------------------------------------------------------------ fused passes
# kernel variants of complex128 one-state passes: rows of 8 with 3-qubit groups (the default for
# 2^10 .. 2^13 tiles), split rows of 16, one lane per row of 16, and the persistent prefetching kernel
KERNELS = {"rows8": (("rows8", 1),), "split": (("rows8", 0), ("split_rows", 1)),
           "classic": (("rows8", 0), ("split_rows", 0)), "stream": (("rows8", 0), ("split_rows", 0), ("stream", 1)),
           "rows8c64": (("rows8", 1), ("rows8_c64", 1)), "layout": (("rows8", 1), ("layout", 1)),
           "layoutc64": (("rows8", 1), ("rows8_c64", 1), ("layout", 1)), "default": ()}


@pytest.mark.parametrize("dtype,tile,low,groups,kernel", [
    ("c128", 12, 6, 12, "split"), ("c128", 12, 6, 12, "classic"), ("c128", 12, 6, 12, "stream"),
    ("c128", 11, 5, 12, "split"), ("c128", 11, 3, 6, "stream"), ("c128", 12, 3, 4, "split"),
    ("c128", 12, 3, 4, "classic"), ("c128", 13, 4, 24, "stream"), ("c128", 13, 4, 24, "classic"),
    ("c128", 8, 2, 2, "split"), ("c64", 13, 6, 12, "split"), ("c64", 14, 4, 24, "split"),
    ("c128", 10, 5, 12, "split"), ("c128", 10, 6, 3, "split"), ("c128", 10, 3, 24, "split"),
    ("c64", 12, 5, 12, "split"), ("c128", 11, 5, 12, "rows8"), ("c128", 10, 3, 4, "rows8"),
    ("c128", 10, 6, 24, "rows8"), ("c128", 12, 4, 2, "rows8"), ("c128", 13, 5, 12, "rows8"),
    ("c128", 11, 2, 24, "rows8"), ("c128", 12, 6, 12, "rows8"), ("c64", 12, 5, 12, "rows8c64"),
    ("c64", 10, 3, 4, "rows8c64"), ("c64", 13, 6, 24, "rows8c64"), ("c64", 11, 2, 24, "rows8c64"),
    ("c128", 11, 5, 12, "layout"), ("c128", 10, 3, 4, "layout"), ("c128", 13, 4, 24, "layout"),
    ("c128", 12, 6, 2, "layout"), ("c64", 12, 5, 12, "layoutc64"), ("c64", 11, 2, 24, "layoutc64")])
def test_fused_matches_oracle(dtype, tile, low, groups, kernel):
    rng = np.random.Generator(np.random.PCG64(tile * 100 + low))
    n = 17
    prog = random_gate_program(n, 400, rng, max_multi=3)
    psi0 = rand_state(n, rng)
    cpu = CpuState(n)
    cpu.load(psi0)
    cpu.apply(prog)
    dev = gpu_state(n, dtype)
    for k, v in (("tile_bits", tile), ("low_bits", low), ("max_groups", groups)) + KERNELS[kernel]:
        dev.set_option(k, v)
    dev.load(psi0)
    dev.apply(prog)
    assert np.abs(dev.store() - cpu.store()).max() < TOL[dtype]
    assert dev.profile()["fused_pass"]["launches"] > 0


@pytest.mark.parametrize("kernel", ["rows8", "split", "classic", "stream", "layout"])
def test_fused_equals_unfused_at_24_qubits(kernel):
    """The two GPU paths on the HBM workload's structure at 24 qubits, and the oracle."""
    n = 24
    prog = compile_circuit(workloads.random_u3_cx(n, 6, 15, seed=31, measure=False))
    a = gpu_state(n)
    for k, v in KERNELS[kernel]:
        a.set_option(k, v)
    a.apply(prog)
    fa = a.store()
    a.close()
    b = gpu_state(n)
    b.set_option("fusion", 0)
    b.apply(prog)
    fb = b.store()
    b.close()
    assert np.abs(fa - fb).max() < 1e-12
    cpu = CpuState(n)
    cpu.apply(prog)
    assert np.abs(fa - cpu.store()).max() < 1e-12


def test_layout_fewer_passes_and_standard_order_between_calls():
    """Option "layout" on the brickwork structure at 22 qubits: fewer fused passes than the plain
    planner, and every call leaves the state in the standard qubit order (two calls in a row, then
    expectations and sampling on the result, against the oracle)."""
    n = 22
    prog = compile_circuit(workloads.random_u3_cx(n, 6, 8, seed=41, measure=False))
    cpu = CpuState(n)
    cpu.apply(prog)
    cpu.apply(prog)
    want = cpu.store()
    launches = {}
    for lay in (0, 1):
        dev = gpu_state(n)
        dev.set_option("layout", lay)
        dev.apply(prog)
        launches[lay] = dev.profile(reset=True)["fused_pass"]["launches"]
        dev.apply(prog)
        got = dev.store()
        assert np.abs(got - want).max() < 1e-12
        idx = np.arange(1 << n)
        p = np.abs(want) ** 2
        z = dev.expect([("Z", [3]), ("ZZ", [0, n - 1])])
        want_z = [np.sum(p * (1 - 2 * ((idx >> 3) & 1))), np.sum(p * (1 - 2 * ((idx ^ (idx >> (n - 1))) & 1)))]
        assert np.allclose(z, want_z, atol=1e-12)
        dev.close()
    plans = {lay: _native.plan_json(prog, n, "c128", layout=bool(lay)) for lay in (0, 1)}
    for lay in (0, 1):
        assert launches[lay] == sum(1 for it in plans[lay]["items"] if it["fused"])
    assert launches[1] < launches[0], launches


@pytest.mark.parametrize("dtype,kernel", [("c128", "rows8"), ("c128", "layout"), ("c64", "rows8c64"), ("c64", "default"),
                                          ("c128", "split")])
def test_two_target_dense_gates_fused(dtype, kernel):
    """Dense 2-target gates (SWAP, iSWAP, RXX, RYY, controlled SWAP) run inside the row-of-8 passes
    as 4x4 ops (no per-gate sweeps), against the oracle; complex64 with default options switches to
    the row-of-8 kernel for such gate lists; the split-row kernel keeps them as per-gate sweeps
    (multi_reg)."""
    n = 18
    prog = compile_circuit(workloads.two_target_dense(n, 6, 77))
    cpu = CpuState(n)
    cpu.apply(prog)
    dev = gpu_state(n, dtype)
    for k, v in KERNELS[kernel]:
        dev.set_option(k, v)
    dev.apply(prog)
    got = dev.store()
    prof = dev.profile()
    dev.close()
    assert np.abs(got - cpu.store()).max() < TOL[dtype]
    if kernel == "split":
        assert prof["multi_reg"]["launches"] > 0
    else:
        assert "multi_reg" not in prof and prof["fused_pass"]["launches"] > 0


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("k,fusion", [(5, 1), (6, 1), (8, 1), (5, 0), (4, 1)])
def test_dense_blocks_option(dtype, k, fusion):
    """Option dense_blocks: deep runs on blocks of k qubits collapse into one dense k-target gate
    per block (matrix cores), the rest (a shorter tail block) stays gates; against the oracle, with
    and without fused passes, and one matrix-core launch per collapsed block."""
    n = 18
    prog = compile_circuit(workloads.deep_blocks(n, k, 6, seed=60 + k))
    cpu = CpuState(n)
    cpu.apply(prog)
    dev = gpu_state(n, dtype)
    dev.set_option("fusion", fusion)
    dev.set_option("dense_blocks", k)
    dev.set_option("dense_min_gates", 8)
    dev.apply(prog)
    got = dev.store()
    prof = dev.profile()
    dev.close()
    print(f"[dense_blocks k={k} {dtype} fusion={fusion}] err {np.abs(got - cpu.store()).max():.3g}", flush=True)
    assert np.abs(got - cpu.store()).max() < TOL[dtype]
    assert prof["multi_mfma"]["launches"] == n // k + (1 if n % k >= 4 else 0)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_dense_blocks_with_layout_planner(dtype):
    """Collapsed blocks between layout-aware passes (option layout): the dense gates run on the
    qubits' current positions; deep 5-qubit blocks then a brickwork spreading over all qubits, twice,
    against the oracle."""
    from pyquafu_amd.circuits import QuantumCircuit

    n = 18
    qc = QuantumCircuit(n)
    for rep in range(2):
        for g in workloads.deep_blocks(n, 5, 4, seed=90 + rep).gates:
            qc << g
        for g in workloads.random_u3_cx(n, 2, 3, seed=95 + rep, measure=False).gates:
            qc << g
    prog = compile_circuit(qc)
    cpu = CpuState(n)
    cpu.apply(prog)
    dev = gpu_state(n, dtype)
    for key, v in (("tile_bits", 10), ("layout", 1)) + ((("rows8_c64", 1),) if dtype == "c64" else ()):
        dev.set_option(key, v)
    dev.apply(prog)
    got = dev.store()
    prof = dev.profile()
    dev.close()
    assert np.abs(got - cpu.store()).max() < TOL[dtype]
    assert prof["multi_mfma"]["launches"] == 6


def test_dense_blocks_defaults():
    """Default options collapse deep 5-qubit runs (dense_blocks 5, 24 non-diagonal gates); with
    dense_blocks = 0, or mfma = 0, the same circuit runs in fused passes only."""
    n = 16
    prog = compile_circuit(workloads.deep_blocks(n, 5, 4, seed=71))
    cpu = CpuState(n)
    cpu.apply(prog)
    for opts, dense in (((), True), ((("dense_blocks", 0),), False), ((("mfma", 0),), False)):
        dev = gpu_state(n)
        for key, v in opts:
            dev.set_option(key, v)
        dev.apply(prog)
        got = dev.store()
        prof = dev.profile()
        dev.close()
        assert np.abs(got - cpu.store()).max() < TOL["c128"]
        assert ("multi_mfma" in prof) == dense and "multi_reg" not in prof and "multi_lds" not in prof


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_single_qubit_runs_merged_on_gpu(dtype):
    """Option merge_1q (default on) against off and the oracle: 8 layers of U3 on every qubit with
    CX(2i, 2i+1) after every other layer; the merged circuit needs no more passes."""
    n = 16
    prog = compile_circuit(workloads.u3_runs(n, 8, 2, seed=81))
    cpu = CpuState(n)
    cpu.apply(prog)
    passes = {}
    for merge in (1, 0):
        dev = gpu_state(n, dtype)
        dev.set_option("merge_1q", merge)
        dev.apply(prog)
        got = dev.store()
        passes[merge] = dev.profile()["fused_pass"]["launches"]
        dev.close()
        assert np.abs(got - cpu.store()).max() < TOL[dtype]
    assert passes[1] <= passes[0]


def test_dense_blocks_in_run_between_measurements():
    """qfvm_run collapses the gate runs between mid-circuit measurements (default options): the
    deep-block circuit twice with measurements of qubits 2 and 7 in between, against the oracle's
    run with the same uniforms."""
    from pyquafu_amd.simulators.program import MEASURE, ProgramBuilder

    n = 16
    gates = compile_circuit(workloads.deep_blocks(n, 5, 4, seed=72))
    b = ProgramBuilder()
    for rep in range(2):
        for r in gates.ops:
            q = gates.idx[r["q_off"]: r["q_off"] + r["n_qubits"]].tolist()
            m = None
            if r["m_off"] >= 0:
                d = 1 << (int(r["n_qubits"]) - int(r["n_ctrl"]))
                a = gates.mat[r["m_off"]: r["m_off"] + 2 * d * d]
                m = (a[0::2] + 1j * a[1::2]).reshape(d, d)
            b.add(int(r["code"]), q, n_ctrl=int(r["n_ctrl"]), matrix=m, param=float(r["param"]))
        if rep == 0:
            b.add(MEASURE, [2], cbits=[0])
            b.add(MEASURE, [7], cbits=[1])
    prog = b.program(n_qubits=n, n_cbits=2)
    u = np.array([0.3, 0.8])
    cpu = CpuState(n)
    want_creg = cpu.run(prog, u, 2)
    dev = gpu_state(n)
    got_creg = dev.run(prog, u, 2)
    prof = dev.profile()
    got = dev.store()
    dev.close()
    assert np.array_equal(got_creg, want_creg)
    assert np.abs(got - cpu.store()).max() < TOL["c128"]
    assert prof["multi_mfma"]["launches"] == 6  # three 5-qubit blocks per run of gates


def test_option_validation():
    dev = gpu_state(8)
    with pytest.raises(_native.EngineError):
        dev.set_option("tile_bits", 14)  # 2^14 complex128 amplitudes exceed the LDS
    with pytest.raises(_native.EngineError):
        dev.set_option("no_such_option", 1)
    with pytest.raises(_native.EngineError):
        dev.set_option("dense_blocks", 3)


def test_gather_and_inner():
    rng = np.random.Generator(np.random.PCG64(1800))
    n = 12
    a = rand_state(n, rng)
    b = rand_state(n, rng)
    da = _native.DeviceState(n, "c128")
    db = _native.DeviceState(n, "c64")
    da.load(a)
    db.load(b)
    idx = rng.integers(0, 1 << n, size=300)
    assert np.array_equal(da.gather(idx), a[idx])
    b64 = b.astype(np.complex64).astype(np.complex128)
    assert np.abs(db.gather(idx) - b64[idx]).max() == 0.0
    assert abs(da.inner(db) - np.vdot(a, b64)) < 1e-12
    assert abs(db.inner(da) - np.vdot(b64, a)) < 1e-12
    with pytest.raises(_native.EngineError):
        da.gather([1 << n])
    dc = _native.DeviceState(n - 1)
    with pytest.raises(_native.EngineError):
        da.inner(dc)
    for d in (da, db, dc):
        d.close()


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-12), ("c64", 2e-6)])
def test_qft_closed_form_22_qubits(dtype, tol):
    n = 22
    x = 0x2B5E31
    dev = _native.DeviceState(n, dtype)
    dev.apply(compile_circuit(workloads.qft(n, x)))
    got = dev.store()
    dev.close()
    want = dense.qft_closed_form(n, x)
    assert np.abs(got - want).max() < tol


def diagonal_heavy_program(n: int, n_ops: int, rng):
    """Mostly diagonal gates (phases, controlled phases, diagonal matrices with controls anywhere)
    with X gates (relabelings) and a few dense gates between them: long diagonal runs."""
    from pyquafu_amd.simulators.program import CP, CX, CZ, GENERAL, P, RZ, RZZ, S, T, X, ProgramBuilder

    b = ProgramBuilder()
    for _ in range(n_ops):
        k = int(rng.integers(0, 10))
        if k < 2:
            b.add(int(rng.choice([S, T, P, RZ])), [int(rng.integers(n))], param=float(rng.uniform(-7, 7)))
        elif k < 5:
            q = rng.choice(n, 2, replace=False)
            b.add(int(rng.choice([CP, CZ, RZZ])), [int(x) for x in q], n_ctrl=1, param=float(rng.uniform(-7, 7)))
        elif k < 7:
            nc = int(rng.integers(0, 3))
            q = rng.choice(n, nc + 1, replace=False)
            d = np.exp(1j * rng.uniform(-4, 4, 2))
            b.add(GENERAL, [int(x) for x in q], n_ctrl=nc, matrix=np.diag(d))
        elif k < 9:
            q = rng.choice(n, 2, replace=False)
            if rng.integers(2):
                b.add(X, [int(q[0])])
            else:
                b.add(CX, [int(x) for x in q], n_ctrl=1)
        else:
            from tests.progs import random_unitary
            b.add(GENERAL, [int(rng.integers(n))], matrix=random_unitary(2, rng))
    return b.program(n_qubits=n, n_cbits=n)


@pytest.mark.parametrize("n,dtype,kernel", [(9, "c128", "split"), (13, "c128", "split"), (16, "c128", "split"),
                                            (13, "c128", "classic"), (16, "c128", "stream"), (13, "c64", "split"),
                                            (10, "c128", "rows8"), (13, "c128", "rows8"), (16, "c128", "rows8"),
                                            (16, "c64", "rows8c64")])
def test_diagonal_runs_match_oracle(n, dtype, kernel):
    rng = np.random.Generator(np.random.PCG64(1900 + n))
    prog = diagonal_heavy_program(n, 400, rng)
    psi0 = rand_state(n, rng)
    cpu = CpuState(n)
    cpu.load(psi0)
    cpu.apply(prog)
    dev = _native.DeviceState(n, dtype)
    for k, v in KERNELS[kernel]:
        dev.set_option(k, v)
    dev.load(psi0)
    dev.apply(prog)
    got = dev.store()
    dev.close()
    tol = 1e-12 if dtype == "c128" else 2e-5
    assert np.abs(got - cpu.store()).max() < tol


# ------------------------------------------------------------------ marginal probabilities
@pytest.mark.parametrize("n,k,dtype", [(5, 1, "c128"), (9, 4, "c128"), (14, 12, "c128"), (16, 7, "c64"), (13, 13, "c128")])
def test_marginals_match_numpy(n, k, dtype):
    """qfvm_marginals over k qubits in arbitrary order (13 goes through the host fallback)."""
    rng = np.random.Generator(np.random.PCG64(2300 + n * 10 + k))
    psi = rand_state(n, rng)
    qubits = [int(q) for q in rng.permutation(n)[:k]]
    dev = gpu_state(n, dtype)
    dev.load(psi)
    got = dev.marginals(qubits)
    want = dense.marginal(dev.store(), n, qubits)
    assert np.abs(got - want).max() < (1e-14 if dtype == "c128" else 1e-7)
    assert abs(got.sum() - 1.0) < (1e-12 if dtype == "c128" else 1e-5)


def test_result_probabilities_use_device_marginals():
    """SimuResult.probabilities from the device marginals equals the reference's host route
    (reverse the qubit order, partial trace over the measured qubits, order by cbit)."""
    from pyquafu_amd.results.results import SimuResult

    rng = np.random.Generator(np.random.PCG64(2400))
    n = 10
    qc = workloads.random_u3_cx(n, 3, 2, seed=2401, measure=False)
    qc.measure([7, 2, 9, 4], [2, 0, 3, 1])
    res = pq.simulate(qc, shots=10)
    assert "probabilities" in res._meta_data  # pylint: disable=protected-access
    host = SimuResult({k: v for k, v in res._meta_data.items() if k != "probabilities"})  # pylint: disable=protected-access
    assert np.abs(res.probabilities - host.probabilities).max() < 1e-14
    del rng


# ------------------------------------------------------------------ full-size properties
def test_full_size_qft_closed_form_30_qubits():
    """The 30-qubit QFT of a basis state against its closed form on 100,000 sampled amplitudes
    (complex128 and complex64), plus the norm."""
    n, x = 30, 0x2B5E_71C3 & ((1 << 30) - 1)
    prog = compile_circuit(workloads.qft(n, x))
    ys = np.random.Generator(np.random.PCG64(3000)).integers(0, 1 << n, size=100000, dtype=np.int64)
    want = workloads.qft_amplitudes(n, x, ys)
    for dtype, tol in (("c128", 1e-15), ("c64", 2e-10)):
        dev = _native.DeviceState(n, dtype)
        try:
            dev.apply(prog)
            got = dev.gather(ys)
            assert np.abs(got - want).max() < tol, dtype
            assert abs(dev.norm2() - 1.0) < (1e-12 if dtype == "c128" else 1e-5)
        finally:
            dev.close()


def test_full_size_round_trip_30_qubits():
    """The 30-qubit benchmark circuit (2,960 gates) followed by its inverse returns |0...0>: a
    size-independent check of every fused pass at the benchmark's shape."""
    n = 30
    qc = workloads.random_u3_cx(n, 40, 15, seed=20250002, measure=False)
    fwd, back = compile_circuit(qc), compile_circuit(qc.dagger())
    dev = _native.DeviceState(n, "c128")
    try:
        dev.apply(fwd)
        dev.apply(back)
        a0 = dev.gather([0])[0]
        assert abs(a0 - 1.0) < 1e-11
        assert abs(dev.norm2() - 1.0) < 1e-11
    finally:
        dev.close()


# ------------------------------------------------------------------ dense gates on the matrix cores
@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("k,where,nc", [(4, "low", 0), (4, "mixed", 1), (5, "low", 0), (5, "high", 1),
                                        (5, "mixed", 2), (6, "low", 1), (6, "high", 0), (6, "mixed", 2),
                                        (7, "low", 0), (7, "high", 1), (7, "mixed", 2), (8, "low", 0),
                                        (8, "mixed", 1), (8, "high", 3)])
def test_dense_mfma_matches_oracle(k, where, nc, dtype):
    """Dense gates on 4 to 8 targets run on the matrix cores (option mfma, the default:
    v_mfma_f64_16x16x4f64 at complex128, v_mfma_f32_16x16x4f32 at complex64; 7 and 8 targets with the
    matrix streamed from memory); against the oracle and against the VALU kernels (mfma = 0), with
    the targets among the lowest qubits (member-fast tile order), the highest, or spread, and with
    controls (8 targets and 3 controls at 15 qubits leave 16 groups: a partly filled tile)."""
    from pyquafu_amd.simulators.program import MULTI, ProgramBuilder
    from tests.progs import random_unitary

    n = 15
    rng = np.random.Generator(np.random.PCG64(1700 + 10 * k + nc + (0 if where == "low" else 3 if where == "high" else 6)))
    if where == "low":
        targets = [int(x) for x in rng.permutation(k)]
    elif where == "high":
        targets = [int(x) for x in rng.permutation(np.arange(n - k, n))]
    else:
        targets = [int(x) for x in rng.choice(n, k, replace=False)]
    rest = [q for q in range(n) if q not in targets]
    ctrls = [int(x) for x in rng.choice(rest, nc, replace=False)]
    b = ProgramBuilder()
    for _ in range(2):
        b.add(MULTI, ctrls + targets, n_ctrl=nc, matrix=random_unitary(1 << k, rng))
    prog = b.program(n_qubits=n)
    psi0 = rand_state(n, rng)
    cpu = CpuState(n)
    cpu.load(psi0)
    cpu.apply(prog)
    want = cpu.store()
    got = {}
    for mfma in (1, 0):
        dev = gpu_state(n, dtype)
        dev.set_option("mfma", mfma)
        dev.load(psi0)
        dev.apply(prog)
        got[mfma] = dev.store()
        prof = dev.profile()
        assert ("multi_mfma" in prof) == bool(mfma)
    assert np.abs(got[1] - want).max() < TOL[dtype]
    assert np.abs(got[1] - got[0]).max() < TOL[dtype]
