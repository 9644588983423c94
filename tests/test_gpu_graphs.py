"""Graph mode of qfvm_adjoint (option "graphs"): calls captured into hipGraphs and replayed through
cached executable graphs, through the C-ABI, against eager calls of the same engine (same kernels in
the same order, so the results agree exactly) and against the oracle. Covered: replays with new
parameters and eager work between them, alternating structures on one handle, a buffer growth
after graphs were cached, the empty Hamiltonian, calls that must stay eager (dense multi-target
gates) and the batch runner over two steps."""

import numpy as np
import pytest

from oracle import gradients as og
from pyquafu_amd import _native, batch, workloads
from pyquafu_amd.algorithms.gradients.gradients import adjoint_program, hamiltonian_terms
from pyquafu_amd.algorithms.hamiltonian import Hamiltonian, PauliOp
from pyquafu_amd.circuits import QuantumCircuit
from pyquafu_amd.elements import CRXGate, CZGate, HGate, Parameter, RYGate, RZGate
from tests.progs import random_pauli_terms, rotation_program

pytestmark = pytest.mark.gpu


def layered(n, layers, seed):
    """RY / RZ on every qubit, controlled RX on neighbours, CZ ring, a Hadamard per layer: a fixed
    structure whose angles come from `seed`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    qc = QuantumCircuit(n)
    k = 0
    for layer in range(layers):
        for q in range(n):
            qc << RYGate(q, Parameter(f"a{k}", float(rng.uniform(-np.pi, np.pi))))
            qc << RZGate(q, Parameter(f"b{k}", float(rng.uniform(-np.pi, np.pi))))
            k += 1
        for q in range(0, n - 1, 2):
            qc << CRXGate(q, q + 1, Parameter(f"c{k}_{q}", float(rng.uniform(-np.pi, np.pi))))
        for q in range(n):
            qc << CZGate(q, (q + 1) % n)
        qc << HGate(layer % n)
    return qc


def tfim(n):
    return Hamiltonian([PauliOp(f"Z{i} Z{(i + 1) % n}") for i in range(n)] + [PauliOp(f"X{i}", 0.7) for i in range(n)])


def call_args(qc, ham):
    prog, ops, gens, _ = adjoint_program(qc)
    terms, coeffs = hamiltonian_terms(ham)
    return prog, ops, gens, terms, coeffs


def run(dev, args):
    dev.init_zero()
    return dev.adjoint(*args)


def graph_counts(dev):
    p = dev.profile(reset=False)
    return {k: p.get(k, {}).get("launches", 0) for k in ("graph_launch", "graph_update", "graph_instantiate")}


@pytest.mark.parametrize("dtype,n", [("c128", 12), ("c64", 14), ("c128", 5)])
def test_replays_match_eager(dtype, n):
    eager = _native.DeviceState(n, dtype)
    graph = _native.DeviceState(n, dtype)
    try:
        graph.set_option("graphs", 1)
        ham = tfim(n)
        for rep in range(5):
            args = call_args(layered(n, 3, 700 + rep), ham)
            g0, e0 = run(eager, args)
            g1, e1 = run(graph, args)
            assert np.array_equal(g1, g0), (rep, np.abs(g1 - g0).max())
            assert e1 == e0
            if dtype == "c128":
                want, e_want = og.adjoint(args[0], n, *args[1:])
                assert abs(e1 - e_want) < 1e-11
                assert np.abs(g1 - want).max() < 1e-10
            # eager work on the same handle between replays
            graph.init_zero()
            graph.apply(args[0])
            assert abs(graph.norm2() - 1.0) < (1e-12 if dtype == "c128" else 1e-5)
        c = graph_counts(graph)
        assert c["graph_launch"] >= 4
        assert c["graph_update"] >= 3
        assert c["graph_instantiate"] <= 2
    finally:
        eager.close()
        graph.close()


def test_alternating_structures_and_growth():
    n = 11
    eager = _native.DeviceState(n, "c128")
    graph = _native.DeviceState(n, "c128")
    try:
        graph.set_option("graphs", 1)
        ham = tfim(n)
        small = [layered(n, 2, 800 + k) for k in range(3)]
        other = [layered(n, 4, 810 + k) for k in range(3)]
        # buffers are sized by the first structures; the 9-layer circuit grows the gradient
        # buffers after graphs were cached, which drops them
        seq = small[:1] + other[:1] + [small[1], other[1], small[2], other[2], layered(n, 9, 820), small[0], other[0], small[1]]
        for k, qc in enumerate(seq):
            a = call_args(qc, ham)
            g0, e0 = run(eager, a)
            g1, e1 = run(graph, a)
            assert np.array_equal(g1, g0), (k, np.abs(g1 - g0).max())
            assert e1 == e0
        c = graph_counts(graph)
        assert c["graph_update"] >= 3
        assert c["graph_launch"] >= 6
    finally:
        eager.close()
        graph.close()


def test_empty_hamiltonian_and_no_gradients():
    n = 9
    dev = _native.DeviceState(n, "c128")
    try:
        dev.set_option("graphs", 1)
        qc = layered(n, 2, 830)
        prog, ops, gens, _ = adjoint_program(qc)
        for _ in range(3):
            g, e = run(dev, (prog, ops, gens, [], []))
            assert e == 0.0
            assert np.array_equal(g, np.zeros(len(ops)))
        terms, coeffs = hamiltonian_terms(tfim(n))
        want, e_want = og.adjoint(prog, n, [], "", terms, coeffs)
        for _ in range(3):
            g, e = run(dev, (prog, [], "", terms, coeffs))
            assert g.size == 0
            assert abs(e - e_want) < 1e-11
    finally:
        dev.close()


def test_multi_target_gates_stay_eager():
    n = 10
    rng = np.random.Generator(np.random.PCG64(840))
    make, thetas, ops, gens = rotation_program(n, 50, rng)
    terms, coeffs = random_pauli_terms(n, 4, rng)
    prog = make(thetas)
    want, e_want = og.adjoint(prog, n, ops, gens, terms, coeffs)
    dev = _native.DeviceState(n, "c128")
    try:
        dev.set_option("graphs", 1)
        for _ in range(2):
            g, e = run(dev, (prog, ops, gens, terms, coeffs))
            assert abs(e - e_want) < 1e-11
            assert np.abs(g - want).max() < 1e-10
        assert graph_counts(dev)["graph_launch"] == 0
    finally:
        dev.close()


def test_batch_runner_graph_steps_match_eager():
    circuits = workloads.batch_circuits(24, seed=850, n_range=(10, 14), depth_range=(2, 5))
    items = []
    for qc, _ in circuits:
        items.append(batch.compile_item(qc, Hamiltonian([PauliOp(f"Z{i}") for i in range(qc.num)])))
    eager = batch.BatchRunner(dtype="c64", workers=3, graphs=False)
    graph = batch.BatchRunner(dtype="c64", workers=3, graphs=True)
    try:
        want, _ = eager.run(items)
        for _ in range(2):
            got, _ = graph.run(items)
            for a, b in zip(got, want):
                assert np.array_equal(a.grads, b.grads)
                assert a.expectation == b.expectation
    finally:
        eager.close()
        graph.close()


def test_graph_replays_with_checkpoints():
    n = 12
    eager = _native.DeviceState(n, "c128")
    graph = _native.DeviceState(n, "c128")
    try:
        for d in (eager, graph):
            d.set_option("checkpoints", 3)
        graph.set_option("graphs", 1)
        ham = tfim(n)
        for rep in range(4):
            args = call_args(layered(n, 4, 900 + rep), ham)
            g0, e0 = run(eager, args)
            g1, e1 = run(graph, args)
            assert np.array_equal(g1, g0)
            assert e1 == e0
        assert graph_counts(graph)["graph_update"] >= 2
    finally:
        eager.close()
        graph.close()


def test_value_dependent_plan_gets_its_own_executable():
    """RX(0) plans as a diagonal gate, RX(0.3) as a dense one: one structure key, two different
    captured graphs. The second call must not update the first call's executable in place (that
    needs the same kernels, grids and LDS sizes); it instantiates a graph of its own, and every
    call matches the eager engine and the oracle."""
    from pyquafu_amd.elements import RXGate

    n = 11

    def circuit(theta):
        qc = QuantumCircuit(n)
        for q in range(n):
            qc << RYGate(q, Parameter(f"y{q}", 0.1 * (q + 1)))
        for q in range(n):
            qc << RXGate(q, Parameter(f"x{q}", theta))
        for q in range(n):
            qc << CZGate(q, (q + 1) % n)
        return qc

    eager = _native.DeviceState(n, "c128")
    graph = _native.DeviceState(n, "c128")
    try:
        graph.set_option("graphs", 1)
        ham = tfim(n)
        for theta in (0.0, 0.3, 0.0, 0.3):
            args = call_args(circuit(theta), ham)
            g0, e0 = run(eager, args)
            g1, e1 = run(graph, args)
            assert np.array_equal(g1, g0) and e1 == e0, theta
            want, e_want = og.adjoint(args[0], n, *args[1:])
            assert np.abs(g1 - want).max() < 1e-10
        c = graph_counts(graph)
        # (the first call on a fresh handle grows its buffers and so runs eagerly)
        assert c["graph_launch"] >= 3
        assert c["graph_instantiate"] >= 2  # the 0.3 capture never updates the 0.0 executable
    finally:
        eager.close()
        graph.close()


def test_layout_option_in_adjoint_eager_and_graphs():
    """Option "layout" also plans the adjoint's forward passes (each call ends in the standard qubit
    order): eager against the default planner and the oracle, and graph replays against eager calls
    with the same option (same kernels in the same order: equal bits)."""
    n = 15
    ham = tfim(n)
    plain = _native.DeviceState(n, "c128")
    eager = _native.DeviceState(n, "c128")
    graph = _native.DeviceState(n, "c128")
    try:
        eager.set_option("layout", 1)
        graph.set_option("layout", 1)
        graph.set_option("graphs", 1)
        for rep in range(3):
            args = call_args(layered(n, 4, 900 + rep), ham)
            g0, e0 = run(plain, args)
            g1, e1 = run(eager, args)
            g2, e2 = run(graph, args)
            assert np.abs(g1 - g0).max() < 1e-12 and abs(e1 - e0) < 1e-12
            assert np.array_equal(g2, g1) and e2 == e1
            want, e_want = og.adjoint(args[0], n, *args[1:])
            assert abs(e1 - e_want) < 1e-11
            assert np.abs(g1 - want).max() < 1e-10
        assert graph_counts(graph)["graph_launch"] >= 2
    finally:
        plain.close()
        eager.close()
        graph.close()
