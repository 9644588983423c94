"""Adjoint gradients and Hamiltonian expectations on the GPU, through the C-ABI, against the oracle
(oracle/gradients.py) and against parameter shift / finite differences on the engine itself."""

import math

import numpy as np
import pytest

from oracle import gradients as og
from pyquafu_amd import _native, workloads
from tests.progs import random_pauli_terms, rotation_program

pytestmark = pytest.mark.gpu


def engine_adjoint(prog, n, ops, gens, terms, coeffs, dtype="c128", options=()):
    dev = _native.DeviceState(n, dtype)
    try:
        for k, v in options:
            dev.set_option(k, v)
        dev.init_zero()
        return dev.adjoint(prog, ops, gens, terms, coeffs)
    finally:
        dev.close()


@pytest.mark.parametrize("n,fusion", [(3, 1), (5, 0), (6, 1), (8, 0), (9, 1), (12, 1)])
def test_adjoint_matches_oracle(n, fusion):
    rng = np.random.Generator(np.random.PCG64(1200 + n))
    make, thetas, ops, gens = rotation_program(n, 60, rng)
    terms, coeffs = random_pauli_terms(n, 6, rng)
    prog = make(thetas)
    want, e_want = og.adjoint(prog, n, ops, gens, terms, coeffs)
    got, e = engine_adjoint(prog, n, ops, gens, terms, coeffs, options=[("fusion", fusion)])
    assert abs(e - e_want) < 1e-11
    assert np.abs(got - want).max() < 1e-10


@pytest.mark.parametrize("tile,low,groups", [(6, 2, 2), (8, 3, 4), (10, 6, 24), (13, 6, 12)])
def test_adjoint_tile_shapes(tile, low, groups):
    # small tiles and few groups force many passes, stages next to every kind of group and X / Y
    # markers on qubits far above the low bits
    n = 13
    rng = np.random.Generator(np.random.PCG64(1300 + tile))
    make, thetas, ops, gens = rotation_program(n, 80, rng)
    terms, coeffs = random_pauli_terms(n, 5, rng)
    prog = make(thetas)
    want, e_want = og.adjoint(prog, n, ops, gens, terms, coeffs)
    got, e = engine_adjoint(prog, n, ops, gens, terms, coeffs,
                            options=[("tile_bits", tile), ("low_bits", low), ("max_groups", groups)])
    assert abs(e - e_want) < 1e-11
    assert np.abs(got - want).max() < 1e-10


@pytest.mark.parametrize("split,tile", [(1, 11), (1, 12), (1, 13), (0, 12), (0, 13)])
def test_adjoint_kernel_variants(split, tile):
    """The reverse sweep on the split-row two-state kernel (two lanes per row; option tile_bits = b + 1
    for two-state passes) and on the one-lane kernel, against the oracle."""
    n = 15
    rng = np.random.Generator(np.random.PCG64(1350 + 10 * split + tile))
    make, thetas, ops, gens = rotation_program(n, 90, rng)
    terms, coeffs = random_pauli_terms(n, 6, rng)
    prog = make(thetas)
    want, e_want = og.adjoint(prog, n, ops, gens, terms, coeffs)
    got, e = engine_adjoint(prog, n, ops, gens, terms, coeffs, options=[("split_rows", split), ("tile_bits", tile)])
    assert abs(e - e_want) < 1e-11
    assert np.abs(got - want).max() < 1e-10


@pytest.mark.parametrize("split,tile", [(1, 11), (1, 12), (1, 13), (0, 12)])
def test_adjoint_complex64_kernel_variants(split, tile):
    """complex64 reverse sweeps on the split-row two-state kernel (2^10 .. 2^12 tiles per state) and on
    the one-lane kernel, against the oracle at the complex64 tolerance."""
    n = 15
    rng = np.random.Generator(np.random.PCG64(1450 + 10 * split + tile))
    make, thetas, ops, gens = rotation_program(n, 90, rng)
    terms, coeffs = random_pauli_terms(n, 6, rng)
    prog = make(thetas)
    want, e_want = og.adjoint(prog, n, ops, gens, terms, coeffs)
    got, e = engine_adjoint(prog, n, ops, gens, terms, coeffs, dtype="c64",
                            options=[("split_rows", split), ("tile_bits", tile)])
    assert abs(e - e_want) < 1e-4
    assert np.abs(got - want).max() < 1e-4


def test_adjoint_complex64():
    n = 10
    rng = np.random.Generator(np.random.PCG64(1400))
    make, thetas, ops, gens = rotation_program(n, 50, rng)
    terms, coeffs = random_pauli_terms(n, 4, rng)
    prog = make(thetas)
    want, e_want = og.adjoint(prog, n, ops, gens, terms, coeffs)
    got, e = engine_adjoint(prog, n, ops, gens, terms, coeffs, dtype="c64")
    assert abs(e - e_want) < 1e-4
    assert np.abs(got - want).max() < 1e-4


def test_expect_hamiltonian_matches_oracle():
    n = 11
    rng = np.random.Generator(np.random.PCG64(1500))
    terms, coeffs = random_pauli_terms(n, 20, rng)
    coeffs = [c * (1 + 0.5j) for c in coeffs]  # the real part of a complex-weighted sum
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    dev = _native.DeviceState(n)
    dev.load(psi)
    got = dev.expect_hamiltonian(terms, coeffs)
    dev.close()
    assert abs(got - og.expectation(psi, n, terms, coeffs)) < 1e-12


def test_adjoint_equals_engine_parameter_shift_on_ansatz():
    # hardware-efficient ansatz (RY, RZ layers + CZ ring) with the TFIM Hamiltonian, 12 qubits
    from pyquafu_amd.algorithms.gradients.gradients import adjoint_program, hamiltonian_terms
    from pyquafu_amd.simulators.program import compile_circuit

    n, layers = 12, 2
    qc, values = workloads.hea_ansatz(n, layers, parametrised=True)
    ham = workloads.tfim_hamiltonian(n)
    terms, coeffs = hamiltonian_terms(ham)
    prog, ops, gens, _ = adjoint_program(qc)
    dev = _native.DeviceState(n)
    dev.init_zero()
    grads, e = dev.adjoint(prog, ops, gens, terms, coeffs)
    ps = np.zeros(len(ops))
    for k, row in enumerate(ops):
        vals = []
        for shift in (np.pi / 2, -np.pi / 2):
            p2 = compile_circuit(qc)
            p2.ops["param"][row] += shift
            dev.init_zero()
            dev.apply(p2)
            vals.append(dev.expect_hamiltonian(terms, coeffs))
        ps[k] = (vals[0] - vals[1]) / 2
    dev.close()
    assert np.abs(grads - ps).max() < 1e-10


def _reference_gradient_circuit():
    # the circuit of the reference's test_gradient (tests/quafu/algorithms/varational_test.py:68-89)
    from pyquafu_amd.circuits import QuantumCircuit
    from pyquafu_amd.elements import CXGate, HGate, Parameter, RYGate

    pq = QuantumCircuit(4)
    theta = [Parameter("theta_%d" % i, i + 1) for i in range(4)]
    for i in range(4):
        pq << HGate(i)
    for i in range(3):
        pq << CXGate(i, i + 1)
    for i in range(4):
        pq << RYGate(i, theta[i])
    for i in range(3):
        pq << CXGate(i, i + 1)
    for i in range(4):
        pq << RYGate(i, theta[i])
    pq << RYGate(2, theta[0] * theta[1] - 3 * theta[0])
    pq.get_parameter_grads()
    return pq


def test_reference_gradient_scenario():
    from pyquafu_amd.algorithms import Hamiltonian, PauliOp
    from pyquafu_amd.algorithms.gradients import grad_adjoint, grad_finit_diff, grad_para_shift

    pq = _reference_gradient_circuit()
    hamil = Hamiltonian([PauliOp("X1 X2"), PauliOp("Z0 Z3")])
    for values in (None, [0.1, 0.73, 2.1, 3.2]):
        if values is not None:
            pq._update_params(values)
        fd = grad_finit_diff(pq, hamil, h=1e-6)
        ps = grad_para_shift(pq, hamil)
        ad = grad_adjoint(pq, hamil)
        assert len(ad) == 4
        for i in range(4):
            assert math.isclose(ad[i], ps[i], rel_tol=1e-9, abs_tol=1e-12)
            assert abs(fd[i] - ps[i]) < 1e-5


def test_reference_controlled_adjoint_scenario():
    # the circuit of the reference's test_ctrl_adjoint (varational_test.py:107-127)
    from pyquafu_amd.algorithms import Hamiltonian, PauliOp
    from pyquafu_amd.algorithms.gradients import grad_adjoint, grad_finit_diff
    from pyquafu_amd.circuits import QuantumCircuit
    from pyquafu_amd.elements import CRYGate, CXGate, HGate, Parameter

    pq = QuantumCircuit(4)
    theta = [Parameter("theta_%d" % i, i + 1) for i in range(4)]
    for i in range(4):
        pq << HGate(i)
    for i in range(3):
        pq << CXGate(i, i + 1)
        pq << CRYGate(i, i + 1, theta[i])
    for i in range(3):
        pq << CXGate(i, i + 1)
    for i in range(3):
        pq << CRYGate(i + 1, i, theta[i])
    pq.get_parameter_grads()
    hamil = Hamiltonian([PauliOp("X1 X2"), PauliOp("Z0 Z3")])
    fd = grad_finit_diff(pq, hamil, h=1e-6)
    ad = grad_adjoint(pq, hamil)
    assert len(ad) == len(fd) == 3
    for i in range(3):
        assert abs(fd[i] - ad[i]) < 1e-5


def test_adjoint_errors():
    n = 3
    rng = np.random.Generator(np.random.PCG64(1600))
    make, thetas, ops, gens = rotation_program(n, 10, rng)
    prog = make(thetas)
    dev = _native.DeviceState(n)
    with pytest.raises(_native.EngineError):
        dev.adjoint(prog, [len(prog.ops)], "X", [("Z", [0])], [1.0])  # op index out of range
    with pytest.raises(_native.EngineError):
        dev.adjoint(prog, ops[:1], "Q", [("Z", [0])], [1.0])  # unknown generator
    with pytest.raises(_native.EngineError):
        dev.adjoint(prog, ops[:1], gens[:1], [("ZZ", [0, 0])], [1.0])  # repeated qubit in a term
    dev.close()


def test_torch_expectation_function_backward():
    import torch

    from pyquafu_amd.algorithms.gradients import grad_para_shift_device
    from pyquafu_amd.algorithms.interface.torch import expectation

    n, layers = 8, 2
    qc, values = workloads.hea_ansatz(n, layers, parametrised=True)
    ham = workloads.tfim_hamiltonian(n)
    params = torch.tensor(values, dtype=torch.float64, requires_grad=True)
    loss = 3.0 * expectation(qc, ham, params)
    loss.backward()
    ps = np.array(grad_para_shift_device(qc, ham))
    assert np.abs(params.grad.numpy() - 3.0 * ps).max() < 1e-10
    # the loss equals the oracle's <H>
    from oracle import dense
    from pyquafu_amd.algorithms.gradients.gradients import hamiltonian_terms
    from pyquafu_amd.simulators.program import compile_circuit

    psi, _ = dense.run_dense(compile_circuit(qc), n)
    terms, coeffs = hamiltonian_terms(ham)
    assert abs(loss.item() / 3.0 - og.expectation(psi, n, terms, coeffs)) < 1e-12
    # the same through forward-state checkpoints
    params2 = torch.tensor(values, dtype=torch.float64, requires_grad=True)
    loss2 = 3.0 * expectation(qc, ham, params2, checkpoints=3)
    loss2.backward()
    assert abs(loss2.item() - loss.item()) < 1e-12
    assert np.abs(params2.grad.numpy() - params.grad.numpy()).max() < 1e-12


def test_execute_circuits_jacobians_agree():
    import torch

    from pyquafu_amd.algorithms.interface.torch import TorchTransformer

    n, layers = 4, 1
    qc, values = workloads.hea_ansatz(n, layers, parametrised=True)
    rng = np.random.Generator(np.random.PCG64(1700))
    batch = torch.tensor(rng.uniform(-np.pi, np.pi, size=(2, values.size)), dtype=torch.float64, requires_grad=True)
    dy = torch.tensor(rng.normal(size=(2, n)), dtype=torch.float64)
    grads = {}
    for method in ("internal", "adjoint"):
        x = batch.detach().clone().requires_grad_(True)
        out = TorchTransformer.execute(qc, x, method=method)
        (out * dy).sum().backward()
        grads[method] = x.grad.numpy()
        # forward: <Z_i> per row, against the dense oracle
        from oracle import dense
        from pyquafu_amd.simulators.program import compile_circuit

        for r in range(2):
            qc._update_params(batch[r].detach().numpy().tolist())
            psi, _ = dense.run_dense(compile_circuit(qc), n)
            want = [og.expectation(psi, n, [("Z", [i])], [1.0]) for i in range(n)]
            assert np.abs(out[r].detach().numpy() - want).max() < 1e-12
    assert np.abs(grads["internal"] - grads["adjoint"]).max() < 1e-9


def test_batch_runner_matches_single_calls():
    from pyquafu_amd.batch import Autotuner, BatchRunner, compile_item

    circuits = workloads.batch_circuits(count=24, seed=77, n_range=(6, 12), depth_range=(2, 5))
    items = [compile_item(qc, workloads.z_sum(qc.num)) for qc, _ in circuits]
    runner = BatchRunner(dtype="c128", workers=4, tuner=Autotuner(probe_layers=2))
    runner.warm([it.n for it in items])
    results, wall = runner.run(items)
    runner.close()
    assert len(results) == len(items) and wall > 0
    for it, r in zip(items, results):
        dev = _native.DeviceState(it.n)
        dev.init_zero()
        g, e = dev.adjoint(it.prog, it.ops, it.gens, it.terms, it.coeffs)
        dev.close()
        want, e_want = og.adjoint(it.prog, it.n, it.ops, it.gens, it.terms, it.coeffs)
        assert abs(r.expectation - e_want) < 1e-11 and abs(e - e_want) < 1e-11
        assert np.abs(r.grads - want).max() < 1e-10 and np.abs(g - want).max() < 1e-10


@pytest.mark.parametrize("dtype,K,n", [("c128", 1, 12), ("c128", 3, 12), ("c128", 7, 9), ("c64", 4, 13)])
def test_adjoint_checkpoints(dtype, K, n):
    """Option "checkpoints": the forward sweep keeps K copies of psi and the reverse sweep restores
    psi from them instead of continuing the uncompute. Gradients against the oracle and against the
    run without checkpoints; the handle ends in the state before the first differentiated op."""
    rng = np.random.Generator(np.random.PCG64(1500 + 10 * K + n))
    make, thetas, ops, gens = rotation_program(n, 160, rng)
    terms, coeffs = random_pauli_terms(n, 5, rng)
    prog = make(thetas)
    want, e_want = og.adjoint(prog, n, ops, gens, terms, coeffs)
    tol = 1e-10 if dtype == "c128" else 1e-4
    plain = _native.DeviceState(n, dtype)
    ck = _native.DeviceState(n, dtype)
    try:
        ck.set_option("checkpoints", K)
        plain.init_zero()
        g0, e0 = plain.adjoint(prog, ops, gens, terms, coeffs)
        for _ in range(2):  # the second call reuses the checkpoint buffers
            ck.init_zero()
            g1, e1 = ck.adjoint(prog, ops, gens, terms, coeffs)
            assert abs(e1 - e_want) < tol
            assert np.abs(g1 - want).max() < tol
            assert np.abs(g1 - g0).max() < tol
        p = ck.profile()
        assert p["checkpoint_copy"]["launches"] == 2 * K
        assert p["checkpoint_restore"]["launches"] == 2 * K
        assert np.abs(ck.store() - plain.store()).max() < (1e-12 if dtype == "c128" else 1e-5)
    finally:
        plain.close()
        ck.close()


def test_adjoint_checkpoints_more_than_ops():
    """More checkpoints than op boundaries: one per boundary at most."""
    n = 8
    rng = np.random.Generator(np.random.PCG64(1590))
    make, thetas, ops, gens = rotation_program(n, 6, rng)
    terms, coeffs = random_pauli_terms(n, 3, rng)
    prog = make(thetas)
    want, e_want = og.adjoint(prog, n, ops, gens, terms, coeffs)
    got, e = engine_adjoint(prog, n, ops, gens, terms, coeffs, options=[("checkpoints", 64)])
    assert abs(e - e_want) < 1e-11
    assert np.abs(got - want).max() < 1e-10


def test_qnn_module_wrapper_forward_backward():
    """QuantumNeuralNetwork + ModuleWrapper (interface/torch.py:92-130): inputs concatenated with
    the trainable weights form each row's parameters; outputs <Z_i> equal the oracle's and the
    weight gradient equals the sum over rows of the input-row gradients' weight part."""
    import torch

    from oracle import dense
    from pyquafu_amd.algorithms.interface.torch import ModuleWrapper, QuantumNeuralNetwork
    from pyquafu_amd.elements import CZGate, RXGate, RYGate
    from pyquafu_amd.elements.parameters import Parameter
    from pyquafu_amd.simulators.program import compile_circuit

    n = 3
    layers = [RXGate(q, Parameter(f"x{q}", 0.0, tunable=False)) for q in range(n)]
    layers += [RYGate(q, Parameter(f"w{q}", 0.1 * q)) for q in range(n)] + [CZGate(0, 1), CZGate(1, 2)]
    qnn = QuantumNeuralNetwork(n, layers)
    assert qnn.num_tunable_parameters == n and tuple(qnn.weights.shape) == (1, n)
    model = ModuleWrapper(qnn)
    rng = np.random.Generator(np.random.PCG64(1800))
    x = torch.tensor(rng.uniform(-1, 1, size=(2, n)), dtype=torch.float64)
    out = model(x)
    assert tuple(out.shape) == (2, n)
    w = model.weights.detach().numpy()[0]
    for r in range(2):
        qnn.circ._update_params(list(x[r].numpy()) + list(w))  # pylint: disable=protected-access
        psi, _ = dense.run_dense(compile_circuit(qnn.circ), n)
        want = [og.expectation(psi, n, [("Z", [i])], [1.0]) for i in range(n)]
        assert np.abs(out[r].detach().numpy() - want).max() < 1e-12
    out.sum().backward()
    assert model.weights.grad is not None and tuple(model.weights.grad.shape) == (1, n)
    assert np.isfinite(model.weights.grad.numpy()).all()


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-10), ("c64", 1e-4)])
def test_adjoint_with_two_target_dense_gates(dtype, tol):
    """RY / RZ parameters between SWAP, iSWAP and fixed-angle RXX gates (dense 2-target matrices):
    the forward passes hold them as 4x4 ops (complex64 switches to the row-of-8 kernel for them),
    the reverse sweep applies their inverses per gate; against the oracle's adjoint."""
    from pyquafu_amd.algorithms.gradients.gradients import adjoint_program, hamiltonian_terms
    from pyquafu_amd.circuits import QuantumCircuit
    from pyquafu_amd.elements import ISwapGate, Parameter, RXXGate, RYGate, RZGate, SwapGate

    n = 14
    rng = np.random.Generator(np.random.PCG64(1500))
    qc = QuantumCircuit(n)
    k = 0
    for layer in range(3):
        for q in range(n):
            qc << RYGate(q, Parameter(f"a{k}", float(rng.uniform(-np.pi, np.pi))))
            qc << RZGate(q, Parameter(f"b{k}", float(rng.uniform(-np.pi, np.pi))))
            k += 1
        for q in range(layer % 2, n - 1, 2):
            kind = (q // 2) % 3
            if kind == 0:
                qc << SwapGate(q, q + 1)
            elif kind == 1:
                qc << ISwapGate(q, q + 1)
            else:
                qc << RXXGate(q, q + 1, float(rng.uniform(0, 2 * np.pi)))
    ham = workloads.tfim_hamiltonian(n)
    prog, ops, gens, _ = adjoint_program(qc)
    terms, coeffs = hamiltonian_terms(ham)
    want, e_want = og.adjoint(prog, n, ops, gens, terms, coeffs)
    got, e = engine_adjoint(prog, n, ops, gens, terms, coeffs, dtype=dtype)
    assert abs(e - e_want) < tol
    assert np.abs(got - want).max() < tol
