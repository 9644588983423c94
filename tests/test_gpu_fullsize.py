"""GPU parity at the benchmark workloads' real shapes, through the C-ABI, against the CPU oracle.

* the 30-qubit complex128 benchmark circuit (random U3 + CX brickwork + 15 CRZ per layer, PCG64
  seed 20250002): its first four layers (296 gates) on the GPU, with the default and with the
  layout-aware planner, and on the oracle (the restatement of qfvm's OpenMP engine), every one of
  the 2^30 amplitudes compared, 1e-12 max-abs;
* the 24-qubit, 8-layer hardware-efficient ansatz with the periodic TFIM: all 384 adjoint partials
  and <H> against the oracle's grad_adjoint (gradiant.py:144-173) at 1e-10, plus the torch
  autograd.Function path;
* the batched uneven workload in complex64 (one circuit of every size 10..22 from the seeded
  256-circuit batch) through BatchRunner, eager and graph-captured, against the oracle;
* the 33-qubit QFT of a seeded basis state in complex128 and complex64: sampled amplitudes against
  the closed form, and 10^6 shots against floor(u 2^n) (the exact answer for a uniform
  distribution) except within rounding of a cell boundary.
"""

import numpy as np
import pytest

from oracle.cpu import CpuState
from pyquafu_amd import _native, workloads
from pyquafu_amd.simulators.program import compile_circuit

pytestmark = pytest.mark.gpu


def _max_abs_chunked(a: np.ndarray, b: np.ndarray, chunk: int = 1 << 24) -> float:
    err = 0.0
    for k in range(0, a.size, chunk):
        err = max(err, float(np.abs(a[k:k + chunk] - b[k:k + chunk]).max()))
    return err


def test_benchmark_circuit_30q_first_layers_match_oracle():
    n, layers = 30, 4
    qc = workloads.random_u3_cx(n, layers, 15, seed=20250002, measure=False)
    prog = compile_circuit(qc)
    assert prog.n_ops == layers * (n + n // 2 + (n - 1) // 2 + 15)
    got = {}
    for lay in (0, 1):  # the default planner and the layout-aware one (option "layout")
        print(f"[30q] {prog.n_ops} gates on the GPU, layout {lay}", flush=True)
        dev = _native.DeviceState(n, "c128")
        try:
            dev.set_option("layout", lay)
            dev.apply(prog)
            got[lay] = dev.store()
        finally:
            dev.close()
    print("[30q] oracle", flush=True)
    cpu = CpuState(n)
    cpu.apply(prog)
    for lay in (0, 1):
        err = _max_abs_chunked(got[lay], cpu.view())
        print(f"[30q] layout {lay}: max-abs over 2^30 amplitudes vs oracle: {err:.3e}", flush=True)
        assert err < 1e-12
    del cpu, got


def test_adjoint_24q_8_layers_matches_oracle():
    import torch

    from pyquafu_amd.algorithms.gradients.gradients import adjoint_program, hamiltonian_terms
    from pyquafu_amd.algorithms.interface.torch import expectation

    n, layers = 24, 8
    qc, values = workloads.hea_ansatz(n, layers, parametrised=True)
    ham = workloads.tfim_hamiltonian(n)
    terms, coeffs = hamiltonian_terms(ham)
    prog, ops, gens, _ = adjoint_program(qc)
    assert len(ops) == 2 * n * layers
    dev = _native.DeviceState(n, "c128")
    try:
        grads, e = dev.adjoint(prog, ops, gens, terms, coeffs)
    finally:
        dev.close()
    print("[24q] oracle grad_adjoint", flush=True)
    want, e_want = CpuState(n).adjoint(prog, ops, gens, terms, coeffs)
    gerr = float(np.abs(grads - want).max())
    print(f"[24q] <H> {e:.12f} vs {e_want:.12f}; max |grad - oracle| over {len(ops)} partials {gerr:.3e}",
          flush=True)
    assert abs(e - e_want) < 1e-10
    assert gerr < 1e-10
    # the same gradient through torch.autograd (ExpectationFunction)
    params = torch.tensor(values, dtype=torch.float64, requires_grad=True)
    loss = expectation(qc, ham, params)
    loss.backward()
    assert abs(loss.item() - e_want) < 1e-10
    assert float(np.abs(params.grad.numpy() - want).max()) < 1e-10


@pytest.mark.parametrize("graphs", [False, True])
def test_batch_c64_matches_oracle(graphs):
    from pyquafu_amd.batch import Autotuner, BatchRunner, compile_item

    circuits = workloads.batch_circuits(count=256)
    first = {}
    for k, (qc, _) in enumerate(circuits):
        first.setdefault(qc.num, k)
    assert sorted(first) == list(range(10, 23))
    picked = [circuits[first[n]][0] for n in sorted(first)]
    items = [compile_item(qc, workloads.z_sum(qc.num)) for qc in picked]
    runner = BatchRunner(dtype="c64", workers=4, tuner=Autotuner(), graphs=graphs)
    try:
        runner.warm([it.n for it in items])
        runner.run(items)  # graphs: instantiate; the second run replays
        results, _ = runner.run(items)
    finally:
        runner.close()
    worst = 0.0
    for it, r in zip(items, results):
        g_want, e_want = CpuState(it.n).adjoint(it.prog, it.ops, it.gens, it.terms, it.coeffs)
        # complex64: the observable's norm bound (sum of |coefficients|) times 1e-5
        tol = 1e-5 * float(np.abs(np.asarray(it.coeffs)).sum())
        gerr = float(np.abs(r.grads - g_want).max())
        eerr = abs(r.expectation - e_want)
        worst = max(worst, gerr / tol, eerr / tol)
        print(f"[batch c64 graphs={graphs}] n={it.n} ops={len(it.ops)} |dE|={eerr:.2e} max|dgrad|={gerr:.2e} "
              f"tol={tol:.1e}", flush=True)
        assert eerr < tol and gerr < tol
    print(f"[batch c64] worst error / tolerance {worst:.3f}", flush=True)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_qft_33q_closed_form_and_sampler(dtype):
    n = 33
    rng = np.random.Generator(np.random.PCG64(20250005))
    x = int(rng.integers(0, 1 << n))
    ys = rng.integers(0, 1 << n, size=100000, dtype=np.int64)
    prog = compile_circuit(workloads.qft(n, x))
    want = workloads.qft_amplitudes(n, x, ys)
    dev = _native.DeviceState(n, dtype)
    try:
        dev.apply(prog)
        got = dev.gather(ys)
        err = float(np.abs(got - want).max())
        print(f"[qft33 {dtype}] max-abs over 10^5 sampled amplitudes vs closed form {err:.3e}", flush=True)
        assert err < (1e-15 if dtype == "c128" else 1e-10)
        u = np.random.Generator(np.random.PCG64(20250006)).random(1000000)
        idx = dev.sample(u).astype(np.int64)
        total = dev.norm2()
    finally:
        dev.close()
    # |amp|^2 = total 2^-n for every basis state, so the exact draw is floor(u 2^n / total).
    # complex128: every |amp|^2 is exactly 2^-33, every running sum is exact, and the sampler (qfvm's
    # sequential rule, evaluated exactly) must give floor(u 2^n) except on a boundary itself.
    # complex64: |amp|^2 are not dyadic, so each of the k additions up to index k rounds, by at most
    # half an ulp of a sum below 1 (2^-53), i.e. k 2^-20 cells at n = 33, and the amplitudes carry
    # the complex64 error (relative ~1e-6, the size of the norm error): the draw may move by up to
    # k (2^-20 + 1e-6) cells.
    cell = u * (float(1 << n) / total)
    m = np.floor(cell).astype(np.int64)
    off = idx != m
    dist = np.abs(cell - np.rint(cell))
    print(f"[qft33 {dtype}] norm^2 - 1 = {total - 1.0:.3e}; {int(off.sum())} of {u.size} draws off "
          f"floor(u 2^n / norm^2); largest move {int(np.abs(idx - m).max())} cells", flush=True)
    if dtype == "c128":
        assert np.all(np.abs(idx - m) <= 1)
        assert np.all(dist[off] < 1e-3)
    else:
        assert np.all(np.abs(idx - m) <= m * (2.0**-20 + 1e-6) + 1)


def test_qft_34q_c64_closed_form_and_sampler():
    """The largest state one MI355X holds in complex64: 34 qubits (128 GiB). QFT of a seeded basis
    state: 10^5 sampled amplitudes against the closed form, 10^6 draws against floor(u 2^n / norm^2)
    within the complex64 rounding bound (k additions of at most half an ulp of a sum below 1 each,
    2^(n-53) cells per addition, plus the amplitudes' relative error)."""
    n = 34
    rng = np.random.Generator(np.random.PCG64(20250008))
    x = int(rng.integers(0, 1 << n))
    ys = rng.integers(0, 1 << n, size=100000, dtype=np.int64)
    prog = compile_circuit(workloads.qft(n, x))
    want = workloads.qft_amplitudes(n, x, ys)
    print(f"[qft34 c64] {prog.n_ops} gates", flush=True)
    dev = _native.DeviceState(n, "c64")
    try:
        dev.apply(prog)
        got = dev.gather(ys)
        err = float(np.abs(got - want).max())
        print(f"[qft34 c64] max-abs over 10^5 sampled amplitudes vs closed form {err:.3e}", flush=True)
        assert err < 1e-10
        u = np.random.Generator(np.random.PCG64(20250009)).random(1000000)
        idx = dev.sample(u).astype(np.int64)
        total = dev.norm2()
    finally:
        dev.close()
    cell = u * (float(1 << n) / total)
    m = np.floor(cell).astype(np.int64)
    print(f"[qft34 c64] norm^2 - 1 = {total - 1.0:.3e}; largest move {int(np.abs(idx - m).max())} cells", flush=True)
    assert abs(total - 1.0) < 1e-5
    assert np.all(np.abs(idx - m) <= m * (2.0 ** (n - 53) + 1e-6) + 1)
