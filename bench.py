"""Benchmark of the statevector hot path on one MI355X.

Workload (default): 30 qubits, complex128 (16 GiB state), depth 40; per layer a U3 on every qubit
(angles from a seeded PCG64), a CX brickwork (15 even + 14 odd pairs) and 15 CRZ gates on random
ordered qubit pairs: 2,960 gates. One step = the whole circuit applied to the HBM-resident state.

Printed (one JSON line):
  value          HBM GB/s in the pass model: sum over kernel launches of the bytes each launch must
                 read + write (2 * 16 B * amplitudes it touches) / wall time of the step
  roofline       the dominant kernel's algorithmic bytes per launch / its mean HIP-event duration,
                 against MI355X's 8 TB/s HBM peak; stream_ref times a plain in-place streaming kernel
                 (torch neg_) over a buffer of the state's size, and frac_of_stream = achieved / its rate
  layout         the same circuit with the layout-aware planner (option "layout"): time per step,
                 fused passes, speedup over the default planner
  merge_1q       runs of U3 layers with one-qubit runs merged (default) and not (option merge_1q)
  gate_equiv     the unfused per-gate byte model (sum over gates of 2 * 16 B * 2^(n - controls)) /
                 step time: comparable with the CPU, may exceed the HBM peak once gates are fused
  cpu_baseline   the CPU oracle (a restatement of qfvm's OpenMP engine, built with its flags) on
                 the first 4 layers of the same circuit, every host core, median of 3 runs, same unit
                 as gate_equiv; nproc and the CPU model are recorded
  parity_vs_oracle  max-abs difference over all 2^n amplitudes between the GPU and the oracle after
                 those 4 layers

roofline.traffic is the HBM bytes per launch of the dominant kernel read from rocprofv3 FETCH_SIZE
and WRITE_SIZE runs, each scaled by the same counter on a plain streaming kernel of known size
(`--copy-calibration`); the figure is kept in profiles/<kernel>_<n>q_<dtype>_traffic.txt and is null
when no such file exists for the workload.
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time


import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "HBM GB/s (% peak) on 30q c128 gates; adjoint grads/sec vs param-shift, 1 GPU"
PEAK_GBPS = 8000.0


def gate_equiv_bytes(prog, n: int, amp_bytes: int) -> float:
    """Sum over gates of 2 * s * 2^(n - controls) (the reference's one-sweep-per-gate model)."""
    from pyquafu_amd.simulators import program as P

    ctrl_of = {P.CX: 1, P.CZ: 1, P.CP: 1, P.CCX: 2}
    tot = 0.0
    for r in prog.ops:
        code = int(r["code"])
        if code in (P.MEASURE, P.RESET, P.CIF):
            continue
        if code == P.RZZ:
            tot += 2 * amp_bytes * (2 * 2.0 ** (n - 1) + 2.0**n)
            continue
        c = int(r["n_ctrl"]) if code in (P.GENERAL, P.MULTI) else ctrl_of.get(code, 0)
        tot += 2 * amp_bytes * 2.0 ** (n - c)
    return tot


def host_cpu():
    """(logical cores, CPU model) of this host, as nproc and lscpu report them."""
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.lower().startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    return os.cpu_count() or 1, model


def cpu_baseline(n: int, layers: int, crz: int, gpu_state=None, runs: int = 3):
    """The oracle (the reference's CPU algorithm with its flags: one OpenMP sweep per gate) on the
    first `layers` layers of the benchmark circuit, every host core, median of `runs` runs from |0>;
    with `gpu_state`, the max-abs difference of every amplitude against it."""
    cores, model = host_cpu()
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from oracle.cpu import CpuState
    from pyquafu_amd import workloads
    from pyquafu_amd.simulators.program import compile_circuit

    sub = compile_circuit(workloads.random_u3_cx(n, layers, crz, seed=20250002, measure=False))
    s = CpuState(n)
    times = []
    for r in range(runs):
        s.init_basis(0)
        t0 = time.perf_counter()
        s.apply(sub)
        times.append(time.perf_counter() - t0)
        print(f"[bench] CPU baseline run {r + 1}/{runs}: {times[-1]:.2f} s", file=sys.stderr, flush=True)
    dt = float(np.median(times))
    parity = None
    if gpu_state is not None:
        ref = s.view()
        err = 0.0
        for k in range(0, ref.size, 1 << 24):
            err = max(err, float(np.abs(gpu_state[k:k + (1 << 24)] - ref[k:k + (1 << 24)]).max()))
        parity = {"layers": layers, "gates": int(sub.n_ops), "amplitudes_compared": int(ref.size), "max_abs": err}
    by = gate_equiv_bytes(sub, n, 16)
    del s
    threads = int(os.environ["OMP_NUM_THREADS"])
    return {
        "value": round(by / dt / 1e9, 3),
        "unit": "GB/s gate-equivalent",
        "cores": threads,
        "kind": "port",
        "sample": f"first {layers} layers ({sub.n_ops} gates) of the same {n}-qubit circuit, complex128, from "
                  f"|0>, median of {runs} runs ({', '.join(f'{t:.2f}' for t in times)} s), one OpenMP sweep per gate",
        "seconds": round(dt, 3),
        "host": {"nproc": cores, "model": model, "omp_threads": threads,
                 "omp_proc_bind": os.environ.get("OMP_PROC_BIND")},
    }, parity


def traffic_from_profile(kernel: str, n: int, dtype: str):
    """HBM bytes per launch recorded by the counter runs for this kernel and workload, or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{n}q_{dtype}_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split("=")[1])
    return None


def copy_calibration() -> int:
    """A plain streaming kernel of known size (torch.neg over 4 GiB of float64: reads 4 GiB and
    writes 4 GiB per launch), run 3 times, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    nbytes = 1 << 32
    src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"op": "torch.neg float64", "read_bytes_per_launch": nbytes,
                                           "write_bytes_per_launch": nbytes, "launches": 3}}), flush=True)
    return 0


def stream_reference(nbytes: int) -> dict:
    """A plain streaming kernel over a buffer the size of the state: torch's in-place neg_ on
    float64 (reads and writes every byte once), timed with events on torch's current stream, best of
    3, in a child process (torch's own HIP runtime does not share a process with the engine's). Its
    rate is the practical HBM ceiling a pass is compared with."""
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--stream-reference", str(nbytes)],
                       capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError(f"stream reference failed: {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def stream_reference_child(nbytes: int) -> int:
    import torch

    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    x.neg_()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    best = float("inf")
    for _ in range(3):
        a.record()
        x.neg_()
        b.record()
        b.synchronize()
        best = min(best, a.elapsed_time(b))
    del x
    gbps = 2 * nbytes / (best / 1e3) / 1e9
    print(json.dumps({"op": "torch neg_ in place, float64", "bytes": 2 * nbytes, "ms": round(best, 4),
                      "GBps": round(gbps, 2), "frac": round(gbps / PEAK_GBPS, 4)}), flush=True)
    return 0


def adjoint_leg(n: int, layers: int, fd_params: int = 8):
    """Adjoint gradient vs parameter shift on the same engine: 24-qubit complex128 hardware-efficient
    ansatz (per layer RY and RZ on every qubit, CZ ring; parameters U[-pi, pi], PCG64 seed 20250003),
    loss <H> of the periodic TFIM H = sum Z_i Z_i+1 + 0.7 sum X_i. One full gradient = the loss and
    all 2 * n * layers partials. Parameter shift runs all 2P shifted circuits (each: reset, the
    fused forward passes, one Hamiltonian pass); central differences (h = 1e-6) on the first
    `fd_params` parameters are a sanity bound."""
    from pyquafu_amd import _native, workloads
    from pyquafu_amd.algorithms.gradients.gradients import adjoint_program, hamiltonian_terms

    qc, values = workloads.hea_ansatz(n, layers, parametrised=True)
    ham = workloads.tfim_hamiltonian(n)
    terms, coeffs = hamiltonian_terms(ham)
    prog, ops, gens, _ = adjoint_program(qc)
    dev = _native.DeviceState(n, "c128")

    def adjoint_once():
        dev.init_zero()
        return dev.adjoint(prog, ops, gens, terms, coeffs)

    adjoint_once()
    times = []
    for _ in range(5):
        t0 = time.perf_counter()
        grads, e = adjoint_once()
        times.append(time.perf_counter() - t0)
    t_adj = float(np.median(times))
    dev.set_profiling(True)
    adjoint_once()
    prof = dev.profile(reset=True)
    dev.set_profiling(False)
    # the same gradient with 4 forward-state checkpoints restored during the reverse sweep
    dev.set_option("checkpoints", 4)
    adjoint_once()
    ck_times = []
    for _ in range(5):
        t0 = time.perf_counter()
        ck_grads, _ = adjoint_once()
        ck_times.append(time.perf_counter() - t0)
    dev.set_option("checkpoints", 0)
    t_ck = float(np.median(ck_times))

    def energy(p):
        dev.init_zero()
        dev.apply(p)
        return dev.expect_hamiltonian(terms, coeffs)

    energy(prog)
    t0 = time.perf_counter()
    ps = np.zeros(len(ops))
    for k, row in enumerate(ops):
        theta = float(prog.ops["param"][row])
        prog.ops["param"][row] = theta + np.pi / 2
        plus = energy(prog)
        prog.ops["param"][row] = theta - np.pi / 2
        minus = energy(prog)
        prog.ops["param"][row] = theta
        ps[k] = (plus - minus) / 2
    t_ps = time.perf_counter() - t0
    fd = np.zeros(fd_params)
    for k in range(fd_params):
        row = ops[k]
        theta = float(prog.ops["param"][row])
        prog.ops["param"][row] = theta + 1e-6
        plus = energy(prog)
        prog.ops["param"][row] = theta - 1e-6
        minus = energy(prog)
        prog.ops["param"][row] = theta
        fd[k] = (plus - minus) / 2e-6
    dev.close()
    return {
        "workload": f"{n}q c128 hardware-efficient ansatz, {layers} layers (RY+RZ on every qubit, CZ ring), "
                    f"TFIM h=0.7 periodic, {len(ops)} parameters",
        "adjoint_grads_per_s": round(1.0 / t_adj, 3),
        "adjoint_ms": round(t_adj * 1e3, 3),
        "param_shift_grads_per_s": round(1.0 / t_ps, 4),
        "param_shift_ms": round(t_ps * 1e3, 1),
        "speedup_vs_param_shift": round(t_ps / t_adj, 1),
        "max_abs_diff_vs_param_shift": float(np.abs(grads - ps).max()),
        "max_abs_diff_vs_central_fd": float(np.abs(grads[:fd_params] - fd).max()),
        "expectation": e,
        "checkpoints": {"count": 4, "adjoint_ms": round(t_ck * 1e3, 3),
                        "max_abs_diff_vs_plain": float(np.abs(ck_grads - grads).max())},
        "kernels": {k: {"launches": v["launches"], "ms": round(v["ms"], 3)} for k, v in prof.items()},
    }


def qft_leg(n: int, shots: int, samples: int = 100000):
    """Precision study: the swap-free QFT (H, then CP(2 pi / 2^k) from the later qubits, for each
    target in textbook order) of a seeded basis state |x> (x ~ U{0 .. 2^n - 1}, PCG64 seed
    20250005), run in complex64 and complex128 side by side. Errors: max-abs over `samples` seeded
    random amplitudes against the closed form 2^(-n/2) exp(2 pi i x rev(y) / 2^n) (integer phase,
    float64 exp), and the fidelity |<psi64|psi128>|^2 over the whole state. Throughput: pass-model
    GB/s per dtype and the sampler's shots/s on the complex128 state."""
    from pyquafu_amd import _native, workloads
    from pyquafu_amd.simulators.program import compile_circuit

    rng = np.random.Generator(np.random.PCG64(20250005))
    x = int(rng.integers(0, 1 << n))
    ys = rng.integers(0, 1 << n, size=samples, dtype=np.int64)
    prog = compile_circuit(workloads.qft(n, x))
    want = workloads.qft_amplitudes(n, x, ys)
    out = {"workload": f"{n}q QFT of |x>, x={x}, {prog.n_ops} gates", "dtypes": {}}
    devs = {}
    for dtype in ("c64", "c128"):
        try:
            dev = _native.DeviceState(n, dtype)
        except _native.EngineError as e:
            out["dtypes"][dtype] = {"skipped": str(e)}
            continue
        dev.init_zero()
        dev.apply(prog)  # warm-up (plans, op upload)
        dev.init_zero()
        dev.synchronize()
        dev.profile(reset=True)
        t0 = time.perf_counter()
        dev.apply(prog)
        dev.synchronize()
        dt = time.perf_counter() - t0
        c = dev.profile(reset=True)
        by = sum(v["bytes"] for v in c.values())
        got = dev.gather(ys)
        out["dtypes"][dtype] = {
            "ms": round(dt * 1e3, 2),
            "launches": int(sum(v["launches"] for v in c.values())),
            "GBps": round(by / dt / 1e9, 1),
            "frac_of_peak": round(by / dt / 1e9 / PEAK_GBPS, 4),
            "max_abs_err_sampled": float(np.abs(got - want).max()),
            "norm_error": abs(dev.norm2() - 1.0),
        }
        devs[dtype] = dev
    if len(devs) == 2:
        ov = devs["c64"].inner(devs["c128"])
        out["fidelity_c64_vs_c128"] = float(abs(ov) ** 2)
    if "c128" in devs and shots:
        u = np.random.Generator(np.random.PCG64(20250006)).random(shots)
        dev = devs["c128"]
        dev.sample(u[:1000])
        t0 = time.perf_counter()
        idx = dev.sample(u)
        dt = time.perf_counter() - t0
        out["sampler"] = {"shots": shots, "ms": round(dt * 1e3, 2), "shots_per_s": round(shots / dt, 1),
                          "distinct": int(np.unique(idx).size)}
    for dev in devs.values():
        dev.close()
    return out


def batch_leg(count: int, workers: int, graphs: bool = True, steps: int = 3, pin_qubits: int = 21):
    """Batched uneven workload: `count` seeded circuits (PCG64 seed 20250004), n ~ U{10..22}, depth ~
    U{4..16} layers of RY and RZ on every qubit plus a CZ ring, complex64; each evaluates <sum_i Z_i>
    and its full adjoint gradient. One engine call per step (qfvm_batch_run): `workers` engine
    threads, one handle (one HIP stream) per qubit count each; tile shapes and fusion width from the
    autotuner. Compilation, tuning and one untimed step (plans, graphs) come first; then `steps`
    training steps, each rebinding every circuit's parameters in place (new values) and running the
    batch. Reported: circuits/s and latencies of the steps, the host time of one rebinding of all
    circuits against the engine call's wall time, and the same with graphs off."""
    from pyquafu_amd import workloads
    from pyquafu_amd.batch import Autotuner, BatchRunner, compile_item

    t0 = time.perf_counter()
    circuits = workloads.batch_circuits(count=count)
    items = [compile_item(qc, workloads.z_sum(qc.num)) for qc, _ in circuits]
    t_compile = time.perf_counter() - t0
    tuner = Autotuner()
    rng = np.random.Generator(np.random.PCG64(20250006))
    step_values = [[rng.uniform(-np.pi, np.pi, size=it.pp.n_vars) for it in items] for _ in range(steps)]

    def leg(graph_mode: bool):
        runner = BatchRunner(dtype="c64", workers=workers, tuner=tuner, graphs=graph_mode, pin_qubits=pin_qubits)
        t0 = time.perf_counter()
        runner.warm([it.n for it in items])
        t_warm = time.perf_counter() - t0
        for it, (_, v) in zip(items, circuits):
            it.bind(v)
        runner.run(items)  # untimed: plans, graphs
        walls, binds, lats = [], [], []
        for vals in step_values:
            t0 = time.perf_counter()
            for it, v in zip(items, vals):
                it.bind(v)
            binds.append(time.perf_counter() - t0)
            results, wall = runner.run(items)
            walls.append(wall)
            lats.extend(r.latency_s for r in results)
        runner.close()
        lat = np.array(lats) * 1e3
        wall = float(np.median(walls))
        return {
            "circuits_per_s": round(count / wall, 1),
            "wall_ms": round(wall * 1e3, 2),
            "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
            "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
            "bind_ms": round(float(np.median(binds)) * 1e3, 2),
            "tune_s": round(t_warm, 2),
        }, results

    main, results = leg(graphs)
    extra = {}
    if graphs:  # the same steps eager (never graphs under a profiler: --no-graph-batch)
        other, other_results = leg(False)
        other["max_abs_grad_diff"] = max(float(np.abs(a.grads - b.grads).max(initial=0.0))
                                         for a, b in zip(results, other_results))
        extra["graphs_off"] = other
    return {
        "workload": f"{count} circuits, n 10-22, depth 4-16 (RY+RZ layers, CZ ring), c64, <sum Z> + adjoint gradient",
        "graphs": bool(graphs),
        **main,
        "steps": steps,
        "workers": workers,
        "pin_qubits": pin_qubits,
        "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "4 (HIP default)"),
        "parameters": int(sum(len(it.ops) for it in items)),
        "host_compile_s": round(t_compile, 2),
        "tiles": {str(k[0]): list(v) for k, v in sorted(tuner.choice.items())},
        **extra,
    }


def dense_leg(n: int = 28, reps: int = 3):
    """Dense gates on k = 4 to 8 targets (seeded Haar-random unitaries, PCG64 seed 20250006;
    targets the k lowest, the k highest, or qubits 8 .. 8+k-1) on an n-qubit state, complex128 and
    complex64, on the matrix cores (option mfma = 1: FP64 / FP32 MFMA) and on the vector-ALU kernels
    (mfma = 0; for k = 7 and 8 only with the middle targets, one timed call). Per gate: mean
    HIP-event duration, algorithmic bytes 2 * s * 2^n and flops 8 * 2^k * 2^n."""
    from pyquafu_amd import _native
    from pyquafu_amd.simulators.program import MULTI, ProgramBuilder

    rng = np.random.Generator(np.random.PCG64(20250006))
    out = {"workload": f"{n}q, one dense k-target gate", "rows": []}
    mats = {}
    for k in (4, 5, 6, 7, 8):
        d = 1 << k
        a = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
        q, r = np.linalg.qr(a)
        mats[k] = q * (np.diag(r) / np.abs(np.diag(r)))
    for dtype, amp in (("c128", 16), ("c64", 8)):
        dev = _native.DeviceState(n, dtype)
        dev.set_profiling(True)
        for k in (4, 5, 6, 7, 8):
            d = 1 << k
            for where, tg in (("low", list(range(k))), ("mid", list(range(8, 8 + k))), ("high", list(range(n - k, n)))):
                b = ProgramBuilder()
                b.add(MULTI, tg, matrix=mats[k])
                prog = b.program(n_qubits=n)
                for mfma in (1, 0):
                    if not mfma and k >= 7 and where != "mid":
                        continue
                    dev.set_option("mfma", mfma)
                    dev.init_zero()
                    dev.apply(prog)
                    dev.synchronize()
                    dev.profile(reset=True)
                    for _ in range(reps if mfma or k < 7 else 1):
                        dev.apply(prog)
                    dev.synchronize()
                    prof = dev.profile(reset=True)
                    name = "multi_mfma" if mfma else ("multi_reg" if k <= 5 else "multi_lds")
                    ms = prof[name]["ms"] / prof[name]["launches"]
                    out["rows"].append({"dtype": dtype, "k": k, "targets": where, "kernel": name, "ms": round(ms, 3),
                                        "GBps": round(2 * amp * 2 ** n / ms / 1e6, 1),
                                        "TFLOPs": round(8 * d * 2 ** n / ms / 1e9, 2)})
        dev.close()
    out["two_target_circuit"] = two_target_leg(n)
    out["dense_blocks"] = blocks_leg(n)
    return out


def two_target_leg(n: int, layers: int = 8, reps: int = 2):
    """A circuit of dense 2-target gates (workloads.two_target_dense, PCG64 seed 20250007) on n qubits in both
    precisions: those gates as 4x4 ops inside the row-of-8 passes (default options) against one
    per-gate sweep each (option rows8 = 0). Wall time per circuit, best of `reps`."""
    from pyquafu_amd import _native, workloads
    from pyquafu_amd.simulators.program import compile_circuit

    prog = compile_circuit(workloads.two_target_dense(n, layers, 20250007))
    res = {"workload": f"{n}q, {layers} layers of U3 + {n // 2} SWAP/iSWAP/RXX/RYY + 1 CSWAP", "gates": int(prog.n_ops)}
    for dtype in ("c128", "c64"):
        dev = _native.DeviceState(n, dtype)
        row = {}
        for name, rows8 in (("fused", 1), ("per_gate", 0)):
            dev.set_option("rows8", rows8)
            dev.init_zero()
            dev.apply(prog)
            dev.synchronize()
            best = float("inf")
            for _ in range(reps):
                dev.profile(reset=True)
                t0 = time.perf_counter()
                dev.apply(prog)
                dev.synchronize()
                best = min(best, time.perf_counter() - t0)
            prof = dev.profile(reset=True)
            row[name] = {"ms": round(best * 1e3, 2), "launches": {k: v["launches"] for k, v in prof.items()}}
        row["speedup"] = round(row["per_gate"]["ms"] / row["fused"]["ms"], 2)
        res[dtype] = row
        dev.close()
    return res


def blocks_leg(n: int, blocks=(5, 6), layer_counts=(1, 2, 4, 8, 16), reps: int = 2):
    """Option dense_blocks on deep runs of gates on few qubits (workloads.deep_blocks, PCG64 seed
    20250008: per layer U3 on every qubit, then a CX chain inside each block of k consecutive
    qubits) on n qubits: the gates in fused passes (default options) against each block's run
    collapsed into one dense k-target gate on the matrix cores (dense_blocks = k, dense_min_gates = 1,
    so every block collapses), for 1 to 16 layers and both precisions. Wall time per circuit, the
    host's collapse included, best of `reps`, and the kernels' HIP-event time per circuit (mean of
    the `reps` calls)."""
    from pyquafu_amd import _native, workloads
    from pyquafu_amd.simulators.program import compile_circuit

    res = {"workload": f"{n}q deep_blocks: per layer U3 on every qubit + CX chain in each block of k qubits",
           "rows": []}
    for dtype in ("c128", "c64"):
        dev = _native.DeviceState(n, dtype)
        dev.set_profiling(True)
        for k in blocks:
            for layers in layer_counts:
                prog = compile_circuit(workloads.deep_blocks(n, k, layers))
                row = {"dtype": dtype, "k": k, "layers": layers, "gates": int(prog.n_ops)}
                for name, opts in (("passes", (("dense_blocks", 0),)),
                                   ("dense", (("dense_blocks", k), ("dense_min_gates", 1)))):
                    for key, v in opts:
                        dev.set_option(key, v)
                    dev.init_zero()
                    dev.apply(prog)
                    dev.synchronize()
                    best = float("inf")
                    for _ in range(reps):
                        dev.profile(reset=True)
                        t0 = time.perf_counter()
                        dev.apply(prog)
                        dev.synchronize()
                        best = min(best, time.perf_counter() - t0)
                    dev.profile(reset=True)
                    for _ in range(reps):
                        dev.apply(prog)
                    dev.synchronize()
                    prof = dev.profile(reset=True)
                    row[name] = {"ms": round(best * 1e3, 2),
                                 "kernel_ms": round(sum(v["ms"] for v in prof.values()) / reps, 2),
                                 "launches": {key: v["launches"] // reps for key, v in prof.items()}}
                row["speedup"] = round(row["passes"]["ms"] / row["dense"]["ms"], 2)
                res["rows"].append(row)
        dev.close()
    return res


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=30)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--crz", type=int, default=15)
    ap.add_argument("--dtype", default="c128", choices=["c128", "c64"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-layers", type=int, default=4,
                    help="layers of the circuit the CPU baseline times (and the GPU state is compared on)")
    ap.add_argument("--no-fusion", action="store_true", help="one kernel per gate (the unfused path)")
    ap.add_argument("--tile-bits", type=int, default=0)
    ap.add_argument("--low-bits", type=int, default=0)
    ap.add_argument("--max-groups", type=int, default=0)
    ap.add_argument("--stream-reference", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only the streaming calibration kernel for the HBM counters")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-vs-parameter-shift leg")
    ap.add_argument("--adjoint-n", type=int, default=24)
    ap.add_argument("--adjoint-layers", type=int, default=8)
    ap.add_argument("--no-qft", action="store_true", help="skip the 33-qubit QFT precision leg")
    ap.add_argument("--no-dense", action="store_true", help="skip the dense k-target gate leg (matrix cores)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched uneven-circuit leg")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--batch-workers", type=int, default=8)
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="GPU_MAX_HW_QUEUES for this process (1..32; 0 leaves the environment alone): HIP maps "
                         "the batch leg's engine streams onto that many hardware queues, and kernels that share "
                         "a queue run one after another")
    ap.add_argument("--batch-pin", type=int, default=21,
                    help="circuits of at least this many qubits run one after another on one worker (0: off)")
    ap.add_argument("--no-graph-batch", action="store_true",
                    help="run the batch leg eagerly only, without hipGraph capture (profiled runs use this)")
    ap.add_argument("--qft-n", type=int, default=33)
    ap.add_argument("--qft-shots", type=int, default=1000000)
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the statevector "
              "engine holds the whole state on one MI355X; no multi-GPU figure is produced)", file=sys.stderr)
        return 2
    if args.hw_queues:  # read once at HIP initialisation, which the first engine call below performs
        if not 1 <= args.hw_queues <= 32:
            print("--hw-queues must be in 1..32", file=sys.stderr)
            return 2
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    if args.copy_calibration:
        return copy_calibration()
    if args.stream_reference:
        return stream_reference_child(args.stream_reference)

    from pyquafu_amd import _native, workloads
    from pyquafu_amd.simulators.program import compile_circuit

    n = args.n
    amp = 16 if args.dtype == "c128" else 8
    qc = workloads.random_u3_cx(n, args.depth, args.crz, seed=20250002, measure=False)
    prog = compile_circuit(qc)
    print(f"[bench] {n}q {args.dtype} depth {args.depth}: {prog.n_ops} gates", file=sys.stderr, flush=True)

    dev = _native.DeviceState(n, args.dtype)
    dev.set_option("fusion", 0 if args.no_fusion else 1)
    for key in ("tile_bits", "low_bits", "max_groups"):
        if getattr(args, key):
            dev.set_option(key, getattr(args, key))
    dev.init_zero()
    for k in range(args.warmup):
        dev.apply(prog)
        print(f"[bench] warmup {k + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    dev.synchronize()
    dev.profile(reset=True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        dev.apply(prog)
        print(f"[bench] step {k + 1}/{args.steps} done", file=sys.stderr, flush=True)
    dev.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    counters = dev.profile(reset=True)
    pass_bytes = sum(v["bytes"] for v in counters.values()) / args.steps
    launches = sum(v["launches"] for v in counters.values()) / args.steps

    # one more step with HIP events around every launch, for the per-kernel roofline
    dev.set_profiling(True)
    dev.apply(prog)
    dev.synchronize()
    prof = dev.profile(reset=True)
    dev.set_profiling(False)
    dom_name, dom = max(prof.items(), key=lambda kv: kv[1]["ms"])
    per_launch_bytes = dom["bytes"] / dom["launches"]
    per_launch_s = dom["ms"] / dom["launches"] / 1e3
    achieved = per_launch_bytes / per_launch_s / 1e9
    norm_err = abs(dev.norm2() - 1.0)

    # expect_statevec of the periodic TFIM (n ZZ + n X terms) on the resident final state: all terms
    # in a few LDS-tiled passes, one host sync
    tfim = [("ZZ", [i, (i + 1) % n]) for i in range(n)] + [("X", [i]) for i in range(n)]
    dev.expect(tfim)
    dev.synchronize()
    dev.profile(reset=True)
    t0 = time.perf_counter()
    tfim_vals = dev.expect(tfim)
    t_tfim = time.perf_counter() - t0
    tfim_prof = dev.profile(reset=True)
    expect_tfim = {
        "terms": len(tfim),
        "ms": round(t_tfim * 1e3, 3),
        "launches": {k: v["launches"] for k, v in tfim_prof.items()},
        "state_reads": int(sum(v["launches"] for k, v in tfim_prof.items() if k in ("expect_tile", "pauli_expect"))),
        "energy_h0.7": float(sum(tfim_vals[:n]) + 0.7 * sum(tfim_vals[n:])),
    }

    # the layout-aware planner (option "layout") on the same circuit: fewer passes, more gates each
    dev.set_option("layout", 1)
    dev.apply(prog)
    dev.synchronize()
    dev.profile(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dev.apply(prog)
    dev.synchronize()
    dt_lay = (time.perf_counter() - t0) / args.steps
    lay_prof = dev.profile(reset=True)
    dev.set_option("layout", 0)
    layout_leg = {
        "option": "layout",
        "ms_per_step": round(dt_lay * 1e3, 3),
        "fused_passes_per_step": lay_prof.get("fused_pass", {}).get("launches", 0) / args.steps,
        "speedup_vs_default": round(dt / dt_lay, 4),
        "pass_model_GBps": round(sum(v["bytes"] for v in lay_prof.values()) / args.steps / dt_lay / 1e9, 2),
    }

    # option merge_1q (runs of one-qubit gates merged, default on) on runs of U3 layers
    runs_leg = {"workload": f"{n}q {args.dtype} workloads.u3_runs: 8 layers of U3 on every qubit, CX(2i, 2i+1) "
                            "after every other layer (PCG64 seed 20250009)"}
    rprog = compile_circuit(workloads.u3_runs(n, 8, 2))
    for merge in (1, 0):
        dev.set_option("merge_1q", merge)
        dev.apply(rprog)
        dev.synchronize()
        dev.profile(reset=True)
        t0 = time.perf_counter()
        dev.apply(rprog)
        dev.synchronize()
        rp = dev.profile(reset=True)
        runs_leg["merged" if merge else "unmerged"] = {
            "ms": round((time.perf_counter() - t0) * 1e3, 2),
            "fused_passes": rp.get("fused_pass", {}).get("launches", 0)}
    dev.set_option("merge_1q", 1)
    runs_leg["speedup"] = round(runs_leg["unmerged"]["ms"] / runs_leg["merged"]["ms"], 2)

    ge = gate_equiv_bytes(prog, n, amp)
    out = {
        "metric": METRIC,
        "value": round(pass_bytes / dt / 1e9, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic: seeded random U3 + CX brickwork + CRZ circuit (PCG64 seed 20250002), state from |0>",
        "config": {
            "workload": f"{n}q {args.dtype} random U3+CX brickwork+{args.crz} CRZ/layer, depth {args.depth}",
            "n_qubits": n,
            "depth": args.depth,
            "gates": int(prog.n_ops),
            "kernel_launches_per_step": launches,
            "pass_bytes_per_step": pass_bytes,
            "parallelism": "single GPU",
        },
        "frac_of_peak": round(pass_bytes / dt / 1e9 / PEAK_GBPS, 4),
        "gate_equiv": {"GBps": round(ge / dt / 1e9, 2), "bytes_per_step": ge},
        "roofline": {
            "bound": "hbm",
            "kernel": dom_name,
            "achieved": round(achieved, 2),
            "peak": PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_GBPS, 4),
            "traffic": traffic_from_profile(dom_name, n, args.dtype),
            "bytes_per_launch": per_launch_bytes,
            "mean_launch_ms": round(per_launch_s * 1e3, 4),
        },
        "kernels": {k: {"launches": v["launches"], "ms": round(v["ms"], 3)} for k, v in prof.items()},
        "norm_error": norm_err,
        "expect_tfim": expect_tfim,
        "layout": layout_leg,
        "merge_1q": runs_leg,
    }
    dev.close()
    print("[bench] streaming reference", file=sys.stderr, flush=True)
    sref = stream_reference(amp * (1 << n))
    out["roofline"]["stream_ref"] = sref
    out["roofline"]["frac_of_stream"] = round(achieved / sref["GBps"], 4)
    if not args.no_adjoint:
        print(f"[bench] adjoint leg: {args.adjoint_n}q, {args.adjoint_layers} layers", file=sys.stderr, flush=True)
        out["adjoint"] = adjoint_leg(args.adjoint_n, args.adjoint_layers)
    else:
        out["adjoint"] = None
    if not args.no_batch:
        print(f"[bench] batch leg: {args.batch_count} circuits", file=sys.stderr, flush=True)
        out["batch"] = batch_leg(args.batch_count, args.batch_workers, not args.no_graph_batch,
                                 pin_qubits=args.batch_pin)
    else:
        out["batch"] = None
    if not args.no_qft:
        print(f"[bench] QFT precision leg: {args.qft_n}q, c64 and c128", file=sys.stderr, flush=True)
        out["precision"] = qft_leg(args.qft_n, args.qft_shots)
    else:
        out["precision"] = None
    if not args.no_dense:
        print("[bench] dense gate leg: k = 4 to 8 on 28 qubits", file=sys.stderr, flush=True)
        out["dense"] = dense_leg()
    else:
        out["dense"] = None
    if not args.no_cpu:
        # the same first layers on the GPU, every amplitude read back for the comparison
        sub = compile_circuit(workloads.random_u3_cx(n, args.cpu_layers, args.crz, seed=20250002, measure=False))
        gdev = _native.DeviceState(n, args.dtype)
        gdev.apply(sub)
        gstate = gdev.store()
        gdev.close()
        print(f"[bench] CPU baseline: {args.cpu_layers} layers ({sub.n_ops} gates) on the oracle", file=sys.stderr,
              flush=True)
        out["cpu_baseline"], out["parity_vs_oracle"] = cpu_baseline(n, args.cpu_layers, args.crz, gstate)
        del gstate
    else:
        out["cpu_baseline"] = None
        out["parity_vs_oracle"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
