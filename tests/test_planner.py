"""The fusion planner on the CPU: the plan the engine would execute (qfvm_plan_json) is replayed
op by op in NumPy and must reproduce the oracle's state. This checks the planner's reordering,
grouping and control / phase conversion without a GPU; the kernel that executes plans is checked
against the oracle by the GPU parity tests."""

import numpy as np
import pytest

from oracle import dense
from oracle.cpu import CpuState
from pyquafu_amd import _native, workloads
from pyquafu_amd.simulators.program import compile_circuit
from tests.conftest import rand_state
from tests.progs import random_gate_program

F_DENSE, F_REAL, F_DIAG, F_SWAP, F_PHASE, F_FLUSH, F_GRAD, F_RELABEL = range(8)
F_DENSE4 = 9
G_MULTI = 4


def move_bits(psi, n, Q, out_bit):
    """A pass's store with its tile qubits permuted: the bit at position Q[k] of every index moves to
    position Q[out_bit[k]] (the amplitudes stay at the tile's addresses)."""
    idx = np.arange(1 << n, dtype=np.int64)
    qmask = sum(1 << q for q in Q)
    dst = idx & ~qmask
    for k, q in enumerate(Q):
        dst |= ((idx >> q) & 1) << Q[out_bit[k]]
    new = np.empty_like(psi)
    new[dst] = psi
    return new


def replay(plan, prog, n, psi, collapsed=False):
    """Executes a plan's fused ops as controlled one-qubit gates on qubit positions, in plan order,
    and each pass's tile permutation (layout-aware plans). Non-fused items are dense gates with the
    matrix the plan gives them: the program's MULTI ops in order (their matrices checked against
    the program's), or, for plans with collapsed runs of gates, whatever the collapse formed."""
    psi = np.array(psi, dtype=complex)
    multi_rows = [] if collapsed else [r for r in prog.ops if r["code"] == 1]
    if any(it["fused"] and it["group_bits"] == 3 for it in plan["items"]):
        # row-of-8 plans run dense 2-target gates inside their groups (kind F_DENSE4)
        multi_rows = [r for r in multi_rows if int(r["n_qubits"]) - int(r["n_ctrl"]) != 2]
    mi = 0
    idx = np.arange(1 << n, dtype=np.int64)
    for it in plan["items"]:
        if not it["fused"]:
            assert it["kind"] == G_MULTI
            d = 1 << len(it["targets"])
            u = np.array([complex(*z) for z in it["m"]]).reshape(d, d)
            if not collapsed:
                row = multi_rows[mi]
                mi += 1
                q = prog.idx[row["q_off"]: row["q_off"] + row["n_qubits"]].tolist()
                nc = int(row["n_ctrl"])
                assert len(it["ctrls"]) == nc and len(it["targets"]) == len(q) - nc
                a = prog.mat[row["m_off"]: row["m_off"] + 2 * d * d]
                assert np.array_equal(u, (a[0::2] + 1j * a[1::2]).reshape(d, d))
            psi = dense.apply_unitary(psi, n, it["ctrls"], it["targets"], u)
            continue
        Q = it["Q"]
        assert Q == sorted(Q) and len(set(Q)) == len(Q)
        gb = it["group_bits"]
        assert gb in (3, 4)
        for G in it["groups"]:
            S = G["S"][:gb]
            assert len(set(S)) == gb and set(S) <= set(Q) and all(x == -1 for x in G["S"][gb:])
            for op in G["ops"]:
                if op["kind"] == F_FLUSH:  # moves data between registers only; no logical effect
                    assert 0 <= op["g"] < gb
                    continue
                m = [complex(*z) for z in op["m"]]
                cm = int(op["cme"])
                assert op["cmg"] < (1 << gb)
                for j in range(gb):
                    if (op["cmg"] >> j) & 1:
                        cm |= 1 << S[j]
                if op["kind"] == F_DENSE4:  # 4x4 on (S[g], S[g2]), index bit 0 <-> S[g]
                    assert gb == 3 and len(m) == 16 and op["g"] != op["g2"]
                    t = [S[op["g"]], S[op["g2"]]]
                    assert not any((cm >> q) & 1 for q in t)
                    ctrls = [q for q in range(n) if (cm >> q) & 1]
                    psi = dense.apply_unitary(psi, n, ctrls, t, np.array(m).reshape(4, 4))
                    continue
                if op["kind"] == F_PHASE:
                    q = op["g"]
                    assert q not in S
                else:
                    q = S[op["g"]]
                assert not (cm >> q) & 1
                ok = (idx & cm) == cm
                bit = (idx >> q) & 1
                if op["kind"] in (F_DIAG, F_PHASE):
                    f = np.where(bit == 1, m[3], m[0])
                    psi = np.where(ok, psi * f, psi)
                    continue
                i0 = idx[(bit == 0) & ok]
                i1 = i0 | (1 << q)
                a, b = psi[i0].copy(), psi[i1].copy()
                if op["kind"] in (F_SWAP, F_RELABEL):  # a relabeling is an X folded into parity masks
                    psi[i0], psi[i1] = b, a
                else:
                    if op["kind"] == F_REAL:
                        m = [z.real for z in m]
                    psi[i0] = m[0] * a + m[1] * b
                    psi[i1] = m[2] * a + m[3] * b
        if it["out_bit"]:
            assert sorted(it["out_bit"]) == list(range(len(Q)))
            psi = move_bits(psi, n, Q, it["out_bit"])
    assert mi == len(multi_rows)
    return psi


def check(prog, n, **kw):
    plan = _native.plan_json(prog, n, **kw)
    rng = np.random.Generator(np.random.PCG64(n))
    psi0 = rand_state(n, rng)
    cpu = CpuState(n)
    cpu.load(psi0)
    cpu.apply(prog)
    got = replay(plan, prog, n, psi0, collapsed=bool(kw.get("dense_blocks")))
    assert np.abs(got - cpu.store()).max() < 1e-12
    return plan


@pytest.mark.parametrize("n,tile,low,groups", [(6, 0, 0, 0), (9, 6, 2, 2), (10, 8, 4, 3), (12, 0, 0, 0),
                                                (13, 7, 3, 12), (14, 12, 6, 24), (16, 9, 4, 0), (17, 13, 5, 6)])
def test_random_programs_replay(n, tile, low, groups):
    rng = np.random.Generator(np.random.PCG64(500 + n))
    prog = random_gate_program(n, 300, rng, max_multi=3)
    check(prog, n, tile_bits=tile, low_bits=low, max_groups=groups)


@pytest.mark.parametrize("n,tile,low,groups", [(11, 10, 4, 0), (12, 11, 5, 0), (14, 12, 6, 24), (16, 10, 3, 4),
                                                (17, 13, 5, 6), (18, 11, 5, 12)])
def test_layout_plans_replay(n, tile, low, groups):
    """Layout-aware plans (option "layout"): gates run on moving positions, tiles are stored with
    their qubits permuted, and the plan ends in the standard order (the replay is compared with the
    oracle's state index by index)."""
    rng = np.random.Generator(np.random.PCG64(700 + n))
    prog = random_gate_program(n, 300, rng, max_multi=3)
    plan = check(prog, n, tile_bits=tile, low_bits=low, max_groups=groups, layout=True)
    assert any(it.get("out_bit") for it in plan["items"] if it["fused"])


def test_layout_plan_brickwork_fewer_passes():
    n = 20
    prog = compile_circuit(workloads.random_u3_cx(n, 8, 6, seed=11, measure=False))
    plain = _native.plan_json(prog, n, "c128")
    moved = check(prog, n, layout=True)
    p0 = sum(1 for it in plain["items"] if it["fused"])
    p1 = sum(1 for it in moved["items"] if it["fused"])
    assert p1 < 0.85 * p0, (p0, p1)
    # every gate runs exactly once; the passes at the end carry no gates, only the restoring moves
    ops = sum(1 for it in moved["items"] if it["fused"] for g in it["groups"] for op in g["ops"] if op["kind"] != F_FLUSH)
    assert ops == prog.n_ops


def test_brickwork_plan_is_compact():
    n = 20
    prog = compile_circuit(workloads.random_u3_cx(n, 6, 5, seed=9, measure=False))
    plan = check(prog, n)
    passes = [it for it in plan["items"] if it["fused"]]
    assert all(it["fused"] for it in plan["items"])
    # 6 layers of a 20-qubit brickwork need at least 6 * 14 / 6 passes with 6 free tile qubits
    assert len(passes) <= 6 * 4, len(passes)
    total_ops = sum(1 for it in passes for g in it["groups"] for op in g["ops"] if op["kind"] != F_FLUSH)
    assert total_ops == prog.n_ops


def test_qft_and_ansatz_plans():
    n = 14
    prog = compile_circuit(workloads.qft(n, 12345))
    check(prog, n)
    qc, _ = workloads.hea_ansatz(n, 3, parametrised=False)
    check(compile_circuit(qc), n)


def test_plan_depends_on_parameter_values():
    """The fused plan of RX(0) differs from that of RX(0.3) (a zero-angle rotation is diagonal and
    takes no tile slot), so a cache keyed on program structure alone must also compare plans: graph
    mode compares the captured graphs' shapes before it updates an executable."""
    from pyquafu_amd import _native
    from pyquafu_amd.simulators.program import RX, ProgramBuilder

    def prog(theta):
        b = ProgramBuilder()
        for q in range(14):
            b.add(RX, [q], param=theta)
        return b.program(n_qubits=14)

    p0 = _native.plan_json(prog(0.0), 14, "c128")
    p1 = _native.plan_json(prog(0.3), 14, "c128")
    assert p0 != p1


@pytest.mark.parametrize("seed", range(12))
def test_layout_plans_replay_random_shapes(seed):
    """Layout-aware plans over random sizes and tile shapes (tile 10..13 bits, 2..6 low bits, 1..24
    groups), random programs with controlled and multi-target gates, replayed against the oracle."""
    rng = np.random.Generator(np.random.PCG64(9000 + seed))
    tile = int(rng.integers(10, 14))
    n = int(rng.integers(tile + 1, 17))
    low = int(rng.integers(2, 7))
    groups = int(rng.integers(1, 25))
    prog = random_gate_program(n, 200, rng, max_multi=3)
    plan = check(prog, n, tile_bits=tile, low_bits=low, max_groups=groups, layout=True)
    fused = [it for it in plan["items"] if it["fused"]]
    assert all(len(it["Q"]) == tile for it in fused)


def test_dense_blocks_collapse_deep_runs():
    """Deep runs on 5-qubit blocks collapse into one dense 5-target gate per block, on the block's
    qubits in ascending order; the replayed plan reproduces the oracle."""
    n = 13
    prog = compile_circuit(workloads.deep_blocks(n, 5, 6))
    plan = check(prog, n, dense_blocks=5, dense_min_gates=8)
    dense_items = [it for it in plan["items"] if not it["fused"]]
    assert sorted(it["targets"] for it in dense_items) == [[0, 1, 2, 3, 4], [5, 6, 7, 8, 9]]
    # the 3-qubit tail stays gates (a dense gate needs 4 or more qubits), run in fused passes
    assert any(it["fused"] for it in plan["items"])
    # a threshold above the runs' gate count leaves the plan as it was
    assert _native.plan_json(prog, n, dense_blocks=5, dense_min_gates=10**6) == _native.plan_json(prog, n)


@pytest.mark.parametrize("k", [4, 5, 6, 8])
@pytest.mark.parametrize("seed", range(4))
def test_dense_blocks_random_programs(k, seed):
    """Random gate programs (controls, diagonal and dense multi-target gates) with every qualifying
    run collapsed: the order the collapse emits keeps the result."""
    rng = np.random.Generator(np.random.PCG64(900 + seed))
    n = int(rng.integers(7, 12))
    prog = random_gate_program(n, 80, rng)
    plan = check(prog, n, dense_blocks=k, dense_min_gates=int(rng.integers(1, 4)))
    assert all(len(it["targets"]) <= max(k, 6) for it in plan["items"] if not it["fused"])


def test_dense_blocks_option_range():
    prog = compile_circuit(workloads.deep_blocks(8, 4, 2))
    for bad in (1, 3, 9):
        with pytest.raises(_native.EngineError):
            _native.plan_json(prog, 8, dense_blocks=bad)


def test_single_qubit_runs_merged():
    """Runs of one-qubit gates on one qubit (U3 layers, H / T / S / RX chains) become one op each in
    the plan; a CX between them ends a run. The replayed plan reproduces the oracle."""
    from pyquafu_amd.circuits import QuantumCircuit
    from pyquafu_amd.elements import U3Gate

    n = 9
    rng = np.random.Generator(np.random.PCG64(5))
    qc = QuantumCircuit(n)
    for layer in range(6):
        for q in range(n):
            qc << U3Gate(q, *map(float, rng.uniform(0, 2 * np.pi, 3)))
        if layer == 2:
            qc.cx(0, 1)
    for q in range(n):
        qc.h(q)
        qc.t(q)
        qc.s(q)
        qc.rz(q, 0.3)
    prog = compile_circuit(qc)
    plan = check(prog, n)
    ops = [op for it in plan["items"] if it["fused"] for G in it["groups"] for op in G["ops"]
           if op["kind"] not in (F_FLUSH, F_RELABEL)]
    # qubits 0 and 1: runs before and after the CX (2 each); the others: one run each; plus the CX
    assert len(ops) <= 2 * 2 + (n - 2) + 1


def test_dense_blocks_in_layout_plans():
    """Collapsed blocks between layout-aware passes run on the qubits' current positions: deep
    5-qubit blocks, then a brickwork over all qubits, twice; the replayed plan reproduces the
    oracle."""
    from pyquafu_amd.circuits import QuantumCircuit

    n = 14
    qc = QuantumCircuit(n)
    for rep in range(2):
        for g in workloads.deep_blocks(n, 5, 4, seed=90 + rep).gates:
            qc << g
        for g in workloads.random_u3_cx(n, 2, 3, seed=95 + rep, measure=False).gates:
            qc << g
    plan = check(compile_circuit(qc), n, tile_bits=10, layout=True, dense_blocks=5)
    # per repetition: qubits 0..4, 5..9 and the 4-qubit tail 10..13
    assert sum(1 for it in plan["items"] if not it["fused"]) == 6
