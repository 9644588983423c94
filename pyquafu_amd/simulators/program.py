"""Circuit -> flat program compiler (host side of the C-ABI boundary).

Restates what the reference's pybind layer does when it builds its C++ `Circuit`
(src/qfvm/circuit.hpp:98-135) from a Python circuit, instruction by instruction through
`from_pyops` (src/qfvm/instructions.hpp:229-299):

  * `barrier`, `delay`, `id` are dropped;
  * `measure` records (qubit, cbit) pairs in order; any other instruction after a measure clears
    the final-measure flag;
  * an instruction with a `circuit` attribute (CircuitWrapper) is flattened recursively; its own
    measures run as instructions but do not enter the pair list (the reference moves only the
    inner instructions);
  * a gate whose `name.lower()` is in qfvm's native set (OPMAP, src/qfvm/qasm.hpp:43-46) becomes a
    named op with its first parameter; any other gate becomes GENERAL (one target) or MULTI, carrying
    `_get_targ_matrix(reverse_order=True)` for controlled gates and `matrix` otherwise;
  * positions are `pos` (controls first), the control count is `len(ctrls)`;
  * `cif` carries its cbits, condition and a compiled body (without wrapper flattening, as
    from_pyops recurses without it).

The result is three flat arrays laid out as `qfvm_op` / `qfvm_program` in include/qfvm.h.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

# op codes (include/qfvm.h)
GENERAL, MULTI, X, Y, Z, H, S, SDG, T, TDG, P, RX, RY, RZ, CX, CZ, CP, CCX, RZZ, MEASURE, RESET, CIF, KRAUS, UCHAN = \
    range(24)

NAMED = {
    "x": X, "y": Y, "z": Z, "h": H, "s": S, "sdg": SDG, "t": T, "tdg": TDG, "p": P, "rx": RX, "ry": RY,
    "rz": RZ, "cnot": CX, "cx": CX, "cz": CZ, "cp": CP, "ccx": CCX, "rzz": RZZ,
}
PARAMETRISED = {P, RX, RY, RZ, CP, RZZ}
SKIPPED = {"barrier", "delay", "id"}

OP_DTYPE = np.dtype(
    [
        ("code", "<i4"),
        ("n_ctrl", "<i4"),
        ("n_qubits", "<i4"),
        ("n_cbits", "<i4"),
        ("q_off", "<i8"),
        ("c_off", "<i8"),
        ("m_off", "<i8"),
        ("param", "<f8"),
        ("cond", "<i8"),
        ("body", "<i8"),
    ]
)
assert OP_DTYPE.itemsize == 64


@dataclass
class Program:
    """Flat op stream plus the facts simulate_circuit needs about the circuit."""

    ops: np.ndarray
    idx: np.ndarray
    mat: np.ndarray
    n_qubits: int = 0
    n_cbits: int = 0
    measure_vec: List[Tuple[int, int]] = field(default_factory=list)
    final_measure: bool = True
    n_random: int = 0  # MEASURE, RESET and channel ops in the stream (uniforms one shot may need)

    @property
    def n_ops(self) -> int:
        return len(self.ops)

    @property
    def has_classical(self) -> bool:
        return bool(np.isin(self.ops["code"], [MEASURE, RESET, CIF, KRAUS, UCHAN]).any()) if len(self.ops) else False


class ProgramBuilder:
    def __init__(self):
        self.rows = []
        self.idx: List[int] = []
        self.mat: List[np.ndarray] = []
        self.mat_len = 0

    def _push_idx(self, values) -> int:
        off = len(self.idx)
        self.idx.extend(int(v) for v in values)
        return off

    def _push_mat(self, m: np.ndarray) -> int:
        a = np.ascontiguousarray(np.asarray(m, dtype=np.complex128)).view(np.float64).ravel()
        off = self.mat_len
        self.mat.append(a)
        self.mat_len += a.size
        return off

    def add(self, code, qubits=(), n_ctrl=0, cbits=(), matrix=None, param=0.0, cond=0, body=0) -> int:
        row = (
            code,
            n_ctrl,
            len(qubits),
            len(cbits),
            self._push_idx(qubits),
            self._push_idx(cbits),
            -1 if matrix is None else self._push_mat(matrix),
            float(param),
            int(cond),
            int(body),
        )
        self.rows.append(row)
        return len(self.rows) - 1

    def set_body(self, row: int, body: int) -> None:
        r = list(self.rows[row])
        r[9] = body
        self.rows[row] = tuple(r)

    def program(self, **kw) -> Program:
        ops = np.array(self.rows, dtype=OP_DTYPE) if self.rows else np.zeros(0, dtype=OP_DTYPE)
        idx = np.array(self.idx, dtype=np.int32)
        mat = np.concatenate(self.mat) if self.mat else np.zeros(0, dtype=np.float64)
        return Program(ops, idx, mat, **kw)


def gate_matrix(op) -> np.ndarray:
    """The matrix a GENERAL / MULTI op carries: the target matrix (reverse_order=True) of a
    controlled gate, else `matrix`."""
    if hasattr(op, "_targ_matrix"):
        m = op._get_targ_matrix(reverse_order=True)  # pylint: disable=protected-access
    else:
        m = op.matrix
    return np.asarray(m, dtype=np.complex128)


def _emit_pyop(b: ProgramBuilder, op) -> str:
    """from_pyops (instructions.hpp:229-299) for one non-wrapper instruction; returns its kind."""
    name = op.name.lower()
    if name in SKIPPED:
        return "skip"
    if name == "measure":
        b.add(MEASURE, qubits=list(op.qbits), cbits=list(op.cbits))
        return "measure"
    if name == "reset":
        b.add(RESET, qubits=list(op.pos))
        return "reset"
    if hasattr(op, "gatelist"):  # noise channel (elements/noise.py)
        mats = [np.asarray(g.matrix, dtype=np.complex128) for g in op.gatelist]
        if any(m.shape != (2, 2) for m in mats):
            raise ValueError(f"channel {op.name}: only one-qubit operators are supported")
        if hasattr(op, "probs"):
            probs = np.asarray(op.probs, dtype=np.float64)
            if probs.shape != (len(mats),):
                raise ValueError(f"channel {op.name}: one probability per gate")
            b.add(UCHAN, qubits=list(op.pos), matrix=np.concatenate([m.reshape(-1) for m in mats] +
                                                                    [probs.astype(np.complex128)]),
                  body=len(mats))
        else:
            b.add(KRAUS, qubits=list(op.pos), matrix=np.concatenate([m.reshape(-1) for m in mats]),
                  body=len(mats))
        return "channel"
    if name == "cif":
        row = b.add(CIF, cbits=list(op.cbits), cond=int(op.condition))
        start = len(b.rows)
        if isinstance(op.instructions, list):
            for sub in op.instructions:
                _emit_pyop(b, sub)
        b.set_body(row, len(b.rows) - start)
        return "cif"
    pos = [int(p) for p in op.pos]
    paras = op._paras  # pylint: disable=protected-access
    n_ctrl = len(op.ctrls) if hasattr(op, "ctrls") else 0
    code = NAMED.get(name)
    if code is not None:
        b.add(code, qubits=pos, n_ctrl=n_ctrl, param=float(paras[0]) if paras else 0.0)
        return "gate"
    m = gate_matrix(op)
    n_targ = len(pos) - n_ctrl
    if n_targ < 1:
        raise RuntimeError("Invalid target number")
    if m.shape != (1 << n_targ, 1 << n_targ):
        raise ValueError(f"gate {op.name}: matrix shape {m.shape} does not match {n_targ} target qubit(s)")
    b.add(GENERAL if n_targ == 1 else MULTI, qubits=pos, n_ctrl=n_ctrl, matrix=m)
    return "gate"


def _emit_instructions(b: ProgramBuilder, instructions, state: dict, top: bool) -> None:
    for op in instructions:
        if hasattr(op, "circuit"):
            _emit_instructions(b, op.circuit.instructions, state, top=False)
            continue
        kind = _emit_pyop(b, op)
        if not top or kind == "skip":
            continue
        if kind == "measure":
            state["measured"] = True
            state["measure_vec"].extend(zip(op.qbits, op.cbits))
        elif state["measured"]:
            state["final_measure"] = False


def compile_circuit(qc) -> Program:
    """Compile a QuantumCircuit (or anything with `num`, `cbits_num`, `instructions`)."""
    b = ProgramBuilder()
    state = {"measured": False, "measure_vec": [], "final_measure": True}
    _emit_instructions(b, qc.instructions, state, top=True)
    prog = b.program(
        n_qubits=int(qc.num),
        n_cbits=int(qc.cbits_num),
        measure_vec=[(int(q), int(c)) for q, c in state["measure_vec"]],
        final_measure=state["final_measure"],
    )
    prog.n_random = int(np.isin(prog.ops["code"], [MEASURE, RESET, KRAUS, UCHAN]).sum()) if prog.n_ops else 0
    return prog


def compile_ops(ops, n_qubits: int) -> Program:
    """Compile a plain list of gates (applyop_statevec and the gradient code use this)."""
    return compile_gates(ops, n_qubits)[0]


def compile_gates(gates, n_qubits: int):
    """Compile a list of gates; also returns, per gate, the index of the op row it became (-1 for
    barriers, delays and identities)."""
    b = ProgramBuilder()
    rows = []
    for op in gates:
        r = len(b.rows)
        kind = _emit_pyop(b, op)
        rows.append(r if kind == "gate" else -1)
    return b.program(n_qubits=n_qubits), rows
