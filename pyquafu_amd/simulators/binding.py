"""Parameter rebinding: a circuit compiled once, new parameter values written into its program.

The reference rebuilds its C++ circuit from the Python gates on every run: `_update_params`
(quantum_circuit.py:304) sets the variables, then the pybind layer walks every instruction through
`from_pyops` (src/qfvm/instructions.hpp:229-299). Here the flat program (program.py) is compiled
once together with a slot map -- which op rows read which parameter -- and `bind(values)` rewrites
only those rows:

  * a named rotation (P, RX, RY, RZ, CP, RZZ) whose angle is one variable takes
    `values[var]` in its `param` field, for all such rows in one numpy assignment;
  * any other row that depends on a variable (an expression angle, a GENERAL / MULTI matrix such
    as U3 or a controlled rotation) is re-derived from its gate after the variables are set, the
    way compile_gates derived it, into the same `param` field or matrix slot.

The program's arrays are updated in place, so engine call arguments built from it once
(`_native.AdjointCall`, `as_cprogram`) see the new values, and the engine's plan cache reuses the
plan of the unchanged structure.
"""

from __future__ import annotations

from typing import List, Sequence

import numpy as np

from ..elements.parameters import Parameter, ParameterExpression
from .program import MULTI, GENERAL, PARAMETRISED, compile_gates, gate_matrix


def _symbolic(op) -> bool:
    return any(isinstance(p, (Parameter, ParameterExpression)) for p in getattr(op, "paras", []) or [])


class ParametricProgram:
    """`qc`'s gates compiled once; `variables` in the order of `qc.variables` (the order
    `_update_params` and the gradient code use)."""

    def __init__(self, qc, gates=None):
        gates = list(qc.gates if gates is None else gates)
        self.qc = qc
        self.gates = gates
        self.prog, self.rows = compile_gates(gates, qc.num)
        self.variables: List[Parameter] = list(qc.variables)
        index = {v: k for k, v in enumerate(self.variables)}
        ops = self.prog.ops
        fast_rows, fast_vars, slow = [], [], []
        for g, (op, r) in enumerate(zip(gates, self.rows)):
            if r < 0 or not _symbolic(op):
                continue
            code = int(ops["code"][r])
            para = op.paras[0]
            if code in PARAMETRISED and isinstance(para, Parameter) and para in index and len(op.paras) == 1:
                fast_rows.append(r)
                fast_vars.append(index[para])
            elif code in PARAMETRISED or code in (GENERAL, MULTI):
                slow.append((g, r))
            else:
                raise ValueError(f"gate {op.name}: parameters on an op without angle or matrix slot")
        self.fast_rows = np.asarray(fast_rows, dtype=np.int64)
        self.fast_vars = np.asarray(fast_vars, dtype=np.int64)
        self.slow = slow

    @property
    def n_vars(self) -> int:
        return len(self.variables)

    def bind(self, values: Sequence[float]):
        """Writes `values` (one per variable) into the program; returns the program."""
        v = np.asarray(values, dtype=np.float64).reshape(-1)
        if v.size != len(self.variables):
            raise ValueError(f"{v.size} values for {len(self.variables)} parameters")
        ops = self.prog.ops
        if self.fast_rows.size:
            ops["param"][self.fast_rows] = v[self.fast_vars]
        if self.slow:
            for var, x in zip(self.variables, v):
                var.value = float(x)
            mat = self.prog.mat
            for g, r in self.slow:
                op = self.gates[g]
                code = int(ops["code"][r])
                if code in PARAMETRISED:
                    ops["param"][r] = float(op._paras[0])  # pylint: disable=protected-access
                    continue
                m = np.ascontiguousarray(gate_matrix(op), dtype=np.complex128).view(np.float64).ravel()
                off = int(ops["m_off"][r])
                mat[off: off + m.size] = m
        else:
            for var, x in zip(self.variables, v):  # keep the circuit's own values in step
                var.value = float(x)
        return self.prog
