"""Non-gate instructions: Barrier, Delay, Reset, Measure, Cif.

Attribute names follow quafu/elements/instruction.py:25-183 and classical_element.py:22-44,
because the program compiler reads them the way the reference's from_pyops does
(src/qfvm/instructions.hpp:229-299): `name`, `pos`, `qbits`, `cbits`, `condition`,
`instructions`.
"""

from __future__ import annotations

from typing import Dict, List


class Instruction:
    """Base class of every instruction. `pos` are the qubits it acts on."""

    ins_classes: Dict[str, type] = {}
    name = "instruction"

    def __init__(self, pos, paras=None, **kwargs):  # pylint: disable=unused-argument
        self.pos = pos
        self.paras = [] if paras is None else paras
        self._symbol = None

    @classmethod
    def register_ins(cls, subclass, name: str = None):
        name = str(subclass.__name__).lower() if name is None else name
        if name in cls.ins_classes:
            raise ValueError(f"Name {name} already exists.")
        cls.ins_classes[name] = subclass
        return subclass

    @property
    def named_pos(self) -> Dict:
        return {"pos": self.pos}

    @property
    def named_paras(self) -> Dict:
        return {}


class Barrier(Instruction):
    name = "barrier"

    def __init__(self, pos):
        super().__init__(pos)
        self.symbol = "||"

    def __repr__(self):
        return "Barrier"

    def to_qasm(self, _=False):
        return "barrier " + ",".join(f"q[{p}]" for p in range(min(self.pos), max(self.pos) + 1))


class Delay(Instruction):
    """Idle time on one qubit; the statevector engine skips it (instructions.hpp:242)."""

    name = "delay"

    def __init__(self, pos: int, duration: int, unit="ns"):
        if not isinstance(duration, int):
            raise TypeError("duration must be int")
        super().__init__([pos] if isinstance(pos, int) else pos)
        self.duration = duration
        self.unit = unit
        self.symbol = f"Delay({duration}{unit})"


class Reset(Instruction):
    name = "reset"

    def __init__(self, pos):
        super().__init__(list(pos))

    def __repr__(self):
        return "Reset"


class Measure(Instruction):
    name = "measure"

    def __init__(self, bitmap: dict):
        super().__init__(list(bitmap.keys()))
        self.qbits = list(bitmap.keys())
        self.cbits = list(bitmap.values())

    def to_qasm(self, _=False):
        return "".join(f"measure q[{q}] -> meas[{c}];\n" for q, c in zip(self.qbits, self.cbits))


class Cif(Instruction):
    """Classically conditioned block: runs `instructions` when creg[cbits] (big-endian) == condition."""

    name = "cif"

    def __init__(self, cbits: List[int], condition: int, instructions=None):
        self.cbits = cbits
        self.condition = condition
        self.instructions = instructions
        super().__init__(pos=-1)

    def set_ins(self, instructions):
        self.instructions = instructions


for _cls in (Barrier, Delay, Reset, Measure, Cif):
    Instruction.register_ins(_cls)
