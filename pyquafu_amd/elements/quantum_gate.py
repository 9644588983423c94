"""Gate base classes: QuantumGate, ControlledGate, ControlledU, CircuitWrapper.

Semantics follow quafu/elements/quantum_gate.py:52-548. What the engine consumes is fixed by the
reference's pybind reader (src/qfvm/instructions.hpp:229-299):
  * `name.lower()` picks a named kernel when it is in qfvm's native set (qasm.hpp:43-46);
  * otherwise a controlled gate hands over `_get_targ_matrix(reverse_order=True)` and any other
    gate hands over `matrix` (which for >1 qubit is `reorder_matrix(raw, pos)`);
  * `pos` lists controls first, then targets; `len(ctrls)` is the control count.
"""

from __future__ import annotations

import copy
from typing import Callable, Dict, List, Optional, Union

import numpy as np

from .instruction import Instruction
from .matrices import reorder_matrix
from .parameters import ParameterType, extract_float

HERMITIAN = ["id", "x", "y", "z", "h", "w", "cx", "cz", "cy", "cnot", "swap", "mcx", "mxy", "mcz"]
ROTATION = ["rx", "ry", "rz", "p", "rxx", "ryy", "rzz", "cp"]
_paired = {k: k + "dg" for k in ["sx", "sy", "s", "t", "sw"]}
PAIRED = {**_paired, **{v: k for k, v in _paired.items()}}

MatrixType = Union[np.ndarray, Callable]


class QuantumGate(Instruction):
    gate_classes: Dict[str, type] = {}

    def __init__(
        self,
        name: str,
        pos: List[int],
        paras: Union[List[ParameterType], None] = None,
        matrix: Optional[MatrixType] = None,
    ):
        super().__init__(pos, paras=[] if paras is None else paras)
        self._name = name
        self._symbol = None
        self._raw_matrix = matrix

    def __repr__(self):
        return f"{self.__class__.__name__}"

    @property
    def name(self):
        return self._name

    @name.setter
    def name(self, __name):
        self._name = __name

    @property
    def pos(self):
        return self._pos

    @pos.setter
    def pos(self, __pos):
        self._pos = copy.deepcopy(__pos)

    @classmethod
    def register_gate(cls, subclass, name: str = None):
        name = str(subclass.__name__).lower() if name is None else name
        if name.endswith("gate"):
            name = name[:-4]
        if name in cls.gate_classes:
            raise ValueError(f"Name {name} already exists.")
        cls.gate_classes[name] = subclass
        Instruction.register_ins(subclass, name)
        return subclass

    @classmethod
    def register(cls, name: str = None):
        def wrapper(subclass):
            return cls.register_gate(subclass, name)

        return wrapper

    @property
    def _paras(self):
        return extract_float(self.paras)

    @property
    def symbol(self) -> str:
        if len(self.paras) > 0:
            return f"{self.name}({','.join(f'{p:.3f}' for p in self._paras)})"
        return f"{self.name}"

    @symbol.setter
    def symbol(self, symbol: str):
        self._symbol = symbol

    @property
    def matrix(self):
        raw = self._raw_matrix
        if callable(raw):
            raw = raw(self._paras)
        if len(self.pos) > 1:
            return reorder_matrix(raw, self.pos)
        return raw

    def _get_raw_matrix(self, reverse_order=False):
        raw = self._raw_matrix
        if callable(raw):
            raw = raw(self._paras)
        if reverse_order and len(self.pos) > 1:
            return reorder_matrix(raw, np.arange(len(self.pos))[::-1])
        return raw

    def to_qasm(self, with_para=False) -> str:
        qstr = f"{self.name.lower()}"
        if self.paras:
            qstr += f"({','.join(f'{p}' for p in self._paras)})"
        return qstr + " " + ",".join(f"q[{p}]" for p in self.pos)

    def update_params(self, paras):
        if paras is None:
            return
        self.paras = paras if isinstance(paras, list) else [paras]

    def power(self, n) -> "QuantumGate":
        """n-fold product of this gate (quantum_gate.py:194-243)."""
        name = self.name.lower()
        name = "sz" if name == "s" else name
        order4 = ["sx", "sy", "s", "sw"]
        order4 += [_ + "dg" for _ in order4]
        if name in HERMITIAN:
            if n % 2 == 0:
                return self.gate_classes["id"](self.pos[0])
            return copy.deepcopy(self)
        if name in order4:
            if n % 4 == 0:
                return self.gate_classes["id"](self.pos[0])
            if n % 4 == 1:
                return copy.deepcopy(self)
            if n % 4 == 2:
                return self.gate_classes["z" if name == "s" else name[1:]](self.pos[0])
            return self.gate_classes[PAIRED[name]](self.pos[0])
        if name in ["t", "tdg"]:
            theta = -np.pi / 4 if name.endswith("dg") else np.pi / 4
            return self.gate_classes["rz"](self.pos[0], theta * n)
        if name in ROTATION:
            return self.gate_classes[name](*(self.pos + [self.paras[0] * n]))
        raw = self._raw_matrix
        if callable(raw):

            def raw_matrix_func(paras, _f=raw):
                return np.linalg.matrix_power(_f(paras), n)

            new_raw = raw_matrix_func
        else:
            new_raw = np.linalg.matrix_power(raw, n)
        return QuantumGate(self.name + f"^{n}", self.pos, self.paras, new_raw)

    def dagger(self) -> "QuantumGate":
        """Hermitian conjugate on the same qubits (quantum_gate.py:262-283)."""
        name = self.name
        if name in HERMITIAN:
            return copy.deepcopy(self)
        if name in ROTATION:
            return self.gate_classes[name](*(self.pos + [-self.paras[0]]))
        if name in PAIRED:
            return self.gate_classes[PAIRED[name]](self.pos[0])
        raw = self._raw_matrix
        if callable(raw):

            def raw_matrix_func(paras, _f=raw):
                return _f(paras).conj().T

            new_raw = raw_matrix_func
        else:
            new_raw = raw.conj().T
        return QuantumGate(self.name + "^†", self.pos, self.paras, new_raw)

    def ctrl_by(self, ctrls: Union[int, List[int]]) -> "ControlledGate":
        """Controlled version of this gate (quantum_gate.py:285-330)."""
        ctrls = [ctrls] if not isinstance(ctrls, list) else ctrls
        pos = [self.pos] if not isinstance(self.pos, list) else self.pos
        name = self.name.lower()
        if isinstance(self, ControlledGate):
            if self.ctrls == ctrls:
                raise ValueError("Can't not add the same control qubit.")
            ctrls = list(set(self.ctrls) | set(ctrls))
        elif set(ctrls) & set(pos):
            raise ValueError("Control qubits should not be overlap with target qubits.")
        if self.name in ["X", "Y", "Z", "RX", "RY", "RZ"]:
            args = self.pos + self.paras
            if len(ctrls) == 1:
                return QuantumGate.gate_classes["c" + name](ctrls[0], *args)
            return QuantumGate.gate_classes["mc" + name](ctrls, *args)
        if self.name in ["CX", "CY", "CZ", "CRX", "CRY", "CRZ"]:
            return QuantumGate.gate_classes["m" + name](ctrls, *(self.targs + self.paras))
        if isinstance(self, ControlledGate):
            return ControlledGate(
                "MC" + self._targ_name, self._targ_name, ctrls + self.ctrls, self.targs, self.paras,
                self._targ_matrix,
            )
        return ControlledU("MC" + name, ctrls, self)


class ControlledGate(QuantumGate):
    """A gate acting as `targ_matrix` on targs when every control qubit is 1."""

    def __init__(
        self,
        name: str,
        targ_name: str,
        ctrls: List[int],
        targs: List[int],
        paras=None,
        targ_matrix: MatrixType = None,
    ):
        self.ctrls = copy.deepcopy(ctrls)
        self.targs = copy.deepcopy(targs)
        self._targ_name = targ_name
        super().__init__(name, ctrls + targs, [] if paras is None else paras, targ_matrix)
        self._targ_matrix = targ_matrix
        self._raw_matrix = self._rawmatfunc

    def _rawmatfunc(self, paras):
        targ_dim = 2 ** len(self.targs)
        dim = 2 ** len(self.pos)
        raw = np.zeros((dim, dim), dtype=complex)
        tm = self._targ_matrix(paras) if callable(self._targ_matrix) else self._targ_matrix
        if tm.shape[0] != targ_dim:
            raise ValueError("Dimension dismatch")
        cdim = dim - targ_dim
        for i in range(cdim):
            raw[i, i] = 1.0
        raw[cdim:, cdim:] = tm
        return raw

    def power(self, n) -> "ControlledGate":
        if self._targ_name in ["RX", "RY", "RZ", "P", "RZZ", "RXX", "RYY"]:
            return ControlledGate(self.name, self._targ_name, self.ctrls, self.targs, [self.paras[0] * n],
                                  self._targ_matrix)
        tm = self._targ_matrix
        if callable(tm):

            def f(paras, _f=tm):
                return np.linalg.matrix_power(_f(paras), n)

            new_tm = f
        else:
            new_tm = np.linalg.matrix_power(tm, n)
        return ControlledGate(self.name + f"^{n}", self._targ_name + f"^{n}", self.ctrls, self.targs, self.paras,
                              new_tm)

    def dagger(self) -> "ControlledGate":
        if self._targ_name in ["RX", "RY", "RZ", "P", "RZZ", "RXX", "RYY"]:
            return ControlledGate(self.name, self._targ_name, self.ctrls, self.targs, [-self.paras[0]],
                                  self._targ_matrix)
        tm = self._targ_matrix
        if callable(tm):

            def f(paras, _f=tm):
                return _f(paras).conj().T

            new_tm = f
        else:
            new_tm = tm.conj().T
        return ControlledGate(self.name + "^†", self._targ_name + "^†", self.ctrls, self.targs, self.paras, new_tm)

    @property
    def symbol(self):
        if len(self.paras) > 0:
            return f"{self._targ_name}({','.join(f'{p:.3f}' for p in self._paras)})"
        return f"{self._targ_name}"

    @property
    def ct_nums(self):
        return len(self.ctrls), len(self.targs), len(self.ctrls) + len(self.targs)

    def _get_targ_matrix(self, reverse_order=False):
        tm = self._targ_matrix(self._paras) if callable(self._targ_matrix) else self._targ_matrix
        if reverse_order and len(self.targs) > 1:
            return reorder_matrix(tm, np.array(range(len(self.targs))[::-1]))
        return tm

    @property
    def named_pos(self) -> Dict:
        return {"ctrls": self.ctrls, "targs": self.targs}


class ControlledU(ControlledGate):
    def __init__(self, name, ctrls: List[int], U: QuantumGate):
        self.targ_gate = U
        super().__init__(name, U.name, ctrls, U.pos, U.paras, targ_matrix=U._raw_matrix)


class CircuitWrapper(QuantumGate):
    """A sub-circuit used as one instruction; the engine flattens it (circuit.hpp:110-115)."""

    # pylint: disable=super-init-not-called
    def __init__(self, name: str, circ, qbits: Union[None, List] = None):
        self.name = name
        self.pos = list(range(circ.num))
        self.circuit = copy.deepcopy(circ)
        self.paras = []
        if qbits:
            self._reallocate(qbits)

    def _reallocate(self, qbits):
        num = max(self.circuit.num - 1, max(qbits)) + 1
        self.pos = qbits
        self.circuit._reallocate(num, qbits)

    @property
    def symbol(self):
        return f"{self.name}"

    def add_controls(self, ctrls=None) -> QuantumGate:
        return ControlledCircuitWrapper("MC" + self.name, self, [] if ctrls is None else ctrls)

    def power(self, n: int):
        self.name += f"^{n}"
        self.circuit = self.circuit.power(n)
        return self

    def dagger(self):
        self.name += "^†"
        self.circuit = self.circuit.dagger()
        return self


class ControlledCircuitWrapper(CircuitWrapper):
    # pylint: disable=super-init-not-called
    def __init__(self, name: str, circwrp: CircuitWrapper, ctrls: List[int]):
        self.name = name
        self.ctrls = ctrls
        self.targs = circwrp.pos
        self.circuit = circwrp.circuit.add_controls(len(ctrls), ctrls, self.targs)
        self.pos = list(range(self.circuit.num))
        self.paras = []
        self._targ_name = circwrp.name
