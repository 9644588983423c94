"""QuantumCircuit (quafu/circuits/quantum_circuit.py:50-1320), the statevector path's subset.

Kept: construction, the gate helpers, `<<`, measure / reset / cif, parameters and their
gradients (get_parameter_grads, _calc_parameter_grads, variables, _update_params), wrap / unwrap,
dagger, power, add_controls, join. Left out (not on the statevector path): pulses, noise channels,
XY resonance, QASM text I/O, drawing and layering.
"""

from __future__ import annotations

import copy
from contextlib import contextmanager
from typing import Any, Iterable, List, Optional, Union

import numpy as np

from ..elements import element_gates as qeg
from ..elements.instruction import Barrier, Cif, Delay, Instruction, Measure, Reset
from ..elements.noise import KrausChannel, UnitaryChannel
from ..elements.parameters import Parameter, ParameterExpression, ParameterType
from ..elements.quantum_gate import CircuitWrapper, ControlledGate, QuantumGate
from ..exceptions import CircuitError


class QuantumCircuit:
    def __init__(self, qnum: int, cnum: Optional[int] = None, name="", **kwargs):  # pylint: disable=unused-argument
        self.name = name
        self._num = qnum if qnum > 0 else 0
        cnum = self._num if cnum is None else cnum
        self._cnum = cnum if cnum > 0 else 0
        self._gates = []
        self.instructions = []
        self._measures = []
        self.executable_on_backend = True
        self._parameterized_gates = []
        self._parameter_grads = {}
        self._variables = []
        self._has_wrap = False
        self._def_para_name = "theta_inner"
        self._para_id = 0

    # ------------------------------------------------------------------ sizes
    @property
    def num(self):
        return self._num

    @num.setter
    def num(self, num: int):
        self._num = num if num > 0 else 0
        self._cnum = self._num

    @property
    def cbits_num(self):
        return self._cnum

    @property
    def parameterized_gates(self):
        if not self._parameterized_gates:
            self._parameterized_gates = [g for g in self.gates if len(g.paras) != 0]
        return self._parameterized_gates

    @property
    def measures(self):
        measures = {}
        for meas in self._measures:
            measures.update(dict(zip(meas.qbits, meas.cbits)))
        return measures

    @measures.setter
    def measures(self, measures: dict):
        self._measures = [Measure(measures)]

    @property
    def gates(self):
        return self._gates

    @gates.setter
    def gates(self, gates: list):
        self._gates = gates

    @property
    def used_qubits(self) -> List[int]:
        used = set()
        for ins in self.instructions:
            if isinstance(ins, Cif) or ins.pos == -1:
                continue
            p = ins.pos if isinstance(ins.pos, list) else [ins.pos]
            used.update(p)
        used.update(self.measures.keys())
        return sorted(used)

    # ------------------------------------------------------------------ adding
    def __lshift__(self, operation: Instruction):
        max_pos = max(operation.pos) if isinstance(operation.pos, Iterable) else operation.pos
        if max_pos >= self.num:
            raise CircuitError("Operation act on qubit that not allocated")
        self.add_ins(operation)
        if isinstance(operation, CircuitWrapper):
            self._has_wrap = True
        return self

    def add_gates(self, gates: list):
        for gate in gates:
            self.add_ins(gate)

    def add_gate(self, gate: QuantumGate):
        pos = np.array(gate.pos)
        if np.any(pos >= self.num):
            raise CircuitError(f"Gate position out of range: {gate.pos}")
        self.gates.append(gate)

    def add_ins(self, ins: Instruction):
        if isinstance(ins, (QuantumGate, Delay, Barrier, KrausChannel)):
            self.add_gate(ins)
        self.instructions.append(ins)

    NOISE_CHANNELS = ("bitflip", "dephasing", "depolarizing", "ampdamping")

    def add_noise(self, channel: str, channel_args, qubits: Optional[List[int]] = None,
                  gates: Optional[List[str]] = None, checkgates: bool = True):
        """After every gate that acts on one of `qubits` (and whose lowercased name is in `gates`,
        when given), insert `channel(q, *channel_args)` for each of its positions q in `qubits`
        (quantum_circuit.py:181-245). As there, an empty `qubits` adds nothing."""
        qubits = [] if qubits is None else list(qubits)
        gates = [] if gates is None else list(gates)
        if channel not in self.NOISE_CHANNELS:
            raise ValueError("Invalid channel name")
        if checkgates:
            for g in gates:
                if g not in QuantumGate.gate_classes:
                    raise ValueError("Invalid gate name")
        cls = Instruction.ins_classes[channel]
        instructions, kept = [], []
        for op in self.instructions:
            instructions.append(op)
            if isinstance(op, (QuantumGate, Delay, Barrier, KrausChannel)):
                kept.append(op)
            if not isinstance(op, QuantumGate):
                continue
            if gates and op.name.lower() not in gates:
                continue
            for q in op.pos:
                if q in qubits:
                    instructions.append(cls(q, *channel_args))
                    kept.append(cls(q, *channel_args))
        self.instructions = instructions
        self._gates = kept
        return self

    @property
    def noised(self):
        if self._has_wrap:
            self.unwrap()
        return any(isinstance(op, KrausChannel) for op in self.instructions)

    # ------------------------------------------------------------------ parameters
    def get_parameter_grads(self):
        if self._has_wrap:
            self.unwrap()
        self._parameter_grads = {}
        for i, op in enumerate(self.gates):
            for j, para in enumerate(op.paras):
                if isinstance(para, Parameter):
                    self._parameter_grads.setdefault(para, []).append([(i, j), 1.0])
                elif isinstance(para, ParameterExpression):
                    para_grads = para.grad()
                    for var in para._variables:
                        self._parameter_grads.setdefault(var, []).append([(i, j), para_grads[para._variables[var]]])
        self._variables = list(self._parameter_grads)
        return self._parameter_grads

    def _calc_parameter_grads(self):
        for var, grads in self._parameter_grads.items():
            for item in grads:
                i, j = item[0]
                para = self.gates[i].paras[j]
                if isinstance(para, ParameterExpression) and not isinstance(para, Parameter):
                    item[1] = para.grad()[para._variables[var]]
        return self._parameter_grads

    @property
    def variables(self):
        self._variables = list(self.get_parameter_grads().keys())
        return self._variables

    def _update_params(self, values, order: Union[None, List] = None, tunable_only: bool = False):
        order = [] if order is None else order
        if len(values) != len(self.variables):
            raise CircuitError("The size of input values must be the same to the parameters")
        for i, v in enumerate(values):
            val = values[order[i]] if order else v
            if tunable_only and not self._variables[i].tunable:
                continue
            self._variables[i].value = val

    def update_params(self, paras_list: List[Any]):
        if len(paras_list) != len(self.parameterized_gates):
            raise CircuitError("`params_list` must have the same size with parameterized gates")
        for gate, paras in zip(self.parameterized_gates, paras_list):
            gate.update_params(paras)

    def angles2parameters(self):
        for g in self.gates:
            if all(isinstance(x, float) for x in g.paras):
                for i, para in enumerate(g.paras):
                    g.paras[i] = Parameter(f"{self._def_para_name}_{self._para_id}", value=para)
                    self._para_id += 1
        self.get_parameter_grads()

    # ------------------------------------------------------------------ structure
    def _reallocate(self, num, qbits: List[int]):
        assert self.num == len(qbits)
        if max(qbits) > num:
            raise CircuitError("Bad allocation")
        self.num = num
        qmap = dict(zip(range(len(qbits)), qbits))
        for op in self.instructions:
            for i, p in enumerate(op.pos):
                op.pos[i] = qmap[p]
            if isinstance(op, ControlledGate):
                for i, c in enumerate(op.ctrls):
                    op.ctrls[i] = qmap[c]
                for i, t in enumerate(op.targs):
                    op.targs[i] = qmap[t]

    def add_controls(self, ctrlnum, ctrls=None, targs=None, inplace=False) -> "QuantumCircuit":
        ctrls = [] if ctrls is None else ctrls
        targs = [] if targs is None else targs
        if len(ctrls + targs) == 0:
            ctrls = list(np.arange(ctrlnum) + self.num)
            num = self.num + ctrlnum
            instrs = self.instructions
        else:
            if len(ctrls) == 0 or len(targs) == 0:
                raise ValueError("Must provide both ctrls and targs")
            assert len(targs) == self.num
            assert len(ctrls) == ctrlnum
            num = max(ctrls + targs) + 1
            temp = self if inplace else copy.deepcopy(self)
            temp._reallocate(num, targs)
            instrs = temp.instructions
        if inplace:
            self.instructions = [op.ctrl_by(ctrls) if isinstance(op, QuantumGate) else op for op in instrs]
            self._gates = [op for op in self.instructions if isinstance(op, QuantumGate)]
            return self
        qc = QuantumCircuit(num)
        for op in instrs:
            if isinstance(op, QuantumGate):
                qc << op.ctrl_by(ctrls)
        return qc

    def join(self, qc, qbits=None, inplace=True) -> "QuantumCircuit":
        qbits = [] if qbits is None else qbits
        rnum = qc.num if isinstance(qc, QuantumCircuit) else qc.circuit.num
        num = max(self.num, rnum) if not qbits else max(self.num - 1, max(qbits)) + 1
        if inplace:
            self.num = num
            nqc = self
        else:
            nqc = QuantumCircuit(num)
            for op in self.instructions:
                nqc << op
        if isinstance(qc, QuantumCircuit):
            if qbits:
                newqc = copy.deepcopy(qc)
                newqc._reallocate(num, qbits)
                for op in newqc.instructions:
                    nqc << op
            else:
                for op in qc.instructions:
                    nqc << op
        else:
            if qbits:
                qc._reallocate(qbits)
            nqc << qc
        return nqc

    def power(self, n, inplace=False):
        if inplace:
            base = list(self.instructions)
            for _ in range(n - 1):
                self.instructions += base
            return self
        nq = QuantumCircuit(self.num)
        for _ in range(n):
            for op in self.instructions:
                nq << op
        nq.measures = self.measures
        return nq

    def dagger(self, inplace=False):
        if inplace:
            self.instructions = [op.dagger() if isinstance(op, QuantumGate) else op
                                 for op in self.instructions[::-1]]
            self._gates = [op for op in self.instructions if isinstance(op, QuantumGate)]
            return self
        nq = QuantumCircuit(self.num)
        for op in self.instructions[::-1]:
            nq << (op.dagger() if isinstance(op, QuantumGate) else op)
        nq.measures = self.measures
        return nq

    def wrap(self, qbits=None):
        name = self.name if self.name else "Oracle"
        return CircuitWrapper(name, self, [] if qbits is None else qbits)

    # ------------------------------------------------------------------ OpenQASM
    def from_openqasm(self, openqasm: str):
        """Replace this circuit by the OpenQASM 2.0 program (quantum_circuit.py:557-566,
        qfasm_convertor.py:38-58): registers flattened in declaration order."""
        from ..qfasm import parse_qasm  # pylint: disable=import-outside-toplevel

        new = parse_qasm(openqasm)
        self._num, self._cnum = new.num, new.cbits_num
        self.instructions = new.instructions
        self._gates = new.gates
        self._measures = new._measures  # pylint: disable=protected-access
        self._parameterized_gates = []
        self.openqasm = openqasm
        self.executable_on_backend = new.executable_on_backend
        return self

    def to_openqasm(self, with_para=False) -> str:
        """OpenQASM 2.0 text of the gates and final measures (quantum_circuit.py:568-600)."""
        valid = set(QuantumGate.gate_classes.keys()) | {"barrier", "delay", "reset"}
        qasm = 'OPENQASM 2.0;\ninclude "qelib1.inc";\n'
        qasm += f"qreg q[{self.num}];\n"
        qasm += f"creg meas[{self.cbits_num}];\n"
        for gate in self.gates:
            if gate.name.lower() not in valid:
                raise NotImplementedError(
                    f"gate {gate.name} can not convert to qasm2 directly, you may decompose it manuallly")
            qasm += gate.to_qasm(with_para) + ";\n"
        for q, c in self.measures.items():
            qasm += f"measure q[{q}] -> meas[{c}];\n"
        self.openqasm = qasm
        return qasm

    def unwrap(self):
        instructions, gates = [], []
        for op in self.instructions:
            if isinstance(op, CircuitWrapper):
                for op_ in op.circuit.unwrap().instructions:
                    instructions.append(op_)
                    if isinstance(op_, (QuantumGate, Delay, Barrier, KrausChannel)):
                        gates.append(op_)
            else:
                instructions.append(op)
                if isinstance(op, (QuantumGate, Delay, Barrier, KrausChannel)):
                    gates.append(op)
        self.instructions = instructions
        self._gates = gates
        self._has_wrap = False
        return self

    # ------------------------------------------------------------------ gate helpers
    def id(self, pos: int):  # noqa: A003
        self.add_ins(qeg.IdGate(pos))
        return self

    def h(self, pos: int):
        self.add_ins(qeg.HGate(pos))
        return self

    def x(self, pos: int):
        self.add_ins(qeg.XGate(pos))
        return self

    def y(self, pos: int):
        self.add_ins(qeg.YGate(pos))
        return self

    def z(self, pos: int):
        self.add_ins(qeg.ZGate(pos))
        return self

    def t(self, pos: int):
        self.add_ins(qeg.TGate(pos))
        return self

    def tdg(self, pos: int):
        self.add_ins(qeg.TdgGate(pos))
        return self

    def s(self, pos: int):
        self.add_ins(qeg.SGate(pos))
        return self

    def sdg(self, pos: int):
        self.add_ins(qeg.SdgGate(pos))
        return self

    def sx(self, pos: int):
        self.add_ins(qeg.SXGate(pos))
        return self

    def sxdg(self, pos: int):
        self.add_ins(qeg.SXdgGate(pos))
        return self

    def sy(self, pos: int):
        self.add_ins(qeg.SYGate(pos))
        return self

    def sydg(self, pos: int):
        self.add_ins(qeg.SYdgGate(pos))
        return self

    def w(self, pos: int):
        self.add_ins(qeg.WGate(pos))
        return self

    def sw(self, pos: int):
        self.add_ins(qeg.SWGate(pos))
        return self

    def rx(self, pos: int, para: ParameterType):
        self.add_ins(qeg.RXGate(pos, para))
        return self

    def ry(self, pos: int, para: ParameterType):
        self.add_ins(qeg.RYGate(pos, para))
        return self

    def rz(self, pos: int, para: ParameterType):
        self.add_ins(qeg.RZGate(pos, para))
        return self

    def p(self, pos: int, para: ParameterType):
        self.add_ins(qeg.PhaseGate(pos, para))
        return self

    def cnot(self, ctrl: int, tar: int):
        self.add_ins(qeg.CXGate(ctrl, tar))
        return self

    def cx(self, ctrl: int, tar: int):
        return self.cnot(ctrl=ctrl, tar=tar)

    def cy(self, ctrl: int, tar: int):
        self.add_ins(qeg.CYGate(ctrl, tar))
        return self

    def cz(self, ctrl: int, tar: int):
        self.add_ins(qeg.CZGate(ctrl, tar))
        return self

    def cs(self, ctrl: int, tar: int):
        self.add_ins(qeg.CSGate(ctrl, tar))
        return self

    def ct(self, ctrl: int, tar: int):
        self.add_ins(qeg.CTGate(ctrl, tar))
        return self

    def cp(self, ctrl: int, tar: int, para: ParameterType):
        self.add_ins(qeg.CPGate(ctrl, tar, para))
        return self

    def swap(self, q1: int, q2: int):
        self.add_ins(qeg.SwapGate(q1, q2))
        return self

    def iswap(self, q1: int, q2: int):
        self.add_ins(qeg.ISwapGate(q1, q2))
        return self

    def toffoli(self, ctrl1: int, ctrl2: int, targ: int):
        self.add_ins(qeg.ToffoliGate(ctrl1, ctrl2, targ))
        return self

    def fredkin(self, ctrl: int, targ1: int, targ2: int):
        self.add_ins(qeg.FredkinGate(ctrl, targ1, targ2))
        return self

    def barrier(self, qlist: List[int] = None):
        self.add_ins(Barrier(list(range(self.num)) if qlist is None else qlist))
        return self

    def rxx(self, q1: int, q2: int, theta):
        self.add_ins(qeg.RXXGate(q1, q2, theta))

    def ryy(self, q1: int, q2: int, theta):
        self.add_ins(qeg.RYYGate(q1, q2, theta))

    def rzz(self, q1: int, q2: int, theta):
        self.add_ins(qeg.RZZGate(q1, q2, theta))

    def mcx(self, ctrls: List[int], targ: int):
        self.add_ins(qeg.MCXGate(ctrls, targ))

    def mcy(self, ctrls: List[int], targ: int):
        self.add_ins(qeg.MCYGate(ctrls, targ))

    def mcz(self, ctrls: List[int], targ: int):
        self.add_ins(qeg.MCZGate(ctrls, targ))

    def delay(self, pos, duration, unit="ns"):
        self.add_ins(Delay(pos, duration, unit=unit))
        return self

    def reset(self, qlist: List[int] = None):
        self.add_ins(Reset(list(range(self.num)) if qlist is None else qlist))
        self.executable_on_backend = False
        return self

    def measure(self, pos: List[int] = None, cbits: List[int] = None) -> None:
        """Measure `pos` into `cbits` (quantum_circuit.py:1234-1269)."""
        if pos is None:
            pos = list(range(self.num))
        if np.any(np.array(pos) >= self.num):
            raise ValueError("Index out of range.")
        e_num = len(self.measures)
        n_num = len(pos)
        if n_num > len(set(pos)):
            raise ValueError("Measured qubits not uniquely assigned.")
        if cbits:
            if len(set(cbits)) != len(cbits):
                raise ValueError("Classical bits not uniquely assigned.")
            if len(cbits) != n_num:
                raise ValueError("Number of measured bits should equal to the number of classical bits")
        else:
            cbits = list(range(e_num, e_num + n_num))
        for cbit in cbits:
            if cbit < 0 or cbit > self.cbits_num:
                raise ValueError("Cbits index out of range.")
        measure = Measure(dict(zip(pos, cbits)))
        self._measures.append(measure)
        self.add_ins(measure)

    @contextmanager
    def cif(self, cbits: List[int], condition: int):
        """`with qc.cif(cbits, condition): ...` (quantum_circuit.py:1271-1320)."""
        if len(set(cbits)) != len(cbits):
            raise ValueError("Classical bits not uniquely assigned.")
        if max(cbits) > self.cbits_num - 1 or min(cbits) < 0:
            raise ValueError("Classical bits index out of range.")
        if condition < 0:
            raise ValueError("Classical should be a positive integer.")
        self.executable_on_backend = False
        self.add_ins(Cif(cbits=cbits, condition=condition))
        yield
        instructions = []
        for i in range(len(self.instructions) - 1, -1, -1):
            ins = self.instructions[i]
            if isinstance(ins, Cif) and ins.instructions is None:
                instructions.reverse()
                ins.set_ins(instructions)
                self.instructions = self.instructions[0: i + 1]
                return
            instructions.append(ins)
