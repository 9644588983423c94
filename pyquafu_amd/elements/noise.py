"""Noise channels (quafu/elements/noise.py:26-163).

A KrausChannel carries Kraus operators as one-qubit QuantumGates; a UnitaryChannel carries
unitaries and their probabilities and picks one per trajectory (UnitaryChannel.gen_gate,
noise.py:64-68). The engine applies both as program ops (include/qfvm.h QFVM_OP_KRAUS /
QFVM_OP_UCHAN), each drawing one uniform.
"""

from __future__ import annotations

from typing import List, Optional

import numpy as np

from .element_gates import IdGate, XGate, YGate, ZGate
from .instruction import Instruction
from .quantum_gate import QuantumGate


class KrausChannel(Instruction):
    def __init__(self, name, pos: int, gatelist: Optional[List[QuantumGate]] = None):
        super().__init__([pos])
        self._name = name
        self.gatelist = [] if gatelist is None else gatelist

    @property
    def name(self):
        return self._name

    @name.setter
    def name(self, value):
        self._name = value

    def to_qasm(self, with_para=False):
        raise ValueError("Can not convert noise channel to qasm")

    @property
    def symbol(self):
        if self.paras:
            return f"{self.name}({','.join(f'{p:.3f}' for p in self.paras)})"
        return f"{self.name}"

    def __repr__(self):
        return self.symbol


class UnitaryChannel(KrausChannel):
    def __init__(self, name, pos: int, gatelist=None, probs=None):
        super().__init__(name, pos, gatelist)
        self.probs = [] if probs is None else probs

    def gen_gate(self, rng=None):
        """One gate drawn with the channel's probabilities (numpy's choice rule)."""
        rng = np.random.default_rng() if rng is None else rng
        return self.gatelist[int(rng.choice(len(self.gatelist), p=self.probs))]


class BitFlip(UnitaryChannel):
    def __init__(self, pos: int, p):
        super().__init__("BitFlip", pos, [XGate(pos), IdGate(pos)], [p, 1 - p])
        self.paras = [p]


class Dephasing(UnitaryChannel):
    def __init__(self, pos: int, p):
        super().__init__("Dephasing", pos, [ZGate(pos), IdGate(pos)], [p, 1 - p])
        self.paras = [p]


class Depolarizing(UnitaryChannel):
    def __init__(self, pos: int, p):
        super().__init__("Depolarizing", pos, [XGate(pos), YGate(pos), ZGate(pos), IdGate(pos)],
                         [p / 3, p / 3, p / 3, 1 - p])
        self.paras = [p]


class AmplitudeDamping(KrausChannel):
    def __init__(self, pos: int, p):
        k0 = np.array([[1.0, 0.0], [0.0, np.sqrt(1 - p)]], dtype=complex)
        k1 = np.array([[0.0, np.sqrt(p)], [0.0, 0.0]], dtype=complex)
        super().__init__("AmpDamping", pos, [QuantumGate("AMPDAMP0", [pos], [], k0),
                                             QuantumGate("AMPDAMP1", [pos], [], k1)])
        self.paras = [p]


class Decoherence(KrausChannel):
    def __init__(self, pos, t, T1, T2):
        k0 = np.array([[1.0, 0.0], [0.0, np.exp(-t / T2)]], dtype=complex)
        k1 = np.array([[0.0, np.sqrt(1 - np.exp(-t / T1))], [0.0, 0.0]], dtype=complex)
        k2 = np.array([[1.0, 0.0], [0.0, np.sqrt(np.exp(-t / T1) - np.exp(-2 * t / T2))]], dtype=complex)
        super().__init__("Decoherence", pos, [QuantumGate("DECAY0", [pos], [], k0),
                                              QuantumGate("DECAY1", [pos], [], k1),
                                              QuantumGate("DECAY2", [pos], [], k2)])
        self.paras = [t, T1, T2]


# registry names used by QuantumCircuit.add_noise (the reference registers the channels under
# their lowercased `name`: bitflip, dephasing, depolarizing, ampdamping, decoherence)
for _cls, _name in [(BitFlip, "bitflip"), (Dephasing, "dephasing"), (Depolarizing, "depolarizing"),
                    (AmplitudeDamping, "ampdamping"), (Decoherence, "decoherence")]:
    if _name not in Instruction.ins_classes:
        Instruction.register_ins(_cls, _name)
