"""Concrete gate classes (quafu/elements/element_gates/element_gates.py:46-476).

Names, positions and matrices are kept identical to the reference because the engine dispatches
on `name.lower()` and reads the same matrices. One reference quirk is kept on purpose so results
match it: CSGate carries the X matrix as its target matrix (element_gates.py:330-345), so a
simulated CS acts as CX.
"""

from typing import List

from . import matrices as mat
from .quantum_gate import ControlledGate, QuantumGate


def wrap_para(matfunc):
    def wrap_func(paras):
        return matfunc(*paras)

    return wrap_func


def _fixed(cls_name, gate_name, matrix, reg_name=None):
    def __init__(self, pos: int):
        self.pos = [pos]

    cls = type(cls_name, (QuantumGate,), {"name": gate_name, "_raw_matrix": matrix, "paras": [],
                                          "__init__": __init__})
    return QuantumGate.register_gate(cls, reg_name)


IdGate = _fixed("IdGate", "ID", mat.IdMatrix)
XGate = _fixed("XGate", "X", mat.XMatrix)
YGate = _fixed("YGate", "Y", mat.YMatrix)
ZGate = _fixed("ZGate", "Z", mat.ZMatrix)
SGate = _fixed("SGate", "S", mat.SMatrix)
SdgGate = _fixed("SdgGate", "Sdg", mat.SMatrix.conj().T)
TGate = _fixed("TGate", "T", mat.TMatrix)
TdgGate = _fixed("TdgGate", "Tdg", mat.TMatrix.conj().T)
SXGate = _fixed("SXGate", "SX", mat.SXMatrix)
SXdgGate = _fixed("SXdgGate", "SXdg", mat.SXMatrix.conj().T)
SYGate = _fixed("SYGate", "SY", mat.SYMatrix)
SYdgGate = _fixed("SYdgGate", "SYdg", mat.SYMatrix.conj().T)
HGate = _fixed("HGate", "H", mat.HMatrix)
WGate = _fixed("WGate", "W", mat.WMatrix)
SWGate = _fixed("SWGate", "SW", mat.SWMatrix)
SWdgGate = _fixed("SWdgGate", "SWdg", mat.SWMatrix.conj().T)


@QuantumGate.register()
class RXGate(QuantumGate):
    def __init__(self, pos: int, theta):
        super().__init__("RX", [pos], [theta], wrap_para(mat.rx_mat))


@QuantumGate.register()
class RYGate(QuantumGate):
    def __init__(self, pos: int, theta):
        super().__init__("RY", [pos], [theta], wrap_para(mat.ry_mat))


@QuantumGate.register()
class RZGate(QuantumGate):
    def __init__(self, pos: int, theta):
        super().__init__("RZ", [pos], [theta], wrap_para(mat.rz_mat))


@QuantumGate.register()
class U3Gate(QuantumGate):
    def __init__(self, pos: int, theta, phi, _lambda):
        super().__init__("U3", [pos], [theta, phi, _lambda], matrix=wrap_para(mat.u3matrix))


@QuantumGate.register(name="p")
class PhaseGate(QuantumGate):
    def __init__(self, pos: int, _lambda):
        super().__init__("P", [pos], [_lambda], wrap_para(mat.pmatrix))


@QuantumGate.register()
class RXXGate(QuantumGate):
    def __init__(self, q1: int, q2: int, theta):
        super().__init__("RXX", [q1, q2], [theta], wrap_para(mat.rxx_mat))


@QuantumGate.register()
class RYYGate(QuantumGate):
    def __init__(self, q1: int, q2: int, theta):
        super().__init__("RYY", [q1, q2], [theta], wrap_para(mat.ryy_mat))


@QuantumGate.register()
class RZZGate(QuantumGate):
    def __init__(self, q1: int, q2: int, theta):
        super().__init__("RZZ", [q1, q2], [theta], wrap_para(mat.rzz_mat))


def _c11(cls_name, gate_name, targ_name, targ_matrix, raw_matrix):
    def __init__(self, ctrl: int, targ: int):
        assert ctrl != targ
        self.ctrls = [ctrl]
        self.targs = [targ]
        self.pos = self.ctrls + self.targs

    cls = type(cls_name, (ControlledGate,), {"name": gate_name, "_targ_name": targ_name,
                                             "_targ_matrix": targ_matrix, "_raw_matrix": raw_matrix,
                                             "paras": [], "__init__": __init__})
    return QuantumGate.register_gate(cls)


CXGate = _c11("CXGate", "CX", "X", mat.XMatrix, mat.CXMatrix)
CYGate = _c11("CYGate", "CY", "Y", mat.YMatrix, mat.CYMatrix)
CZGate = _c11("CZGate", "CZ", "Z", mat.ZMatrix, mat.CZMatrix)
CSGate = _c11("CSGate", "CS", "S", mat.XMatrix, mat.CXMatrix)  # reference quirk, kept
CTGate = _c11("CTGate", "CT", "T", mat.TMatrix, mat.CTMatrix)


@QuantumGate.register()
class CPGate(ControlledGate):
    def __init__(self, ctrl: int, targ: int, _lambda):
        super().__init__("CP", "P", [ctrl], [targ], [_lambda], wrap_para(mat.pmatrix))


@QuantumGate.register()
class CRXGate(ControlledGate):
    def __init__(self, ctrl: int, targ: int, theta):
        super().__init__("CRX", "RX", [ctrl], [targ], [theta], wrap_para(mat.rx_mat))


@QuantumGate.register()
class CRYGate(ControlledGate):
    def __init__(self, ctrl: int, targ: int, theta):
        super().__init__("CRY", "RY", [ctrl], [targ], [theta], wrap_para(mat.ry_mat))


@QuantumGate.register()
class CRZGate(ControlledGate):
    def __init__(self, ctrl: int, targ: int, theta):
        super().__init__("CRZ", "RZ", [ctrl], [targ], [theta], wrap_para(mat.rz_mat))


@QuantumGate.register()
class MCXGate(ControlledGate):
    def __init__(self, ctrls: List[int], targ: int):
        super().__init__("MCX", "X", ctrls, [targ], [], mat.XMatrix)


@QuantumGate.register()
class MCYGate(ControlledGate):
    def __init__(self, ctrls: List[int], targ: int):
        super().__init__("MCY", "Y", ctrls, [targ], [], mat.YMatrix)


@QuantumGate.register()
class MCZGate(ControlledGate):
    def __init__(self, ctrls: List[int], targ: int):
        super().__init__("MCZ", "Z", ctrls, [targ], [], mat.ZMatrix)


@QuantumGate.register()
class MCRXGate(ControlledGate):
    def __init__(self, ctrls: List[int], targ: int, theta):
        super().__init__("MCRX", "RX", ctrls, [targ], [theta], wrap_para(mat.rx_mat))


@QuantumGate.register()
class MCRYGate(ControlledGate):
    def __init__(self, ctrls: List[int], targ: int, theta):
        super().__init__("MCRY", "RY", ctrls, [targ], [theta], wrap_para(mat.ry_mat))


@QuantumGate.register()
class MCRZGate(ControlledGate):
    def __init__(self, ctrls: List[int], targ: int, theta):
        super().__init__("MCRZ", "RZ", ctrls, [targ], [theta], wrap_para(mat.rz_mat))


@QuantumGate.register()
class CCXGate(ControlledGate):
    def __init__(self, ctrl1: int, ctrl2: int, targ: int):
        super().__init__("CCX", "X", [ctrl1, ctrl2], [targ], [], mat.XMatrix)


@QuantumGate.register()
class CSwapGate(ControlledGate):
    def __init__(self, ctrl: int, targ1: int, targ2: int):
        super().__init__("CSWAP", "SWAP", [ctrl], [targ1, targ2], [], mat.SwapMatrix)


@QuantumGate.register()
class SwapGate(QuantumGate):
    name = "SWAP"
    matrix = mat.SwapMatrix
    _raw_matrix = mat.SwapMatrix
    paras = []

    def __init__(self, q1: int, q2: int):
        self.pos = [q1, q2]

    @property
    def symbol(self):
        return "x"


@QuantumGate.register()
class ISwapGate(QuantumGate):
    name = "ISWAP"
    matrix = mat.ISwapMatrix
    _raw_matrix = mat.ISwapMatrix
    paras = []

    def __init__(self, q1: int, q2: int):
        self.pos = [q1, q2]


QuantumGate.gate_classes["cnot"] = CXGate
QuantumGate.gate_classes["toffoli"] = CCXGate
QuantumGate.gate_classes["fredon"] = CSwapGate
FredkinGate = CSwapGate
ToffoliGate = CCXGate

__all__ = [
    "XGate", "YGate", "ZGate", "IdGate", "WGate", "HGate", "SGate", "SdgGate", "TGate", "TdgGate",
    "SXGate", "SXdgGate", "SYGate", "SYdgGate", "SWGate", "SWdgGate", "RXGate", "RYGate", "RZGate",
    "PhaseGate", "RXXGate", "RYYGate", "RZZGate", "SwapGate", "ISwapGate", "CXGate", "CYGate", "CZGate",
    "CSGate", "CTGate", "CPGate", "CRXGate", "CRYGate", "CRZGate", "ToffoliGate", "FredkinGate", "MCXGate",
    "MCYGate", "MCZGate", "MCRXGate", "MCRYGate", "MCRZGate", "U3Gate", "CCXGate", "CSwapGate",
]
