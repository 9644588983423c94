"""OpenQASM 2.0 ingress: text -> QuantumCircuit (quafu/qfasm/qfasm_parser.py, qfasm_convertor.py).

A hand-written tokenizer and recursive-descent parser with the reference parser's semantics:

  * quantum / classical registers are laid out in declaration order into one flat qubit / cbit
    index space (QfasmParser.update_sym_tab, qfasm_parser.py:145-180);
  * a gate named in the native table (qfasm_utils.py:128-166) becomes that gate class; `U(t,p,l)`
    becomes rz(l), ry(t), rz(p) and `CX` becomes cx (handle_gateins, qfasm_parser.py:185-244);
    mcx / mcy / mcz take every argument but the last as controls;
  * any other gate is expanded through its definition, recursively: the user's `gate` blocks and the
    standard library preloaded from `library.QELIB1` (the lexer pushes qelib1.inc first,
    qfasm_lexer.py:56-58, and `include "qelib1.inc";` is then a no-op);
  * register arguments broadcast (all register arguments of one call must have the same size);
  * `measure a -> c`, `reset a`, `barrier ...` and `if (creg == n) op` become Measure, Reset, Barrier
    and Cif; `name = number;` declares a named Parameter (p_statement_defparam).

Parameter expressions: numbers, pi, gate parameters and named parameters, + - * / ^ (^ right
associative, unary minus binding tighter than ^, qfasm_parser.py:84-89) and sin cos tan exp ln sqrt
acos atan asin.
"""

from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional

from ..elements import element_gates as qeg
from ..elements.instruction import Barrier, Cif, Delay, Measure, Reset
from ..elements.parameters import Parameter, ParameterExpression
from .exceptions import LexerError, ParserError
from .library import NATIVE_SIGNATURES, QELIB1

# native gate table (qfasm_utils.py:128-166); "xy" (pulse-level XYResonance) is not simulated here
NATIVE = {
    "x": qeg.XGate, "y": qeg.YGate, "z": qeg.ZGate, "h": qeg.HGate, "s": qeg.SGate, "sdg": qeg.SdgGate,
    "t": qeg.TGate, "tdg": qeg.TdgGate, "rx": qeg.RXGate, "ry": qeg.RYGate, "rz": qeg.RZGate, "id": qeg.IdGate,
    "sx": qeg.SXGate, "sy": qeg.SYGate, "w": qeg.WGate, "sw": qeg.SWGate, "p": qeg.PhaseGate, "cx": qeg.CXGate,
    "cnot": qeg.CXGate, "cp": qeg.CPGate, "swap": qeg.SwapGate, "rxx": qeg.RXXGate, "ryy": qeg.RYYGate,
    "rzz": qeg.RZZGate, "cy": qeg.CYGate, "cz": qeg.CZGate, "cs": qeg.CSGate, "ct": qeg.CTGate,
    "ccx": qeg.ToffoliGate, "toffoli": qeg.ToffoliGate, "cswap": qeg.FredkinGate, "fredkin": qeg.FredkinGate,
    "mcx": qeg.MCXGate, "mcy": qeg.MCYGate, "mcz": qeg.MCZGate, "delay": Delay,
}
MULTI_CONTROL = ("mcx", "mcy", "mcz")
RESERVED = {"creg", "qreg", "pi", "measure", "include", "barrier", "gate", "opaque", "reset", "if", "OPENQASM"}
FUNCS = {"sin": math.sin, "cos": math.cos, "tan": math.tan, "exp": math.exp, "ln": math.log, "sqrt": math.sqrt,
         "acos": math.acos, "atan": math.atan, "asin": math.asin}
PFUNCS = {"sin": "sin", "cos": "cos", "tan": "tan", "exp": "exp", "ln": "log", "sqrt": "sqrt",
          "acos": "arccos", "atan": "arctan", "asin": "arcsin"}

_TOKEN = re.compile(
    r"""
    (?P<ws>[ \t\r\n]+) | (?P<comment>//[^\n]*) |
    (?P<real>(?:\d+\.\d*|\.\d+)(?:[eE][-+]?\d+)?|\d+[eE][-+]?\d+) | (?P<int>\d+) |
    (?P<string>"[^"\n]*") |
    (?P<id>OPENQASM|U|CX|[a-z][A-Za-z0-9_]*) |
    (?P<op>->|==|[;,()\[\]{}+\-*/^=])
    """,
    re.VERBOSE,
)


@dataclass
class Tok:
    kind: str
    text: str
    line: int


def tokenize(text: str) -> List[Tok]:
    out: List[Tok] = []
    pos, line = 0, 1
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if m is None:
            raise LexerError(f"Illegal character {text[pos]!r} at line {line}")
        kind = m.lastgroup
        s = m.group()
        if kind not in ("ws", "comment"):
            out.append(Tok(kind, s, line))
        line += s.count("\n")
        pos = m.end()
        if kind == "id" and s == "include":
            # the file name must be one double-quoted string on the line (qfasm_lexer.py:93-110)
            rest = re.match(r"[ \t]*([^\s;]*)", text[pos:])
            name = rest.group(1)
            if name.startswith("'"):
                raise LexerError(f"Illegal character {name[0]!r} at line {line}")
            if not re.fullmatch(r'"[^"\n]*"', name):
                raise LexerError("Invalid include: need a quoted string as filename.")
    out.append(Tok("eof", "", line))
    return out


# ------------------------------------------------------------------ expression tree
@dataclass
class Num:
    v: float


@dataclass
class Name:
    name: str
    line: int


@dataclass
class Neg:
    a: object


@dataclass
class Bin:
    op: str
    a: object
    b: object


@dataclass
class Call:
    fn: str
    a: object


@dataclass
class Arg:
    """A register or a register element: name, index (None = whole register)."""

    name: str
    index: Optional[int]
    line: int


@dataclass
class GateCall:
    name: str
    params: List[object]
    args: List[Arg]
    line: int


@dataclass
class GateDef:
    name: str
    params: List[str]
    qargs: List[str]
    body: Optional[List[GateCall]]  # None for opaque


@dataclass
class Register:
    start: int
    size: int
    quantum: bool


@dataclass
class State:
    gates: Dict[str, GateDef] = field(default_factory=dict)
    regs: Dict[str, Register] = field(default_factory=dict)
    params: Dict[str, Parameter] = field(default_factory=dict)
    nq: int = 0
    nc: int = 0
    out: list = field(default_factory=list)  # instructions, in order
    measured: bool = False
    dynamic: bool = False  # an operation after a measure, or a cif (not runnable on hardware)


class Parser:
    def __init__(self, text: str, state: State, filename: str = "<qasm>"):
        self.toks = tokenize(text)
        self.i = 0
        self.st = state
        self.filename = filename

    # ---- token helpers
    def peek(self, k: int = 0) -> Tok:
        return self.toks[min(self.i + k, len(self.toks) - 1)]

    def take(self) -> Tok:
        t = self.toks[self.i]
        self.i = min(self.i + 1, len(self.toks) - 1)
        return t

    def expect(self, text: str, what: str = "") -> Tok:
        t = self.take()
        if t.text != text:
            raise ParserError(f"Expecting '{text}'{' ' + what if what else ''} at line {t.line}, found {t.text!r}")
        return t

    def ident(self) -> Tok:
        t = self.take()
        if t.kind != "id" or t.text in RESERVED:
            raise ParserError(f"Expecting an identifier at line {t.line}, found {t.text!r}")
        return t

    def integer(self) -> int:
        t = self.take()
        if t.kind != "int":
            raise ParserError(f"Expecting an integer at line {t.line}, found {t.text!r}")
        return int(t.text)

    # ---- program
    def parse(self) -> None:
        if self.peek().text == "OPENQASM":
            self.take()
            v = self.take()
            if v.kind not in ("real", "int") or float(v.text) != 2.0:
                raise ParserError(f"Only OPENQASM 2.0 is supported (line {v.line})")
            self.expect(";", "after the version")
        while self.peek().kind != "eof":
            self.statement()

    def statement(self) -> None:
        t = self.peek()
        if t.text == "include":
            self.take()
            f = self.take()
            if f.kind != "string":
                raise LexerError("Invalid include: need a quoted string as filename.")
            if self.peek().text != ";":
                raise LexerError(f'Expecting ";" for INCLUDE at line {f.line}')
            self.take()
            self.include(f.text[1:-1])
        elif t.text in ("qreg", "creg"):
            self.take()
            name = self.ident()
            self.expect("[")
            size = self.integer()
            self.expect("]")
            self.expect(";")
            self.declare_reg(name, size, t.text == "qreg")
        elif t.text in ("gate", "opaque"):
            self.take()
            self.gate_def(opaque=t.text == "opaque")
        elif t.text == "if":
            self.take()
            self.expect("(")
            reg = self.ident()
            if self.peek().text != "==":
                raise ParserError(f"Illegal IF statement, expecting '==' at line {reg.line}")
            self.take()
            value = self.integer()
            self.expect(")")
            op = self.qop()
            self.apply_if(reg, value, op)
        elif t.kind == "id" and self.peek(1).text == "=":
            name = self.ident()
            self.take()
            v = self.take()
            if v.kind not in ("int", "real"):
                raise ParserError(f"Expecting 'INT' or 'FLOAT' behind '=' at line {v.line}")
            self.expect(";")
            if name.text in self.st.params:
                raise ParserError(f"Duplicate declaration for parameter {name.text} at line {name.line}")
            self.st.params[name.text] = Parameter(name.text, float(v.text) if v.kind == "real" else int(v.text))
        else:
            self.apply(self.qop())

    def include(self, filename: str) -> None:
        if filename == "qelib1.inc":
            return  # preloaded
        if not os.path.isfile(filename):
            raise LexerError(f"Include file {filename} cannot be found, check the path")
        with open(filename, encoding="utf-8") as fh:
            Parser(fh.read(), self.st, filename).parse()

    def declare_reg(self, name: Tok, size: int, quantum: bool) -> None:
        if name.text in self.st.regs or name.text in self.st.gates:
            raise ParserError(f"Duplicate declaration for {name.text} at line {name.line}")
        if quantum:
            self.st.regs[name.text] = Register(self.st.nq, size, True)
            self.st.nq += size
        else:
            self.st.regs[name.text] = Register(self.st.nc, size, False)
            self.st.nc += size

    def id_list(self) -> List[Tok]:
        ids = [self.ident()]
        while self.peek().text == ",":
            self.take()
            ids.append(self.ident())
        return ids

    def gate_def(self, opaque: bool) -> None:
        name = self.ident()
        params: List[str] = []
        if self.peek().text == "(":
            self.take()
            if self.peek().text != ")":
                params = [t.text for t in self.id_list()]
            self.expect(")")
        qargs = [t.text for t in self.id_list()]
        if len(set(params)) != len(params) or len(set(qargs)) != len(qargs):
            raise ParserError(f"Duplicate argument in the declaration of gate {name.text} at line {name.line}")
        body: Optional[List[GateCall]] = None
        if opaque:
            self.expect(";")
        else:
            self.expect("{")
            body = []
            while self.peek().text != "}":
                call = self.qop(in_gate=True)
                if call.name == name.text:
                    raise ParserError(f"The gate {name.text} calls itself at line {call.line}")
                for a in call.args:
                    if a.name not in qargs or a.index is not None:
                        raise ParserError(f"The qubit {a.name} is undefined in gate qubit parameters at line {a.line}")
                if call.name != "barrier":
                    self.check_call(call)
                for p in call.params:
                    self.check_expr(p, params)
                body.append(call)
            self.expect("}")
        if name.text in self.st.gates or name.text in self.st.regs:
            raise ParserError(f"Duplicate declaration for {name.text} at line {name.line}")
        self.st.gates[name.text] = GateDef(name.text, params, qargs, body)

    def check_expr(self, e, params: List[str]) -> None:
        if isinstance(e, Name):
            if e.name not in params and e.name not in self.st.params:
                raise ParserError(f"The parameter {e.name} is undefined at line {e.line}")
        elif isinstance(e, Neg):
            self.check_expr(e.a, params)
        elif isinstance(e, Call):
            self.check_expr(e.a, params)
        elif isinstance(e, Bin):
            self.check_expr(e.a, params)
            self.check_expr(e.b, params)

    def check_call(self, call: GateCall) -> None:
        if call.name in MULTI_CONTROL:
            return
        g = self.st.gates.get(call.name)
        if g is None:
            raise ParserError(f"The gate {call.name} is undefined at line {call.line}")
        if len(call.args) != len(g.qargs):
            raise ParserError(f"The number of qubits of gate {call.name} is inconsistent with its declaration "
                              f"at line {call.line}")
        if len(call.params) != len(g.params):
            raise ParserError(f"The number of parameters of gate {call.name} is inconsistent with its declaration "
                              f"at line {call.line}")

    def qop(self, in_gate: bool = False) -> GateCall:
        t = self.take()
        if t.kind != "id":
            raise ParserError(f"Expecting an operation at line {t.line}, found {t.text!r}")
        name = t.text
        if name == "measure" and not in_gate:
            a = self.arg()
            self.expect("->", "in measure")
            c = self.arg()
            self.expect(";")
            return GateCall("measure", [], [a, c], t.line)
        params: List[object] = []
        if self.peek().text == "(":
            self.take()
            if self.peek().text != ")":
                params.append(self.expr())
                while self.peek().text == ",":
                    self.take()
                    params.append(self.expr())
            self.expect(")")
        args = [self.arg()]
        while self.peek().text == ",":
            self.take()
            args.append(self.arg())
        self.expect(";")
        return GateCall(name, params, args, t.line)

    def arg(self) -> Arg:
        t = self.ident()
        index = None
        if self.peek().text == "[":
            self.take()
            index = self.integer()
            self.expect("]")
        return Arg(t.text, index, t.line)

    # ---- expressions
    def expr(self):
        return self.additive()

    def additive(self):
        a = self.multiplicative()
        while self.peek().text in ("+", "-"):
            op = self.take().text
            a = Bin(op, a, self.multiplicative())
        return a

    def multiplicative(self):
        a = self.power()
        while self.peek().text in ("*", "/"):
            op = self.take().text
            a = Bin(op, a, self.power())
        return a

    def power(self):
        a = self.unary()
        if self.peek().text == "^":
            self.take()
            return Bin("^", a, self.power())
        return a

    def unary(self):
        if self.peek().text == "-":
            self.take()
            return Neg(self.unary())
        if self.peek().text == "+":
            self.take()
            return self.unary()
        return self.primary()

    def primary(self):
        t = self.take()
        if t.kind in ("int", "real"):
            return Num(float(t.text) if t.kind == "real" else int(t.text))
        if t.text == "pi":
            return Num(math.pi)
        if t.text == "(":
            e = self.expr()
            self.expect(")")
            return e
        if t.kind == "id" and t.text in FUNCS and self.peek().text == "(":
            self.take()
            e = self.expr()
            self.expect(")")
            return Call(t.text, e)
        if t.kind == "id" and t.text not in RESERVED:
            return Name(t.text, t.line)
        raise ParserError(f"Unexpected {t.text!r} in an expression at line {t.line}")

    def value(self, e, env: Dict[str, object]):
        if isinstance(e, Num):
            return e.v
        if isinstance(e, Name):
            if e.name in env:
                return env[e.name]
            if e.name in self.st.params:
                return self.st.params[e.name]
            raise ParserError(f"The parameter {e.name} is undefined at line {e.line}")
        if isinstance(e, Neg):
            return -self.value(e.a, env)
        if isinstance(e, Call):
            v = self.value(e.a, env)
            if isinstance(v, ParameterExpression):
                return getattr(v, PFUNCS[e.fn])()
            return FUNCS[e.fn](v)
        a, b = self.value(e.a, env), self.value(e.b, env)
        if e.op == "+":
            return a + b
        if e.op == "-":
            return a - b
        if e.op == "*":
            return a * b
        if e.op == "/":
            return a / b
        return a**b

    # ---- semantics
    def qubits_of(self, a: Arg, quantum: bool = True) -> List[int]:
        reg = self.st.regs.get(a.name)
        kind = "qubit" if quantum else "classical bit"
        if reg is None or reg.quantum != quantum:
            raise ParserError(f"The {kind} {a.name} is undefined in {kind} register at line {a.line}")
        if a.index is None:
            return list(range(reg.start, reg.start + reg.size))
        if a.index >= reg.size:
            raise ParserError(f"{kind.capitalize()} arrays {a.name} are out of bounds at line {a.line}")
        return [reg.start + a.index]

    def expand(self, call: GateCall, env: Dict[str, object], qmap: Optional[Dict[str, int]] = None) -> list:
        """Gate call -> instructions. At top level (qmap None) arguments are registers and may
        broadcast; inside a gate body they are the body's qubit names."""
        if call.name == "barrier":
            if qmap is not None:
                return [Barrier([qmap[a.name] for a in call.args])]
            qs: List[int] = []
            for a in call.args:
                qs.extend(self.qubits_of(a))
            return [Barrier(qs)]
        if call.name == "reset":
            return [Reset(self.qubits_of(a)) for a in call.args]
        if call.name == "measure":
            qs, cs = self.qubits_of(call.args[0]), self.qubits_of(call.args[1], quantum=False)
            if len(qs) != len(cs):
                raise ParserError(f"MEASURE: the num of qubit and clbit doesn't match at line {call.line}")
            return [Measure(dict(zip(qs, cs)))]
        if call.name not in MULTI_CONTROL:
            self.check_call(call)
        params = [self.value(p, env) for p in call.params]
        if qmap is not None:
            rows = [[qmap[a.name] for a in call.args]]
        else:
            cols = [self.qubits_of(a) for a in call.args]
            width = max(len(c) for c in cols)
            for a, c in zip(call.args, cols):
                if len(c) != width and (a.index is None or width == 1):
                    raise ParserError(f"The num of qreg's qubit is inconsistent at line {call.line}")
            rows = [[c[0] if len(c) == 1 else c[k] for c in cols] for k in range(width)]
        out = []
        for qs in rows:
            if len(set(qs)) != len(qs):
                raise ParserError(f"Qubit used as different argument when call gate {call.name} at line {call.line}")
            out.extend(self.gate_instructions(call, qs, params))
        return out

    def gate_instructions(self, call: GateCall, qs: List[int], params: list) -> list:
        name = call.name
        if name == "U":
            theta, phi, lam = params
            return [qeg.RZGate(qs[0], lam), qeg.RYGate(qs[0], theta), qeg.RZGate(qs[0], phi)]
        if name == "CX":
            return [qeg.CXGate(qs[0], qs[1])]
        if name in MULTI_CONTROL:
            if params:
                raise ParserError(f"{name} takes no parameters (line {call.line})")
            return [NATIVE[name](qs[:-1], qs[-1])]
        if name in NATIVE:
            return [NATIVE[name](*qs, *params)]
        g = self.st.gates[name]
        if g.body is None:
            raise ParserError(f"The opaque gate {name} cannot be simulated (line {call.line})")
        env = dict(zip(g.params, params))
        qmap = dict(zip(g.qargs, qs))
        out = []
        for sub in g.body:
            out.extend(self.expand(sub, env, qmap))
        return out

    def apply(self, call: GateCall) -> None:
        if call.name not in ("measure", "barrier") and self.st.measured:
            self.st.dynamic = True
        if call.name == "measure":
            self.st.measured = True
        self.st.out.extend(self.expand(call, {}))

    def apply_if(self, reg: Tok, value: int, call: GateCall) -> None:
        r = self.st.regs.get(reg.text)
        if r is None or r.quantum:
            raise ParserError(f"The classical bit {reg.text} is undefined in classical bit register at line {reg.line}")
        self.st.dynamic = True
        cbits = list(range(r.start, r.start + r.size))
        self.st.out.append(Cif(cbits=cbits, condition=value, instructions=self.expand(call, {})))


def _library_state() -> State:
    st = State()
    st.gates["U"] = GateDef("U", ["theta", "phi", "lambda"], ["q"], [])
    st.gates["CX"] = GateDef("CX", [], ["c", "t"], [])
    for name, n_par, n_q in NATIVE_SIGNATURES:
        st.gates[name] = GateDef(name, [f"p{k}" for k in range(n_par)], [f"q{k}" for k in range(n_q)], [])
    Parser(QELIB1, st, "qelib1.inc").parse()
    return st


def parse_qasm(text: str):
    """OpenQASM 2.0 text -> QuantumCircuit (qasm_to_quafu, qfasm_convertor.py:20-35)."""
    from ..circuits.quantum_circuit import QuantumCircuit  # pylint: disable=import-outside-toplevel

    st = _library_state()
    Parser(text, st).parse()
    qc = QuantumCircuit(st.nq, st.nc)
    qc._cnum = st.nc  # pylint: disable=protected-access  # no creg: no classical bits
    for ins in st.out:
        qc.add_ins(ins)
        if isinstance(ins, Measure):
            qc._measures.append(ins)  # pylint: disable=protected-access
    qc.openqasm = text
    qc.executable_on_backend = not st.dynamic
    return qc


__all__ = ["parse_qasm", "tokenize", "NATIVE"]
