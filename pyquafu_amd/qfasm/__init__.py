"""OpenQASM 2.0 front end (quafu.qfasm)."""

from .exceptions import LexerError, ParserError, QfasmError
from .parser import parse_qasm


def qasm_to_quafu(openqasm: str):
    """OpenQASM 2.0 text -> QuantumCircuit (qfasm_convertor.py:20-35)."""
    return parse_qasm(openqasm)


__all__ = ["qasm_to_quafu", "parse_qasm", "LexerError", "ParserError", "QfasmError"]
