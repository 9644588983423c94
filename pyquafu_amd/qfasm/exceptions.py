"""Errors of the OpenQASM front end (quafu/qfasm/exceptions.py)."""

from ..exceptions import QuafuError


class QfasmError(QuafuError):
    """Base error of the QASM front end."""


class LexerError(QfasmError):
    """Malformed input text: an illegal character, a bad include."""


class ParserError(QfasmError):
    """Well-formed tokens that do not make a valid program."""
