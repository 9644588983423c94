"""Exception types of the host API (names mirror quafu/exceptions/quafu_error.py:17-39)."""


class QuafuError(Exception):
    """Base error of the host API."""


class CircuitError(QuafuError):
    """Raised when a circuit is built or used incorrectly."""


class ServerError(QuafuError):
    """Kept for name compatibility; nothing in the statevector path raises it."""


class CompileError(QuafuError):
    """Kept for name compatibility; nothing in the statevector path raises it."""
