from .results import Result, SimuResult, permutebits, ptrace

__all__ = ["Result", "SimuResult", "permutebits", "ptrace"]
