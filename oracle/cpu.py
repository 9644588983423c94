"""ORACLE -- test infrastructure, not product code.

ctypes wrapper around oracle/_build/libqfvm_oracle.so, the CPU restatement of the reference's qfvm
engine (src/qfvm/statevector.hpp). Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg import this module. It reads any program object that has `.ops` (structured array laid out
as `qfvm_op` in include/qfvm.h), `.idx` (int32) and `.mat` (float64).
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libqfvm_oracle.so")

# Layout of `qfvm_op` (include/qfvm.h), stated independently of the product's copy.
OP_DTYPE = np.dtype(
    [
        ("code", "<i4"),
        ("n_ctrl", "<i4"),
        ("n_qubits", "<i4"),
        ("n_cbits", "<i4"),
        ("q_off", "<i8"),
        ("c_off", "<i8"),
        ("m_off", "<i8"),
        ("param", "<f8"),
        ("cond", "<i8"),
        ("body", "<i8"),
    ]
)
assert OP_DTYPE.itemsize == 64


class _Program(ctypes.Structure):
    _fields_ = [
        ("ops", ctypes.c_void_p),
        ("n_ops", ctypes.c_int64),
        ("idx", ctypes.c_void_p),
        ("n_idx", ctypes.c_int64),
        ("mat", ctypes.c_void_p),
        ("n_mat", ctypes.c_int64),
    ]


def build() -> None:
    """Compile the oracle library with its Makefile (reference flags)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32
        L.qfvm_last_error.restype = ctypes.c_char_p
        L.qfvm_create.argtypes = [i32, i32, ctypes.POINTER(vp)]
        L.qfvm_destroy.argtypes = [vp]
        L.qfvm_init_zero.argtypes = [vp]
        L.qfvm_init_basis.argtypes = [vp, ctypes.c_uint64]
        L.qfvm_load.argtypes = [vp, vp, ctypes.c_uint64]
        L.qfvm_store.argtypes = [vp, vp, ctypes.c_uint64]
        L.qfvm_apply.argtypes = [vp, ctypes.POINTER(_Program)]
        L.qfvm_run.argtypes = [vp, ctypes.POINTER(_Program), vp, i64, vp, i32, i32]
        L.qfvm_expect_paulis.argtypes = [vp, vp, vp, vp, vp, i64, vp]
        L.qfvm_probabilities.argtypes = [vp, vp, ctypes.c_uint64]
        L.qfvm_sample.argtypes = [vp, vp, i64, vp]
        L.qfvm_sample_sequential.argtypes = [vp, vp, i64, vp]
        L.qfvm_norm2.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.qfvm_oracle_data.restype = ctypes.c_void_p
        L.qfvm_oracle_data.argtypes = [vp]
        L.qfvm_adjoint.argtypes = [vp, ctypes.POINTER(_Program), vp, vp, i64, vp, vp, vp, vp, vp, i64, vp,
                                   ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _check(rc: int) -> None:
    if rc != 0:
        raise OracleError(rc, lib().qfvm_last_error().decode())


def _as_program(prog):
    ops = np.ascontiguousarray(prog.ops).view(OP_DTYPE) if len(prog.ops) else np.zeros(0, OP_DTYPE)
    idx = np.ascontiguousarray(prog.idx, dtype=np.int32)
    mat = np.ascontiguousarray(prog.mat, dtype=np.float64)
    keep = (ops, idx, mat)
    p = _Program(
        ops.ctypes.data if len(ops) else None,
        len(ops),
        idx.ctypes.data if len(idx) else None,
        len(idx),
        mat.ctypes.data if len(mat) else None,
        len(mat),
    )
    return p, keep


class CpuState:
    """One complex128 state in host memory, driven through the oracle's C-ABI."""

    def __init__(self, n: int):
        h = ctypes.c_void_p()
        _check(lib().qfvm_create(n, 2, ctypes.byref(h)))
        self.h = h
        self.n = n

    def __del__(self):
        if getattr(self, "h", None):
            lib().qfvm_destroy(self.h)
            self.h = None

    def load(self, psi: np.ndarray) -> None:
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        _check(lib().qfvm_load(self.h, psi.ctypes.data, psi.size))

    def init_basis(self, index: int) -> None:
        _check(lib().qfvm_init_basis(self.h, index))

    def store(self) -> np.ndarray:
        out = np.empty(1 << self.n, dtype=np.complex128)
        _check(lib().qfvm_store(self.h, out.ctypes.data, out.size))
        return out

    def view(self) -> np.ndarray:
        """The host state itself (complex128, no copy); valid while this object lives and until the
        next call that reallocates it."""
        ptr = lib().qfvm_oracle_data(self.h)
        buf = (ctypes.c_double * (2 << self.n)).from_address(ptr)
        return np.frombuffer(buf, dtype=np.complex128)

    def apply(self, prog) -> None:
        p, keep = _as_program(prog)
        _check(lib().qfvm_apply(self.h, ctypes.byref(p)))
        del keep

    def run(self, prog, uniforms, n_creg: int, skip_measure: bool = False) -> np.ndarray:
        p, keep = _as_program(prog)
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        creg = np.zeros(max(n_creg, 1), dtype=np.int32)
        _check(
            lib().qfvm_run(
                self.h, ctypes.byref(p), u.ctypes.data if u.size else None, u.size, creg.ctypes.data,
                n_creg, int(skip_measure)
            )
        )
        del keep
        return creg[:n_creg]

    def expect(self, terms) -> np.ndarray:
        """terms: iterable of (paulistr, positions)."""
        chars, pos, off, ln = [], [], [], []
        for s, p in terms:
            off.append(len(pos))
            ln.append(len(p))
            chars.extend(s.encode())
            pos.extend(p)
        c = np.array(chars + [0], dtype=np.uint8)
        ps = np.array(pos + [0], dtype=np.int32)
        o = np.array(off, dtype=np.int64)
        l_ = np.array(ln, dtype=np.int32)
        out = np.zeros(len(off), dtype=np.float64)
        _check(lib().qfvm_expect_paulis(self.h, c.ctypes.data, ps.ctypes.data, o.ctypes.data, l_.ctypes.data,
                                        len(off), out.ctypes.data))
        return out

    def probabilities(self) -> np.ndarray:
        out = np.empty(1 << self.n, dtype=np.float64)
        _check(lib().qfvm_probabilities(self.h, out.ctypes.data, out.size))
        return out

    def sample(self, uniforms) -> np.ndarray:
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        out = np.empty(u.size, dtype=np.uint64)
        _check(lib().qfvm_sample(self.h, u.ctypes.data, u.size, out.ctypes.data))
        return out

    def sample_sequential(self, uniforms) -> np.ndarray:
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        out = np.empty(u.size, dtype=np.uint64)
        _check(lib().qfvm_sample_sequential(self.h, u.ctypes.data, u.size, out.ctypes.data))
        return out

    def adjoint(self, prog, grad_ops, grad_gens: str, terms, coeffs):
        """grad_adjoint restated on the oracle's kernels (qfvm_adjoint in qfvm_cpu.cpp); terms as
        for `expect`, one complex coefficient per term. Returns (grads, <H>)."""
        p, keep = _as_program(prog)
        ops = np.ascontiguousarray(grad_ops, dtype=np.int64)
        gens = np.frombuffer(grad_gens.encode() + b"\0", dtype=np.uint8)
        chars, pos, off, ln = [], [], [], []
        for s, q in terms:
            off.append(len(pos))
            ln.append(len(q))
            chars.extend(s.encode())
            pos.extend(int(x) for x in q)
        c = np.array(chars + [0], dtype=np.uint8)
        ps = np.array(pos + [0], dtype=np.int32)
        o = np.array(off + [0], dtype=np.int64)
        l_ = np.array(ln + [0], dtype=np.int32)
        cf = np.asarray(coeffs, dtype=np.complex128).reshape(-1)
        cfa = np.zeros(2 * max(cf.size, 1))
        cfa[0: 2 * cf.size: 2] = cf.real
        cfa[1: 2 * cf.size: 2] = cf.imag
        grads = np.zeros(max(ops.size, 1))
        e = ctypes.c_double(0.0)
        _check(lib().qfvm_adjoint(self.h, ctypes.byref(p), ops.ctypes.data, gens.ctypes.data, ops.size,
                                  c.ctypes.data, ps.ctypes.data, o.ctypes.data, l_.ctypes.data, cfa.ctypes.data,
                                  len(off), grads.ctypes.data, ctypes.byref(e)))
        del keep
        return grads[: ops.size], e.value

    def norm2(self) -> float:
        v = ctypes.c_double()
        _check(lib().qfvm_norm2(self.h, ctypes.byref(v)))
        return v.value


def sample_on(psi: np.ndarray, uniforms) -> np.ndarray:
    """measure_samples' indices (qfvm_sample: sequential running sum, binary search per shot) for an
    explicit host state, e.g. a state read back from the GPU (complex64 states widen exactly)."""
    n = int(np.log2(psi.size))
    s = CpuState(n)
    s.load(psi)
    return s.sample(uniforms)
