"""ctypes binding of libqfvm_hip.so, the MI355X engine behind the C-ABI in include/qfvm.h.

There is no CPU fallback: if the library is missing or cannot reach a GPU, every call raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

from .simulators.program import OP_DTYPE, Program

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqfvm_hip.so")

C64, C128 = 1, 2
_DTYPES = {"c64": C64, "complex64": C64, C64: C64, "c128": C128, "complex128": C128, C128: C128}

# status codes (qfvm_status)
OK, ERR_VALUE, ERR_INDEX, ERR_MEMORY, ERR_DEVICE, ERR_RUNTIME = range(6)


class EngineError(RuntimeError):
    """A failure reported by the HIP engine."""

    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class CProgram(ctypes.Structure):
    _fields_ = [
        ("ops", ctypes.c_void_p),
        ("n_ops", ctypes.c_int64),
        ("idx", ctypes.c_void_p),
        ("n_idx", ctypes.c_int64),
        ("mat", ctypes.c_void_p),
        ("n_mat", ctypes.c_int64),
    ]


class CBatchItem(ctypes.Structure):
    """qfvm_batch_item (include/qfvm.h)."""

    _fields_ = [
        ("prog", ctypes.POINTER(CProgram)),
        ("n_qubits", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("grad_ops", ctypes.c_void_p),
        ("grad_gen", ctypes.c_void_p),
        ("n_grads", ctypes.c_int64),
        ("chars", ctypes.c_void_p),
        ("pos", ctypes.c_void_p),
        ("off", ctypes.c_void_p),
        ("len", ctypes.c_void_p),
        ("coef", ctypes.c_void_p),
        ("n_terms", ctypes.c_int64),
        ("grads", ctypes.c_void_p),
        ("expectation", ctypes.c_void_p),
        ("seconds", ctypes.c_void_p),
    ]


_lib = None
_lib_lock = threading.Lock()

# (name, restype, argtypes) for every symbol include/qfvm.h declares
_vp, _i32, _i64, _u64, _dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
SIGNATURES = {
    "qfvm_last_error": (ctypes.c_char_p, []),
    "qfvm_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "qfvm_create": (ctypes.c_int, [_i32, _i32, ctypes.POINTER(_vp)]),
    "qfvm_destroy": (ctypes.c_int, [_vp]),
    "qfvm_num_qubits": (ctypes.c_int, [_vp, ctypes.POINTER(_i32)]),
    "qfvm_dtype_of": (ctypes.c_int, [_vp, ctypes.POINTER(_i32)]),
    "qfvm_init_zero": (ctypes.c_int, [_vp]),
    "qfvm_init_basis": (ctypes.c_int, [_vp, _u64]),
    "qfvm_load": (ctypes.c_int, [_vp, _vp, _u64]),
    "qfvm_store": (ctypes.c_int, [_vp, _vp, _u64]),
    "qfvm_apply": (ctypes.c_int, [_vp, ctypes.POINTER(CProgram)]),
    "qfvm_run": (ctypes.c_int, [_vp, ctypes.POINTER(CProgram), _vp, _i64, _vp, _i32, _i32]),
    "qfvm_expect_paulis": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "qfvm_expect_paulis_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "qfvm_expect_hamiltonian": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "qfvm_adjoint": (ctypes.c_int, [_vp, ctypes.POINTER(CProgram), _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _i64,
                                    _vp, _vp]),
    "qfvm_gather": (ctypes.c_int, [_vp, _vp, _i64, _vp]),
    "qfvm_inner": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(_dbl), ctypes.POINTER(_dbl)]),
    "qfvm_probabilities": (ctypes.c_int, [_vp, _vp, _u64]),
    "qfvm_marginals": (ctypes.c_int, [_vp, _vp, _i32, _vp]),
    "qfvm_sample": (ctypes.c_int, [_vp, _vp, _i64, _vp]),
    "qfvm_norm2": (ctypes.c_int, [_vp, ctypes.POINTER(_dbl)]),
    "qfvm_synchronize": (ctypes.c_int, [_vp]),
    "qfvm_set_profiling": (ctypes.c_int, [_vp, _i32]),
    "qfvm_profile_json": (ctypes.c_int, [_vp, ctypes.c_char_p, _i64, ctypes.POINTER(_i64), _i32]),
    "qfvm_set_option": (ctypes.c_int, [_vp, ctypes.c_char_p, _i64]),
    "qfvm_batch_create": (ctypes.c_int, [_i32, _i32, ctypes.POINTER(_vp)]),
    "qfvm_batch_destroy": (ctypes.c_int, [_vp]),
    "qfvm_batch_set_option": (ctypes.c_int, [_vp, _i32, ctypes.c_char_p, _i64]),
    "qfvm_batch_run": (ctypes.c_int, [_vp, ctypes.POINTER(CBatchItem), _i64]),
    "qfvm_plan_json": (ctypes.c_int, [_i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32, ctypes.POINTER(CProgram), ctypes.c_char_p, _i64,
                                      ctypes.POINTER(_i64)]),
}


def plan_json(prog: Program, n: int, dtype: str = "c128", tile_bits: int = 0, low_bits: int = 0,
              max_groups: int = 0, pass_cost: int = 0, layout: bool = False, dense_blocks: int = 0,
              dense_min_gates: int = 0) -> dict:
    """The fused execution plan the engine would run for `prog` (host only; no GPU needed);
    `layout`, `dense_blocks` and `dense_min_gates` plan as the options of those names do (with
    `dense_blocks` = 0, the default here, no run is collapsed)."""
    import json  # pylint: disable=import-outside-toplevel

    cp, keep = as_cprogram(prog)
    ln = ctypes.c_int64(0)
    args = (int(n), _DTYPES[dtype], int(tile_bits), int(low_bits), int(max_groups), int(pass_cost), int(bool(layout)),
            int(dense_blocks), int(dense_min_gates), ctypes.byref(cp))
    check(lib().qfvm_plan_json(*args, None, 0, ctypes.byref(ln)))
    buf = ctypes.create_string_buffer(ln.value + 1)
    check(lib().qfvm_plan_json(*args, buf, ln.value + 1, ctypes.byref(ln)))
    del keep
    return json.loads(buf.value.decode())


def lib():
    """Load the engine library (once). Raises if it was not built."""
    global _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"HIP engine not built: {LIB_PATH} is missing (run __graft_entry__.build() or "
                    "python -m pyquafu_amd.build)"
                )
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        msg = lib().qfvm_last_error().decode(errors="replace")
        raise EngineError(rc, msg)


def device_count() -> int:
    n = ctypes.c_int(0)
    check(lib().qfvm_device_count(ctypes.byref(n)))
    return n.value


def as_cprogram(prog: Program):
    """A qfvm_program view of `prog`; keep the returned tuple alive while it is used."""
    ops = np.ascontiguousarray(prog.ops, dtype=OP_DTYPE)
    idx = np.ascontiguousarray(prog.idx, dtype=np.int32)
    mat = np.ascontiguousarray(prog.mat, dtype=np.float64)
    cp = CProgram(
        ops.ctypes.data if len(ops) else None, len(ops),
        idx.ctypes.data if len(idx) else None, len(idx),
        mat.ctypes.data if len(mat) else None, len(mat),
    )
    return cp, (ops, idx, mat)


def pauli_arrays(terms):
    """Flatten [(paulistr, positions), ...] for qfvm_expect_paulis."""
    chars, pos, off, ln = [], [], [], []
    for s, p in terms:
        if len(s) != len(p):
            raise ValueError("pauli string and position list differ in length")
        off.append(len(pos))
        ln.append(len(p))
        chars.extend(s.encode())
        pos.extend(int(q) for q in p)
    return (
        np.array(chars + [0], dtype=np.uint8),
        np.array(pos + [0], dtype=np.int32),
        np.array(off, dtype=np.int64),
        np.array(ln, dtype=np.int32),
    )


def _coeff_array(coeffs, n: int) -> np.ndarray:
    cf = np.asarray(coeffs, dtype=np.complex128).reshape(-1)
    if cf.size != n:
        raise ValueError("one coefficient per Pauli term")
    out = np.zeros(2 * max(n, 1), dtype=np.float64)
    out[0: 2 * n: 2] = cf.real
    out[1: 2 * n: 2] = cf.imag
    return out


class AdjointCall:
    """The C arguments of one qfvm_adjoint call, built once so that repeated calls (batches,
    training loops) skip the conversion."""

    def __init__(self, prog: Program, grad_ops, grad_gens: str, terms, coeffs):
        self.cp, self._keep = as_cprogram(prog)
        self._ops = np.ascontiguousarray(grad_ops, dtype=np.int64)
        if len(grad_gens) != self._ops.size:
            raise ValueError("one generator letter per differentiated op")
        self._gens = np.frombuffer(grad_gens.encode() + b"\0", dtype=np.uint8)
        self._terms = pauli_arrays(terms)
        c, p, o, l_ = self._terms
        self._cf = _coeff_array(coeffs, len(o))
        self.n_grads = int(self._ops.size)
        self.args = (self._ops.ctypes.data, self._gens.ctypes.data, self.n_grads, c.ctypes.data, p.ctypes.data,
                     o.ctypes.data, l_.ctypes.data, self._cf.ctypes.data, len(o))


class DeviceState:
    """One statevector resident in HBM, owned by the engine."""

    def __init__(self, n_qubits: int, dtype="c128"):
        if dtype not in _DTYPES:
            raise ValueError(f"unknown dtype {dtype!r}")
        h = ctypes.c_void_p()
        check(lib().qfvm_create(int(n_qubits), _DTYPES[dtype], ctypes.byref(h)))
        self._h = h
        self.n = int(n_qubits)
        self.dtype = "c64" if _DTYPES[dtype] == C64 else "c128"

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().qfvm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # pylint: disable=broad-except
            pass

    def init_zero(self) -> None:
        check(lib().qfvm_init_zero(self._h))

    def init_basis(self, index: int) -> None:
        check(lib().qfvm_init_basis(self._h, int(index)))

    def load(self, psi: np.ndarray) -> None:
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        check(lib().qfvm_load(self._h, psi.ctypes.data, psi.size))

    def store(self, out: np.ndarray = None) -> np.ndarray:
        if out is None:
            out = np.empty(1 << self.n, dtype=np.complex128)
        if out.dtype != np.complex128 or not out.flags.c_contiguous or out.size != (1 << self.n):
            raise ValueError("output must be a contiguous complex128 array of length 2^n")
        check(lib().qfvm_store(self._h, out.ctypes.data, out.size))
        return out

    def apply(self, prog: Program) -> None:
        cp, keep = as_cprogram(prog)
        check(lib().qfvm_apply(self._h, ctypes.byref(cp)))
        del keep

    def run(self, prog: Program, uniforms, n_creg: int, skip_measure: bool = False) -> np.ndarray:
        cp, keep = as_cprogram(prog)
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        creg = np.zeros(max(n_creg, 1), dtype=np.int32)
        check(lib().qfvm_run(self._h, ctypes.byref(cp), u.ctypes.data if u.size else None, u.size,
                             creg.ctypes.data, int(n_creg), int(skip_measure)))
        del keep
        return creg[:n_creg]

    def expect(self, terms) -> np.ndarray:
        c, p, o, l_ = pauli_arrays(terms)
        out = np.zeros(len(o), dtype=np.float64)
        if len(o):
            check(lib().qfvm_expect_paulis(self._h, c.ctypes.data, p.ctypes.data, o.ctypes.data, l_.ctypes.data,
                                           len(o), out.ctypes.data))
        return out

    def expect_hamiltonian(self, terms, coeffs) -> float:
        """<psi|H|psi> for H = sum_t coeffs[t] * P_t, terms as for `expect`."""
        c, p, o, l_ = pauli_arrays(terms)
        cf = _coeff_array(coeffs, len(o))
        out = ctypes.c_double(0.0)
        check(lib().qfvm_expect_hamiltonian(self._h, c.ctypes.data, p.ctypes.data, o.ctypes.data, l_.ctypes.data,
                                            cf.ctypes.data, len(o), ctypes.byref(out)))
        return out.value

    def adjoint(self, prog: Program, grad_ops, grad_gens: str, terms, coeffs):
        """Reverse-mode gradient of <H> after `prog` (qfvm_adjoint). Returns (grads, expectation)."""
        return self.adjoint_prepared(AdjointCall(prog, grad_ops, grad_gens, terms, coeffs))

    def adjoint_prepared(self, call: "AdjointCall"):
        """qfvm_adjoint with arguments converted once (see AdjointCall)."""
        grads = np.zeros(max(call.n_grads, 1), dtype=np.float64)
        e = ctypes.c_double(0.0)
        check(lib().qfvm_adjoint(self._h, ctypes.byref(call.cp), *call.args, grads.ctypes.data, ctypes.byref(e)))
        return grads[: call.n_grads], e.value

    def gather(self, indices) -> np.ndarray:
        """Amplitudes at `indices` (complex128)."""
        idx = np.ascontiguousarray(indices, dtype=np.uint64)
        out = np.empty(idx.size, dtype=np.complex128)
        if idx.size:
            check(lib().qfvm_gather(self._h, idx.ctypes.data, idx.size, out.ctypes.data))
        return out

    def inner(self, other: "DeviceState") -> complex:
        """<self|other> on the device."""
        re, im = ctypes.c_double(0.0), ctypes.c_double(0.0)
        check(lib().qfvm_inner(self._h, other._h, ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    def probabilities(self) -> np.ndarray:
        out = np.empty(1 << self.n, dtype=np.float64)
        check(lib().qfvm_probabilities(self._h, out.ctypes.data, out.size))
        return out

    def marginals(self, qubits) -> np.ndarray:
        """out[m] = sum of |amp_i|^2 over i with bit qubits[j] of i equal to bit j of m."""
        q = np.ascontiguousarray(qubits, dtype=np.int32)
        out = np.empty(1 << q.size, dtype=np.float64)
        check(lib().qfvm_marginals(self._h, q.ctypes.data if q.size else None, q.size, out.ctypes.data))
        return out

    def sample(self, uniforms) -> np.ndarray:
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        out = np.empty(u.size, dtype=np.uint64)
        if u.size:
            check(lib().qfvm_sample(self._h, u.ctypes.data, u.size, out.ctypes.data))
        return out

    def norm2(self) -> float:
        v = ctypes.c_double()
        check(lib().qfvm_norm2(self._h, ctypes.byref(v)))
        return v.value

    def synchronize(self) -> None:
        check(lib().qfvm_synchronize(self._h))

    def set_option(self, key: str, value: int) -> None:
        check(lib().qfvm_set_option(self._h, key.encode(), int(value)))

    def set_profiling(self, enable: bool) -> None:
        check(lib().qfvm_set_profiling(self._h, int(bool(enable))))

    def profile(self, reset: bool = True) -> dict:
        """Per-kernel {launches, ms, bytes} since the last reset (ms needs profiling enabled)."""
        import json  # pylint: disable=import-outside-toplevel

        n = ctypes.c_int64(0)
        check(lib().qfvm_profile_json(self._h, None, 0, ctypes.byref(n), 0))
        buf = ctypes.create_string_buffer(n.value + 1)
        check(lib().qfvm_profile_json(self._h, buf, n.value + 1, ctypes.byref(n), int(bool(reset))))
        return json.loads(buf.value.decode())


class StatePool:
    """Reuses one device state per (n, dtype) so repeated simulate() calls do not reallocate."""

    def __init__(self):
        self._states = {}
        self._lock = threading.Lock()

    def get(self, n: int, dtype: str = "c128") -> DeviceState:
        key = (int(n), "c64" if _DTYPES[dtype] == C64 else "c128")
        with self._lock:
            st = self._states.get(key)
            if st is None:
                # keep at most one large state around
                if key[0] >= 26:
                    for k in [k for k in self._states if k[0] >= 26]:
                        self._states.pop(k).close()
                st = DeviceState(*key)
                self._states[key] = st
            return st

    def clear(self) -> None:
        with self._lock:
            for st in self._states.values():
                st.close()
            self._states.clear()


pool = StatePool()


class PreparedBatch:
    """The qfvm_batch_item array of a list of (n_qubits, AdjointCall), with output buffers. Built
    once; rebinding parameters in the calls' programs in place (ParametricProgram.bind) needs no
    rebuild."""

    def __init__(self, calls):
        self.calls = list(calls)
        self.n = len(self.calls)
        sizes = [c.n_grads for _, c in self.calls]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.grads = np.zeros(max(1, int(self.offsets[-1])), dtype=np.float64)
        self.expectations = np.zeros(max(1, self.n), dtype=np.float64)
        self.seconds = np.zeros(max(1, self.n), dtype=np.float64)
        self.items = (CBatchItem * max(1, self.n))()
        g0, e0, t0 = self.grads.ctypes.data, self.expectations.ctypes.data, self.seconds.ctypes.data
        for k, (nq, c) in enumerate(self.calls):
            ops, gens, n_grads, ch, pos, off, ln, cf, n_terms = c.args
            it = self.items[k]
            it.prog = ctypes.pointer(c.cp)
            it.n_qubits = int(nq)
            it.grad_ops, it.grad_gen, it.n_grads = ops, gens, n_grads
            it.chars, it.pos, it.off, it.len, it.coef, it.n_terms = ch, pos, off, ln, cf, n_terms
            it.grads = g0 + 8 * int(self.offsets[k])
            it.expectation = e0 + 8 * k
            it.seconds = t0 + 8 * k

    def grads_of(self, k: int) -> np.ndarray:
        return self.grads[self.offsets[k]: self.offsets[k + 1]]


class NativeBatch:
    """qfvm_batch: worker threads inside the engine evaluate many circuits' <H> and adjoint
    gradients per call, with their handles (and plan / graph caches) kept between calls."""

    def __init__(self, dtype: str = "c64", workers: int = 4):
        h = ctypes.c_void_p()
        check(lib().qfvm_batch_create(_DTYPES[dtype], int(workers), ctypes.byref(h)))
        self._h = h
        self.dtype = "c64" if _DTYPES[dtype] == C64 else "c128"
        self.workers = int(workers)

    def set_option(self, key: str, value: int, n_qubits: int = 0) -> None:
        check(lib().qfvm_batch_set_option(self._h, int(n_qubits), key.encode(), int(value)))

    def run(self, prepared: PreparedBatch) -> None:
        """Evaluates every item; results land in prepared.grads / expectations / seconds."""
        if prepared.n:
            check(lib().qfvm_batch_run(self._h, prepared.items, prepared.n))

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().qfvm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # pylint: disable=broad-except
            pass
