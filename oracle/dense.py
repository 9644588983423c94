"""ORACLE -- test infrastructure, not product code.

An independent NumPy statevector simulator for small n (<= ~14), used to pin the C++ oracle
(oracle/qfvm_cpu.cpp) and the host-side program compiler. It shares no index arithmetic with
either: gates are applied by tensor contraction (np.tensordot on the qubit axes of a
[2]*n-shaped state), controls by slicing, and the named gates use the textbook matrices of
quafu/elements/matrices/mat_lib.py:18-197, restated here.

Input: a program with `.ops` (qfvm_op records), `.idx` (int32 pool), `.mat` (float64 pool), as
described in include/qfvm.h. Little-endian: qubit q is bit q of the flat index
(doc/source/tutorial.rst:333-334).
"""

from __future__ import annotations

import numpy as np

# op codes of include/qfvm.h
GENERAL, MULTI, X, Y, Z, H, S, SDG, T, TDG, P, RX, RY, RZ, CX, CZ, CP, CCX, RZZ, MEASURE, RESET, CIF, KRAUS, UCHAN = range(24)

I2 = np.eye(2, dtype=complex)
XM = np.array([[0, 1], [1, 0]], dtype=complex)
YM = np.array([[0, -1j], [1j, 0]], dtype=complex)
ZM = np.array([[1, 0], [0, -1]], dtype=complex)
HM = np.array([[1, 1], [1, -1]], dtype=complex) / np.sqrt(2)
SM = np.diag([1, 1j]).astype(complex)
TM = np.diag([1, np.exp(1j * np.pi / 4)]).astype(complex)


def rx(t):
    c, s = np.cos(t / 2), np.sin(t / 2)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=complex)


def ry(t):
    c, s = np.cos(t / 2), np.sin(t / 2)
    return np.array([[c, -s], [s, c]], dtype=complex)


def rz(t):
    return np.diag([np.exp(-0.5j * t), np.exp(0.5j * t)]).astype(complex)


def pm(t):
    return np.diag([1, np.exp(1j * t)]).astype(complex)


def u3(theta, phi, lam):
    """OpenQASM 3 U (mat_lib.py:186-197)."""
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return np.array(
        [[c, -np.exp(1j * lam) * s], [np.exp(1j * phi) * s, np.exp(1j * (phi + lam)) * c]], dtype=complex
    )


def rzz_lsb(t):
    """exp(-i t/2 Z(x)Z) on (q0, q1); a diagonal, so its endianness does not matter."""
    d = [np.exp(-0.5j * t), np.exp(0.5j * t), np.exp(0.5j * t), np.exp(-0.5j * t)]
    return np.diag(d).astype(complex)


def apply_unitary(psi: np.ndarray, n: int, ctrls, targs, U: np.ndarray) -> np.ndarray:
    """psi -> (|1..1><1..1|_ctrls (x) U_targs + rest (x) I) psi.

    U's row/column index bit j belongs to targs[j] (the engine's convention,
    statevector.hpp:954-963)."""
    k = len(targs)
    t = np.array(psi, dtype=complex).reshape([2] * n)  # axis a <-> qubit n-1-a

    def ax(q):
        return n - 1 - q

    sel = [slice(None)] * n
    for c in ctrls:
        sel[ax(c)] = 1
    sel = tuple(sel)
    sub = t[sel]
    cax = {ax(c) for c in ctrls}
    remaining = [a for a in range(n) if a not in cax]
    tax = [remaining.index(ax(q)) for q in targs]
    Ut = U.reshape([2] * (2 * k))  # axes: r_{k-1} .. r_0, c_{k-1} .. c_0
    res = np.tensordot(Ut, sub, axes=([2 * k - 1 - j for j in range(k)], tax))
    res = np.moveaxis(res, [k - 1 - j for j in range(k)], tax)
    t[sel] = res
    return t.reshape(-1)


def marginal(psi: np.ndarray, n: int, qbits) -> np.ndarray:
    """probs[m], bit b of m <-> qbits[b]."""
    p = (np.abs(psi) ** 2).reshape([2] * n)
    order = [n - 1 - q for q in reversed(qbits)]
    rest = [a for a in range(n) if a not in order]
    p = np.transpose(p, order + rest).reshape(1 << len(qbits), -1)
    return p.sum(axis=1)


def discrete_pick(p: np.ndarray, u: float) -> int:
    """std::discrete_distribution as in libstdc++: first i with u <= cumsum(p/sum)[i], last := 1."""
    if len(p) < 2:
        return 0
    cp = np.cumsum(p / p.sum())
    cp[-1] = 1.0
    return int(np.searchsorted(cp, u, side="left"))


def project(psi: np.ndarray, n: int, qbits, final_state: int, meas_state: int, prob: float) -> np.ndarray:
    idx = np.arange(1 << n, dtype=np.int64)
    m = np.zeros_like(idx)
    for b, q in enumerate(qbits):
        m |= ((idx >> q) & 1) << b
    out = np.where(m == meas_state, psi / np.sqrt(prob), 0.0).astype(complex)
    if final_state != meas_state:
        fm = sum(((final_state >> b) & 1) << q for b, q in enumerate(qbits))
        mm = sum(((meas_state >> b) & 1) << q for b, q in enumerate(qbits))
        qm = sum(1 << q for q in qbits)
        src = (idx & ~qm) | mm
        dst = (idx & ~qm) | fm
        sel = (idx & qm) == mm
        res = out.copy()
        res[dst[sel]] = out[src[sel]]
        res[src[sel]] = out[dst[sel]]
        out = res
    return out


class DenseRunner:
    def __init__(self, prog, n: int, psi=None, uniforms=(), n_creg: int = 0, skip_measure=False):
        self.p = prog
        self.n = n
        if psi is None:
            psi = np.zeros(1 << n, dtype=complex)
            psi[0] = 1.0
        self.psi = np.array(psi, dtype=complex)
        self.u = list(uniforms)
        self.ui = 0
        self.creg = np.zeros(n_creg, dtype=np.int32)
        self.skip = skip_measure

    def _q(self, op):
        return [int(x) for x in self.p.idx[op["q_off"]: op["q_off"] + op["n_qubits"]]]

    def _c(self, op):
        return [int(x) for x in self.p.idx[op["c_off"]: op["c_off"] + op["n_cbits"]]]

    def _m(self, op, dim):
        a = self.p.mat[op["m_off"]: op["m_off"] + 2 * dim * dim]
        return (a[0::2] + 1j * a[1::2]).reshape(dim, dim)

    def gate_parts(self, op):
        """(controls, targets, target matrix) of one gate op."""
        q = self._q(op)
        c, th = int(op["code"]), float(op["param"])
        if c in (GENERAL, MULTI):
            nc = int(op["n_ctrl"])
            nt = len(q) - nc
            return q[:nc], q[nc:], self._m(op, 1 << nt)
        one = {X: XM, Y: YM, Z: ZM, H: HM, S: SM, SDG: SM.conj().T, T: TM, TDG: TM.conj().T}
        if c in one:
            return [], [q[0]], one[c]
        table = {P: ([], [0], pm), RX: ([], [0], rx), RY: ([], [0], ry), RZ: ([], [0], rz), CP: ([0], [1], pm)}
        if c in table:
            cs, ts, f = table[c]
            return [q[k] for k in cs], [q[k] for k in ts], f(th)
        fixed = {CX: ([0], [1], XM), CZ: ([0], [1], ZM), CCX: ([0, 1], [2], XM)}
        if c in fixed:
            cs, ts, m = fixed[c]
            return [q[k] for k in cs], [q[k] for k in ts], m
        if c == RZZ:
            return [], [q[0], q[1]], rzz_lsb(th)
        raise ValueError(f"not a gate: {c}")

    def gate(self, op):
        ctrls, targs, m = self.gate_parts(op)
        self.psi = apply_unitary(self.psi, self.n, ctrls, targs, m)

    def channel(self, op):
        """Noise channel on one qubit, with the next uniform u. Unitary channel: numpy's choice rule
        (cumulative probabilities over their total, first entry above u; noise.py:64-68). Kraus
        channel: SVSimulator._apply_op (simulator.py:59-74): K_k psi for k = 0, 1, ...; running sum
        of |K_k psi| / |psi|; the first k with u below it gives K_k psi / |K_k psi|, none leaves psi."""
        q = self._q(op)[0]
        nk = int(op["body"])
        a = self.p.mat[op["m_off"]: op["m_off"] + 8 * nk]
        mats = (a[0::2] + 1j * a[1::2]).reshape(nk, 2, 2)
        u = self.u[self.ui]
        self.ui += 1
        if int(op["code"]) == UCHAN:
            probs = self.p.mat[op["m_off"] + 8 * nk: op["m_off"] + 10 * nk: 2]
            cdf = np.cumsum(probs)
            cdf = cdf / cdf[-1]
            k = int(np.searchsorted(cdf, u, side="right"))
            self.psi = apply_unitary(self.psi, self.n, [], [q], mats[min(k, nk - 1)])
            return
        norm0 = np.linalg.norm(self.psi)
        s = 0.0
        for k in range(nk):
            t = apply_unitary(self.psi, self.n, [], [q], mats[k])
            norm1 = np.linalg.norm(t)
            s += norm1 / norm0
            if u < s:
                self.psi = t / norm1
                return

    def exec(self, begin: int, end: int) -> None:
        k = begin
        ops = self.p.ops
        while k < end:
            op = ops[k]
            c = int(op["code"])
            if c == MEASURE:
                k += 1
                if self.skip:
                    continue
                q, cb = self._q(op), self._c(op)
                pr = marginal(self.psi, self.n, q)
                m = discrete_pick(pr, self.u[self.ui])
                self.ui += 1
                self.psi = project(self.psi, self.n, q, m, m, pr[m])
                for j in range(len(q)):
                    self.creg[cb[j]] = (m >> j) & 1
            elif c == RESET:
                k += 1
                q = self._q(op)
                pr = marginal(self.psi, self.n, q)
                m = discrete_pick(pr, self.u[self.ui])
                self.ui += 1
                self.psi = project(self.psi, self.n, q, 0, m, pr[m])
            elif c in (KRAUS, UCHAN):
                k += 1
                self.channel(op)
            elif c == CIF:
                cb = self._c(op)
                out = 0
                for x in cb:
                    out = out * 2 + int(self.creg[x])
                end_b = k + 1 + int(op["body"])
                if out == int(op["cond"]):
                    self.exec(k + 1, end_b)
                k = end_b
            else:
                self.gate(op)
                k += 1


def run_dense(prog, n: int, psi=None, uniforms=(), n_creg: int = 0, skip_measure=False):
    r = DenseRunner(prog, n, psi, uniforms, n_creg, skip_measure)
    r.exec(0, len(prog.ops))
    return r.psi, r.creg


def pauli_matrix_le(paulistr: str, pos, n: int) -> np.ndarray:
    """Dense little-endian Pauli-string matrix (qubit q = bit q), built with np.kron from qubit n-1
    down to 0 (the convention PauliOp.get_matrix(big_endian=False) produces,
    quafu/algorithms/hamiltonian.py:64-89)."""
    ops = {"I": I2, "X": XM, "Y": YM, "Z": ZM}
    facs = [I2] * n
    for s, q in zip(paulistr, pos):
        facs[q] = facs[q] @ ops[s]
    m = np.array([[1.0 + 0j]])
    for q in reversed(range(n)):
        m = np.kron(m, facs[q])
    return m


def qft_closed_form(n: int, x: int) -> np.ndarray:
    """Amplitudes of the swap-free QFT (H then CP(2pi/2^k) from later qubits, for each target in
    textbook order) applied to |x>: amp(y) = 2^{-n/2} exp(2 pi i x rev(y) / 2^n), with integer phase
    arithmetic mod 2^n."""
    y = np.arange(1 << n, dtype=np.int64)
    rev = np.zeros_like(y)
    for b in range(n):
        rev |= ((y >> b) & 1) << (n - 1 - b)
    ph = (np.uint64(x) * rev.astype(np.uint64)) & np.uint64((1 << n) - 1)  # exact: wraps mod 2^64
    return np.exp(2j * np.pi * (ph.astype(np.float64) / (1 << n))) / np.sqrt(1 << n)
