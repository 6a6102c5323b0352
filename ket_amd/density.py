"""Density-matrix simulation: `Circuit.dm()` (src/circuit.ts:4549-4579) and
the `DensityMatrix` read-outs (src/density.ts:274-380).

rho lives on the device as a 2n-qubit state whose index is (row << n) | col,
the key layout of ket's sparse map (src/density.ts:5-16), so every channel
is an ordinary tape op on the library's pass kernels:
- a gate U on qubits qs is U (x) conj(U) on (qs + n, qs) as one op, so ket's
  prune (|v|^2 < 1e-14) follows the whole U rho U^dagger, as dmSet does;
  permutations (cnot, swap, toffoli, cswap) go to the row and then the
  column qubits; unitaries wider than 3 qubits likewise as two ops;
- depolarize1 (src/density.ts:193-224) is the real 4x4 map on (q + n, q)
  with (1 - 2p/3, 2p/3) between |00> and |11> and 1 - 4p/3 on |01>, |10>;
- depolarize2 (:239-272) is (1 - p) I + p/15 sum_P P (x) conj(P) on
  (a + n, b + n, a, b);
- csrswap keeps ket's block-wise form (:521-570): with V = sqrt(iSWAP),
  V (x) conj(V) where the row control is 1, conj(V) (x) V where only the
  column control is 1, identity where both are 0.
ket's depolarize2 accumulates with dmAcc, which drops a partial sum that
falls below the prune threshold; here the sum is formed whole and pruned
once (they differ only for entries of magnitude ~1e-7).
"""
import math

import numpy as np

from . import gates as G
from ._native import State
from .tape import pack

# src/density.ts:453-457
DM_DEVICE_NOISE = {"aria-1": {"p1": 0.0003, "p2": 0.005}, "forte-1": {"p1": 0.0001, "p2": 0.002},
                   "harmony": {"p1": 0.001, "p2": 0.015}}

_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
_I2 = np.eye(2, dtype=np.complex128)
_SRISW = np.array([[1, 0, 0, 0], [0, 2 ** -0.5, 1j * 2 ** -0.5, 0], [0, 1j * 2 ** -0.5, 2 ** -0.5, 0], [0, 0, 0, 1]])


def _mat(g, dim):
    a = np.array(g, dtype=np.complex128)
    return a.reshape(dim, dim)


def _depolarize1_map(p):
    sa, sb, cf = 1 - 2 * p / 3, 2 * p / 3, 1 - 4 * p / 3
    return np.array([[sa, 0, 0, sb], [0, cf, 0, 0], [0, 0, cf, 0], [sb, 0, 0, sa]], dtype=np.complex128)


def _depolarize2_map(p):
    paulis = [_I2, _X, _Y, _Z]
    m = (1 - p) * np.eye(16, dtype=np.complex128)
    for i in range(4):
        for j in range(4):
            if i == 0 and j == 0:
                continue
            P = np.kron(paulis[i], paulis[j])
            m = m + (p / 15) * np.kron(P, P.conj())
    return m


def _csrswap_map():
    """64x64 on (c_row, a_row, b_row, c_col, a_col, b_col), MSB first."""
    V = _SRISW
    blocks = {(1, 0): np.kron(V, V.conj()), (1, 1): np.kron(V, V.conj()), (0, 1): np.kron(V.conj(), V),
              (0, 0): np.eye(16, dtype=np.complex128)}
    m = np.zeros((64, 64), dtype=np.complex128)
    for cr in (0, 1):
        for cc in (0, 1):
            blk = blocks[(cr, cc)]
            for ro in range(16):            # (a_row, b_row, a_col, b_col) local index of the 16x16 block
                for co in range(16):
                    if blk[ro, co] == 0:
                        continue
                    ar, br, ac, bc = (ro >> 3) & 1, (ro >> 2) & 1, (ro >> 1) & 1, ro & 1
                    ar2, br2, ac2, bc2 = (co >> 3) & 1, (co >> 2) & 1, (co >> 1) & 1, co & 1
                    R = (cr << 5) | (ar << 4) | (br << 3) | (cc << 2) | (ac << 1) | bc
                    C = (cr << 5) | (ar2 << 4) | (br2 << 3) | (cc << 2) | (ac2 << 1) | bc2
                    m[R, C] = blk[ro, co]
    return m


def dm_ops(n, flat, p1=0.0, p2=0.0):
    """ket FlatOps -> flat ops on the 2n-qubit vectorised rho."""
    out = []

    def uni(qs, m):
        out.append(("unitary", tuple(qs), [list(r) for r in m]))

    def fused(qs, m):
        uni([q + n for q in qs] + list(qs), np.kron(m, m.conj()))

    for op in flat:
        k = op[0]
        if k == "single":
            fused([op[1]], _mat(op[2], 2))
            if p1:
                uni([op[1] + n, op[1]], _depolarize1_map(p1))
        elif k in ("cnot", "swap"):
            out.append((k, op[1] + n, op[2] + n))
            out.append((k, op[1], op[2]))
            if p2:
                uni([op[1] + n, op[2] + n, op[1], op[2]], _depolarize2_map(p2))
        elif k == "two":
            fused([op[1], op[2]], _mat(op[3], 4))
            if p2:
                uni([op[1] + n, op[2] + n, op[1], op[2]], _depolarize2_map(p2))
        elif k == "controlled":
            g = _mat(op[3], 2)
            cg = np.eye(4, dtype=np.complex128)
            cg[2:, 2:] = g
            fused([op[1], op[2]], cg)
            if p2:
                uni([op[1] + n, op[2] + n, op[1], op[2]], _depolarize2_map(p2))
        elif k in ("toffoli", "cswap"):
            out.append((k, op[1] + n, op[2] + n, op[3] + n))
            out.append((k, op[1], op[2], op[3]))
        elif k == "csrswap":
            c, a, b = op[1], op[2], op[3]
            uni([c + n, a + n, b + n, c, a, b], _csrswap_map())
        elif k == "unitary":
            qs = list(op[1])
            m = _mat(op[2], 1 << len(qs))
            if len(qs) <= 3:
                fused(qs, m)
            else:
                uni([q + n for q in qs], m)
                uni(qs, m.conj())
        # barrier: nothing
    return out


def _bitstring(idx, n):
    return format(idx, "b").zfill(n)[::-1] if n else ""


class DensityMatrix:
    """src/density.ts:274-380.  Holds the dense rho read back from the device."""

    def __init__(self, qubits, rho):
        self.qubits = qubits
        self._rho = rho              # [row, col], index bit q = qubit q, pruned entries 0

    def get(self, row, col):
        return complex(self._rho[int(row), int(col)])

    def probabilities(self):
        d = self._rho.diagonal().real
        return {_bitstring(int(r), self.qubits): float(d[r]) for r in np.nonzero(d > 1e-14)[0]}

    def purity(self):
        a = self._rho
        return float(np.sum(a.real * a.real + a.imag * a.imag))

    def entropy(self):
        dim = 1 << self.qubits
        if dim > 4096:
            from .circuit import RangeError
            raise RangeError(f"entropy(): circuit too large (n={self.qubits}, dim={dim})")
        lam = np.linalg.eigvalsh(self._rho)
        s = 0.0
        for v in lam:
            if v > 1e-14:
                s -= v * math.log2(v)
        return float(s)

    def blochAngles(self, q):
        n = self.qubits
        idx = np.arange(1 << n)
        d = self._rho.diagonal().real
        bit = (idx >> q) & 1
        rho00 = float(d[bit == 0].sum())
        rho11 = float(d[bit == 1].sum())
        lo = idx[bit == 0]
        off = self._rho[lo, lo | (1 << q)]
        rz, rx, ry = rho00 - rho11, 2 * float(off.real.sum()), -2 * float(off.imag.sum())
        return {"theta": math.acos(max(-1.0, min(1.0, rz))), "phi": math.atan2(ry, rx)}


def run_dm(n, flat, noise=None, precision="f64", prune=True):
    """runDM (src/density.ts:475-588) on the device."""
    p1 = (noise or {}).get("p1", 0) or 0
    p2 = (noise or {}).get("p2", 0) or 0
    ops = dm_ops(n, flat, p1, p2)
    st = State(2 * n, precision, prune=prune)
    try:
        st.init_basis(0)
        if ops:
            st.apply(pack(ops))
        rho = st.amplitudes().reshape(1 << n, 1 << n).copy()
    finally:
        st.free()
    p = rho.real ** 2 + rho.imag ** 2
    rho[p < 1e-14] = 0
    return DensityMatrix(n, rho)
