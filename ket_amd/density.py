"""Density-matrix simulation: `Circuit.dm()` (src/circuit.ts:4549-4579) and
the `DensityMatrix` read-outs (src/density.ts:274-380).

rho lives on the device (ksv_dm_*): dense, entry key (row << n) | col as in
ket's sparse map (src/density.ts:5-16), with every entry's place in the
map's insertion order beside it.  Every channel is ket's arithmetic step by
step, so depolarize2's partial-sum pruning (dmAcc, src/density.ts:29-35),
which depends on that order, gives ket's values; the order also fixes the
summation order of purity() and blochAngles(), which ket sums in map order.
"""
import math

import numpy as np

from ._native import DensityState
from .tape import pack

# src/density.ts:453-457
DM_DEVICE_NOISE = {"aria-1": {"p1": 0.0003, "p2": 0.005}, "forte-1": {"p1": 0.0001, "p2": 0.002},
                   "harmony": {"p1": 0.001, "p2": 0.015}}


def _bitstring(idx, n):
    return format(idx, "b").zfill(n)[::-1] if n else ""


def _seq_sum(x):
    """x[0] + x[1] + ... left to right, every addition rounded (a JavaScript loop)."""
    return float(np.add.accumulate(x)[-1]) if len(x) else 0.0


class DensityMatrix:
    """src/density.ts:274-380.  Holds rho read back from the device and the keys
    of its present entries in ket's map order."""

    def __init__(self, qubits, rho, order):
        self.qubits = qubits
        self._rho = rho              # [row, col], index bit q = qubit q, absent entries 0
        self._order = order          # keys (row << n) | col, map order

    def get(self, row, col):
        return complex(self._rho[int(row), int(col)])

    def probabilities(self):
        d = self._rho.diagonal().real
        return {_bitstring(int(r), self.qubits): float(d[r]) for r in np.nonzero(d > 1e-14)[0]}

    def _entries(self):
        n = self.qubits
        k = self._order
        r, c = k >> np.uint64(n), k & np.uint64((1 << n) - 1)
        return r, c, self._rho.reshape(-1)[k.astype(np.int64)]

    def purity(self):
        """sum of |rho_rc|^2 over the entries in map order (src/density.ts:297-301)"""
        _, _, v = self._entries()
        return _seq_sum(v.real * v.real + v.imag * v.imag)

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
        """Reduced state of qubit q summed over the entries in map order
        (src/density.ts:352-370)."""
        r, c, v = self._entries()
        m = np.uint64(1 << q)
        diag = r == c
        low = (r & m) == 0
        rho00 = _seq_sum(v.real[diag & low])
        rho11 = _seq_sum(v.real[diag & ~low])
        off = ~diag & low & (c == (r | m))
        rz, rx, ry = rho00 - rho11, 2 * _seq_sum(v.real[off]), -2 * _seq_sum(v.imag[off])
        return {"theta": math.acos(max(-1.0, min(1.0, rz))), "phi": math.atan2(ry, rx)}


def run_dm(n, flat, noise=None):
    """runDM (src/density.ts:475-588) on the device."""
    p1 = (noise or {}).get("p1", 0) or 0
    p2 = (noise or {}).get("p2", 0) or 0
    d = DensityState(n)
    try:
        ops, mats = pack(flat)
        if len(ops):
            d.apply((ops, mats), p1, p2)
        rho, order = d.read()
    finally:
        d.free()
    return DensityMatrix(n, rho, order)
