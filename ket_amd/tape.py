"""Packing of ket FlatOps into the ksv_op tape of include/ketsv.h.

A FlatOp (src/circuit.ts:47) is one of the op tuples built by Circuit:
  ('single', q, gate2, meta)            ('cnot', control, target)
  ('swap', a, b)                        ('two', a, b, gate4, meta)
  ('controlled', control, target, gate2, meta)
  ('toffoli', c1, c2, target)           ('cswap', control, a, b)
  ('csrswap', control, a, b)            ('unitary', qubits, matrix)
  ('barrier', qubits)
Each becomes one tape record (barriers included, so tape index == flat index);
matrices go to a parallel float64 array of interleaved (re, im).
"""
import numpy as np

from ._native import (KSV_BARRIER, KSV_CNOT, KSV_CONTROLLED, KSV_CSRSWAP, KSV_CSWAP, KSV_MAX_OPERANDS, KSV_SINGLE,
                      KSV_SWAP, KSV_TOFFOLI, KSV_TWO, KSV_UNITARY, OP_DTYPE)

_KIND = {"single": KSV_SINGLE, "cnot": KSV_CNOT, "swap": KSV_SWAP, "two": KSV_TWO, "controlled": KSV_CONTROLLED,
         "toffoli": KSV_TOFFOLI, "cswap": KSV_CSWAP, "csrswap": KSV_CSRSWAP, "unitary": KSV_UNITARY,
         "barrier": KSV_BARRIER}


def _flat_matrix(m):
    """Row-major complex entries (a tuple of complex, or rows of complex) -> list of floats."""
    out = []
    if m and isinstance(m[0], (list, tuple)):
        for row in m:
            for v in row:
                v = complex(v)
                out.append(v.real)
                out.append(v.imag)
    else:
        for v in m:
            v = complex(v)
            out.append(v.real)
            out.append(v.imag)
    return out


def pack(flat_ops):
    """flat ops -> (ops: structured array of OP_DTYPE, mats: float64 array).

    Matrices get their offsets in op order; gates of 4 and of 16 entries
    (the 2x2 and 4x4 gate tuples) are converted to doubles in bulk."""
    n = len(flat_ops)
    ops = np.zeros(n, dtype=OP_DTYPE)
    kinds = ops["kind"]
    nqs = ops["nq"]
    qs = ops["q"]
    matoff = ops["mat"]
    total = 0
    by_size = {4: ([], []), 16: ([], [])}   # entries -> (offsets, matrices)
    other = []                                # (offset, matrix) of any other shape
    cache = {}
    for i, op in enumerate(flat_ops):
        k = op[0]
        if k == "single" or k == "controlled" or k == "two":
            if k == "single":
                q = (op[1],)
                g = op[2]
            else:
                q = (op[1], op[2])
                g = op[3]
            key = id(g)
            hit = cache.get(key)
            if hit is None or hit[0] is not g:
                off = total
                flat = isinstance(g, tuple) and len(g) in by_size and not isinstance(g[0], (list, tuple))
                if flat:
                    by_size[len(g)][0].append(off)
                    by_size[len(g)][1].append(g)
                    total += 2 * len(g)
                else:
                    m = _flat_matrix(g)
                    other.append((off, m))
                    total += len(m)
                cache[key] = (g, off)
            else:
                off = hit[1]
            matoff[i] = off
        elif k == "unitary":
            q = tuple(op[1])
            if len(q) > KSV_MAX_OPERANDS:
                raise TypeError(f"unitary on {len(q)} qubits exceeds the {KSV_MAX_OPERANDS}-qubit limit")
            matoff[i] = total
            m = _flat_matrix(op[2])
            other.append((total, m))
            total += len(m)
        elif k == "barrier":
            q = ()
            matoff[i] = -1
        elif k in ("cnot", "swap"):
            q = (op[1], op[2])
            matoff[i] = -1
        elif k in ("toffoli", "cswap", "csrswap"):
            q = (op[1], op[2], op[3])
            matoff[i] = -1
        else:
            raise TypeError(f"op kind '{k}' cannot be packed into a pure tape")
        kinds[i] = _KIND[k]
        nqs[i] = len(q)
        for j, v in enumerate(q):
            qs[i, j] = v
    mats = np.empty(total, dtype=np.float64)
    for size, (offs, gs) in by_size.items():
        if offs:
            vals = np.array(gs, dtype=np.complex128).view(np.float64)          # rows of 2 * size doubles
            mats[np.asarray(offs, dtype=np.int64)[:, None] + np.arange(2 * size)[None, :]] = vals
    for off, m in other:
        mats[off:off + len(m)] = m
    return ops, mats
