"""ctypes binding of the CPU oracle (libketsv_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker — never by the ket_amd package.
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, c_char_p, c_double, c_int, c_size_t, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libketsv_oracle.so")
MAXQ = 12


class OrcGate(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 12), ("nq", ctypes.c_int32), ("q", ctypes.c_int32 * MAXQ),
                ("p", c_double * 3), ("mat", ctypes.c_int64)]


def build():
    subprocess.run(["make", "-B", "-s", "-C", HERE], check=True)


_lib = None


OMP_PATH = os.path.join(HERE, "libketsv_oracle_omp.so")
_omp = None


def lib_omp():
    """The same oracle with its apply loops split over all host cores (OpenMP);
    bit-identical results, used as the all-core dense CPU baseline."""
    global _omp
    if _omp is None:
        if not os.path.exists(OMP_PATH):
            build()
        _omp = _configure(ctypes.CDLL(OMP_PATH))
    return _omp


def apply_tape_omp(n, psi, ops, mats, prune=True):
    m = mats if len(mats) else np.zeros(1)
    rc = lib_omp().orc_apply(n, _p(psi), _p(ops), len(ops), _p(m), 1 if prune else 0)
    if rc:
        raise RuntimeError(f"oracle rejected the tape (code {rc})")
    return psi


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _configure(ctypes.CDLL(LIB_PATH))
    return _lib


def _configure(L):
    if True:
        L.orc_gate_matrix.argtypes = [c_char_p, POINTER(c_double), POINTER(c_double)]
        L.orc_gate_matrix.restype = c_int
        L.orc_apply.argtypes = [c_int, c_void_p, c_void_p, c_size_t, c_void_p, c_int]
        L.orc_apply.restype = c_int
        L.orc_apply_named.argtypes = [c_int, c_void_p, c_void_p, c_size_t, c_void_p, c_int]
        L.orc_apply_named.restype = c_int
        L.orc_init_basis.argtypes = [c_int, c_void_p, c_uint64]
        L.orc_probabilities.argtypes = [c_int, c_void_p, c_double, c_void_p, c_void_p]
        L.orc_probabilities.restype = c_size_t
        L.orc_prng.argtypes = [c_uint32, c_size_t, c_void_p]
        L.orc_sample.argtypes = [c_int, c_void_p, c_void_p, c_size_t, c_void_p]
        L.orc_sample.restype = c_int
        L.orc_sample_sv.argtypes = [c_int, c_void_p, c_double]
        L.orc_sample_sv.restype = c_uint64
        L.orc_marginals.argtypes = [c_int, c_void_p, c_void_p]
        L.orc_bloch.argtypes = [c_int, c_void_p, c_int, POINTER(c_double), POINTER(c_double)]
        L.orc_expectation.argtypes = [c_int, c_void_p, c_char_p]
        L.orc_expectation.restype = c_double
        L.orc_vqe.argtypes = [c_int, c_void_p, POINTER(c_char_p), c_void_p, c_size_t]
        L.orc_vqe.restype = c_double
        L.orc_expect_direct.argtypes = [c_int, c_void_p, c_uint64, c_uint64]
        L.orc_expect_direct.restype = c_double
        L.orc_collapse.argtypes = [c_int, c_void_p, c_int, c_double]
        L.orc_collapse.restype = c_int
        L.orc_gate_struct_size.restype = c_size_t
        L.orc_op_struct_size.restype = c_size_t
        assert L.orc_gate_struct_size() == ctypes.sizeof(OrcGate)
        assert L.orc_op_struct_size() == 64
    return L


def _p(a):
    return c_void_p(a.ctypes.data)


def gate_matrix(name, params=()):
    ps = (c_double * 3)(*(list(params) + [0.0] * (3 - len(params))))
    out = (c_double * 32)()
    dim = lib().orc_gate_matrix(name.encode(), ps, out)
    if dim < 0:
        raise KeyError(name)
    v = np.array(out[: 2 * dim * dim])
    return (v[0::2] + 1j * v[1::2]).reshape(dim, dim)


def zero_state(n, index=0):
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[index] = 1
    return psi


# method name -> number of qubit operands (the rest of the arguments are angles)
NQ = {}
for _m in ("id h x y z s si sdg t ti tdg v vi srn srndg rx ry rz vz r2 r4 r8 u1 p u2 u3 gpi gpi2").split():
    NQ[_m] = 1
for _m in ("cnot cx swap xx yy zz xy iswap srswap ms cy cz ch crx cry crz cr2 cr4 cr8 cu1 cu2 cu3 cs ct csdg ctdg "
           "csrn").split():
    NQ[_m] = 2
for _m in ("ccx cswap csrswap").split():
    NQ[_m] = 3


def apply_named(n, psi, calls, prune=True):
    """Apply a list of (method, args) in ket's Circuit-method form to psi (in place).

    ('unitary', (matrix, q0, q1, ...)) carries its matrix first.
    """
    gates = (OrcGate * max(len(calls), 1))()
    mats = []
    for i, (name, args) in enumerate(calls):
        g = gates[i]
        g.name = name.encode()
        g.mat = -1
        if name == "unitary":
            m = np.asarray(args[0], dtype=np.complex128)
            qs = args[1:]
            g.mat = len(mats)
            for v in m.reshape(-1):
                mats.extend([v.real, v.imag])
        elif name == "barrier":
            qs = args
        else:
            k = NQ[name]
            qs = args[len(args) - k:]
            for j, v in enumerate(args[: len(args) - k]):
                g.p[j] = float(v)
        g.nq = len(qs)
        for j, q in enumerate(qs):
            g.q[j] = int(q)
    m = np.asarray(mats if mats else [0.0], dtype=np.float64)
    rc = lib().orc_apply_named(n, _p(psi), ctypes.addressof(gates), len(calls), _p(m), 1 if prune else 0)
    if rc:
        raise RuntimeError(f"oracle rejected the circuit (code {rc})")
    return psi


def apply_tape(n, psi, ops, mats, prune=True):
    m = mats if len(mats) else np.zeros(1)
    rc = lib().orc_apply(n, _p(psi), _p(ops), len(ops), _p(m), 1 if prune else 0)
    if rc:
        raise RuntimeError(f"oracle rejected the tape (code {rc})")
    return psi


def probabilities(n, psi, thresh=1e-14):
    idx = np.empty(1 << n, dtype=np.uint64)
    p = np.empty(1 << n, dtype=np.float64)
    k = lib().orc_probabilities(n, _p(psi), thresh, _p(idx), _p(p))
    return idx[:k], p[:k]


def prng(seed, count):
    out = np.empty(count, dtype=np.float64)
    lib().orc_prng(seed & 0xFFFFFFFF, count, _p(out))
    return out


def sample(n, psi, r):
    r = np.ascontiguousarray(r, dtype=np.float64)
    out = np.empty(len(r), dtype=np.uint64)
    lib().orc_sample(n, _p(psi), _p(r), len(r), _p(out))
    return out


def sample_sv(n, psi, r):
    """sampleSV (src/circuit.ts:74-81): the per-shot path's final draw."""
    return int(lib().orc_sample_sv(n, _p(psi), float(r)))


def marginals(n, psi):
    out = np.empty(n, dtype=np.float64)
    lib().orc_marginals(n, _p(psi), _p(out))
    return out


def bloch(n, psi, q):
    th, ph = c_double(), c_double()
    lib().orc_bloch(n, _p(psi), q, ctypes.byref(th), ctypes.byref(ph))
    return th.value, ph.value


def expectation(n, psi, pauli):
    return lib().orc_expectation(n, _p(psi), pauli.upper().encode())


def vqe(n, psi, terms):
    ops = (c_char_p * len(terms))(*[t["ops"].upper().encode() for t in terms])
    c = np.array([t["coeff"] for t in terms], dtype=np.float64)
    return lib().orc_vqe(n, _p(psi), ops, _p(c), len(terms))


def expect_direct(n, psi, x, z):
    return lib().orc_expect_direct(n, _p(psi), x, z)


def collapse(n, psi, q, r):
    return lib().orc_collapse(n, _p(psi), q, r)


# ---------------------------------------------------------------------------
# run()'s per-shot path, restated (src/circuit.ts:164-207 applyOps,
# :289-304 dep1/dep2, :55-81 collapseQubit/sampleSV, :3455-3481 run's
# per-shot loop, :813-819 makePrng).  Calls are ket Circuit-method tuples
# plus ("measure", (q, creg, bit)), ("reset", (q,)) and
# ("if", (creg, value, inner_calls)).
# ---------------------------------------------------------------------------
_SINGLE = {m for m, k in NQ.items() if k == 1}
_TWO_KIND = {"cnot", "cx", "swap", "xx", "yy", "zz", "xy", "iswap", "srswap", "ms", "cy", "cz", "ch", "crx",
             "cry", "crz", "cr2", "cr4", "cr8", "cu1", "cu2", "cu3", "cs", "ct", "csdg", "ctdg", "csrn"}
# the 15 non-identity two-qubit Paulis in ket's order (src/circuit.ts:273-278)
_PAIRS = [(None, "x"), (None, "y"), (None, "z"), ("x", None), ("x", "x"), ("x", "y"), ("x", "z"),
          ("y", None), ("y", "x"), ("y", "y"), ("y", "z"), ("z", None), ("z", "x"), ("z", "y"), ("z", "z")]


def xorshift(seed):
    """makePrng (src/circuit.ts:813-819): xorshift32 <<13 >>>17 <<5, r = s / 2^32, seed 0 -> 1."""
    s = (int(seed) & 0xFFFFFFFF) or 1
    while True:
        s ^= (s << 13) & 0xFFFFFFFF
        s ^= s >> 17
        s ^= (s << 5) & 0xFFFFFFFF
        yield s / 4294967296.0


def creg_value(bits):
    """cregValue (src/circuit.ts:164-166): acc | (1 << i) over the set bits, with
    JavaScript's 32-bit << (count mod 32) and | (a signed 32-bit result)."""
    v = 0
    for i, b in enumerate(bits):
        if b:
            v |= 1 << (i & 31)
    return v - (1 << 32) if v & (1 << 31) else v


def _apply_ops(n, psi, calls, bits, rng, p1, p2, pm, prune):
    """applyOps (src/circuit.ts:169-207) on one shot's state, recursive for `if`:
    ("if", (creg, value, inner_calls)) runs inner_calls when cregValue == value."""
    for name, args in calls:
        if name == "measure":
            q, reg, b = args
            o = collapse(n, psi, q, next(rng))
            if pm and next(rng) < pm:
                o = 1 - o
            if reg in bits:
                bits[reg][b] = o == 1
            continue
        if name == "reset":
            (q,) = args
            if collapse(n, psi, q, next(rng)) == 1:
                apply_named(n, psi, [("x", (q,))], prune)
            continue
        if name == "if":
            reg, value, inner = args
            if creg_value(bits.get(reg, [])) == value:
                _apply_ops(n, psi, inner, bits, rng, p1, p2, pm, prune)
            continue
        if name == "barrier":
            continue
        apply_named(n, psi, [(name, args)], prune)
        qs = args[len(args) - NQ[name]:] if name in NQ else ()
        if name in _SINGLE and p1:                      # dep1
            rand = next(rng)
            if rand < p1:
                r = rand / p1
                apply_named(n, psi, [("x" if r < 1 / 3 else "y" if r < 2 / 3 else "z", (qs[0],))], prune)
        elif name in _TWO_KIND and p2:                  # dep2
            rand = next(rng)
            if rand < p2:
                pa, pb = _PAIRS[min(int(rand / p2 * 15), 14)]
                if pa:
                    apply_named(n, psi, [(pa, (qs[0],))], prune)
                if pb:
                    apply_named(n, psi, [(pb, (qs[1],))], prune)


def run_shots(n, calls, shots, seed, noise=None, init=0, cregs=None, prune=True):
    """Per-shot simulation (run()'s per-shot loop, src/circuit.ts:3455-3481);
    returns (counts {index: n} in first-seen order, creg bit counts)."""
    rng = xorshift(seed)
    p1 = (noise or {}).get("p1", 0) or 0
    p2 = (noise or {}).get("p2", 0) or 0
    pm = (noise or {}).get("pMeas", 0) or 0
    cregs = dict(cregs or {})
    counts, creg_counts = {}, {k: [0] * v for k, v in cregs.items()}
    for _ in range(shots):
        psi = zero_state(n, init)
        bits = {k: [False] * v for k, v in cregs.items()}
        _apply_ops(n, psi, calls, bits, rng, p1, p2, pm, prune)
        idx = sample_sv(n, psi, next(rng))
        if pm:
            for q in range(n):
                if next(rng) < pm:
                    idx ^= 1 << q
        counts[idx] = counts.get(idx, 0) + 1
        for k, v in bits.items():
            for j, b in enumerate(v):
                creg_counts[k][j] += 1 if b else 0
    return counts, creg_counts


# ---------------------------------------------------------------------------
# Density matrix, restated (src/density.ts:475-588 runDM, :193-272 the
# depolarising channels, :274-380 the read-outs) on a dense 2^n x 2^n rho.
# Every op ends with dmSet's prune: entries with |v|^2 < 1e-14 become 0.
# ---------------------------------------------------------------------------
_SRISW = np.array([[1, 0, 0, 0], [0, 2 ** -0.5, 1j * 2 ** -0.5, 0], [0, 1j * 2 ** -0.5, 2 ** -0.5, 0], [0, 0, 0, 1]])
# [flip_a, flip_b, z_a, z_b] of the 15 non-identity two-qubit Paulis (src/density.ts:226-233)
_PAULI15 = [(0, 1, 0, 0), (0, 1, 0, 1), (0, 0, 0, 1), (1, 0, 0, 0), (1, 1, 0, 0), (1, 1, 0, 1), (1, 0, 0, 1),
            (1, 0, 1, 0), (1, 1, 1, 0), (1, 1, 1, 1), (1, 0, 1, 1), (0, 0, 1, 0), (0, 1, 1, 0), (0, 1, 1, 1),
            (0, 0, 1, 1)]


def _full_operator(n, name, args):
    """The 2^n x 2^n matrix of one named call, column by column (no pruning)."""
    dim = 1 << n
    U = np.zeros((dim, dim), dtype=np.complex128)
    for j in range(dim):
        psi = zero_state(n, j)
        apply_named(n, psi, [(name, args)], prune=False)
        U[:, j] = psi
    return U


def _dm_prune(rho):
    return np.where(rho.real ** 2 + rho.imag ** 2 < 1e-14, 0, rho)


def _dm_depolarize1(n, rho, q, p):
    sa, sb, cf = 1 - 2 * p / 3, 2 * p / 3, 1 - 4 * p / 3
    idx = np.arange(1 << n)
    b = (idx >> q) & 1
    same = b[:, None] == b[None, :]
    flipped = rho[np.ix_(idx ^ (1 << q), idx ^ (1 << q))]
    return _dm_prune(np.where(same, sa * rho + sb * flipped, cf * rho))


def _dm_depolarize2(n, rho, a, b, p):
    w1, w2 = 1 - p, p / 15
    idx = np.arange(1 << n)
    ba, bb = (idx >> a) & 1, (idx >> b) & 1
    new = w1 * rho
    for fa, fb, za, zb in _PAULI15:
        perm = idx ^ ((fa << a) | (fb << b))
        par = (za * (ba[:, None] ^ ba[None, :])) ^ (zb * (bb[:, None] ^ bb[None, :]))
        new = new + np.where(par == 1, -w2, w2) * rho[np.ix_(perm, perm)]
    return _dm_prune(new)


def _dm_csrswap(n, rho, c, a, b):
    """ket's block-wise controlled-sqrt(iSWAP) (src/density.ts:521-570)."""
    W = _full_operator(n, "unitary", (_SRISW, a, b))
    ctl = ((np.arange(1 << n) >> c) & 1).astype(bool)
    B = {(i, j): np.where(np.outer(ctl == bool(i), ctl == bool(j)), rho, 0) for i in (0, 1) for j in (0, 1)}
    out = B[0, 0].copy()
    out = out + W @ B[1, 1] @ W.conj().T
    out = out + W @ B[1, 0] @ W.conj().T
    out = out + (W @ B[0, 1].T @ W.conj().T).T
    return _dm_prune(out)


def run_dm(n, calls, noise=None):
    """runDM: rho of the calls from |0..0><0..0| with optional p1 / p2 depolarising noise."""
    p1 = (noise or {}).get("p1", 0) or 0
    p2 = (noise or {}).get("p2", 0) or 0
    rho = np.zeros((1 << n, 1 << n), dtype=np.complex128)
    rho[0, 0] = 1
    for name, args in calls:
        if name == "barrier":
            continue
        if name == "csrswap":
            rho = _dm_csrswap(n, rho, *args)
            continue
        U = _full_operator(n, name, args)
        rho = _dm_prune(U @ rho @ U.conj().T)
        qs = args[len(args) - NQ[name]:] if name in NQ else ()
        if name in _SINGLE and p1:
            rho = _dm_depolarize1(n, rho, qs[0], p1)
        elif name in _TWO_KIND and p2:
            rho = _dm_depolarize2(n, rho, qs[0], qs[1], p2)
    return rho
