"""ctypes binding of libketsv.so (the C ABI in include/ketsv.h).

The library is the only compute path: if it is missing, importing a
simulation entry point raises at once (there is no CPU fallback).
"""
import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int32, c_int64, c_size_t, c_uint64, c_void_p

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libketsv.so")

KSV_OK, KSV_ERR_RANGE, KSV_ERR_TYPE, KSV_ERR_DEVICE, KSV_ERR_HANDLE = 0, 1, 2, 3, 4
KSV_F64, KSV_F32 = 0, 1
(KSV_SINGLE, KSV_CNOT, KSV_SWAP, KSV_TWO, KSV_CONTROLLED, KSV_TOFFOLI, KSV_CSWAP, KSV_CSRSWAP,
 KSV_UNITARY, KSV_BARRIER) = range(10)
KSV_MAX_OPERANDS = 12
(KSV_OPT_PRUNE, KSV_OPT_TILE_QUBITS, KSV_OPT_TIMING, KSV_OPT_FUSE, KSV_OPT_PASS_KERNEL, KSV_OPT_ROUND_TILE,
 KSV_OPT_BLOCK_QUBITS, KSV_OPT_BLOCK_TILE, KSV_OPT_BLOCK_THREADS, KSV_OPT_ADJ_TILE,
 KSV_OPT_ADJ_THREADS, KSV_OPT_BLOCK_PIPE) = range(12)

OP_DTYPE = np.dtype([("kind", "<i4"), ("nq", "<i4"), ("q", "<i4", (KSV_MAX_OPERANDS,)), ("mat", "<i8")])
PARAM_DTYPE = np.dtype([("op", "<i4"), ("param", "<i4"), ("coeff", "<f8"), ("dmat", "<i8")])
assert OP_DTYPE.itemsize == 64 and PARAM_DTYPE.itemsize == 24


class RangeError(ValueError):
    """ket's RangeError (qubit index, size or capacity out of range)."""


class KsvStats(ctypes.Structure):
    _fields_ = [("pass_launches", c_uint64), ("pass_ms", c_double), ("pass_bytes", c_double),
                ("other_launches", c_uint64), ("other_ms", c_double), ("other_bytes", c_double),
                ("gates_applied", c_uint64)]


class KsvProgramSummary(ctypes.Structure):
    _fields_ = [("ops", c_uint64), ("passes", c_uint64), ("device_ops", c_uint64),
                ("tile_qubits", c_int32), ("reg_passes", c_int32), ("exchanges", c_uint64),
                ("blk_passes", c_int32), ("pad", c_int32), ("blocks", c_uint64)]


# (name, restype, argtypes) of every entry point in include/ketsv.h
_P = c_void_p
_DP = POINTER(c_double)
_U64P = POINTER(c_uint64)
SIGNATURES = [
    ("ksv_last_error", c_int, [ctypes.c_char_p, c_size_t]),
    ("ksv_device_count", c_int, [POINTER(c_int)]),
    ("ksv_create", c_int, [c_int, c_int, POINTER(c_void_p)]),
    ("ksv_destroy", c_int, [_P]),
    ("ksv_set_option", c_int, [_P, c_int, c_double]),
    ("ksv_init_basis", c_int, [_P, c_uint64]),
    ("ksv_copy_state", c_int, [_P, _P]),
    ("ksv_apply", c_int, [_P, _P, c_size_t, _P, c_size_t]),
    ("ksv_program_create", c_int, [_P, _P, c_size_t, _P, c_size_t, POINTER(c_void_p)]),
    ("ksv_program_run", c_int, [_P, _P]),
    ("ksv_program_info", c_int, [_P, POINTER(KsvProgramSummary)]),
    ("ksv_program_set_matrix", c_int, [_P, c_size_t, _P, c_size_t]),
    ("ksv_program_destroy", c_int, [_P]),
    ("ksv_plan", c_int, [c_int, c_int, _P, c_size_t, _P, c_size_t, c_int, c_int, c_int, c_int, c_int,
                          POINTER(KsvProgramSummary)]),
    ("ksv_plan_blocks", c_int, [c_int, c_int, _P, c_size_t, _P, c_size_t, c_int, c_int, c_int, c_int, c_int, c_int,
                                 c_int, POINTER(KsvProgramSummary)]),
    ("ksv_get_amplitudes", c_int, [_P, c_uint64, c_uint64, _P]),
    ("ksv_set_amplitudes", c_int, [_P, c_uint64, c_uint64, _P]),
    ("ksv_amplitude", c_int, [_P, c_uint64, _P]),
    ("ksv_norm2", c_int, [_P, _DP]),
    ("ksv_probs_filtered", c_int, [_P, c_double, _P, _P, c_size_t, POINTER(c_size_t)]),
    ("ksv_topk", c_int, [_P, c_size_t, _P, _P]),
    ("ksv_marginals", c_int, [_P, _P]),
    ("ksv_reduce_1q", c_int, [_P, c_int, _P]),
    ("ksv_sample", c_int, [_P, _P, c_size_t, _P]),
    ("ksv_sample_noalloc", c_int, [_P, _P, c_size_t, _P]),
    ("ksv_expect_pauli", c_int, [_P, _P, _P, _P, c_size_t, c_int, _P, _DP]),
    ("ksv_adjoint", c_int, [_P, _P, c_size_t, _P, c_size_t, _P, c_size_t, c_size_t, _P, _P, _P, c_size_t, _P, _DP]),
    ("ksv_adjoint_batch_create", c_int, [_P, c_int, _P, c_size_t, _P, c_size_t, _P, c_size_t, c_size_t, _P, _P, _P,
                                         c_size_t, POINTER(c_void_p)]),
    ("ksv_adjoint_batch_run", c_int, [_P, _P, c_int, _P, _P]),
    ("ksv_adjoint_batch_destroy", c_int, [_P]),
    ("ksv_collapse", c_int, [_P, c_int, c_double, POINTER(c_int)]),
    ("ksv_shots_init", c_int, [_P, c_int, c_uint64]),
    ("ksv_shots_init_identity", c_int, [_P, c_int]),
    ("ksv_shots_pauli", c_int, [_P, c_int, c_int, _P, _P, c_size_t]),
    ("ksv_shots_measure", c_int, [_P, c_int, c_int, _P, _P, _P]),
    ("ksv_shots_sample", c_int, [_P, c_int, _P, _P]),
    ("ksv_xorshift32", c_int, [_P, c_size_t, _P]),
    ("ksv_shots_run", c_int, [_P, c_int, _P, c_size_t, _P, c_size_t, _P, c_size_t, _P, _P, c_size_t, _P, c_size_t,
                              _P]),
    ("ksv_qaoa_batch_create", c_int, [c_int, _P, _P, _P, c_int, c_int, POINTER(c_void_p)]),
    ("ksv_qaoa_batch_run", c_int, [_P, _P, _P, c_int, _P, _P, _P]),
    ("ksv_qaoa_batch_amplitudes", c_int, [_P, c_int, c_uint64, c_uint64, _P]),
    ("ksv_qaoa_batch_info", c_int, [_P, _U64P, _U64P]),
    ("ksv_qaoa_batch_destroy", c_int, [_P]),
    ("ksv_batch_last_error", c_int, [ctypes.c_char_p, c_size_t]),
    ("ksv_dm_create", c_int, [c_int, POINTER(c_void_p)]),
    ("ksv_dm_apply", c_int, [_P, _P, c_size_t, _P, c_size_t, c_double, c_double]),
    ("ksv_dm_read", c_int, [_P, _P, _P, _U64P]),
    ("ksv_dm_destroy", c_int, [_P]),
    ("ksv_get_stats", c_int, [_P, POINTER(KsvStats)]),
    ("ksv_reset_stats", c_int, [_P]),
]

_lib = None


def lib():
    """Load libketsv.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"ket_amd: HIP library {LIB_PATH} is not built; run __graft_entry__.build() "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    lib().ksv_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc):
    """Map a status code to ket's error classes."""
    if rc == KSV_OK:
        return
    msg = last_error()
    if rc == KSV_ERR_RANGE:
        raise RangeError(msg)
    if rc == KSV_ERR_TYPE:
        raise TypeError(msg)
    raise RuntimeError(f"libketsv: {msg}")


def device_count():
    c = c_int(0)
    lib().ksv_device_count(ctypes.byref(c))
    return c.value


KERNELS = {"auto": 0, "lds": 1, "reg": 2, "mfma": 3}


def plan(n, tape, precision="f64", kernel="auto", tile=0, rtile=0, fuse=False, prune=True, btile=0, bq=0):
    """Dry-run schedule of a tape (no device needed)."""
    ops, mats = tape
    s = KsvProgramSummary()
    check(lib().ksv_plan_blocks(int(n), KSV_F64 if precision == "f64" else KSV_F32, ptr(ops), len(ops), ptr(mats),
                                len(mats), KERNELS[kernel], int(tile), int(rtile), int(btile), int(bq),
                                1 if fuse else 0, 1 if prune else 0, ctypes.byref(s)))
    return {f: getattr(s, f) for f, _ in KsvProgramSummary._fields_}


SHOT_EVENT_DTYPE = np.dtype([("at", np.uint64), ("kind", np.int32), ("q", np.int32), ("off", np.uint64),
                             ("count", np.uint64)])
KSV_SHOT_PAULI, KSV_SHOT_MEASURE, KSV_SHOT_RESET = 0, 1, 2


def ptr(a):
    """Data pointer of a contiguous numpy array (or None)."""
    if a is None:
        return None
    return c_void_p(a.ctypes.data)


class State:
    """Owning handle of one device statevector (2^n amplitudes in HBM)."""

    def __init__(self, n, precision="f64", prune=True, tile=0, fuse=False, timing=False, kernel="auto",
                 rtile=0, bq=0, btile=0, bthreads=0, adj_tile=0, adj_threads=0, bpipe=None):
        self.n = int(n)
        self.precision = precision
        prec = KSV_F64 if precision == "f64" else KSV_F32
        h = c_void_p()
        check(lib().ksv_create(self.n, prec, ctypes.byref(h)))
        self._h = h
        self.set_option(KSV_OPT_PRUNE, 1 if prune else 0)
        if tile:
            self.set_option(KSV_OPT_TILE_QUBITS, tile)
        if fuse:
            self.set_option(KSV_OPT_FUSE, 1)
        if timing:
            self.set_option(KSV_OPT_TIMING, 1)
        if kernel != "auto":
            self.set_option(KSV_OPT_PASS_KERNEL, KERNELS[kernel])
        if rtile:
            self.set_option(KSV_OPT_ROUND_TILE, rtile)
        if bq:
            self.set_option(KSV_OPT_BLOCK_QUBITS, bq)
        if btile:
            self.set_option(KSV_OPT_BLOCK_TILE, btile)
        if bthreads:
            self.set_option(KSV_OPT_BLOCK_THREADS, bthreads)
        if adj_tile:
            self.set_option(KSV_OPT_ADJ_TILE, adj_tile)
        if adj_threads:
            self.set_option(KSV_OPT_ADJ_THREADS, adj_threads)
        if bpipe is not None:
            self.set_option(KSV_OPT_BLOCK_PIPE, 1 if bpipe else 0)

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("state has been freed")
        return self._h

    def free(self):
        if getattr(self, "_h", None) is not None:
            lib().ksv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def set_option(self, opt, value):
        check(lib().ksv_set_option(self.handle, opt, float(value)))

    def init_basis(self, index=0):
        check(lib().ksv_init_basis(self.handle, int(index)))

    def apply(self, tape):
        ops, mats = tape
        check(lib().ksv_apply(self.handle, ptr(ops), len(ops), ptr(mats), len(mats)))

    def program(self, tape):
        return Program(self, tape)

    def amplitudes(self, offset=0, count=None):
        if count is None:
            count = (1 << self.n) - offset
        out = np.empty(2 * count, dtype=np.float64)
        check(lib().ksv_get_amplitudes(self.handle, offset, count, ptr(out)))
        return out.view(np.complex128)

    def set_amplitudes(self, values, offset=0):
        v = np.ascontiguousarray(values, dtype=np.complex128)
        check(lib().ksv_set_amplitudes(self.handle, offset, len(v), ptr(v.view(np.float64))))

    def amplitude(self, index):
        out = np.empty(2, dtype=np.float64)
        check(lib().ksv_amplitude(self.handle, int(index), ptr(out)))
        return complex(out[0], out[1])

    def norm2(self):
        v = c_double()
        check(lib().ksv_norm2(self.handle, ctypes.byref(v)))
        return v.value

    def probs_filtered(self, thresh=1e-14, cap=None):
        if cap is None:
            cnt = c_size_t()
            check(lib().ksv_probs_filtered(self.handle, thresh, None, None, 0, ctypes.byref(cnt)))
            cap = cnt.value
        idx = np.empty(max(cap, 1), dtype=np.uint64)
        p = np.empty(max(cap, 1), dtype=np.float64)
        cnt = c_size_t()
        check(lib().ksv_probs_filtered(self.handle, thresh, ptr(idx), ptr(p), cap, ctypes.byref(cnt)))
        m = min(cnt.value, cap)
        return idx[:m], p[:m]

    def topk(self, k):
        idx = np.empty(k, dtype=np.uint64)
        p = np.empty(k, dtype=np.float64)
        check(lib().ksv_topk(self.handle, k, ptr(idx), ptr(p)))
        return idx, p

    def marginals(self):
        out = np.empty(self.n, dtype=np.float64)
        check(lib().ksv_marginals(self.handle, ptr(out)))
        return out

    def reduce_1q(self, q):
        out = np.empty(4, dtype=np.float64)
        check(lib().ksv_reduce_1q(self.handle, int(q), ptr(out)))
        return out

    def sample(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        out = np.empty(len(r), dtype=np.uint64)
        check(lib().ksv_sample(self.handle, ptr(r), len(r), ptr(out)))
        return out

    def sample_noalloc(self, r):
        """ksv_sample without any device allocation (fixed scratch: <= 20 qubits,
        <= 65536 draws per call; calls that overlap on one GPU race)."""
        r = np.ascontiguousarray(r, dtype=np.float64)
        out = np.empty(len(r), dtype=np.uint64)
        check(lib().ksv_sample_noalloc(self.handle, ptr(r), len(r), ptr(out)))
        return out

    def expect_pauli(self, xmask, zmask, coeff, mode=0):
        x = np.ascontiguousarray(xmask, dtype=np.uint64)
        z = np.ascontiguousarray(zmask, dtype=np.uint64)
        c = np.ascontiguousarray(coeff, dtype=np.float64)
        out = np.empty(max(len(c), 1), dtype=np.float64)
        tot = c_double()
        check(lib().ksv_expect_pauli(self.handle, ptr(x), ptr(z), ptr(c), len(c), int(mode), ptr(out),
                                     ctypes.byref(tot)))
        return out[:len(c)], tot.value

    def adjoint(self, tape, records, nparam, xmask, zmask, coeff):
        ops, mats = tape
        rec = np.ascontiguousarray(records, dtype=PARAM_DTYPE)
        x = np.ascontiguousarray(xmask, dtype=np.uint64)
        z = np.ascontiguousarray(zmask, dtype=np.uint64)
        c = np.ascontiguousarray(coeff, dtype=np.float64)
        grad = np.zeros(max(nparam, 1), dtype=np.float64)
        e = c_double()
        check(lib().ksv_adjoint(self.handle, ptr(ops), len(ops), ptr(mats), len(mats), ptr(rec), len(rec), nparam,
                                ptr(x), ptr(z), ptr(c), len(c), ptr(grad), ctypes.byref(e)))
        return grad[:nparam], e.value

    def collapse(self, q, r):
        o = c_int()
        check(lib().ksv_collapse(self.handle, int(q), float(r), ctypes.byref(o)))
        return o.value

    # batched trajectories: copies of an n-qubit state, copy t = amplitudes [t 2^n, (t+1) 2^n)
    def shots_init(self, n, index=0):
        check(lib().ksv_shots_init(self.handle, int(n), int(index)))

    def shots_init_identity(self, n):
        check(lib().ksv_shots_init_identity(self.handle, int(n)))

    def shots_pauli(self, n, q, copies, paulis):
        c = np.ascontiguousarray(copies, dtype=np.uint32)
        p = np.ascontiguousarray(paulis, dtype=np.uint8)
        if len(c) != len(p):
            raise TypeError("copies and paulis differ in length")
        check(lib().ksv_shots_pauli(self.handle, int(n), int(q), ptr(c), ptr(p), len(c)))

    def shots_measure(self, n, q, r, reset=None):
        r = np.ascontiguousarray(r, dtype=np.float64)
        if len(r) != 1 << (self.n - n):
            raise TypeError("one draw per copy")
        rs = None if reset is None else np.ascontiguousarray(reset, dtype=np.uint8)
        out = np.empty(len(r), dtype=np.uint8)
        check(lib().ksv_shots_measure(self.handle, int(n), int(q), ptr(r), ptr(rs) if rs is not None else None,
                                      ptr(out)))
        return out

    def shots_sample(self, n, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        if len(r) != 1 << (self.n - n):
            raise TypeError("one draw per copy")
        out = np.empty(len(r), dtype=np.uint64)
        check(lib().ksv_shots_sample(self.handle, int(n), ptr(r), ptr(out)))
        return out

    def shots_run(self, n, tape, events, copies, paulis, r):
        """One batch: the gate tape with SHOT_EVENT_DTYPE events; returns the outcomes array (len(r))."""
        ops, mats = tape
        ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
        mats = np.ascontiguousarray(mats, dtype=np.float64)
        ev = np.ascontiguousarray(events, dtype=SHOT_EVENT_DTYPE)
        c = np.ascontiguousarray(copies, dtype=np.uint32)
        pa = np.ascontiguousarray(paulis, dtype=np.uint8)
        r = np.ascontiguousarray(r, dtype=np.float64)
        out = np.zeros(max(1, len(r)), dtype=np.uint8)
        m = mats if len(mats) else np.zeros(1)
        check(lib().ksv_shots_run(self.handle, int(n), ptr(ops), len(ops), ptr(m), len(mats), ptr(ev), len(ev),
                                  ptr(c), ptr(pa), len(c), ptr(r), len(r), ptr(out)))
        return out[:len(r)]

    def copy_from(self, other):
        check(lib().ksv_copy_state(self.handle, other.handle))

    def stats(self):
        s = KsvStats()
        check(lib().ksv_get_stats(self.handle, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in KsvStats._fields_}

    def reset_stats(self):
        check(lib().ksv_reset_stats(self.handle))


class Program:
    """A scheduled tape resident on the device (ksv_program_*)."""

    def __init__(self, state, tape):
        ops, mats = tape
        h = c_void_p()
        check(lib().ksv_program_create(state.handle, ptr(ops), len(ops), ptr(mats), len(mats), ctypes.byref(h)))
        self._h = h
        self.state = state

    def run(self, state=None):
        st = state or self.state
        check(lib().ksv_program_run(st.handle, self._h))

    def set_matrix(self, op, matrix):
        """Give tape op `op` a new matrix (complex, row-major) without rescheduling."""
        m = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128).reshape(-1)).view(np.float64)
        check(lib().ksv_program_set_matrix(self._h, int(op), ptr(m), len(m)))

    def info(self):
        s = KsvProgramSummary()
        check(lib().ksv_program_info(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in KsvProgramSummary._fields_}

    def free(self):
        if getattr(self, "_h", None) is not None:
            lib().ksv_program_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class AdjointBatch:
    """B parameter sets of one tape on the copies of one device state
    (ksv_adjoint_batch_*): energies and adjoint gradients per set, one graph
    replay per evaluation."""

    def __init__(self, n, copies_log, precision, tape, mats_all, records, nparam, xmask, zmask, coeff,
                 adj_tile=0, adj_threads=0):
        ops, _ = tape
        self.B = 1 << copies_log
        self.nparam = int(nparam)
        self.nmats = mats_all.shape[1]
        self.state = State(n + copies_log, precision, prune=False, adj_tile=adj_tile, adj_threads=adj_threads)
        rec = np.ascontiguousarray(records, dtype=PARAM_DTYPE)
        x = np.ascontiguousarray(xmask, dtype=np.uint64)
        z = np.ascontiguousarray(zmask, dtype=np.uint64)
        c = np.ascontiguousarray(coeff, dtype=np.float64)
        m = np.ascontiguousarray(mats_all, dtype=np.float64)
        h = c_void_p()
        try:
            check(lib().ksv_adjoint_batch_create(self.state.handle, int(n), ptr(ops), len(ops), ptr(m), self.nmats,
                                                 ptr(rec), len(rec), self.nparam, ptr(x), ptr(z), ptr(c), len(c),
                                                 ctypes.byref(h)))
        except Exception:
            self.state.free()
            raise
        self._h = h

    def run(self, mats_all, grad=True):
        m = np.ascontiguousarray(mats_all, dtype=np.float64)
        if m.shape != (self.B, self.nmats):
            raise TypeError(f"matrices must be {self.B} x {self.nmats}")
        e = np.empty(self.B)
        g = np.zeros((self.B, max(1, self.nparam)))
        check(lib().ksv_adjoint_batch_run(self._h, ptr(m), 1 if grad else 0, ptr(e), ptr(g)))
        return (e, g[:, :self.nparam]) if grad else e

    def free(self):
        if getattr(self, "_h", None) is not None:
            lib().ksv_adjoint_batch_destroy(self._h)
            self._h = None
        if getattr(self, "state", None) is not None:
            self.state.free()
            self.state = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _check_batch(rc):
    if rc == KSV_OK:
        return
    buf = ctypes.create_string_buffer(1024)
    lib().ksv_batch_last_error(buf, len(buf))
    msg = buf.value.decode(errors="replace")
    if rc == KSV_ERR_RANGE:
        raise RangeError(msg)
    if rc == KSV_ERR_TYPE:
        raise TypeError(msg)
    raise RuntimeError(f"libketsv: {msg}")


class QaoaBatch:
    """Many Max-Cut QAOA instances (ket's qaoa() + maxCutHamiltonian()) resident
    on the device together; `run` returns <H_C> per instance and, optionally,
    adjoint gradients with respect to every gamma and beta."""

    def __init__(self, instances, p, precision="f64"):
        """instances: list of (n, edges); p: layers."""
        self.count = len(instances)
        self.p = int(p)
        nq = np.array([n for n, _ in instances], dtype=np.int32)
        off = np.zeros(self.count + 1, dtype=np.int64)
        flat = []
        for i, (_, edges) in enumerate(instances):
            for u, v in edges:
                flat.extend((u, v))
            off[i + 1] = off[i] + len(edges)
        ed = np.array(flat if flat else [0, 0], dtype=np.int32)
        h = c_void_p()
        _check_batch(lib().ksv_qaoa_batch_create(self.count, ptr(nq), ptr(off), ptr(ed), self.p,
                                                 KSV_F64 if precision == "f64" else KSV_F32, ctypes.byref(h)))
        self._h = h
        self.nq = nq

    def run(self, gamma, beta, grad=False):
        g = np.ascontiguousarray(gamma, dtype=np.float64).reshape(self.count, self.p)
        b = np.ascontiguousarray(beta, dtype=np.float64).reshape(self.count, self.p)
        e = np.empty(self.count)
        dg = np.empty((self.count, self.p))
        db = np.empty((self.count, self.p))
        _check_batch(lib().ksv_qaoa_batch_run(self._h, ptr(g), ptr(b), 1 if grad else 0, ptr(e), ptr(dg), ptr(db)))
        return (e, dg, db) if grad else e

    def amplitudes(self, inst):
        n = int(self.nq[inst])
        out = np.empty(2 << n, dtype=np.float64)
        _check_batch(lib().ksv_qaoa_batch_amplitudes(self._h, int(inst), 0, 1 << n, ptr(out)))
        return out.view(np.complex128)

    def info(self):
        a, w = c_uint64(), c_uint64()
        _check_batch(lib().ksv_qaoa_batch_info(self._h, ctypes.byref(a), ctypes.byref(w)))
        return {"amplitudes": a.value, "work_items": w.value}

    def free(self):
        if getattr(self, "_h", None) is not None:
            lib().ksv_qaoa_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def xorshift_fill(state, count):
    """`count` draws of the xorshift32 stream from uint32 `state` (advanced in place: a 1-element uint32 array)."""
    out = np.empty(count, dtype=np.float64)
    check(lib().ksv_xorshift32(ptr(state), count, ptr(out)))
    return out


class DensityState:
    """A density matrix on the device in ket's Map semantics (ksv_dm_*): rho as
    (row << n) | col plus every entry's place in ket's Map order."""

    def __init__(self, n):
        self.n = n
        h = c_void_p()
        check(lib().ksv_dm_create(int(n), ctypes.byref(h)))
        self._h = h

    def apply(self, tape, p1=0.0, p2=0.0):
        ops, mats = tape
        ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
        mats = np.ascontiguousarray(mats, dtype=np.float64)
        check(lib().ksv_dm_apply(self._h, ptr(ops), len(ops), ptr(mats) if len(mats) else None, len(mats),
                                 float(p1), float(p2)))

    def read(self):
        """(rho as a dense 2^n x 2^n complex128 array, keys in Map order)."""
        dim = 1 << self.n
        rho = np.empty((dim, dim), dtype=np.complex128)
        order = np.empty(dim * dim, dtype=np.uint64)
        cnt = c_uint64(0)
        check(lib().ksv_dm_read(self._h, ptr(rho), ptr(order), ctypes.byref(cnt)))
        return rho, order[:cnt.value].copy()

    def free(self):
        if self._h:
            lib().ksv_dm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
