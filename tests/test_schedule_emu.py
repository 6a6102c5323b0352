"""Scheduled programs executed by the host emulator (tests/emu) against the
oracle: register-round passes (LDS transposes, slot / thread / global
controls), LDS passes and wide unitaries, with every HBM index and LDS
address checked.  Runs without a GPU."""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

from ket_amd import workloads
from ket_amd._native import OP_DTYPE
from helpers import build, dense_ref, random_calls

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_DIR = os.path.join(HERE, "emu")
EMU_LIB = os.path.join(EMU_DIR, "libketsv_emu.so")
TOL = 1e-12

_lib = None


def emu():
    global _lib
    if _lib is None:
        subprocess.run(["make", "-s", "-C", EMU_DIR], check=True)
        L = ctypes.CDLL(EMU_LIB)
        L.emu_run.restype = ctypes.c_int
        L.emu_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                              ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.emu_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
        L.emu_set_block_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


KERNEL = {"auto": 0, "lds": 1, "reg": 2, "mfma": 3}


def run_emu(n, calls, prune=True, kernel="auto", tile=0, rtile=0, fuse=False, init=0, btile=0, bq=0):
    ops, mats = build(n, calls).tape()
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    mats = np.ascontiguousarray(mats, dtype=np.float64)
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[init] = 1
    stats = np.zeros(6, dtype=np.int64)
    rc = emu().emu_run(n, 1 if prune else 0, ops.ctypes.data, len(ops), mats.ctypes.data, len(mats), KERNEL[kernel],
                       tile, rtile, 1 if fuse else 0, psi.ctypes.data, stats.ctypes.data, btile, bq)
    if rc:
        buf = ctypes.create_string_buffer(512)
        emu().emu_last_error(buf, 512)
        raise AssertionError(f"emulator: rc={rc}: {buf.value.decode()}")
    return psi, stats


@pytest.mark.parametrize("n,rtile", [(11, 11), (12, 12), (13, 12), (13, 13), (14, 11), (15, 12), (16, 13)])
@pytest.mark.parametrize("prune", [True, False])
def test_register_round_every_op(n, rtile, prune):
    calls = random_calls(n, 160, seed=31 * n + rtile, with_unitary=True)
    ref = dense_ref(n, calls, prune=prune)
    got, stats = run_emu(n, calls, prune=prune, kernel="reg", rtile=rtile)
    assert stats[1] > 0
    assert np.max(np.abs(got - ref)) < TOL


@pytest.mark.parametrize("rtile", [11, 12, 13])
@pytest.mark.parametrize("fuse", [False, True])
def test_register_round_random_layers(rtile, fuse):
    n = 15
    calls = workloads.random_layer_calls(n, 6, seed=17)
    ref = dense_ref(n, calls, prune=False)
    got, stats = run_emu(n, calls, prune=False, kernel="reg", rtile=rtile, fuse=fuse)
    assert stats[2] > 0
    assert np.max(np.abs(got - ref)) < TOL


def test_register_round_qft_with_initial_state():
    n = 13
    c, bs, x = workloads.qft_input(n, seed=7)
    from ket_amd.circuit import _flatten  # noqa: F401  (tape comes from the circuit below)
    ops, mats = c.tape()
    ref = np.zeros(1 << n, dtype=np.complex128)
    ref[x] = 1
    import ketsv_oracle as O
    O.apply_tape(n, ref, ops, mats, prune=False)
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[x] = 1
    stats = np.zeros(6, dtype=np.int64)
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    rc = emu().emu_run(n, 0, ops.ctypes.data, len(ops), mats.ctypes.data, len(mats), 2, 0, 12, 0, psi.ctypes.data,
                       stats.ctypes.data, 0, 0)
    assert rc == 0
    assert np.max(np.abs(psi - ref)) < TOL
    k = np.arange(1 << n)
    assert np.max(np.abs(psi - np.exp(2j * np.pi * x * k / (1 << n)) / np.sqrt(1 << n))) < 1e-12


@pytest.mark.parametrize("n,tile", [(9, 8), (14, 10), (15, 12)])
def test_lds_passes(n, tile):
    calls = random_calls(n, 120, seed=5 * n + tile, with_unitary=True)
    ref = dense_ref(n, calls, prune=False)
    got, stats = run_emu(n, calls, prune=False, kernel="lds", tile=tile)
    assert stats[1] == 0
    assert np.max(np.abs(got - ref)) < TOL


def test_small_states_fall_back_to_lds_tiles():
    for n in (1, 2, 5, 10):
        calls = random_calls(n, 60, seed=n)
        ref = dense_ref(n, calls, prune=True)
        got, stats = run_emu(n, calls, prune=True)
        assert stats[1] == 0
        assert np.max(np.abs(got - ref)) < TOL


def test_wide_unitary_steps():
    import random as _r
    from helpers import random_unitary
    rnd = _r.Random(5)
    n = 13
    calls = [("h", (q,)) for q in range(n)]
    calls.append(("unitary", (random_unitary(32, rnd), 0, 4, 2, 7, 12)))
    calls.append(("unitary", (random_unitary(16, rnd), 3, 1, 8, 2)))
    calls.append(("unitary", (random_unitary(8, rnd), 11, 0, 5)))
    calls += [("cnot", (q, (q + 5) % n)) for q in range(n)]
    ref = dense_ref(n, calls, prune=False)
    got, _ = run_emu(n, calls, prune=False, kernel="reg")
    assert np.max(np.abs(got - ref)) < TOL


def test_thirty_qubit_schedule_shape():
    """The benchmark circuit's schedule (dry run): pass count and transposes."""
    from ket_amd._native import plan
    c = workloads.random_layers(30, 40, seed=42)
    info = plan(30, c.tape(), kernel="reg", fuse=True, prune=False)
    assert info["reg_passes"] == info["passes"]
    assert info["passes"] <= 70


@pytest.mark.parametrize("n,btile,bq,nt", [(12, 12, 4, 256), (13, 12, 4, 256), (14, 13, 4, 256), (14, 13, 5, 256),
                                           (15, 12, 4, 256), (11, 11, 4, 256), (13, 11, 4, 256), (12, 12, 4, 512),
                                           (14, 12, 4, 512)])
def test_fused_block_passes_every_op(n, btile, bq, nt):
    calls = random_calls(n, 140, seed=7 * n + bq, with_unitary=True)
    ref = dense_ref(n, calls, prune=False)
    emu().emu_set_block_threads(nt)
    try:
        got, stats = run_emu(n, calls, prune=False, kernel="mfma", btile=btile, bq=bq)
    finally:
        emu().emu_set_block_threads(256)
    assert stats[4] > 0 and stats[5] > 0
    assert np.max(np.abs(got - ref)) < TOL


@pytest.mark.parametrize("n,bq,btile", [(13, 4, 12), (14, 5, 13)])
def test_fused_diagonal_blocks(n, bq, btile):
    """QFT: most blocks hold only controlled phases and run as a phase per row."""
    calls = [("x", (q,)) for q in range(0, n, 3)] + [("h", (0,)), ("ry", (0.4, n - 1))]
    for j in range(n):
        calls.append(("h", (j,)))
        for k in range(j + 1, n):
            calls.append(("cu1", (math.pi / (1 << (k - j)), k, j)))
    calls += [("rz", (0.3, 2)), ("cz", (1, 5)), ("crz", (0.7, 3, 6)), ("s", (4,))]
    ref = dense_ref(n, calls, prune=False)
    got, stats = run_emu(n, calls, prune=False, kernel="mfma", btile=btile, bq=bq)
    assert stats[4] > 0
    assert np.max(np.abs(got - ref)) < TOL


@pytest.mark.parametrize("bq,btile", [(4, 12), (5, 13)])
def test_fused_block_random_layers(bq, btile):
    n = 15
    calls = workloads.random_layer_calls(n, 6, seed=23)
    ref = dense_ref(n, calls, prune=False)
    got, stats = run_emu(n, calls, prune=False, kernel="mfma", btile=btile, bq=bq)
    assert stats[1] == 0 and stats[4] > 0
    assert np.max(np.abs(got - ref)) < TOL


@pytest.mark.parametrize("bq,btile", [(4, 12), (4, 11)])
def test_fused_monomial_blocks(bq, btile):
    """Blocks of permutations and phases only (cnot, x, swap, cz, s, t after a
    layer of rotations, and the brick circuit's trailing cnots) move each row to
    its image times its phase instead of running a GEMM; the emulated kernel
    equals the dense oracle and the plan holds such blocks."""
    n = 14
    calls = workloads.random_layer_calls(n, 3, seed=31)
    for q in range(0, n - 1, 2):
        calls += [("cnot", (q, q + 1)), ("x", (q,)), ("cz", (q + 1, (q + 3) % n))]
    calls += [("swap", (0, 5)), ("s", (3,)), ("t", (7,)), ("cnot", (9, 2)), ("y", (11,))]
    ref = dense_ref(n, calls, prune=False)
    got, stats = run_emu(n, calls, prune=False, kernel="mfma", btile=btile, bq=bq)
    assert stats[4] > 0
    assert np.max(np.abs(got - ref)) < TOL
    ops, mats = build(n, calls).tape()
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    mats = np.ascontiguousarray(mats, dtype=np.float64)
    kinds = np.zeros(3, dtype=np.int64)
    L = emu()
    L.emu_block_kinds.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    assert L.emu_block_kinds(n, ops.ctypes.data, len(ops), mats.ctypes.data, len(mats), btile, bq,
                             kinds.ctypes.data) == 0
    assert kinds[2] > 0, kinds


def test_fused_blocks_fall_back_under_pruning():
    n = 13
    calls = random_calls(n, 80, seed=99)
    ref = dense_ref(n, calls, prune=True)
    got, stats = run_emu(n, calls, prune=True, kernel="mfma")
    assert stats[4] == 0
    assert np.max(np.abs(got - ref)) < TOL


def _param_pair(n, seed, layers=5):
    """Two circuits of one structure with different angles (every gate dense or diagonal as before)."""
    import math
    import random
    rnd = random.Random(seed)

    def build_c(angles):
        calls, k = [], 0
        for _ in range(layers):
            for q in range(n):
                calls.append(("rx", (angles[k], q)))
                calls.append(("rz", (angles[k + 1], q)))
                k += 2
            for q in range(0, n - 1, 2):
                calls.append(("cnot", (q, q + 1)))
            calls.append(("crz", (angles[k], 0, n - 1)))
            calls.append(("xx", (angles[k + 1], 1, n - 2)))
            k += 2
        return calls

    cnt = layers * (2 * n + 2)
    a = [rnd.uniform(0.2, 3.0) for _ in range(cnt)]
    b = [rnd.uniform(0.2, 3.0) for _ in range(cnt)]
    return build_c(a), build_c(b)


@pytest.mark.parametrize("kernel,n,kw", [("reg", 13, {"rtile": 11}), ("mfma", 13, {"btile": 12, "bq": 4}),
                                          ("mfma", 14, {"btile": 13, "bq": 5}), ("lds", 10, {"tile": 8}),
                                          ("auto", 6, {})])
def test_program_update_matches_rebuilt_circuit(kernel, n, kw):
    """ksv_program_set_matrix's host side: a scheduled program given new angles
    equals the oracle on the circuit built with those angles."""
    ca, cb = _param_pair(n, seed=n)
    ta, tb = build(n, ca).tape(), build(n, cb).tape()
    ops = np.ascontiguousarray(ta[0], dtype=OP_DTYPE)
    upd, mats = [], []
    width = 32
    for i in range(len(ops)):
        if ops[i]["mat"] < 0:
            continue
        dim = {0: 2, 3: 4, 4: 2}.get(int(ops[i]["kind"]))
        if dim is None:
            continue
        o = int(tb[0][i]["mat"])
        m = tb[1][o:o + 2 * dim * dim]
        row = np.zeros(width)
        row[:len(m)] = m
        upd.append(i)
        mats.append(row)
    upd = np.ascontiguousarray(upd, dtype=np.int64)
    mats = np.ascontiguousarray(np.asarray(mats), dtype=np.float64)
    tm = np.ascontiguousarray(ta[1], dtype=np.float64)
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1
    stats = np.zeros(6, dtype=np.int64)
    L = emu()
    L.emu_run_updated.restype = ctypes.c_int
    # every row holds one op's matrix; width bounds the matrix each update may read
    rc = L.emu_run_updated(ctypes.c_int(n), ctypes.c_int(0), ctypes.c_void_p(ops.ctypes.data), ctypes.c_size_t(len(ops)),
                           ctypes.c_void_p(tm.ctypes.data), ctypes.c_size_t(len(tm)), ctypes.c_int(KERNEL[kernel]),
                           ctypes.c_int(kw.get("tile", 0)), ctypes.c_int(kw.get("rtile", 0)), ctypes.c_int(0),
                           ctypes.c_void_p(psi.ctypes.data), ctypes.c_void_p(stats.ctypes.data),
                           ctypes.c_int(kw.get("btile", 0)), ctypes.c_int(kw.get("bq", 0)),
                           ctypes.c_void_p(upd.ctypes.data), ctypes.c_void_p(mats.ctypes.data),
                           ctypes.c_size_t(len(upd)), ctypes.c_size_t(width), ctypes.c_size_t(width))
    if rc:
        buf = ctypes.create_string_buffer(512)
        L.emu_last_error(buf, 512)
        raise AssertionError(buf.value.decode())
    ref = dense_ref(n, cb, prune=False)
    assert np.max(np.abs(psi - ref)) < TOL
    # the un-updated program still reproduces circuit A
    got_a, _ = run_emu(n, ca, prune=False, kernel=kernel, **kw)
    assert np.max(np.abs(got_a - dense_ref(n, ca, prune=False))) < TOL
