"""Batched copies in per-copy mode (ksv_shots_run with a copy that fits one
tile: register rounds for 11..13 qubits, an LDS pass below): the plan of one
gate segment with Pauli events on chosen copies,
executed by the host emulator (tests/emu), against the oracle run copy by
copy with the Paulis inserted where they fall.  Runs without a GPU."""
import ctypes

import numpy as np
import pytest

import ketsv_oracle as O
from ket_amd._native import OP_DTYPE
from helpers import build, random_calls
from test_schedule_emu import emu

PAULI = {1: "x", 2: "y", 3: "z"}


def _lib():
    L = emu()
    L.emu_copies.restype = ctypes.c_int
    L.emu_copies.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                             ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p]
    return L


@pytest.mark.parametrize("nsys,b,seed,prune", [(6, 3, 1, True), (8, 2, 2, True), (7, 4, 3, False), (5, 5, 4, True),
                                               (11, 2, 5, True), (12, 2, 6, True), (13, 1, 7, False),
                                               (11, 6, 8, True)])           # 64 copies: two bitmap words per op
def test_copies_with_pauli_events(nsys, b, seed, prune):
    rng = np.random.default_rng(seed)
    calls = random_calls(nsys, 60, seed=seed)
    ops, mats = build(nsys, calls).tape()
    assert len(ops) == len(calls)
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    mats = np.ascontiguousarray(mats, dtype=np.float64)
    copies = 1 << b
    # events: (at, q) positions, each with some copies and their Pauli kinds
    ents = []
    for at in sorted(rng.choice(len(ops) + 1, size=12, replace=False)):
        for q in sorted(set(rng.integers(0, nsys, size=2).tolist())):
            for c in rng.choice(copies, size=int(rng.integers(1, copies + 1)), replace=False):
                ents.append((int(at), int(q), int(rng.integers(1, 4)), int(c)))
    ents.sort(key=lambda e: (e[0], e[1], e[3]))
    at = np.array([e[0] for e in ents], dtype=np.uint64)
    qq = np.array([e[1] for e in ents], dtype=np.int32)
    pk = np.array([e[2] for e in ents], dtype=np.int32)
    cp = np.array([e[3] for e in ents], dtype=np.uint32)
    init = rng.integers(0, 1 << nsys, size=copies)
    psi = np.zeros(copies << nsys, dtype=np.complex128)
    for c in range(copies):
        psi[(c << nsys) | int(init[c])] = 1
    stats = np.zeros(2, dtype=np.int32)
    L = _lib()
    rc = L.emu_copies(nsys + b, nsys, 1 if prune else 0, ops.ctypes.data, len(ops), mats.ctypes.data, len(mats),
                      at.ctypes.data, qq.ctypes.data, pk.ctypes.data, cp.ctypes.data, len(ents), psi.ctypes.data,
                      stats.ctypes.data)
    if rc:
        buf = ctypes.create_string_buffer(512)
        L.emu_last_error(buf, 512)
        raise AssertionError(f"emulator: rc={rc}: {buf.value.decode()}")
    assert stats[0] == 1 and stats[1] == nsys        # one pass, a copy per tile
    for c in range(copies):
        seq = []
        mine = [e for e in ents if e[3] == c]
        k = 0
        for i, call in enumerate(calls + [None]):
            while k < len(mine) and mine[k][0] == i:
                seq.append((PAULI[mine[k][2]], (mine[k][1],)))
                k += 1
            if call is not None:
                seq.append(call)
        want = O.zero_state(nsys, int(init[c]))
        O.apply_named(nsys, want, seq, prune=prune)
        got = psi[c << nsys:(c + 1) << nsys]
        assert np.max(np.abs(got - want)) < 1e-12, c
