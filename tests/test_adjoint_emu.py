"""The adjoint backward sweep (tile passes over phi and lambda) executed by
the host emulator (tests/emu) with the plan libketsv uploads: gradients
against central finite differences of the oracle's energy, phi returned to
the initial state, every index and record slot checked.  Runs without a GPU."""
import ctypes

import numpy as np
import pytest

import ketsv_oracle as O
from ket_amd import workloads
from ket_amd._native import OP_DTYPE, PARAM_DTYPE
from ket_amd.algorithms import efficientSU2, pack, trace
from helpers import pauli_apply, random_ansatz
from test_schedule_emu import emu


def _lib():
    L = emu()
    L.emu_adjoint.restype = ctypes.c_int
    L.emu_adjoint.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                              ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_void_p, ctypes.c_void_p]
    return L


def _tape(ansatz, params):
    flat, records, _, n = trace(ansatz, params)
    ops, mats = pack(flat)
    mats = list(mats)
    rec = np.zeros(len(records), dtype=PARAM_DTYPE)
    for j, (oi, pi, coeff, d) in enumerate(records):
        rec[j] = (oi, pi, coeff, len(mats))
        for v in d:
            mats += [complex(v).real, complex(v).imag]
    return n, np.ascontiguousarray(ops, dtype=OP_DTYPE), np.asarray(mats, dtype=np.float64), rec


def _energy(ansatz, params, H):
    n, ops, mats, _ = _tape(ansatz, params)
    psi = O.zero_state(n)
    O.apply_tape(n, psi, ops, mats, prune=False)
    return float(np.vdot(psi, pauli_apply(n, psi, H)).real)


def emu_gradient(ansatz, params, H, K):
    n, ops, mats, rec = _tape(ansatz, params)
    psi = O.zero_state(n)
    O.apply_tape(n, psi, ops, mats, prune=False)
    lam = np.ascontiguousarray(pauli_apply(n, psi, H))
    phi = np.ascontiguousarray(psi.astype(np.complex128))
    sums = np.zeros(max(len(rec), 1))
    stats = np.zeros(5, dtype=np.int32)
    L = _lib()
    rc = L.emu_adjoint(n, K, ops.ctypes.data, len(ops), mats.ctypes.data, len(mats), rec.ctypes.data, len(rec),
                       phi.ctypes.data, lam.ctypes.data, sums.ctypes.data, stats.ctypes.data)
    if rc:
        buf = ctypes.create_string_buffer(512)
        L.emu_last_error(buf, 512)
        raise AssertionError(f"emulator: rc={rc}: {buf.value.decode()}")
    grad = np.zeros(len(params))
    for j in range(len(rec)):
        grad[rec[j]["param"]] += 2 * rec[j]["coeff"] * sums[j]
    return grad, phi, stats


def _fd(ansatz, params, H, idx, h=1e-5):
    out = []
    for i in idx:
        pp = list(params); pp[i] += h
        pm = list(params); pm[i] -= h
        out.append((_energy(ansatz, pp, H) - _energy(ansatz, pm, H)) / (2 * h))
    return np.array(out)


@pytest.mark.parametrize("n,seed", [(9, 1), (11, 2), (12, 3)])
def test_random_ansatz_gradient(n, seed):
    """Every parametrised gate kind (single, controlled, two-qubit; u3 and cu3 with
    three records, shared parameters) among fixed gates of every kind, 3-qubit
    unitaries in passes and 4-qubit ones as wide steps; K = 8 so tiles, global
    controls and several passes are exercised."""
    nparam = 7
    ans = random_ansatz(n, 90, nparam, seed)
    params = list(np.random.default_rng(seed).uniform(-2, 2, nparam))
    H = workloads.tfim_hamiltonian(n)
    g, phi, stats = emu_gradient(ans, params, H, K=8)
    assert stats[0] >= 2 and stats[4] > 0
    fd = _fd(ans, params, H, range(nparam))
    assert np.max(np.abs(g - fd)) < 1e-7, (g, fd)
    want = np.zeros_like(phi)
    want[0] = 1
    assert np.max(np.abs(phi - want)) < 1e-12


def test_efficient_su2_plan():
    """efficientSU2(14, 3) with 11-qubit tiles: far fewer passes than gates,
    gradient = finite differences on a sample of parameters."""
    n, reps = 14, 3
    ans = efficientSU2(n, reps)
    params = list(workloads.ansatz_params(2 * n * (reps + 1), seed=5))
    H = workloads.tfim_hamiltonian(n)
    g, phi, stats = emu_gradient(ans, params, H, K=11)
    ngates = 2 * n * (reps + 1) + (n - 1) * reps
    assert stats[1] == 0 and stats[0] * 8 < ngates, stats
    # rz follows ry on every qubit: chains of two, the rz half diagonal
    assert stats[3] >= n * (reps + 1) // 2 and stats[4] >= n * (reps + 1) // 2, stats
    idx = [0, 5, 17, 40, 63, len(params) - 1]
    fd = _fd(ans, params, H, idx)
    assert np.max(np.abs(g[idx] - fd)) < 1e-7
    assert abs(phi[0] - 1) < 1e-12


def test_chain_groups_share_round_trips():
    """Adjacent one-qubit chains on distinct targets share one LDS round trip
    (group records): efficientSU2(24, 24) on 11-qubit tiles takes 260 kernel
    steps for its 1752 ops (625 as separate chains); a smaller case with groups
    keeps its gradient equal to finite differences (test_efficient_su2_plan)."""
    L = _lib()
    L.emu_adjoint_plan.restype = ctypes.c_int
    L.emu_adjoint_plan.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                   ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    n, reps = 24, 24
    n_, ops, mats, rec = _tape(efficientSU2(n, reps), list(workloads.ansatz_params(2 * n * (reps + 1), seed=1234)))
    st = np.zeros(4, dtype=np.int32)
    rc = L.emu_adjoint_plan(n_, 11, ops.ctypes.data, len(ops), mats.ctypes.data, len(mats), rec.ctypes.data, len(rec),
                            st.ctypes.data)
    assert rc == 0 and st[1] == 0
    assert st[3] <= 300 and st[3] * 5 < len(ops), st
