"""The JavaScript restatement of ket's sparse Map<bigint> algorithm
(oracle/ket_sparse.js, the ket-algorithm CPU baseline) against the C oracle:
under ket's per-gate pruning both compute every amplitude with the same
operations in the same order, so the final states are equal bit for bit, and
the run() fast path gives the same seeded counts.  Runs without a GPU; skipped
where node is not installed."""
import math
import random

import numpy as np
import pytest

import ketsv_oracle as O
import ket_sparse as K
from helpers import build, random_calls
from ket_amd import workloads
from ket_amd.algorithms import qft

pytestmark = pytest.mark.skipif(K.node() is None, reason="node is not installed")


def _oracle(n, ops, mats):
    psi = O.zero_state(n)
    O.apply_tape(n, psi, ops, mats, prune=True)
    return psi


def test_random_layers_12q_states_and_counts_equal_the_oracle():
    """The 12-qubit 20-layer rotation + CNOT-brick circuit (830 gates): amplitudes
    bit for bit, and 10 000 seeded shots with the same counts in the same
    first-seen order as the oracle's fast-path sampler."""
    n, seed, shots = 12, 42, 10000
    ops, mats = workloads.random_layers(n, 20, seed=seed).tape()
    r = K.run(n, ops, mats, task="run", shots=shots, seed=seed, warmup=0, runs=1, dump=True)
    psi = _oracle(n, ops, mats)
    assert np.array_equal(r["amplitudes"], psi)
    assert r["entries"] == np.count_nonzero(psi)
    draws = O.sample(n, psi, O.prng(seed, shots))
    want = {}
    for i in draws:
        want[int(i)] = want.get(int(i), 0) + 1
    assert [(int(k), c) for k, c in r["counts"]] == list(want.items())


@pytest.mark.parametrize("seed", range(4))
def test_every_gate_kind_bit_for_bit(seed):
    """Random circuits over every gate method (controlled, two-qubit, Toffoli,
    cswap, csrswap, 1-3 qubit unitaries)."""
    n = 9
    ops, mats = build(n, random_calls(n, 160, seed=seed)).tape()
    r = K.run(n, ops, mats, warmup=0, runs=1, dump=True)
    assert np.array_equal(r["amplitudes"], _oracle(n, ops, mats))


def test_pruning_heavy_circuit_bit_for_bit():
    """Small rotation angles put most amplitudes below the 1e-14 threshold; the
    two restatements prune the same entries."""
    n = 10
    rnd = random.Random(3)
    calls = []
    for q in range(n):
        s = rnd.choice([1e-4, 3e-3, 2e-5, 1.0])
        calls += [("rx", (s * rnd.uniform(-1, 1), q)), ("ry", (s * rnd.uniform(-1, 1), q))]
    calls += [("cnot", (a, a + 1)) for a in range(n - 1)] + [("crx", (1e-6, 0, n - 1)), ("h", (n // 2,))]
    ops, mats = build(n, calls).tape()
    r = K.run(n, ops, mats, warmup=0, runs=1, dump=True)
    psi = _oracle(n, ops, mats)
    assert np.count_nonzero(psi == 0) > (1 << n) // 4
    assert np.array_equal(r["amplitudes"], psi)


def test_qft_and_uniform_closed_forms():
    """QFT(6) on |0> and H on every qubit: 2^n equal amplitudes 2^(-n/2)."""
    for c in (qft(6), build(6, [("h", (q,)) for q in range(6)])):
        ops, mats = c.tape()
        r = K.run(6, ops, mats, warmup=0, runs=1, dump=True)
        assert r["entries"] == 64
        assert np.max(np.abs(r["amplitudes"] - 1 / 8)) < 1e-15
    assert math.isfinite(r["median_ms"])
