"""Batched parameter sets (ksv_adjoint_batch_*): many parameter vectors of one
ansatz per evaluation, each on its own copy of one device state, against the
same sets evaluated one by one (energy_and_gradient / vqe, themselves checked
against the oracle, the shift rule and finite differences elsewhere)."""
import numpy as np
import pytest

import ket_amd
from ket_amd import workloads
from ket_amd.algorithms import (efficientSU2, energy_and_gradient, energy_and_gradient_batch, realAmplitudes, vqe,
                                vqe_batch)

from helpers import random_ansatz

pytestmark = pytest.mark.gpu


def _sets(count, nparam, seed):
    rng = np.random.default_rng(seed)
    return [list(rng.uniform(-np.pi, np.pi, nparam)) for _ in range(count)]


@pytest.mark.parametrize("n,reps,count", [(16, 2, 64), (12, 3, 5), (9, 1, 1)])
def test_batch_equals_single_runs_efficientsu2(n, reps, count):
    """B parameter sets of efficientSU2(n, reps) with the transverse-field Ising H:
    every energy and gradient equals the single-set adjoint to 1e-12 (a count that is
    not a power of two is padded)."""
    ket_amd.configure(prune=False)
    ans = efficientSU2(n, reps)
    H = workloads.tfim_hamiltonian(n)
    sets = _sets(count, 2 * n * (reps + 1), seed=n)
    es, gs = energy_and_gradient_batch(ans, H, sets)
    assert len(es) == count and len(gs) == count
    worst_e = worst_g = 0.0
    for ps, e, g in zip(sets, es, gs):
        e1, g1 = energy_and_gradient(ans, H, ps)
        worst_e = max(worst_e, abs(e - e1))
        worst_g = max(worst_g, float(np.max(np.abs(np.array(g) - np.array(g1)))))
    print(f"batch of {count} x efficientSU2({n},{reps}): max |d E| = {worst_e:.2e}, max |d grad| = {worst_g:.2e}")
    assert worst_e < 1e-12 and worst_g < 1e-12
    assert np.max(np.abs(np.array(vqe_batch(ans, H, sets)) - np.array(es))) < 1e-13


@pytest.mark.parametrize("n,seed", [(11, 3), (13, 4)])
def test_batch_random_ansatz_every_gate_kind(n, seed):
    """A random ansatz over every parametrised gate kind (shared parameters, affine
    angles) mixed with fixed gates of every kind, csrswap and 2-3-qubit unitaries
    included: 8 sets against single runs."""
    ket_amd.configure(prune=False)
    nparam = 6
    ans = random_ansatz(n, 90, nparam, seed, wide=False)
    H = workloads.tfim_hamiltonian(n)
    sets = _sets(8, nparam, seed)
    es, gs = energy_and_gradient_batch(ans, H, sets)
    for ps, e, g in zip(sets, es, gs):
        e1, g1 = energy_and_gradient(ans, H, ps)
        assert abs(e - e1) < 1e-12
        assert np.max(np.abs(np.array(g) - np.array(g1))) < 1e-11


def test_batch_recreated_when_a_diagonal_gate_stops_being_diagonal():
    """Sets whose rx / ry angles are all 0 make those gates diagonal at creation; a
    later call with other angles rebuilds the batch and still matches single runs;
    vqe_batch agrees with vqe."""
    ket_amd.configure(prune=False)
    n, reps = 8, 2
    ans = realAmplitudes(n, reps)
    H = workloads.tfim_hamiltonian(n)
    npar = n * (reps + 1)
    zero = [[0.0] * npar for _ in range(4)]
    es0, _ = energy_and_gradient_batch(ans, H, zero)
    assert np.max(np.abs(np.array(es0) - energy_and_gradient(ans, H, zero[0])[0])) < 1e-12
    sets = _sets(4, npar, seed=9)
    es, gs = energy_and_gradient_batch(ans, H, sets)
    for ps, e, g in zip(sets, es, gs):
        e1, g1 = energy_and_gradient(ans, H, ps)
        assert abs(e - e1) < 1e-12 and np.max(np.abs(np.array(g) - np.array(g1))) < 1e-12
        assert abs(vqe(ans(ps), H) - e) < 1e-12


def test_batch_complex64():
    """complex64 copies (FP32 arithmetic, FP64 accumulation) against FP64 single runs."""
    n, reps = 12, 2
    ans = efficientSU2(n, reps)
    H = workloads.tfim_hamiltonian(n)
    sets = _sets(16, 2 * n * (reps + 1), seed=4)
    ket_amd.configure(prune=False)
    ref = [energy_and_gradient(ans, H, ps) for ps in sets]
    ket_amd.configure(precision="f32", prune=False)
    es, gs = energy_and_gradient_batch(ans, H, sets)
    de = max(abs(e - r[0]) for e, r in zip(es, ref))
    dg = max(float(np.max(np.abs(np.array(g) - np.array(r[1])))) for g, r in zip(gs, ref))
    print(f"complex64 batch: max |d E| = {de:.2e}, max |d grad| = {dg:.2e}")
    assert de < 1e-4 and dg < 1e-4
