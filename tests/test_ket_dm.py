"""The JavaScript restatement of ket's runDM over a Map (oracle/ket_dm.js), the
reference the device density matrices are checked against bit for bit, pinned
on the CPU: against the dense numpy restatement (oracle run_dm) where no
partial sum is pruned, within the pruned mass where one is, against a case
worked out by hand in which dmAcc's pruning changes an entry, and against
the Map order ket's loops produce on a small case.  Skipped without node."""
import math

import numpy as np
import pytest

import ketsv_oracle as O
import ket_sparse as K
from helpers import build, random_calls
from ket_amd import Circuit

pytestmark = pytest.mark.skipif(K.node() is None, reason="node is not installed")


def _js(c, noise=None):
    ops, mats = c.tape()
    noise = noise or {}
    return K.run_dm(c.qubits, ops, mats, noise.get("p1", 0.0), noise.get("p2", 0.0))


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("noise", [None, {"p1": 0.03}])
def test_equals_dense_restatement_without_partial_sum_pruning(seed, noise):
    """Every op kind, csrswap with its control in superposition included, with no
    noise or single-qubit noise only (no dmAcc): equal to the dense restatement."""
    n = 4
    calls = random_calls(n, 30, seed=seed) + [("h", (0,)), ("csrswap", (0, 1, 2))]
    rho, order, info = _js(build(n, calls), noise)
    assert info["dropped"] == 0
    assert np.max(np.abs(rho - O.run_dm(n, calls, noise))) < 1e-12
    assert len(order) == np.count_nonzero(rho) == info["entries"]
    assert len(set(order.tolist())) == len(order)


@pytest.mark.parametrize("seed", range(3))
def test_two_qubit_noise_within_the_pruned_mass(seed):
    """harmony noise (p2 = 0.015): dmAcc drops running sums below 1e-7; the result
    stays within the dropped mass of the dense restatement's whole sums."""
    n = 4
    calls = random_calls(n, 30, seed=seed, with_unitary=False)
    rho, _, info = _js(build(n, calls), {"p1": 0.001, "p2": 0.015})
    ref = O.run_dm(n, calls, {"p1": 0.001, "p2": 0.015})
    assert info["dropped"] > 0
    assert np.max(np.abs(rho - ref)) <= info["dropped_abs"] + 1e-12
    assert abs(np.trace(rho).real - 1) < 1e-6


def _pruned_case():
    """x(0), ry(theta, 0) with sin(theta / 2) = 1e-3, then cz(0, 1) with p2 = 0.015."""
    th = 2 * math.asin(1e-3)
    return Circuit(2).x(0).ry(th, 0).cz(0, 1), {"p2": 0.015}


def test_partial_sum_pruning_worked_case():
    """After x and ry, rho has rho_00 = s^2 ~ 1e-6 first in the Map and rho_11 = c^2
    last; cz keeps that order.  depolarize2 walks rho_00 first: its four copies
    into entry (1,1) weigh p/15 s^2 ~ 1e-9 each and arrive while (1,1) does not
    exist yet, so dmAcc drops every one; rho_11 then gets (1 - p) c^2 + three
    times p/15 c^2 (the I(x)Z, Z(x)I, Z(x)Z copies of itself), in that order.  Summing
    whole would add 4 p/15 s^2 more."""
    c, noise = _pruned_case()
    pure, order0, _ = _js(c)
    s2, c2 = pure[0, 0].real, pure[1, 1].real
    assert abs(s2 - 1e-6) < 1e-15 and abs(c2 - (1 - 1e-6)) < 1e-15
    assert order0.tolist()[0] == 0 and order0.tolist()[-1] == 5        # (0,0) first, (1,1) = key 5 last
    rho, _, info = _js(c, noise)
    p = noise["p2"]
    keep, w = 1 - p, p / 15
    want = keep * c2
    for _ in range(3):
        want = want + w * c2
    assert rho[1, 1].real == want and rho[1, 1].imag == 0.0
    whole = O.run_dm(2, [("x", (0,)), ("ry", (2 * math.asin(1e-3), 0)), ("cz", (0, 1))], noise)
    assert abs(whole[1, 1].real - (want + 4 * w * s2)) < 1e-15
    assert info["dropped"] >= 4


def test_map_order_of_a_single_gate():
    """h on qubit 1 of |00><00|: ket's applySingle meets the one context first and
    writes (r0,c0), (r0,c1), (r1,c0), (r1,c1), i.e. keys 0, 2, 8, 10; a cnot keeps
    the order and moves the keys."""
    rho, order, _ = _js(Circuit(2).h(1))
    assert order.tolist() == [0, 2, 8, 10]
    assert np.allclose(rho[np.ix_([0, 2], [0, 2])], 0.5, atol=1e-15)
    _, order, _ = _js(Circuit(2).h(1).cnot(1, 0))
    assert order.tolist() == [0, 3, 12, 15]
