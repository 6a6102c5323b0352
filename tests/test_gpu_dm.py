"""Device density matrices (ksv_dm_*, Circuit.dm()) against the Map-order
restatement of ket's runDM (oracle/ket_dm.js): every entry bit for bit, the
Map order of the present entries identical, and the read-outs that ket sums in
Map order (purity, Bloch angles) bit for bit.  The restatement is pinned on the
CPU in test_ket_dm.py."""
import math

import numpy as np
import pytest

import ket_sparse as K
from helpers import build, random_calls, random_unitary
from ket_amd import Circuit
from ket_amd.circuit import RangeError
from ket_amd.density import DM_DEVICE_NOISE

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(K.node() is None, reason="node is not installed")]


def _both(c, noise=None):
    params = DM_DEVICE_NOISE[noise] if isinstance(noise, str) else (noise or {})
    ops, mats = c.tape()
    rho, order, info = K.run_dm(c.qubits, ops, mats, params.get("p1", 0.0), params.get("p2", 0.0))
    d = c.dm(noise=noise)
    return d, rho, order, info


def _exact(d, rho, order):
    assert np.array_equal(d._rho, rho)
    assert np.array_equal(d._order, order)


def _js_purity(rho, order):
    flat = rho.reshape(-1)
    p = 0.0
    for k in order.tolist():
        v = flat[k]
        p += v.real * v.real + v.imag * v.imag
    return p


def _js_bloch(n, rho, order, q):
    flat = rho.reshape(-1)
    m = 1 << q
    r00 = r11 = re = im = 0.0
    for k in order.tolist():
        r, c, v = k >> n, k & ((1 << n) - 1), flat[k]
        if r == c:
            if r & m == 0:
                r00 += v.real
            else:
                r11 += v.real
        elif r & m == 0 and c == (r | m):
            re += v.real
            im += v.imag
    rz, rx, ry = r00 - r11, 2 * re, -2 * im
    return {"theta": math.acos(max(-1.0, min(1.0, rz))), "phi": math.atan2(ry, rx)}


@pytest.mark.parametrize("n,seed", [(3, 1), (4, 2), (5, 3), (6, 4)])
@pytest.mark.parametrize("noise", [None, "harmony", "aria-1", {"p1": 0.05, "p2": 0.1}])
def test_dm_bit_for_bit(n, seed, noise):
    """Every op kind (1-3 qubit unitaries, toffoli, cswap, csrswap with a superposed
    control) with each noise setting."""
    calls = random_calls(n, 40, seed=seed) + [("h", (0,)), ("csrswap", (0, 1, 2))]
    d, rho, order, info = _both(build(n, calls), noise)
    print(f"n={n} seed={seed} noise={noise}: {len(order)} entries, {info['dropped']} running sums dropped")
    _exact(d, rho, order)
    assert d.purity() == _js_purity(rho, order)
    for q in range(n):
        assert d.blochAngles(q) == _js_bloch(n, rho, order, q)


@pytest.mark.parametrize("k", [4, 5, 6])
def test_dm_wide_unitaries(k):
    """Unitaries on 4, 5 and 6 qubits (one workgroup per group through LDS) between
    noisy two-qubit gates."""
    import random
    rnd = random.Random(k)
    n = 6
    calls = [("h", (q,)) for q in range(n)] + [("cnot", (0, 1)), ("rx", (0.3, 2))]
    calls.append(("unitary", (random_unitary(1 << k, rnd),) + tuple(rnd.sample(range(n), k))))
    calls += [("cz", (2, 3)), ("unitary", (random_unitary(4, rnd), 5, 1)), ("swap", (4, 0))]
    d, rho, order, _ = _both(build(n, calls), "harmony")
    _exact(d, rho, order)


def test_dm_long_circuit_renumbers_the_order():
    """300 gates on 5 qubits with noise: the order keys outgrow 62 bits many times and
    are renumbered on the device; the order stays ket's."""
    n = 5
    calls = random_calls(n, 300, seed=11, with_unitary=False)
    d, rho, order, info = _both(build(n, calls), {"p1": 0.002, "p2": 0.02})
    assert info["dropped"] > 0
    _exact(d, rho, order)


def test_dm_eight_qubits_harmony():
    n = 8
    calls = random_calls(n, 60, seed=8, with_unitary=False)
    d, rho, order, _ = _both(build(n, calls), "harmony")
    _exact(d, rho, order)
    assert d.purity() == _js_purity(rho, order)


def test_dm_partial_sum_pruning_case():
    """The case worked out by hand in test_ket_dm.py: rho_11 is (1 - p) c^2 plus three
    p/15 c^2, the four p/15 s^2 copies dropped."""
    th = 2 * math.asin(1e-3)
    c = Circuit(2).x(0).ry(th, 0).cz(0, 1)
    c2 = c.dm()._rho[1, 1].real
    d = c.dm(noise={"p2": 0.015})
    p = 0.015
    want = (1 - p) * c2
    for _ in range(3):
        want = want + p / 15 * c2
    assert d._rho[1, 1].real == want


def test_dm_limits():
    with pytest.raises(RangeError):
        Circuit(16).dm()
    rnd = __import__("random").Random(1)
    with pytest.raises(RangeError):
        Circuit(7).unitary(random_unitary(128, rnd), 0, 1, 2, 3, 4, 5, 6).dm()
