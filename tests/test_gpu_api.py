"""Public calls not exercised elsewhere, on the device: minimize() (ket's own
cases, src/circuit.test.ts:4237-4320), iqft() (the inverse of qft() and the
conjugate of its amplitudes), stateAsArray() (src/circuit.ts:1579-1593),
exactProbsTop() above ket's 24-qubit map limit, and gradient()'s automatic
choice between the adjoint method and ket's shift rule."""
import math

import numpy as np
import pytest

import ketsv_oracle as O
from helpers import build, random_calls
from ket_amd import Circuit, workloads
from ket_amd._native import State
from ket_amd.algorithms import efficientSU2, gradient, iqft, minimize, qaoa, qft, realAmplitudes, vqe

pytestmark = pytest.mark.gpu

Z1 = [{"coeff": 1, "ops": "Z"}]


def test_minimize_single_qubit_z():
    r = minimize(lambda p: Circuit(1).ry(p[0], 0), Z1, [0.1])
    assert abs(r["energy"] + 1) < 5e-5 and r["converged"]


def test_minimize_two_qubit_hamiltonian():
    """ZZ + 0.5 ZI: ground state |10>, energy -1.5."""
    H = [{"coeff": 1, "ops": "ZZ"}, {"coeff": 0.5, "ops": "ZI"}]
    r = minimize(lambda p: Circuit(2).ry(p[0], 0).ry(p[1], 1), H, [0.1, 0.1], lr=0.2, steps=500)
    assert abs(r["energy"] + 1.5) < 5e-4 and r["converged"]


def test_minimize_budget_and_fixed_point():
    r = minimize(lambda p: Circuit(1).ry(p[0], 0), Z1, [0.1], steps=1)
    assert r["converged"] is False and r["steps"] == 1
    r = minimize(lambda p: Circuit(1).ry(p[0], 0), Z1, [math.pi])
    assert r["converged"] is True and r["steps"] == 0


@pytest.mark.parametrize("make,n,H,kw", [(realAmplitudes, 1, Z1, {}),
                                         (realAmplitudes, 2, [{"coeff": 1, "ops": "ZZ"}], {"lr": 0.2, "steps": 500}),
                                         (efficientSU2, 1, Z1, {})])
def test_minimize_ansatz_families(make, n, H, kw):
    ans = make(n, 1)
    r = minimize(ans, H, [0.1] * ans.paramCount, **kw)
    assert abs(r["energy"] + 1) < 5e-4 and r["converged"]


@pytest.mark.parametrize("n", [3, 6, 12])
def test_iqft_conjugates_and_inverts_qft(n):
    """iqft = qft^dagger and the QFT matrix is symmetric, so iqft|x> is the complex
    conjugate of qft|x>; and iqft applied to qft|x> on the device gives |x> back."""
    rng = np.random.default_rng(n)
    x = int(rng.integers(0, 1 << n))
    bs = format(x, "b").zfill(n)[::-1]
    a = qft(n).statevectorDense(initialState=bs)
    b = iqft(n).statevectorDense(initialState=bs)
    assert np.max(np.abs(b - a.conj())) < 1e-12
    st = State(n, "f64", prune=False)
    st.set_amplitudes(a)
    st.apply(iqft(n).tape())
    back = st.amplitudes()
    st.free()
    want = np.zeros(1 << n, dtype=complex)
    want[x] = 1
    assert np.max(np.abs(back - want)) < 1e-12


def test_state_as_array():
    """Bell state (ket's doc example), then a seeded circuit: the entries with
    |a|^2 >= 1e-10, probability descending, against the oracle."""
    s2 = 1 / math.sqrt(2)
    got = Circuit(2).h(0).cnot(0, 1).stateAsArray()
    assert [e["bitstring"] for e in got] == ["00", "11"]
    for e in got:
        assert abs(e["re"] - s2) < 1e-15 and e["im"] == 0 and abs(e["prob"] - 0.5) < 1e-15 and e["phase"] == 0
    n = 7
    calls = random_calls(n, 50, seed=3)
    got = build(n, calls).stateAsArray()
    psi = O.zero_state(n)
    O.apply_named(n, psi, calls, prune=True)
    p = np.abs(psi) ** 2
    keep = [i for i in np.argsort(-p, kind="stable") if p[i] >= 1e-10]
    assert len(got) == len(keep)
    assert all(got[j]["prob"] >= got[j + 1]["prob"] for j in range(len(got) - 1))
    for e in got:
        i = int(e["bitstring"][::-1], 2)
        assert abs(complex(e["re"], e["im"]) - psi[i]) < 1e-12
        assert abs(e["phase"] - math.atan2(psi[i].imag, psi[i].real)) < 1e-9 or abs(psi[i]) < 1e-6


def test_exact_probs_top_above_the_map_limit():
    """26 qubits (past ket's 2^24-entry map): GHZ has two outcomes of 1/2; a product
    of ry rotations, each qubit leaning to 0 with odds r_q in [0.05, 0.1], has its
    most probable outcomes among the strings with at most two 1s (every 2-flip
    string beats every 3-flip one), where the closed form ranks them."""
    n = 26
    c = Circuit(n).h(0)
    for q in range(n - 1):
        c = c.cnot(q, q + 1)
    top = c.exactProbsTop(2)
    assert set(top) == {"0" * n, "1" * n} and all(abs(v - 0.5) < 1e-12 for v in top.values())
    rng = np.random.default_rng(26)
    r = rng.uniform(0.05, 0.1, n)
    th = 2 * np.arctan(np.sqrt(r))          # sin^2 / cos^2 of theta / 2 = r
    c = Circuit(n)
    for q in range(n):
        c = c.ry(float(th[q]), q)
    p0 = np.cos(th / 2) ** 2
    p1 = np.sin(th / 2) ** 2
    base = float(np.prod(p0))
    cands = {(): base}
    for i in range(n):
        cands[(i,)] = base * p1[i] / p0[i]
        for j in range(i + 1, n):
            cands[(i, j)] = base * p1[i] / p0[i] * p1[j] / p0[j]
    want = sorted(cands.items(), key=lambda kv: -kv[1])[:64]
    got = list(c.exactProbsTop(64).items())
    for (bits, pw), (bs, pg) in zip(want, got):
        assert bs == "".join("1" if q in bits else "0" for q in range(n))
        assert abs(pg - pw) < 1e-12 * max(1.0, pw) + 1e-15


def test_gradient_auto_uses_the_shift_rule_where_it_is_not_exact():
    """qaoa's gamma enters rz(-gamma) on every edge and beta enters rx(2 beta): ket's
    pi/2 shift rule is not the derivative there.  method='auto' gives ket's
    shift-rule numbers (equal to method='psr'), which differ from the exact
    gradient (method='adjoint', itself equal to central differences); on
    efficientSU2, where the rule is exact, 'auto' is the adjoint result."""
    n, p = 6, 2
    edges = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (5, 0)]
    ans = lambda prm: qaoa(n, edges, prm[:p], prm[p:])
    H = workloads.tfim_hamiltonian(n)
    params = [0.4, -0.3, 0.7, 0.2]
    auto = gradient(ans, H, params)
    psr = gradient(ans, H, params, method="psr")
    adj = gradient(ans, H, params, method="adjoint")
    assert auto == psr
    h = 1e-6
    fd = []
    for i in range(len(params)):
        up = [v + h if j == i else v for j, v in enumerate(params)]
        dn = [v - h if j == i else v for j, v in enumerate(params)]
        fd.append((vqe(ans(up), H) - vqe(ans(dn), H)) / (2 * h))
    assert np.max(np.abs(np.array(adj) - np.array(fd))) < 1e-7
    assert np.max(np.abs(np.array(auto) - np.array(adj))) > 1e-3
    e = efficientSU2(4, 1)
    prm = list(np.random.default_rng(1).uniform(-1, 1, e.paramCount))
    H4 = workloads.tfim_hamiltonian(4)
    assert gradient(e, H4, prm) == gradient(e, H4, prm, method="adjoint")
    assert np.max(np.abs(np.array(gradient(e, H4, prm)) - np.array(gradient(e, H4, prm, method="psr")))) < 1e-12
