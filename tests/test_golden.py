"""Known answers of ket's own statevector tests (tests/golden/known_answers.json,
written by tests/golden/make_known_answers.py with the file:line of each):
the CPU oracle reproduces every case (no GPU), and the product path through
ket's Circuit API reproduces every case on the GPU."""
import json
import math
import os

import numpy as np
import pytest

import ketsv_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))["cases"]


def _bind(calls, params):
    out = []
    for name, args in calls:
        a = tuple(params[int(x[1:])] if isinstance(x, str) and x.startswith("$") else x for x in args)
        out.append((name, a))
    return out


def _calls(calls):
    """JSON calls -> (name, args) tuples; an `if` carries its inner calls."""
    out = []
    for name, args in calls:
        if name == "if":
            out.append(("if", (args[0], args[1], _calls(args[2]))))
        else:
            out.append((name, tuple(args)))
    return out


def _check_run(case, counts, creg_counts, shots):
    """counts {index: n}: the outcome frequencies ket's test asserts."""
    n, want = case["n"], case["expected"]
    probs = {format(int(i), "b").zfill(n)[::-1]: c / shots for i, c in counts.items()}
    if "probs" in want:
        for bs, p in want["probs"].items():
            assert abs(probs.get(bs, 0.0) - p) < case["tol"]
    if "bit" in want:
        pos, val, p, tol = want["bit"]
        got = sum(v for bs, v in probs.items() if bs[pos] == val)
        assert abs(got - p) <= tol
    for reg, fr in want.get("cregs", {}).items():
        assert np.max(np.abs(np.array(creg_counts[reg]) / shots - np.array(fr))) < case["tol"]


def _init_index(case):
    bs = case.get("initialState")
    return int(bs[::-1], 2) if bs else 0


def _oracle_state(case, calls=None):
    n = case["n"]
    psi = O.zero_state(n, _init_index(case))
    O.apply_named(n, psi, [(nm, tuple(a)) for nm, a in (calls if calls is not None else case["calls"])], prune=False)
    return psi


def _oracle_energy(case, params):
    psi = _oracle_state(case, _bind(case["calls"], params))
    return O.vqe(case["n"], psi, case["hamiltonian"])


@pytest.mark.parametrize("case", CASES, ids=[c["id"] for c in CASES])
def test_oracle_reproduces_known_answer(case):
    n, q, tol = case["n"], case["query"], case["tol"]
    if q == "amplitudes":
        psi = _oracle_state(case)
        for k, (re, im) in case["expected"].items():
            assert abs(psi[int(k)] - complex(re, im)) < tol
        if len(case["expected"]) < (1 << n):
            rest = np.delete(np.arange(1 << n), [int(k) for k in case["expected"]])
            assert np.max(np.abs(psi[rest])) < tol
    elif q == "exact_probs":
        idx, p = O.probabilities(n, _oracle_state(case))
        got = {format(int(i), "b").zfill(n)[::-1]: float(v) for i, v in zip(idx, p)}
        assert set(got) == set(case["expected"])
        for k, v in case["expected"].items():
            assert abs(got[k] - v) < tol
    elif q == "expectation":
        assert abs(O.expectation(n, _oracle_state(case), case["pauli"]) - case["expected"]) < tol
    elif q == "marginals":
        assert np.max(np.abs(O.marginals(n, _oracle_state(case)) - np.array(case["expected"]))) < tol
    elif q == "run":
        counts, creg_counts = O.run_shots(n, _calls(case["calls"]), case["shots"], case["seed"],
                                          cregs=case["cregs"])
        _check_run(case, counts, creg_counts, case["shots"])
    elif q == "gradient":
        ps = case["params"]
        for i, want in enumerate(case["expected"]):
            plus = list(ps)
            minus = list(ps)
            plus[i] += math.pi / 2
            minus[i] -= math.pi / 2
            g = 0.5 * (_oracle_energy(case, plus) - _oracle_energy(case, minus))
            assert abs(g - want) < tol
    else:
        raise AssertionError(f"unknown query {q}")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[c["id"] for c in CASES])
def test_product_path_reproduces_known_answer(case):
    import ket_amd
    from ket_amd.algorithms import gradient
    from helpers import build
    n, q, tol = case["n"], case["query"], case["tol"]
    init = case.get("initialState")
    if q == "gradient":
        calls = case["calls"]
        got = gradient(lambda ps: build(n, _bind(calls, ps)), case["hamiltonian"], case["params"])
        assert np.max(np.abs(np.array(got) - np.array(case["expected"]))) < tol
        return
    if q == "run":
        c = build(n, _calls(case["calls"]), cregs=case["cregs"])
        d = c.run(shots=case["shots"], seed=case["seed"])
        shots = case["shots"]
        counts = {int(k): round(v * shots) for k, v in d.histogram.items()}
        creg_counts = {k: [round(v * shots) for v in fr] for k, fr in d.cregs.items()}
        _check_run(case, counts, creg_counts, shots)
        want = O.run_shots(n, _calls(case["calls"]), shots, case["seed"], cregs=case["cregs"])
        assert list(counts.items()) == list(want[0].items())       # identical counts, first-seen order
        assert creg_counts == want[1]
        return
    c = build(n, [(nm, tuple(a)) for nm, a in case["calls"]])
    if q == "amplitudes":
        sv = c.statevector(initialState=init) if init else c.statevector()
        for k, (re, im) in case["expected"].items():
            assert abs(sv.get(int(k), 0) - complex(re, im)) < tol
    elif q == "exact_probs":
        got = c.exactProbs()
        assert set(got) == set(case["expected"])
        for k, v in case["expected"].items():
            assert abs(got[k] - v) < tol
    elif q == "expectation":
        assert abs(c.expectation(case["pauli"]) - case["expected"]) < tol
    elif q == "marginals":
        assert np.max(np.abs(np.array(c.marginals()) - np.array(case["expected"]))) < tol
