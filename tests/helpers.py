"""Shared test helpers: random circuits in ket's Circuit-method form."""
import math
import random

import numpy as np

import ket_amd

SINGLE0 = "id h x y z s si sdg t ti tdg v vi srn srndg r2 r4 r8".split()
SINGLE1 = "rx ry rz vz u1 p gpi gpi2".split()
TWO0 = "cnot cx swap iswap srswap cy cz ch cr2 cr4 cr8 cs ct csdg ctdg csrn".split()
TWO1 = "xx yy zz xy crx cry crz cu1".split()
THREE = "ccx cswap csrswap".split()


def random_calls(n, count, seed, kinds="all", with_unitary=True):
    """A seeded random circuit over every gate method of ket's Circuit."""
    rnd = random.Random(seed)
    calls = []
    for _ in range(count):
        r = rnd.random()
        ang = lambda: rnd.uniform(-2 * math.pi, 2 * math.pi)
        if r < 0.25 or n < 2:
            q = rnd.randrange(n)
            if rnd.random() < 0.5:
                calls.append((rnd.choice(SINGLE0), (q,)))
            else:
                name = rnd.choice(SINGLE1 + ["u2", "u3"])
                k = {"u2": 2, "u3": 3}.get(name, 1)
                calls.append((name, tuple(ang() for _ in range(k)) + (q,)))
        elif r < 0.65 or n < 3:
            a, b = rnd.sample(range(n), 2)
            if rnd.random() < 0.5:
                calls.append((rnd.choice(TWO0), (a, b)))
            else:
                name = rnd.choice(TWO1 + ["ms", "cu2", "cu3"])
                k = {"ms": 2, "cu2": 2, "cu3": 3}.get(name, 1)
                calls.append((name, tuple(ang() for _ in range(k)) + (a, b)))
        elif r < 0.9 or not with_unitary:
            a, b, c = rnd.sample(range(n), 3)
            calls.append((rnd.choice(THREE), (a, b, c)))
        else:
            k = rnd.choice([1, 2, 3]) if n >= 3 else 1
            qs = rnd.sample(range(n), k)
            calls.append(("unitary", (random_unitary(1 << k, rnd),) + tuple(qs)))
    return calls


def random_unitary(dim, rnd):
    rs = np.random.RandomState(rnd.randrange(1 << 30))
    z = rs.normal(size=(dim, dim)) + 1j * rs.normal(size=(dim, dim))
    q, r = np.linalg.qr(z)
    q = q * (np.diag(r) / np.abs(np.diag(r)))
    return [[complex(v) for v in row] for row in q]


def _add_calls(c, calls):
    for name, args in calls:
        if name == "unitary":
            c = c.unitary(args[0], *args[1:])
        elif name == "if":
            reg, value, inner = args
            c = c.if_(reg, value, lambda k, inner=inner: _add_calls(k, inner))
        else:
            c = getattr(c, name)(*args)
    return c


def build(n, calls, cregs=None):
    """ket's Circuit from (method, args) calls; ("if", (creg, value, inner_calls))
    becomes Circuit.if(creg, value, builder)."""
    c = ket_amd.Circuit(n)
    for name, size in (cregs or {}).items():
        c = c.creg(name, size)
    return _add_calls(c, calls)


def dense_ref(n, calls, prune=True, init=0):
    import ketsv_oracle as O
    psi = O.zero_state(n, init)
    O.apply_named(n, psi, calls, prune=prune)
    return psi


def shot_calls(n, count, seed, measures=2, resets=1):
    """Seeded random gates with mid-circuit measurements and resets (no `if`)."""
    rng = np.random.default_rng(seed)
    calls = random_calls(n, count, seed=seed)
    for j in range(measures):
        at = int(rng.integers(1, len(calls)))
        calls.insert(at, ("measure", (int(rng.integers(0, n)), "c", j)))
    for _ in range(resets):
        at = int(rng.integers(1, len(calls)))
        calls.insert(at, ("reset", (int(rng.integers(0, n)),)))
    return calls


def random_ansatz(n, count, nparam, seed, wide=True):
    """A seeded ansatz p -> Circuit over ket's parametrised gates (angles affine
    in the parameters, some shared between gates) mixed with fixed gates of
    every kind, including 3-qubit and (with `wide`) 4-qubit unitaries."""
    rnd = random.Random(seed)
    plan = []
    for _ in range(count):
        r = rnd.random()
        qs = rnd.sample(range(n), 4)
        if r < 0.45:
            name = rnd.choice(["rx", "ry", "rz", "p", "u1", "u3", "crx", "cry", "crz", "cu1", "cu3",
                               "xx", "yy", "zz", "xy", "ms"])
            k = {"u3": 3, "cu3": 3, "ms": 2}.get(name, 1)
            width = 1 if name in ("rx", "ry", "rz", "p", "u1", "u3") else 2
            angles = [(rnd.randrange(nparam), rnd.choice([1.0, -0.5, 2.0]), rnd.uniform(-1, 1)) for _ in range(k)]
            plan.append((name, angles, qs[:width]))
        elif r < 0.6:
            plan.append((rnd.choice(SINGLE0), None, qs[:1]))
        elif r < 0.75:
            plan.append((rnd.choice(TWO0), None, qs[:2]))
        elif r < 0.9:
            plan.append((rnd.choice(THREE), None, qs[:3]))
        else:
            k = rnd.choice([2, 3, 4] if wide else [2, 3])
            plan.append(("unitary", random_unitary(1 << k, rnd), qs[:k]))

    def ansatz(p):
        c = ket_amd.Circuit(n)
        for name, extra, qs in plan:
            if name == "unitary":
                c = c.unitary(extra, *qs)
            elif extra is None:
                c = getattr(c, name)(*qs)
            else:
                c = getattr(c, name)(*[coef * p[i] + off for i, coef, off in extra], *qs)
        return c
    return ansatz


def pauli_apply(n, psi, terms):
    """sum_t c_t P_t psi from PauliTerm dicts (ops[0] acts on qubit n-1), by
    applying the single-qubit Pauli matrices factor by factor."""
    mats = {"X": np.array([[0, 1], [1, 0]], complex), "Y": np.array([[0, -1j], [1j, 0]]),
            "Z": np.array([[1, 0], [0, -1]], complex)}
    out = np.zeros_like(psi)
    for t in terms:
        v = psi.reshape([2] * n)          # axis a = qubit n-1-a
        for a, ch in enumerate(t["ops"].upper()):
            if ch != "I":
                v = np.moveaxis(np.tensordot(mats[ch], v, axes=([1], [a])), 0, a)
        out += t["coeff"] * v.reshape(-1)
    return out


def if_calls(n, count, seed, noise_kinds=True):
    """A seeded random circuit with mid-circuit measurements into two multi-bit
    registers, resets, and `if` blocks (some nested) whose bodies hold gates,
    measurements and resets: the per-shot path's draw count then depends on the
    measured values."""
    rnd = random.Random(seed)
    regs = {"a": 3, "b": 2}

    def gates(k):
        return random_calls(n, k, seed=rnd.randrange(1 << 30), with_unitary=False)

    def measure():
        reg = rnd.choice(sorted(regs))
        return ("measure", (rnd.randrange(n), reg, rnd.randrange(regs[reg])))

    calls = []
    for _ in range(count):
        r = rnd.random()
        if r < 0.45:
            calls += gates(rnd.randint(1, 4))
        elif r < 0.7:
            calls.append(measure())
        elif r < 0.78:
            calls.append(("reset", (rnd.randrange(n),)))
        else:
            reg = rnd.choice(sorted(regs))
            body = gates(rnd.randint(1, 3))
            if rnd.random() < 0.4:
                body.append(measure())
            if rnd.random() < 0.3:
                reg2 = rnd.choice(sorted(regs))
                body.append(("if", (reg2, rnd.randrange(1 << regs[reg2]), gates(2))))
            calls.append(("if", (reg, rnd.randrange(1 << regs[reg]), body)))
    return calls, regs
