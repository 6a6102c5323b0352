"""Seeded circuit generators of the benchmark configurations.

random_layers: each layer applies Rx, Ry, Rz on every qubit (angles 2*pi*r,
r drawn from ket's xorshift32 in layer -> qubit -> (x, y, z) order), then a
CNOT brick: cnot(2i, 2i+1) on even layers, cnot(2i+1, 2i+2) on odd layers.
12 qubits x 20 layers = 830 gates; 30 qubits x 40 layers = 4180 gates.
"""
import math

from .algorithms import efficientSU2, maxCutHamiltonian, qaoa, qft, realAmplitudes
from .circuit import Circuit, make_prng


def random_layer_calls(n, layers, seed=42):
    """The circuit as (method, args) calls, in application order."""
    rng = make_prng(seed)
    calls = []
    for l in range(layers):
        for q in range(n):
            tx, ty, tz = 2 * math.pi * rng(), 2 * math.pi * rng(), 2 * math.pi * rng()
            calls.append(("rx", (tx, q)))
            calls.append(("ry", (ty, q)))
            calls.append(("rz", (tz, q)))
        start = 0 if l % 2 == 0 else 1
        for a in range(start, n - 1, 2):
            calls.append(("cnot", (a, a + 1)))
    return calls


def build(n, calls):
    c = Circuit(n)
    for name, args in calls:
        c = getattr(c, name)(*args)
    return c


def random_layers(n, layers, seed=42):
    return build(n, random_layer_calls(n, layers, seed))


def qft_input(n, seed=7):
    """(qft(n), q0-leftmost bitstring of x = floor(2^n r))."""
    r = make_prng(seed)()
    x = int(math.floor((1 << n) * r))
    bs = format(x, "b").zfill(n)[::-1]
    return qft(n), bs, x


def tfim_hamiltonian(n, hx=0.7):
    """H = -sum Z_i Z_{i+1} - hx sum X_i over an open chain (PauliTerm convention)."""
    terms = []
    for i in range(n - 1):
        arr = ["I"] * n
        arr[n - 1 - i] = "Z"
        arr[n - 1 - (i + 1)] = "Z"
        terms.append({"coeff": -1.0, "ops": "".join(arr)})
    for i in range(n):
        arr = ["I"] * n
        arr[n - 1 - i] = "X"
        terms.append({"coeff": -hx, "ops": "".join(arr)})
    return terms


def ansatz_params(count, seed=1234):
    rng = make_prng(seed)
    return [2 * math.pi * rng() for _ in range(count)]


def qaoa_instances(count=512, seed=2026, nmin=14, nmax=22, p=4, edge_prob=0.3):
    """Seeded Erdos-Renyi Max-Cut QAOA instances: list of (n, edges, gamma, beta)."""
    rng = make_prng(seed)
    out = []
    for _ in range(count):
        n = nmin + int(rng() * (nmax - nmin + 1))
        edges = [(u, v) for u in range(n) for v in range(u + 1, n) if rng() < edge_prob]
        gamma = [math.pi * rng() for _ in range(p)]
        beta = [math.pi * rng() for _ in range(p)]
        out.append((n, edges, gamma, beta))
    return out


__all__ = ["random_layer_calls", "random_layers", "build", "qft_input", "tfim_hamiltonian", "ansatz_params",
           "qaoa_instances", "efficientSU2", "realAmplitudes", "qaoa", "maxCutHamiltonian"]
