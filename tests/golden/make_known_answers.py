"""Writes known_answers.json: the known answers that ket's own tests assert
for the statevector path, as data (inputs and expected outputs), each with
the reference file:line it comes from.  ket's tests hold no fixture files;
their answers are closed forms written inline, restated here numerically.

Run: python tests/golden/make_known_answers.py
"""
import cmath
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
S2 = 1 / math.sqrt(2)


def cx(z):
    return [z.real, z.imag]


cases = []

# src/circuit.test.ts:3917-3929 -- n = 3 QFT on |1> (initialState '100')
qft3 = [["h", [2]], ["cu1", [math.pi / 2, 1, 2]], ["cu1", [math.pi / 4, 0, 2]],
        ["h", [1]], ["cu1", [math.pi / 2, 0, 1]], ["h", [0]], ["swap", [0, 2]]]
cases.append({"id": "qft3_on_1", "source": "src/circuit.test.ts:3917-3929", "n": 3, "calls": qft3,
              "initialState": "100", "query": "amplitudes",
              "expected": {str(k): cx(cmath.exp(2j * math.pi * k / 8) / math.sqrt(8)) for k in range(8)},
              "tol": 1e-10})

# src/circuit.test.ts:5603-5609 -- X on q0 from '01' -> |11>
cases.append({"id": "initial_state_x", "source": "src/circuit.test.ts:5603-5609", "n": 2, "calls": [["x", [0]]],
              "initialState": "01", "query": "amplitudes", "expected": {"3": [1.0, 0.0]}, "tol": 1e-8})

# Bell and GHZ exact probabilities, src/circuit.test.ts:270-313 (probabilities 1/2 on 0...0 and 1...1)
cases.append({"id": "bell_probs", "source": "src/circuit.test.ts:270-313", "n": 2,
              "calls": [["h", [0]], ["cnot", [0, 1]]], "query": "exact_probs",
              "expected": {"00": 0.5, "11": 0.5}, "tol": 1e-12})
ghz = [["h", [0]]] + [["cnot", [i, i + 1]] for i in range(5)]
cases.append({"id": "ghz6_probs", "source": "src/circuit.test.ts:270-313", "n": 6, "calls": ghz,
              "query": "exact_probs", "expected": {"000000": 0.5, "111111": 0.5}, "tol": 1e-12})

# src/circuit.test.ts:7064-7097 -- expectation() (q0-leftmost Pauli strings)
bell = [["h", [0]], ["cnot", [0, 1]]]
ghz3 = [["h", [0]], ["cnot", [0, 1]], ["cnot", [1, 2]]]
for i, (n, calls, pauli, want) in enumerate([
        (1, [], "Z", 1.0), (1, [["x", [0]]], "Z", -1.0), (1, [["h", [0]]], "Z", 0.0),
        (1, [["h", [0]]], "X", 1.0), (1, [["x", [0]], ["h", [0]]], "X", -1.0), (1, [], "X", 0.0),
        (1, [["rx", [-math.pi / 2, 0]]], "Y", 1.0), (1, [], "Y", 0.0), (1, [["h", [0]]], "I", 1.0),
        (2, bell, "ZZ", 1.0), (2, bell, "XX", 1.0), (2, bell, "ZI", 0.0), (2, bell, "IZ", 0.0),
        (2, bell, "YY", -1.0), (3, ghz3, "ZZZ", 0.0), (3, ghz3, "ZZI", 1.0)]):
    cases.append({"id": f"expectation_{i}", "source": "src/circuit.test.ts:7064-7097", "n": n, "calls": calls,
                  "query": "expectation", "pauli": pauli, "expected": want, "tol": 1e-8})

# src/circuit.test.ts:3701-3740 -- marginals()
for i, (n, calls, want) in enumerate([
        (1, [], [0.0]), (1, [["x", [0]]], [1.0]), (1, [["h", [0]]], [0.5]), (2, bell, [0.5, 0.5]),
        (2, [["x", [1]]], [0.0, 1.0]), (4, [["h", [q]] for q in range(4)], [0.5] * 4)]):
    cases.append({"id": f"marginals_{i}", "source": "src/circuit.test.ts:3701-3740", "n": n, "calls": calls,
                  "query": "marginals", "expected": want, "tol": 1e-12})

# src/circuit.test.ts:4165-4228 -- gradient() (parameter-shift closed forms); "$k" is parameter k
Z = [{"coeff": 1, "ops": "Z"}]
grads = [
    (1, [["ry", ["$0", 0]]], Z, [0.0], [0.0]),
    (1, [["ry", ["$0", 0]]], Z, [math.pi / 4], [-math.sin(math.pi / 4)]),
    (1, [["ry", ["$0", 0]]], Z, [math.pi], [0.0]),
    (2, [["ry", ["$0", 0]], ["ry", ["$1", 1]]], [{"coeff": 1, "ops": "ZI"}, {"coeff": 1, "ops": "IZ"}],
     [math.pi / 3, math.pi / 6], [-math.sin(math.pi / 3), -math.sin(math.pi / 6)]),
    (1, [["h", [0]], ["rz", ["$0", 0]]], [{"coeff": 1, "ops": "X"}], [math.pi / 4], [-math.sin(math.pi / 4)]),
    (1, [["rx", ["$0", 0]]], [{"coeff": 1, "ops": "Y"}], [math.pi / 4], [-math.cos(math.pi / 4)]),
    (2, [["ry", ["$0", 0]], ["cnot", [0, 1]], ["ry", ["$1", 1]]], [{"coeff": 1, "ops": "ZZ"}],
     [0.5, math.pi / 4], [0.0, -math.sin(math.pi / 4)]),
    (1, [["ry", ["$0", 0]]], [{"coeff": 2, "ops": "Z"}], [math.pi / 5], [-2 * math.sin(math.pi / 5)]),
]
for i, (n, calls, H, params, want) in enumerate(grads):
    cases.append({"id": f"gradient_{i}", "source": "src/circuit.test.ts:4165-4228", "n": n, "calls": calls,
                  "query": "gradient", "hamiltonian": H, "params": params, "expected": want, "tol": 1e-10})

# src/circuit.test.ts:1864-1978 -- run() with mid-circuit measure, reset and if
# (query "run": counts from run({shots, seed}); "probs" are exact outcome
# frequencies, "bit" = [q0-leftmost position, value, frequency, tolerance] and
# "cregs" the register bit frequencies ket asserts)
def meas(q, reg, bit):
    return ["measure", [q, reg, bit]]


def cond(reg, value, calls):
    return ["if", [reg, value, calls]]


two_bits = [["x", [1]], meas(0, "c", 0), meas(1, "c", 1)]
teleport = [["h", [1]], ["cnot", [1, 2]], ["cnot", [0, 1]], ["h", [0]], meas(0, "c0", 0), meas(1, "c1", 0),
            cond("c1", 1, [["x", [2]]]), cond("c0", 1, [["z", [2]]])]
runs = [
    ("if_register_zero_fires", "src/circuit.test.ts:1865-1869", 1, {"c": 1},
     [meas(0, "c", 0), cond("c", 0, [["x", [0]]])], 100, 1, {"probs": {"1": 1.0}}),
    ("if_fires_on_one", "src/circuit.test.ts:1871-1875", 1, {"c": 1},
     [["x", [0]], meas(0, "c", 0), cond("c", 1, [["x", [0]]])], 100, 1, {"probs": {"0": 1.0}}),
    ("if_skips_on_one", "src/circuit.test.ts:1877-1881", 1, {"c": 1},
     [["x", [0]], meas(0, "c", 0), cond("c", 0, [["x", [0]]])], 100, 1, {"probs": {"1": 1.0}}),
    ("if_corrects_collapse", "src/circuit.test.ts:1883-1887", 1, {"c": 1},
     [["h", [0]], meas(0, "c", 0), cond("c", 1, [["x", [0]]])], 1000, 42, {"probs": {"0": 1.0}}),
    ("if_multibit_value_2", "src/circuit.test.ts:1889-1899", 2, {"c": 2},
     two_bits + [cond("c", 2, [["x", [0]]])], 100, 1, {"probs": {"11": 1.0}}),
    ("if_multibit_not_1", "src/circuit.test.ts:1901-1911", 2, {"c": 2},
     two_bits + [cond("c", 1, [["x", [0]]])], 100, 1, {"probs": {"01": 1.0}}),
    ("if_multibit_value_3", "src/circuit.test.ts:1913-1923", 3, {"c": 2},
     [["x", [0]], ["x", [1]], meas(0, "c", 0), meas(1, "c", 1), cond("c", 3, [["x", [2]]])], 100, 1,
     {"probs": {"111": 1.0}}),
    ("teleport_0", "src/circuit.test.ts:1934-1948", 3, {"c0": 1, "c1": 1}, teleport, 1000, 1,
     {"bit": [2, "0", 1.0, 1e-5]}),
    ("teleport_1", "src/circuit.test.ts:1934-1956", 3, {"c0": 1, "c1": 1}, [["x", [0]]] + teleport, 1000, 1,
     {"bit": [2, "1", 1.0, 1e-5]}),
    ("teleport_plus", "src/circuit.test.ts:1934-1963", 3, {"c0": 1, "c1": 1}, [["h", [0]]] + teleport, 1000, 1,
     {"bit": [2, "0", 0.5, 0.05]}),
    ("measure_then_reset", "src/circuit.test.ts:1967-1971", 1, {"c": 1},
     [["x", [0]], meas(0, "c", 0), ["reset", [0]]], 100, 1, {"probs": {"0": 1.0}, "cregs": {"c": [1.0]}}),
    ("reset_then_reencode", "src/circuit.test.ts:1974-1977", 1, {"c": 1},
     [["x", [0]], meas(0, "c", 0), ["reset", [0]], ["x", [0]]], 100, 1, {"probs": {"1": 1.0}}),
]
for cid, src, n, cregs, calls, shots, seed, want in runs:
    cases.append({"id": cid, "source": src, "n": n, "cregs": cregs, "calls": calls, "query": "run",
                  "shots": shots, "seed": seed, "expected": want, "tol": 1e-10})

if __name__ == "__main__":
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump({"about": "known answers of ket's statevector tests, restated as data", "cases": cases}, f,
                  indent=1)
    print(f"{len(cases)} cases")
