"""Wire formats: CircuitJSON and OpenQASM 2.0 / 3.0 (ket_amd.wire).

The CPU cases restate ket's own tests of toQASM / fromQASM / toJSON / fromJSON
(src/circuit.test.ts:2158-2338 toQASM and fromQASM, :2338-2398 the QASM round
trip, :4400-4530 and :6835-6869 the JSON round trips, :5364-5460 QASM 3.0
syntax).  Where ket compares statevectors, the packed tapes are compared
(identical tapes give identical device states) or the oracle's states are;
the GPU cases run loaded circuits on the device against the oracle."""
import json
import math
import random

import numpy as np
import pytest

from ket_amd import Circuit
from ket_amd._native import RangeError
from ket_amd.wire import fmt_angle, parse_angle, qasm_tape

import ketsv_oracle as O
from helpers import build, dense_ref


def tapes_equal(a, b):
    ta, tb = a.tape(), b.tape()
    return np.array_equal(ta[0], tb[0]) and np.array_equal(ta[1], tb[1])


def oracle_state(c, prune=False):
    ops, mats = c.tape()
    psi = O.zero_state(c.qubits)
    O.apply_tape(c.qubits, psi, ops, mats, prune=prune)
    return psi


def prob(c, bs):
    psi = oracle_state(c)
    return abs(psi[int(bs[::-1], 2)]) ** 2


# ---- toQASM (src/circuit.test.ts:2158-2272) -------------------------------------
def test_to_qasm_header_and_gates():
    q = Circuit(2).toQASM()
    assert "OPENQASM 2.0;" in q and 'include "qelib1.inc";' in q and "qreg q[2];" in q
    assert "h q[0];" in Circuit(1).h(0).toQASM()
    assert "cx q[0],q[1];" in Circuit(2).cnot(0, 1).toQASM()
    assert "cx q[0],q[1];" in Circuit(2).cx(0, 1).toQASM()
    q = Circuit(1).si(0).ti(0).v(0).vi(0).toQASM()
    for s in ("sdg q[0];", "tdg q[0];", "sx q[0];", "sxdg q[0];"):
        assert s in q
    q = Circuit(1).r2(0).r4(0).r8(0).toQASM()
    for s in ("rz(pi/2) q[0];", "rz(pi/4) q[0];", "rz(pi/8) q[0];"):
        assert s in q
    q = Circuit(1).rx(math.pi / 2, 0).ry(math.pi / 4, 0).toQASM()
    assert "rx(pi/2) q[0];" in q and "ry(pi/4) q[0];" in q
    q = Circuit(1).u1(math.pi / 4, 0).u2(0, math.pi, 0).u3(math.pi, 0, math.pi, 0).toQASM()
    for s in ("u1(pi/4) q[0];", "u2(0,pi) q[0];", "u3(pi,0,pi) q[0];"):
        assert s in q
    q = Circuit(2).cy(0, 1).cz(0, 1).ch(0, 1).toQASM()
    for s in ("cy q[0],q[1];", "cz q[0],q[1];", "ch q[0],q[1];"):
        assert s in q
    q = Circuit(2).crx(math.pi / 2, 0, 1).cry(math.pi / 4, 0, 1).crz(math.pi / 8, 0, 1).toQASM()
    for s in ("crx(pi/2) q[0],q[1];", "cry(pi/4) q[0],q[1];", "crz(pi/8) q[0],q[1];"):
        assert s in q
    q = Circuit(2).cr2(0, 1).cr4(0, 1).cr8(0, 1).toQASM()
    for s in ("crz(pi/2) q[0],q[1];", "crz(pi/4) q[0],q[1];", "crz(pi/8) q[0],q[1];"):
        assert s in q
    q = Circuit(2).cs(0, 1).csdg(0, 1).ct(0, 1).ctdg(0, 1).toQASM()
    for s in ("cu1(pi/2) q[0],q[1];", "cu1(-pi/2) q[0],q[1];", "cu1(pi/4) q[0],q[1];", "cu1(-pi/4) q[0],q[1];"):
        assert s in q
    assert "ccx q[0],q[1],q[2];" in Circuit(3).ccx(0, 1, 2).toQASM()
    assert "cswap q[0],q[1],q[2];" in Circuit(3).cswap(0, 1, 2).toQASM()
    q = Circuit(1).creg("c", 1).h(0).measure(0, "c", 0).toQASM()
    assert "creg c[1];" in q and "measure q[0] -> c[0];" in q
    assert "reset q[0];" in Circuit(1).reset(0).toQASM()


@pytest.mark.parametrize("make", [lambda: Circuit(1).gpi(0, 0), lambda: Circuit(1).gpi2(0, 0),
                                  lambda: Circuit(2).xx(math.pi / 2, 0, 1),
                                  lambda: Circuit(1).creg("c", 1).if_("c", 1, lambda q: q.x(0)),
                                  lambda: Circuit(3).csrswap(0, 1, 2),
                                  lambda: Circuit(1).unitary([[1, 0], [0, 1]], 0)])
def test_to_qasm_rejects_gates_without_a_form(make):
    with pytest.raises(TypeError):
        make().toQASM()


def test_angle_format_and_grammar():
    """fmtAngle's rational multiples of pi and parseAngle's grammar."""
    assert fmt_angle(0.0) == "0" and fmt_angle(1e-15) == "0"
    assert fmt_angle(math.pi) == "pi" and fmt_angle(-math.pi / 2) == "-pi/2"
    assert fmt_angle(2 * math.pi / 3) == "2*pi/3" and fmt_angle(3 * math.pi) == "3*pi"
    assert fmt_angle(0.7) == "0.7" and fmt_angle(1e-7) == "1e-7"
    for text, want in (("pi/2", math.pi / 2), ("2*pi/3", 2 * math.pi / 3), ("-pi/4", -math.pi / 4),
                       ("(pi+1)/2", (math.pi + 1) / 2), (" 0.25 ", 0.25), ("+3", 3.0), ("--1", 1.0),
                       ("2-3*4", -10.0), ("1.5.2", 1.5)):        # parseFloat keeps the longest prefix
        assert parse_angle(text) == want
    for bad in ("pipi", "1e-3", "pi/0", "", "abc"):
        with pytest.raises(TypeError):
            parse_angle(bad)


# ---- fromQASM (src/circuit.test.ts:2274-2337) -----------------------------------
HDR = 'OPENQASM 2.0;\ninclude "qelib1.inc";\n'


def test_from_qasm_parses_ket_cases():
    c = Circuit.fromQASM(HDR + "qreg q[1];\nh q[0];")
    assert c.qubits == 1 and abs(prob(c, "0") - 0.5) < 1e-12
    c = Circuit.fromQASM(HDR + "qreg q[2];\nh q[0];\ncx q[0],q[1];")
    assert tapes_equal(c, Circuit(2).h(0).cnot(0, 1))
    c = Circuit.fromQASM(HDR + "qreg q[1];\nh q[0];\nsdg q[0];\ns q[0];\nh q[0];")
    assert abs(prob(c, "0") - 1) < 1e-12
    c = Circuit.fromQASM(HDR + "qreg q[1];\nrx(pi/2) q[0];")
    assert abs(prob(c, "0") - 0.5) < 1e-12
    c = Circuit.fromQASM(HDR + "qreg q[1];\ncreg c[1];\nh q[0];\nmeasure q[0] -> c[0];")
    assert c._cregs == {"c": 1} and c.ops[-1] == ("measure", 0, "c", 0)
    c = Circuit.fromQASM('OPENQASM 2.0; // version\ninclude "qelib1.inc"; // library\nqreg q[1]; // 1 qubit\n'
                         "x q[0]; // flip")
    assert abs(prob(c, "1") - 1) < 1e-12
    with pytest.raises(TypeError, match="Unknown QASM gate"):
        Circuit.fromQASM(HDR + "qreg q[1];\nbogus q[0];")
    c = Circuit.fromQASM(HDR + "qreg q[3];\nx q[0];\nx q[1];\nccx q[0],q[1],q[2];")
    assert abs(prob(c, "111") - 1) < 1e-12
    with pytest.raises(RangeError):
        Circuit.fromQASM(HDR + "qreg q[2];\nh q[5];")


def test_from_qasm_3():
    """QASM 3.0: qubit[N] / bit[N], assignment-form measure, block comments, p / sx
    (src/circuit.test.ts:5364-5460)."""
    src = ('\nOPENQASM 3.0;\ninclude "stdgates.inc";\nqubit[2] q;\nbit[2] meas;\nh q[0];\ncx q[0], q[1];\n'
           'barrier q[0], q[1];\nmeas[0] = measure q[0];\nmeas[1] = measure q[1];\n')
    c = Circuit.fromQASM(src)
    assert c.qubits == 2 and c._cregs == {"meas": 2}
    assert [op[0] for op in c.ops] == ["single", "cnot", "barrier", "measure", "measure"]
    assert c.ops[3] == ("measure", 0, "meas", 0) and c.ops[4] == ("measure", 1, "meas", 1)
    two = Circuit.fromQASM(HDR + "qreg q[2];\nh q[0];\ncx q[0], q[1];\n")
    three = Circuit.fromQASM('OPENQASM 3.0;\ninclude "stdgates.inc";\nqubit[2] q;\n/* block\ncomment */\n'
                             "h q[0];\ncx q[0], q[1];\n")
    assert tapes_equal(two, three)
    c = Circuit.fromQASM('OPENQASM 3.0;\nqubit[1] q;\nx q[0];\np(pi) q[0];\n')
    assert abs(prob(c, "1") - 1) < 1e-12
    c = Circuit.fromQASM('OPENQASM 3.0;\nqubit q;\nsx q[0];\nsx q[0];\n')
    assert c.qubits == 1 and abs(prob(c, "1") - 1) < 1e-12


ROUND_TRIPS = [
    lambda: Circuit(1).h(0),
    lambda: Circuit(2).h(0).cnot(0, 1),
    lambda: Circuit(1).rx(0.7, 0).rz(1.3, 0).ry(0.5, 0),
    lambda: Circuit(1).u3(math.pi / 3, math.pi / 4, math.pi / 5, 0),
    lambda: Circuit(2).h(0).cy(0, 1).cz(0, 1).ch(0, 1).crx(math.pi / 3, 0, 1),
    lambda: Circuit(3).h(0).cr2(0, 1).cr4(0, 2).h(1).cr2(1, 2).h(2),
    lambda: Circuit(3).x(0).x(1).ccx(0, 1, 2),
    lambda: Circuit(2).h(0).cs(0, 1).ct(0, 1).csdg(0, 1).ctdg(0, 1),
]


@pytest.mark.parametrize("make", ROUND_TRIPS)
def test_qasm_round_trip_statevector(make):
    """toQASM -> fromQASM keeps the statevector (src/circuit.test.ts:2338-2398; ket's
    bar 1e-8, here 1e-12 on the oracle's states)."""
    c = make()
    rt = Circuit.fromQASM(c.toQASM())
    assert np.max(np.abs(oracle_state(c) - oracle_state(rt))) < 1e-12


# ---- CircuitJSON (src/circuit.test.ts:4400-4530, :6835-6869) ----------------------
def rt(c):
    return Circuit.fromJSON(json.dumps(c.toJSON()))


JSON_CASES = [
    lambda: Circuit(2),
    lambda: Circuit(1).h(0).x(0).y(0).z(0).s(0).si(0).t(0).ti(0).v(0).vi(0).id(0).sdg(0).tdg(0).srn(0).srndg(0),
    lambda: Circuit(1).rx(math.pi / 7, 0).ry(math.pi / 7, 0).rz(math.pi / 7, 0).u1(math.pi / 7, 0)
    .u2(math.pi / 7, math.pi / 7, 0).u3(math.pi / 7, math.pi / 7, math.pi / 7, 0).gpi(math.pi / 7, 0)
    .gpi2(math.pi / 7, 0).vz(math.pi / 5, 0).p(0.3, 0).r2(0).r4(0).r8(0),
    lambda: Circuit(2).h(0).cnot(0, 1).x(0).swap(0, 1),
    lambda: Circuit(2).xx(math.pi / 4, 0, 1).yy(math.pi / 4, 0, 1).zz(math.pi / 4, 0, 1).xy(math.pi / 4, 0, 1)
    .iswap(0, 1).srswap(0, 1).ms(math.pi / 4, math.pi / 4, 0, 1),
    lambda: Circuit(2).h(0).cx(0, 1).cy(0, 1).cz(0, 1).ch(0, 1).crx(0.6, 0, 1).cry(0.6, 0, 1).crz(0.6, 0, 1)
    .cu1(0.6, 0, 1).cu2(0.6, 0.6, 0, 1).cu3(0.6, 0.6, 0.6, 0, 1).cs(0, 1).ct(0, 1).csdg(0, 1).ctdg(0, 1)
    .cr2(0, 1).cr4(0, 1).cr8(0, 1).csrn(0, 1),
    lambda: Circuit(3).x(0).x(1).ccx(0, 1, 2).cswap(0, 1, 2).csrswap(0, 1, 2).barrier(0, 1),
    lambda: Circuit(4).defineGate("bell", Circuit(2).h(0).cnot(0, 1)).gate("bell", 0, 1).gate("bell", 2, 3),
    lambda: Circuit(2).unitary([[2 ** -0.5, 2 ** -0.5], [2 ** -0.5, -2 ** -0.5]], 0).cnot(0, 1),
]


@pytest.mark.parametrize("make", JSON_CASES)
def test_json_round_trip_every_op_kind(make):
    """fromJSON(toJSON(c)) rebuilds every gate from its meta: the packed tape (kinds,
    operands and every matrix entry) is identical, so device states are too."""
    c = make()
    r = rt(c)
    assert r.qubits == c.qubits
    assert tapes_equal(c, r)
    assert r.toJSON() == c.toJSON()


def test_json_schema_registers_and_classical_ops():
    j = Circuit(1).h(0).toJSON()
    assert j["ket"] == 1 and j["ops"] == [{"kind": "single", "q": 0, "meta": {"name": "h"}}]
    j = Circuit(2).creg("c", 2).toJSON()
    assert j["cregs"] == {"c": 2}
    Circuit.fromJSON(j).measure(0, "c", 0)
    c = Circuit(2).creg("c", 1).x(0).measure(0, "c", 0).reset(0).if_("c", 1, lambda q: q.x(1))
    r = rt(c)
    assert r.ops == c.ops or json.dumps(r.toJSON()) == json.dumps(c.toJSON())
    assert r._cregs == {"c": 1}
    p = Circuit(1).rx("theta", 0)
    assert rt(p).params == ["theta"]
    assert tapes_equal(rt(p).bind({"theta": 0.4}), p.bind({"theta": 0.4}))


def test_json_errors():
    with pytest.raises(TypeError, match="schema version"):
        Circuit.fromJSON({"ket": 2, "qubits": 1, "cregs": {}, "gates": {}, "ops": []})
    with pytest.raises(TypeError, match="unknown op kind"):
        Circuit.fromJSON({"ket": 1, "qubits": 1, "cregs": {}, "gates": {}, "ops": [{"kind": "bogus"}]})
    with pytest.raises(TypeError, match="unknown single-qubit gate"):
        Circuit.fromJSON({"ket": 1, "qubits": 1, "cregs": {}, "gates": {},
                          "ops": [{"kind": "single", "q": 0, "meta": {"name": "nope"}}]})


def random_qasm_calls(n, count, seed):
    """Seeded calls over every gate method that has an OpenQASM 2.0 form."""
    rnd = random.Random(seed)
    one = "h x y z s si t ti v vi id r2 r4 r8".split()
    one_p = ["rx", "ry", "rz", "u1", "p"]
    two = "cnot cy cz ch cs ct csdg ctdg cr2 cr4 cr8 swap".split()
    two_p = ["crx", "cry", "crz", "cu1"]
    calls = []
    for _ in range(count):
        r = rnd.random()
        qs = rnd.sample(range(n), 3)
        a = lambda: rnd.uniform(-6, 6)
        if r < 0.25:
            calls.append((rnd.choice(one), (qs[0],)))
        elif r < 0.45:
            name = rnd.choice(one_p + ["u2", "u3"])
            k = {"u2": 2, "u3": 3}.get(name, 1)
            calls.append((name, tuple(a() for _ in range(k)) + (qs[0],)))
        elif r < 0.7:
            calls.append((rnd.choice(two), (qs[0], qs[1])))
        elif r < 0.9:
            name = rnd.choice(two_p + ["cu2", "cu3"])
            k = {"cu2": 2, "cu3": 3}.get(name, 1)
            calls.append((name, tuple(a() for _ in range(k)) + (qs[0], qs[1])))
        else:
            calls.append((rnd.choice(["ccx", "cswap"]), tuple(qs)))
    return calls


def test_qasm_tape_direct():
    """qasm_tape: the packed tape of a QASM text equals the tape of the same circuit
    built through the API when every angle prints exactly (QASM text of the API
    circuit read back)."""
    c = build(8, random_qasm_calls(8, 200, 3))
    n, tape = qasm_tape(c.toQASM())
    assert n == 8
    back = Circuit.fromQASM(c.toQASM())
    assert np.array_equal(tape[0], back.tape()[0]) and np.array_equal(tape[1], back.tape()[1])


@pytest.mark.gpu
@pytest.mark.parametrize("n", [6, 14])
def test_gpu_circuit_from_qasm_and_json(n):
    """On the device: a circuit read from QASM text gives the statevector of the same
    circuit built through the API (within 1e-12: QASM prints angles, and r2/cs-style
    gates come back as rz/cu1), and bit for bit the oracle's state of the loaded
    circuit under ket's pruning; a JSON round trip gives bit-identical states and
    identical run() counts with measure, reset and if."""
    calls = random_qasm_calls(n, 300, 10 + n)
    api = build(n, calls)
    loaded = Circuit.fromQASM(api.toQASM())
    a = api.statevectorDense()
    b = loaded.statevectorDense()
    assert np.max(np.abs(a - b)) < 1e-12
    assert np.array_equal(b, oracle_state(loaded, prune=True))
    assert np.array_equal(rt(api).statevectorDense(), a)
    c = Circuit(2).creg("c", 1).x(0).measure(0, "c", 0).reset(0).if_("c", 1, lambda q: q.x(1))
    before, after = c.run(shots=200, seed=42), rt(c).run(shots=200, seed=42)
    assert before.probs == after.probs == {"01": 1.0}
