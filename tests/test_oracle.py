"""The CPU oracle pinned against closed forms, invariants, brute force and
the known answers of ket's own tests (src/gates.test.ts,
src/circuit.test.ts).  Runs without a GPU."""
import cmath
import math

import numpy as np
import pytest

import ketsv_oracle as O
from helpers import random_calls

S2 = 1 / math.sqrt(2)
ALL1 = "id h x y z s si t ti v vi r2 r4 r8".split()
PAR1 = {"rx": 1, "ry": 1, "rz": 1, "u1": 1, "u2": 2, "u3": 3, "gpi": 1, "gpi2": 1}
PAR2 = {"xx": 1, "yy": 1, "zz": 1, "xy": 1, "ms": 2}


def M(name, *p):
    return O.gate_matrix(name, p)


def close(a, b, tol=1e-11):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) < tol


# ---- gate matrices (src/gates.test.ts:105-455) -----------------------------------
@pytest.mark.parametrize("name", ALL1 + ["iswap", "srswap"])
def test_static_gates_unitary(name):
    m = M(name)
    assert close(m.conj().T @ m, np.eye(len(m)))


@pytest.mark.parametrize("name", list(PAR1) + list(PAR2))
def test_parametric_gates_unitary(name):
    k = PAR1.get(name, PAR2.get(name))
    m = M(name, *[1.23, -0.7, 2.9][:k])
    assert close(m.conj().T @ m, np.eye(len(m)))


def test_gate_identities():
    X, Y, Z, H = M("x"), M("y"), M("z"), M("h")
    assert close(M("u3", math.pi, 0, math.pi), X)
    assert close(M("u3", math.pi / 2, 0, math.pi), H)
    assert close(M("u3", 0.81, 0, 0), M("ry", 0.81))
    assert close(M("u3", 0, 0, 0.37), M("u1", 0.37))
    assert close(M("u2", 0, math.pi), H)
    assert close(M("u1", math.pi / 2), M("s"))
    assert close(M("u1", math.pi / 4), M("t"))
    assert close(M("u1", math.pi), Z)
    assert close(M("v") @ M("v"), X)
    assert close(M("s") @ M("s"), Z)
    assert close(M("t") @ M("t"), M("s"))
    assert close(M("gpi", 0), X)
    assert close(M("gpi", math.pi / 2), Y)
    assert close(M("gpi2", 0), M("rx", math.pi / 2))
    assert close(M("gpi2", math.pi / 2), M("ry", math.pi / 2))
    assert close(M("ms", 0, 0), M("xx", math.pi / 2))
    assert close(M("ms", math.pi / 2, math.pi / 2), M("yy", math.pi / 2))
    assert close(M("iswap"), np.array([[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]]))
    assert close(M("srswap") @ M("srswap"), M("iswap"))
    # Rz(theta) = e^{-i theta/2} U1(theta); R2 = Rz(pi/2)
    assert close(M("rz", 0.9), cmath.exp(-0.45j) * M("u1", 0.9))
    assert close(M("r2"), M("rz", math.pi / 2))
    # ZZ(theta) = CNOT (I (x) Rz(theta)) CNOT, first qubit = MSB
    cn = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]])
    assert close(M("zz", 0.77), cn @ np.kron(np.eye(2), M("rz", 0.77)) @ cn)
    # XX(theta) = exp(-i theta/2 X(x)X)
    xx = np.kron(X, X)
    assert close(M("xx", 0.6), math.cos(0.3) * np.eye(4) - 1j * math.sin(0.3) * xx)
    yy = np.kron(Y, Y)
    assert close(M("yy", 0.6), math.cos(0.3) * np.eye(4) - 1j * math.sin(0.3) * yy)


def test_gate_exact_entries():
    """Exact values quoted by src/gates.test.ts."""
    h = M("h")
    assert h[0, 0] == S2 and h[1, 1] == -S2
    assert M("t")[1, 1] == complex(S2, S2)
    rx = M("rx", math.pi)
    assert abs(rx[0, 1] - (-1j)) < 1e-15 and abs(rx[0, 0]) < 1e-15


# ---- brute force: full 2^n matrices built independently with numpy ---------------
def embed(n, m, qs):
    """Full operator of a 2^k matrix on qubits qs (qs[0] = MSB of the local index)."""
    N = 1 << n
    k = len(qs)
    U = np.zeros((N, N), dtype=complex)
    for col in range(N):
        lc = 0
        for b, q in enumerate(qs):
            lc |= ((col >> q) & 1) << (k - 1 - b)
        base = col
        for q in qs:
            base &= ~(1 << q)
        for lr in range(1 << k):
            row = base
            for b, q in enumerate(qs):
                if (lr >> (k - 1 - b)) & 1:
                    row |= 1 << q
            U[row, col] += m[lr, lc]
    return U


def controlled(m):
    out = np.eye(2 * len(m), dtype=complex)
    out[len(m):, len(m):] = m
    return out


CTRL = {"cy": "y", "cz": "z", "ch": "h", "crx": "rx", "cry": "ry", "crz": "rz", "cr2": "r2", "cr4": "r4",
        "cr8": "r8", "cu1": "u1", "cu2": "u2", "cu3": "u3", "cs": "s", "ct": "t", "csdg": "si", "ctdg": "ti",
        "csrn": "v"}
ALIAS = {"sdg": "si", "tdg": "ti", "srn": "v", "srndg": "vi", "vz": "rz", "p": "u1"}


def full_op(n, name, args):
    X = M("x")
    if name == "unitary":
        return embed(n, np.asarray(args[0]), args[1:])
    k = O.NQ[name]
    ps, qs = args[:len(args) - k], args[len(args) - k:]
    if name in ("cnot", "cx"):
        return embed(n, controlled(X), qs)
    if name == "swap":
        return embed(n, np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]]), qs)
    if name == "ccx":
        return embed(n, controlled(controlled(X)), qs)
    if name == "cswap":
        return embed(n, controlled(np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])), qs)
    if name == "csrswap":
        s = 1 / math.sqrt(2)
        return embed(n, controlled(np.array([[1, 0, 0, 0], [0, s, 1j * s, 0], [0, 1j * s, s, 0], [0, 0, 0, 1]])), qs)
    if name in CTRL:
        return embed(n, controlled(M(CTRL[name], *ps)), qs)
    return embed(n, M(ALIAS.get(name, name), *ps), qs)


@pytest.mark.parametrize("seed", range(6))
def test_oracle_matches_brute_force(seed):
    n = 4
    calls = random_calls(n, 40, seed=seed)
    psi = O.zero_state(n)
    O.apply_named(n, psi, calls, prune=False)
    want = np.zeros(1 << n, dtype=complex)
    want[0] = 1
    for name, args in calls:
        want = full_op(n, name, args) @ want
    assert close(psi, want, 1e-12)


# ---- closed forms --------------------------------------------------------------
def qft_calls(n):
    calls = []
    for j in range(n - 1, -1, -1):
        calls.append(("h", (j,)))
        for k in range(j - 1, -1, -1):
            calls.append(("cu1", (math.pi / 2 ** (j - k), k, j)))
    for i in range(n // 2):
        calls.append(("swap", (i, n - 1 - i)))
    return calls


@pytest.mark.parametrize("n", [3, 4, 6])
def test_qft_is_scaled_inverse_fft(n):
    rs = np.random.RandomState(n)
    psi0 = rs.normal(size=1 << n) + 1j * rs.normal(size=1 << n)
    psi0 /= np.linalg.norm(psi0)
    psi = psi0.copy()
    O.apply_named(n, psi, qft_calls(n), prune=False)
    assert close(psi, math.sqrt(1 << n) * np.fft.ifft(psi0), 1e-12)


def test_qft3_phases():
    """src/circuit.test.ts:3917-3929: QFT|1> amplitudes e^{2 pi i k / 8}/sqrt 8."""
    psi = O.zero_state(3, 1)
    O.apply_named(3, psi, qft_calls(3), prune=False)
    for k in range(8):
        assert abs(psi[k] - cmath.exp(2j * math.pi * k / 8) / math.sqrt(8)) < 1e-12


def test_bell_and_ghz():
    psi = O.zero_state(2)
    O.apply_named(2, psi, [("h", (0,)), ("cnot", (0, 1))])
    idx, p = O.probabilities(2, psi)
    assert idx.tolist() == [0, 3] and close(p, [0.5, 0.5], 1e-15)
    n = 20
    psi = O.zero_state(n)
    O.apply_named(n, psi, [("h", (0,))] + [("cnot", (i, i + 1)) for i in range(n - 1)])
    idx, p = O.probabilities(n, psi)
    assert idx.tolist() == [0, (1 << n) - 1]


def test_prune_semantics():
    """accumulate() drops |a|^2 < 1e-14 (src/statevector.ts:19-27)."""
    psi = O.zero_state(1)
    O.apply_named(1, psi, [("rx", (1e-8, 0))], prune=True)
    assert psi[1] == 0
    psi = O.zero_state(1)
    O.apply_named(1, psi, [("rx", (1e-8, 0))], prune=False)
    assert abs(psi[1]) > 0
    psi = O.zero_state(1)
    O.apply_named(1, psi, [("rx", (1e-6, 0))], prune=True)
    assert abs(psi[1]) > 0


# ---- seeded sampling --------------------------------------------------------------
def test_prng_is_xorshift32():
    def ref(seed, k):
        s = (seed & 0xFFFFFFFF) or 1
        out = []
        for _ in range(k):
            s ^= (s << 13) & 0xFFFFFFFF
            s ^= s >> 17
            s ^= (s << 5) & 0xFFFFFFFF
            out.append(s / 2 ** 32)
        return out

    for seed in (1, 42, 7, 2 ** 32 - 1, 0):
        assert O.prng(seed, 50).tolist() == ref(seed, 50)
    # seed 0 behaves as seed 1 ((seed >>> 0) || 1)
    assert O.prng(0, 5).tolist() == O.prng(1, 5).tolist()


def test_sample_cdf_rules():
    psi = np.zeros(8, dtype=complex)
    psi[1] = math.sqrt(0.25)
    psi[6] = math.sqrt(0.75)
    r = np.array([0.1, 0.25, 0.2500001, 0.9, 0.9999999])
    assert O.sample(3, psi, r).tolist() == [1, 1, 6, 6, 6]
    # total below 1: r beyond it lands on the last entry
    psi[6] = math.sqrt(0.5)
    assert O.sample(3, psi, np.array([0.9]))[0] == 6


def test_sample_histogram_matches_probabilities():
    n = 5
    psi = O.zero_state(n)
    O.apply_named(n, psi, random_calls(n, 30, seed=3))
    r = O.prng(42, 200000)
    s = O.sample(n, psi, r)
    counts = np.bincount(s.astype(np.int64), minlength=1 << n)
    p = np.abs(psi) ** 2
    exp = p * len(r)
    m = exp > 5
    chi2 = np.sum((counts[m] - exp[m]) ** 2 / exp[m])
    assert chi2 < 3 * m.sum()


# ---- read-outs --------------------------------------------------------------------
def test_expectation_known_answers():
    """src/circuit.ts:3720-3726 examples and src/circuit.test.ts:7064-7097."""
    psi = O.zero_state(1)
    assert O.expectation(1, psi, "Z") == 1
    psi = O.zero_state(1, 1)
    assert O.expectation(1, psi, "Z") == -1
    psi = O.zero_state(1)
    O.apply_named(1, psi, [("h", (0,))])
    assert abs(O.expectation(1, psi, "X") - 1) < 1e-12
    bell = O.zero_state(2)
    O.apply_named(2, bell, [("h", (0,)), ("cnot", (0, 1))])
    assert abs(O.expectation(2, bell, "ZZ") - 1) < 1e-12
    assert abs(O.expectation(2, bell, "XX") - 1) < 1e-12
    assert abs(O.expectation(2, bell, "YY") + 1) < 1e-12
    for th in (0.3, 1.7, -2.2):
        psi = O.zero_state(1)
        O.apply_named(1, psi, [("ry", (th, 0))])
        assert abs(O.expectation(1, psi, "Z") - math.cos(th)) < 1e-12
        assert abs(O.expectation(1, psi, "X") - math.sin(th)) < 1e-12
        assert abs(O.expect_direct(1, psi, 1, 0) - math.sin(th)) < 1e-12


def test_direct_expectation_equals_rotation_method():
    n = 5
    psi = O.zero_state(n)
    O.apply_named(n, psi, random_calls(n, 40, seed=8), prune=False)
    for pauli in ("XYZIX", "ZZIII", "YYYYY", "IXIXI"):
        x = sum(1 << q for q, ch in enumerate(pauli) if ch in "XY")
        z = sum(1 << q for q, ch in enumerate(pauli) if ch in "YZ")
        P = np.eye(1, dtype=complex)
        for ch in reversed(pauli):
            P = np.kron(P, {"I": np.eye(2), "X": M("x"), "Y": M("y"), "Z": M("z")}[ch])
        want = np.vdot(psi, P @ psi).real
        assert abs(O.expect_direct(n, psi, x, z) - want) < 1e-12
        assert abs(O.expectation(n, psi, pauli) - want) < 1e-10


def test_vqe_convention_and_gradient_closed_forms():
    """vqe's ops[0] acts on qubit n-1; d<Z>/dtheta = -sin(theta) for ry
    (src/circuit.test.ts:4165-4235)."""
    n = 2
    psi = O.zero_state(n)
    O.apply_named(n, psi, [("x", (0,))])
    assert O.vqe(n, psi, [{"coeff": 1.0, "ops": "IZ"}]) == -1
    assert O.vqe(n, psi, [{"coeff": 1.0, "ops": "ZI"}]) == 1
    assert O.vqe(n, psi, [{"coeff": 2.5, "ops": "II"}, {"coeff": 1e-16, "ops": "ZZ"}]) == 2.5
    for th in (0.4, 1.3):
        def e(t):
            s = O.zero_state(1)
            O.apply_named(1, s, [("ry", (t, 0))])
            return O.vqe(1, s, [{"coeff": 1.0, "ops": "Z"}])
        psr = 0.5 * (e(th + math.pi / 2) - e(th - math.pi / 2))
        assert abs(psr + math.sin(th)) < 1e-12


def test_marginals_bloch_collapse():
    n = 3
    psi = O.zero_state(n)
    O.apply_named(n, psi, [("h", (0,)), ("x", (2,))])
    assert close(O.marginals(n, psi), [0.5, 0, 1], 1e-12)
    th, ph = O.bloch(n, psi, 0)
    assert abs(th - math.pi / 2) < 1e-12 and abs(ph) < 1e-12
    th, ph = O.bloch(n, psi, 2)
    # rz = -0.9999999999999998 here; acos amplifies that rounding to ~2e-8
    assert abs(th - math.pi) < 1e-7
    bell = O.zero_state(2)
    O.apply_named(2, bell, [("h", (0,)), ("cnot", (0, 1))])
    out = O.collapse(2, bell, 0, 0.3)     # p1 = 0.5 > 0.3 -> outcome 1
    assert out == 1 and abs(bell[3] - 1) < 1e-12 and bell[0] == 0


# ---- run()'s per-shot path (noise, measure, reset) ------------------------

def _chi2(obs, exp):
    return sum((o - e) ** 2 / e for o, e in zip(obs, exp) if e > 0)


def test_shots_spam_flip_closed_form():
    """pMeas = 1 flips every reported bit (src/circuit.ts:3472-3478)."""
    counts, cr = O.run_shots(3, [("x", (0,)), ("measure", (0, "c", 0))], 64, 7, {"pMeas": 1}, cregs={"c": 1})
    assert counts == {0b110: 64}
    assert cr == {"c": [0]}


def test_shots_dep1_uniform_paulis():
    """p1 = 1: every 1-qubit gate is followed by X, Y or Z with probability 1/3 each (src/circuit.ts:289-295)."""
    counts, _ = O.run_shots(1, [("id", (0,))], 6000, 11, {"p1": 1})
    assert _chi2([counts.get(0, 0), counts.get(1, 0)], [2000, 4000]) < 10.8      # 1 dof, p = 0.001


def test_shots_dep2_fifteen_paulis():
    """p2 = 1 on cnot from |00>: the 15 two-qubit Paulis are equally likely (src/circuit.ts:273-278, :296-304)."""
    counts, _ = O.run_shots(2, [("cnot", (0, 1))], 7500, 12, {"p2": 1})
    # index bit 0 flips when the control's Pauli is X or Y, bit 1 likewise for the target
    exp = [3 / 15 * 7500, 4 / 15 * 7500, 4 / 15 * 7500, 4 / 15 * 7500]
    assert _chi2([counts.get(i, 0) for i in range(4)], exp) < 16.3              # 3 dof, p = 0.001


def test_shots_measure_and_reset_semantics():
    counts, cr = O.run_shots(2, [("h", (0,)), ("measure", (0, "c", 0)), ("cnot", (0, 1)), ("reset", (0,))], 2000, 3,
                             cregs={"c": 1})
    # after the reset q0 is 0; q1 copies the measured value
    assert set(counts) <= {0b00, 0b10}
    assert counts.get(0b10, 0) == cr["c"][0]
    assert abs(cr["c"][0] / 2000 - 0.5) < 0.05


def test_shots_noise_free_matches_born_rule():
    calls = random_calls(4, 30, seed=5)
    psi = O.zero_state(4)
    O.apply_named(4, psi, calls)
    p = np.abs(psi) ** 2
    counts, _ = O.run_shots(4, calls, 4000, 21)
    obs = [counts.get(i, 0) for i in range(16)]
    keep = p * 4000 >= 5
    assert _chi2(np.array(obs)[keep], (p * 4000)[keep]) < 40


# ---- density matrix (runDM) --------------------------------------------------

def test_dm_pure_circuits_are_projectors():
    """Without noise rho = |psi><psi| of the statevector oracle (prune after every op aside)."""
    for seed in range(3):
        calls = [c for c in random_calls(3, 25, seed=seed, with_unitary=True) if c[0] != "csrswap"]
        psi = O.zero_state(3)
        O.apply_named(3, psi, calls, prune=False)
        rho = O.run_dm(3, calls)
        assert np.max(np.abs(rho - np.outer(psi, psi.conj()))) < 1e-12


def test_dm_depolarize1_closed_forms():
    """src/circuit.test.ts:5164-5170, :5199-5203: off-diagonal (1 - 4p/3)/2; p = 3/4 maximally mixed."""
    p = 0.1
    rho = O.run_dm(1, [("h", (0,))], {"p1": p})
    assert abs(rho[0, 1] - (1 - 4 * p / 3) * 0.5) < 1e-15
    assert abs(rho[0, 0] - 0.5) < 1e-15
    rho = O.run_dm(1, [("h", (0,))], {"p1": 0.75})
    assert np.max(np.abs(rho - 0.5 * np.eye(2))) < 1e-15


def test_dm_depolarize2_is_the_pauli_twirl():
    """(1 - p) rho + p/15 sum_{P != II} P rho P^dagger, built from Pauli matrices directly."""
    X = np.array([[0, 1], [1, 0]]); Y = np.array([[0, -1j], [1j, 0]]); Z = np.diag([1, -1]); I = np.eye(2)
    p = 0.2
    pure = O.run_dm(2, [("h", (0,)), ("cnot", (0, 1))])
    got = O.run_dm(2, [("h", (0,)), ("cnot", (0, 1))], {"p2": p})
    want = (1 - p) * pure
    for A in (I, X, Y, Z):
        for B in (I, X, Y, Z):
            if A is I and B is I:
                continue
            P = np.kron(B, A)          # index bit 0 = qubit 0 (control), bit 1 = qubit 1
            want = want + p / 15 * P @ pure @ P.conj().T
    assert np.max(np.abs(got - want)) < 1e-14
    assert abs(np.trace(got) - 1) < 1e-14


def test_dm_csrswap_with_definite_control_is_unitary():
    """Control definitely 1 (or 0): ket's block form reduces to V rho V^dagger (or identity)."""
    calls = [("x", (0,)), ("h", (1,)), ("ry", (0.7, 2)), ("csrswap", (0, 1, 2))]
    psi = O.zero_state(3)
    O.apply_named(3, psi, calls, prune=False)
    assert np.max(np.abs(O.run_dm(3, calls) - np.outer(psi, psi.conj()))) < 1e-12
    calls[0] = ("id", (0,))
    psi = O.zero_state(3)
    O.apply_named(3, psi, calls, prune=False)
    assert np.max(np.abs(O.run_dm(3, calls) - np.outer(psi, psi.conj()))) < 1e-12


@pytest.mark.parametrize("n,heavy", [(8, False), (10, True), (12, True)])
def test_layer0_restatements_used_at_30_qubits(n, heavy):
    """The two per-index restatements the 30-qubit GPU test relies on equal the
    oracle on whole states: the product-state closed form in exact mode, and ket's
    pruned evolution bit for bit (with angles that make most amplitudes prune)."""
    import random
    from test_gpu_configs import brick_product_amplitudes, ket_pruned_layer0
    from ket_amd import workloads
    if heavy:
        rnd = random.Random(n)
        calls = []
        for q in range(n):
            s = rnd.choice([1e-4, 3e-3, 1.0, 2e-5])
            calls += [("rx", (s * rnd.uniform(-1, 1), q)), ("ry", (s * rnd.uniform(-1, 1), q)),
                      ("rz", (rnd.uniform(0, 6), q))]
        calls += [("cnot", (a, a + 1)) for a in range(0, n - 1, 2)]
    else:
        calls = workloads.random_layer_calls(n, 1, seed=42)
    idx = np.arange(1 << n, dtype=np.uint64)
    exact = O.zero_state(n)
    O.apply_named(n, exact, calls, prune=False)
    pruned = O.zero_state(n)
    O.apply_named(n, pruned, calls, prune=True)
    assert np.max(np.abs(brick_product_amplitudes(n, calls, idx) - exact)) < 1e-14
    got = ket_pruned_layer0(n, calls, idx)
    assert np.array_equal(got, pruned)
    if heavy:
        assert np.count_nonzero(pruned == 0) > (1 << n) // 2


def _ket_cdf(n, psi, thresh):
    """ket's running sums, written as its loops are: p = re*re + im*im per entry,
    cum += p in ascending index (Python floats round every operation)."""
    idx, cum, acc = [], [], 0.0
    for i in range(1 << n):
        re, im = float(psi[i].real), float(psi[i].imag)
        p = re * re + im * im
        if p > thresh:
            acc += p
            idx.append(i)
            cum.append(acc)
    return idx, cum


def test_sample_boundary_draws_follow_the_sequential_cdf():
    """Draws equal to a running sum and its neighbouring doubles: the oracle's
    fast-path search (src/circuit.ts:3425-3450) and sampleSV (:74-81) against
    linear scans over ket's sequential sums."""
    n = 7
    psi = O.zero_state(n)
    O.apply_named(n, psi, random_calls(n, 60, seed=8), prune=True)
    idx, cum = _ket_cdf(n, psi, 1e-14)
    rs = []
    for c in cum[:-1]:
        rs += [c, float(np.nextafter(c, 0)), float(np.nextafter(c, 2))]
    rs = np.array(rs)
    got = O.sample(n, psi, rs)
    cdf = cum[:-1] + [1.0]
    for r, g in zip(rs.tolist(), got.tolist()):
        want = next(i for i, c in zip(idx, cdf) if c >= r)
        assert g == want
    idx0, cum0 = _ket_cdf(n, psi, 0.0)
    for r in rs.tolist():
        want = next((i for i, c in zip(idx0, cum0) if r <= c), idx0[-1])
        assert O.sample_sv(n, psi, r) == want


def test_sample_sv_takes_every_entry():
    """sampleSV walks every map entry, the fast path only p > 1e-14; beyond the
    total both give the last entry; an empty state gives index 0."""
    psi = np.zeros(8, dtype=complex)
    psi[0] = 1e-8                      # p = 1e-16: a map entry, filtered by the fast path
    psi[3] = math.sqrt(0.5)
    psi[5] = math.sqrt(0.5)
    assert O.sample_sv(3, psi, 1e-17) == 0
    assert O.sample(3, psi, np.array([1e-17]))[0] == 3
    assert O.sample_sv(3, psi, 0.7) == 5
    assert O.sample_sv(3, psi, 1.5) == 5
    assert O.sample_sv(3, np.zeros(8, dtype=complex), 0.3) == 0


@pytest.mark.parametrize("prune", [False, True])
def test_openmp_build_is_bit_identical(prune):
    """The oracle's OpenMP build (the all-core CPU baseline) splits only the index
    loops between threads, so every amplitude equals the scalar build's bit for bit,
    on a circuit with every gate kind, wide unitaries and csrswap."""
    from ket_amd.circuit import Circuit
    n = 14
    c = Circuit(n)
    for name, args in random_calls(n, 300, seed=5):
        getattr(c, name)(*args)
    ops, mats = c.tape()
    a = O.zero_state(n)
    O.apply_tape(n, a, ops, mats, prune=prune)
    b = O.zero_state(n)
    O.apply_tape_omp(n, b, ops, mats, prune=prune)
    assert np.array_equal(a, b)


def test_prune_candidate_bound():
    """The register-round kernels test ket's isNegligible (norm2 < 1e-14, products
    and sum rounded on their own, src/complex.ts:24-27) exactly only for outputs
    with max(|re|, |im|) < 1.0000001e-7: at or above that bound the rounded norm is
    at least 1e-14 (fl(T*T) >= 1e-14, and fl(a + b) >= a for b >= 0), checked here
    on the bound itself and on random pairs at and just above it; just below it
    both outcomes occur, so the exact test is needed there."""
    T = 1.0000001e-7
    assert T * T >= 1e-14
    rng = np.random.default_rng(3)
    big = T * (1 + rng.uniform(0, 1e-6, 200000))
    small = rng.uniform(0, 1, 200000) * big
    for re, im in ((big, small), (small, big), (big, np.zeros_like(big)), (-big, -small)):
        n2 = re * re + im * im
        assert np.all(n2 >= 1e-14)
    # below the bound: candidates on both sides of ket's test
    assert 8e-8 * 8e-8 + 8e-8 * 8e-8 >= 1e-14 and 5e-8 * 5e-8 + 5e-8 * 5e-8 < 1e-14
