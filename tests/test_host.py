"""Host-side logic without a GPU: gate matrices, tape lowering, builder
errors, PRNG, Distribution, gradient tracing, and the C ABI exports."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import ket_amd
from ket_amd import Circuit, gates as G, workloads
from ket_amd._native import LIB_PATH, OP_DTYPE, SIGNATURES, RangeError
from ket_amd.circuit import Distribution, make_prng, to_uint32

import ketsv_oracle as O
from helpers import build, random_calls

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HOST = {"id": G.Id, "h": G.H, "x": G.X, "y": G.Y, "z": G.Z, "s": G.S, "si": G.Si, "t": G.T, "ti": G.Ti,
        "v": G.V, "vi": G.Vi, "r2": G.R2, "r4": G.R4, "r8": G.R8, "iswap": G.ISwap, "srswap": G.SrSwap}
HOSTP = {"rx": G.Rx, "ry": G.Ry, "rz": G.Rz, "u1": G.U1, "u2": G.U2, "u3": G.U3, "gpi": G.Gpi, "gpi2": G.Gpi2,
         "xx": G.Xx, "yy": G.Yy, "zz": G.Zz, "xy": G.Xy, "ms": G.Ms}
NPAR = {"u2": 2, "u3": 3, "ms": 2}


@pytest.mark.parametrize("name", list(HOST) + list(HOSTP))
def test_host_matrices_bit_identical_to_oracle(name):
    for trial in range(5):
        ps = [1.3 * trial - 2.1, 0.77 * trial + 0.1, -0.4 * trial][:NPAR.get(name, 1 if name in HOSTP else 0)]
        host = np.array(HOSTP[name](*ps) if name in HOSTP else HOST[name], dtype=complex)
        orc = O.gate_matrix(name, ps).reshape(-1)
        assert np.array_equal(host, orc), name


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("prune", [True, False])
def test_tape_lowering_matches_named_semantics(seed, prune):
    """Circuit -> tape (kinds + matrices) applied by the oracle equals the
    oracle's own restatement of every Circuit method, bit for bit."""
    n = 5
    calls = random_calls(n, 80, seed=seed)
    ops, mats = build(n, calls).tape()
    a = O.zero_state(n)
    O.apply_tape(n, a, ops, mats, prune=prune)
    b = O.zero_state(n)
    O.apply_named(n, b, calls, prune=prune)
    assert np.array_equal(a, b)


def test_tape_layout():
    c = Circuit(4).h(0).cnot(0, 3).swap(1, 2).ccx(0, 1, 2).barrier().h(1)
    ops, mats = c.tape()
    assert ops.dtype == OP_DTYPE and ops.dtype.itemsize == 64
    assert ops["kind"].tolist() == [0, 1, 2, 5, 9, 0]
    assert ops["q"][1, :2].tolist() == [0, 3]
    assert ops["mat"][0] == ops["mat"][5] == 0          # shared H matrix
    assert ops["mat"][1] == -1
    assert len(mats) == 8


def test_builder_errors_and_immutability():
    c = Circuit(3)
    with pytest.raises(RangeError, match="qubit index 3 is out of range for a 3-qubit circuit"):
        c.h(3)
    with pytest.raises(RangeError):
        c.cnot(-1, 0)
    with pytest.raises(TypeError, match="control and target qubits must differ"):
        c.crz(0.1, 2, 2)
    with pytest.raises(TypeError, match="must be 4×4"):
        c.unitary([[1, 0], [0, 1]], 0, 1)
    d = c.h(0)
    assert len(c) == 0 and len(d) == 1
    assert len(d.x(1)) == 2 and len(d) == 1
    with pytest.raises(TypeError, match="initialState"):
        Circuit(2).h(0).simulate(initialState="012")


def test_params_bind_compose_decompose():
    a = Circuit(2).rx("theta", 0).ry("phi", 1).cu1("lam", 0, 1)
    assert a.params == ["lam", "phi", "theta"]
    with pytest.raises(TypeError, match="unbound"):
        a.tape()
    b = a.bind({"theta": 0.3, "phi": 0.2, "lam": 1.1})
    assert b.params == []
    with pytest.raises(TypeError, match="bind: unbound parameter 'phi'"):
        a.bind({"theta": 0.3, "lam": 1.0})
    ops, _ = b.tape()
    assert ops["kind"].tolist() == [0, 0, 4]
    bell = Circuit(2).h(0).cnot(0, 1)
    c = Circuit(4).defineGate("bell", bell).gate("bell", 2, 3).gate("bell", 0, 1)
    flat = c.decompose().ops
    assert [op[0] for op in flat] == ["single", "cnot", "single", "cnot"]
    assert flat[1][1:] == (2, 3) and flat[3][1:] == (0, 1)
    with pytest.raises(TypeError, match="expects 2 qubit"):
        c.gate("bell", 0)
    with pytest.raises(TypeError, match="compose: qubit count mismatch"):
        bell.compose(Circuit(3))
    assert len(bell.compose(bell)) == 4


def test_prng_and_seed_conversion():
    for seed in (42, 1, 7, 123456789, 2 ** 31 + 5):
        r = make_prng(seed)
        assert [r() for _ in range(20)] == O.prng(seed, 20).tolist()
    assert to_uint32(-1) == 2 ** 32 - 1
    assert to_uint32(2.9) == 2
    assert to_uint32(float("nan")) == 0
    r0, r1 = make_prng(0), make_prng(1)
    assert r0() == r1()


def test_distribution_keys():
    d = Distribution(3, 4, {1: 3, 6: 1})
    assert d.probs == {"100": 0.75, "011": 0.25}
    assert d.histogram == {"1": 0.75, "6": 0.25}
    assert d.most == "100"
    assert abs(d.entropy - (-(0.75 * math.log2(0.75) + 0.25 * math.log2(0.25)))) < 1e-15


def test_workload_shapes():
    assert len(workloads.random_layer_calls(12, 20)) == 830
    assert len(workloads.random_layer_calls(30, 40)) == 4180
    assert len(ket_amd.qft(30)) == 480
    ans = ket_amd.efficientSU2(24, 24)
    assert ans.paramCount == 1200
    c = ans(workloads.ansatz_params(1200))
    assert len(c) == 1200 + 24 * 23
    assert len(workloads.tfim_hamiltonian(24)) == 47
    inst = workloads.qaoa_instances(8)
    assert all(14 <= n <= 22 for n, *_ in inst)


def test_gradient_tracing():
    from ket_amd.algorithms import trace
    ra = ket_amd.realAmplitudes(4, 2)
    flat, rec, exact, n = trace(ra, [0.1] * ra.paramCount)
    assert exact and len(rec) == ra.paramCount and n == 4
    edges = [(0, 1), (1, 2)]
    qa = lambda p: ket_amd.qaoa(3, edges, p[:1], p[1:])
    flat, rec, exact, n = trace(qa, [0.3, 0.2])
    assert not exact                          # gamma on every edge, beta doubled
    assert sorted({(r[1], r[2]) for r in rec}) == [(0, -1.0), (1, 2.0)]
    bad = lambda p: Circuit(1).rx(math.cos(p[0]), 0)
    assert trace(bad, [0.3]) is None


def test_c_abi_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "ketsv.h")).read()
    declared = set(re.findall(r"^int (ksv_\w+)\(", hdr, re.M))
    assert len(declared) >= 25
    assert declared == {name for name, _, _ in SIGNATURES}
    lib = ctypes.CDLL(LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name


def test_product_path_fails_loudly_without_device():
    if ket_amd.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError, match="no HIP device"):
        Circuit(2).h(0).statevector()


def test_batched_shot_draw_layout():
    """Draws per shot follow ket's per-shot loop: one per noisy 1q / 2q gate, one
    per measure (+1 with pMeas), one per reset, the final draw, n SPAM draws."""
    from ket_amd.noise import batchable, draws_per_shot
    from ket_amd.circuit import _flatten
    c = Circuit(3).creg("c", 2).h(0).cnot(0, 1).ccx(0, 1, 2).measure(1, "c", 0).reset(2).rz(0.3, 2)
    flat = _flatten(c._ops())
    assert draws_per_shot(flat, 3, 0, 0, 0) == 1 + 1 + 1
    assert draws_per_shot(flat, 3, 1e-3, 0, 0) == 2 + 1 + 1 + 1
    assert draws_per_shot(flat, 3, 1e-3, 1e-2, 0) == 2 + 1 + 1 + 1 + 1
    assert draws_per_shot(flat, 3, 1e-3, 1e-2, 1e-2) == 2 + 1 + 2 + 1 + 1 + 3
    assert batchable(c)
    assert not batchable(c.if_("c", 1, lambda k: k.x(0)))


def _events_loop(flat, dr, p1, p2, pm):
    """Event list of one batch built op by op and column by column (ket's per-shot
    loop, src/circuit.ts:263-304, run for every copy of a column at once)."""
    from ket_amd.noise import _PAIR_A, _PAIR_B, _TAPE_KINDS
    B = dr.shape[0]
    gates, events, copies, codes_all, draws, meas = [], [], [], [], [], []
    k = 0

    def paulis(q, t, codes):
        events.append((len(gates), 0, q, len(copies), len(t)))
        copies.extend(t.tolist())
        codes_all.extend(codes.tolist())

    for op in flat:
        kind = op[0]
        if kind in _TAPE_KINDS:
            gates.append(op)
        if kind == "single" and p1:
            rand = dr[:, k]
            k += 1
            t = np.nonzero(rand < p1)[0]
            if len(t):
                r = rand[t] / p1
                paulis(op[1], t, (1 + (r >= 1 / 3) + (r >= 2 / 3)).astype(np.uint8))
        elif kind in ("cnot", "controlled", "swap", "two") and p2:
            rand = dr[:, k]
            k += 1
            t = np.nonzero(rand < p2)[0]
            if len(t):
                pick = np.minimum(np.floor(rand[t] / p2 * 15).astype(np.int64), 14)
                for q, table in ((op[1], _PAIR_A), (op[2], _PAIR_B)):
                    codes = table[pick]
                    m = codes > 0
                    if m.any():
                        paulis(q, t[m], codes[m])
        elif kind in ("measure", "reset"):
            off = len(draws) * B
            events.append((len(gates), 1 if kind == "measure" else 2, op[1], off, 0))
            draws.append(dr[:, k])
            k += 1
            if kind == "measure":
                flip = None
                if pm:
                    flip = dr[:, k] < pm
                    k += 1
                meas.append((op, off, flip))
    return gates, events, copies, codes_all, draws, meas, k


@pytest.mark.parametrize("p1,p2,pm", [(0.2, 0.3, 0.1), (0.0, 0.5, 0.0), (0.4, 0.0, 0.2), (0.0, 0.0, 0.0)])
def test_batched_event_lists_match_the_column_loop(p1, p2, pm):
    """shot_events (all draw columns at once) gives the same tape, events, copies,
    codes, measurement draws and flips as the column-by-column loop."""
    from ket_amd.noise import draws_per_shot, shot_events
    from ket_amd.circuit import _flatten
    rs = np.random.default_rng(5)
    c = Circuit(4).creg("c", 3)
    for step in range(30):
        kind = rs.integers(6)
        a, b = (int(x) for x in rs.choice(4, 2, replace=False))
        if kind == 0:
            c = c.h(a)
        elif kind == 1:
            c = c.cnot(a, b)
        elif kind == 2:
            c = c.measure(a, "c", int(rs.integers(3)))
        elif kind == 3:
            c = c.reset(a)
        elif kind == 4:
            c = c.ccx(a, b, 3 - a if 3 - a not in (a, b) else ({0, 1, 2, 3} - {a, b}).pop())
        else:
            c = c.rz(0.1 * step, a).swap(a, b)
    c = c.measure(0, "c", 0).measure(1, "c", 1)
    flat = _flatten(c._ops())
    B = 257
    D = draws_per_shot(flat, 4, p1, p2, pm)
    dr = rs.random((B, D))
    g0, ev0, cp0, co0, dr0, me0, k0 = _events_loop(flat, dr, p1, p2, pm)
    g1, ev1, cp1, co1, r1, me1, k1 = shot_events(flat, dr, p1, p2, pm)
    assert g1 == g0 and k1 == k0
    assert [tuple(int(x) for x in e) for e in ev1] == ev0
    assert cp1.tolist() == cp0 and co1.tolist() == co0
    assert r1.tolist() == (np.concatenate(dr0).tolist() if dr0 else [])
    assert len(me1) == len(me0)
    for (o1, f1, x1), (o0, f0, x0) in zip(me1, me0):
        assert o1 == o0 and f1 == f0
        assert (x1 is None and x0 is None) or x1.tolist() == x0.tolist()
    assert any(e[1] == 0 for e in ev0) == bool(p1 or p2)


def test_noise_hits_split_over_threads_match_flatnonzero():
    """The draw-versus-rate comparison split by rows over threads, and the
    8-bytes-at-a-time nonzero scan, give np.flatnonzero's indices in order."""
    from ket_amd.noise import _hits, _sparse_flatnonzero
    rs = np.random.default_rng(9)
    for rows, cols in ((4096, 867), (3001, 13), (5, 3)):
        dr = rs.random((rows, cols))
        thr = np.where(rs.random(cols) < 0.8, 0.01, -1.0)
        ref = np.flatnonzero(dr < thr)
        assert np.array_equal(_hits(dr, thr), ref)
        assert np.array_equal(_sparse_flatnonzero(dr < thr), ref)


def test_prng_fill_continues_the_stream():
    """rng.fill(k) (ksv_xorshift32) gives the next k draws of the same stream."""
    a = make_prng(12345)
    b = make_prng(12345)
    first = [a() for _ in range(5)]
    assert first == [b() for _ in range(5)]
    filled = b.fill(1000)
    assert filled.tolist() == [a() for _ in range(1000)]
    assert a() == b()
    assert make_prng(0).fill(3).tolist() == make_prng(1).fill(3).tolist()     # seed 0 -> 1


def test_prng_long_fill_is_the_serial_stream():
    """A fill long enough to be cut into per-thread chunks (each started by the
    GF(2) jump-ahead) is the serial stream: every draw is one xorshift32 step of
    the one before it, the first follows the seed, and the stream continues."""
    seed = 987654321
    a, b = make_prng(seed), make_prng(seed)
    count = (3 << 20) + 12345
    out = b.fill(count)
    s = np.round(out * 4294967296.0).astype(np.uint64).astype(np.uint32)
    assert np.array_equal(s.astype(np.float64) / 4294967296.0, out)
    v = s[:-1].copy()
    v ^= v << np.uint32(13)
    v ^= v >> np.uint32(17)
    v ^= v << np.uint32(5)
    assert np.array_equal(v, s[1:])
    ref = O.xorshift(seed)
    assert [next(ref) for _ in range(5)] == out[:5].tolist() == [a() for _ in range(5)]
    c = make_prng(int(s[-1]))       # state after the fill is the last draw's state
    assert b() == c()


class _OracleCopies:
    """CPU stand-in for a batched device state (the ksv_shots_* methods), every copy
    simulated by the oracle; lets the host driver's event logic run without a GPU."""

    def __init__(self, n_total, precision="f64", prune=True, tile=0):
        self.n = n_total
        self.prune = prune

    def shots_init(self, n, index=0):
        import ketsv_oracle as O
        self.k = n
        self.psi = np.array([O.zero_state(n, index) for _ in range(1 << (self.n - n))])

    def _apply(self, ops, mats):
        import ketsv_oracle as O
        if len(ops):
            for t in range(len(self.psi)):
                O.apply_tape(self.k, self.psi[t], np.ascontiguousarray(ops, dtype=OP_DTYPE), mats, prune=self.prune)

    def shots_run(self, n, tape, events, copies, paulis, r):
        import ketsv_oracle as O
        ops, mats = tape
        out = np.zeros(len(r), dtype=np.uint8)
        cur = 0
        for e in events:
            self._apply(ops[cur:int(e["at"])], mats)
            cur = int(e["at"])
            q, off = int(e["q"]), int(e["off"])
            if e["kind"] == 0:
                for j in range(off, off + int(e["count"])):
                    O.apply_named(n, self.psi[int(copies[j])], [("xyz"[int(paulis[j]) - 1], (q,))], self.prune)
            else:
                for t in range(len(self.psi)):
                    o = O.collapse(n, self.psi[t], q, float(r[off + t]))
                    out[off + t] = o
                    if e["kind"] == 2 and o == 1:
                        O.apply_named(n, self.psi[t], [("x", (q,))], self.prune)
        self._apply(ops[cur:], mats)
        return out

    def shots_sample(self, n, r):
        import ketsv_oracle as O
        return np.array([int(O.sample(n, self.psi[t], np.array([r[t]]))[0]) for t in range(len(self.psi))],
                        dtype=np.uint64)

    def free(self):
        pass


@pytest.mark.parametrize("noise,seed", [("harmony", 1), ({"p1": 0.05, "p2": 0.1, "pMeas": 0.05}, 2), (None, 3)])
def test_batched_shot_driver_on_the_oracle(monkeypatch, noise, seed):
    """run_batched's draw layout and event list (Pauli errors after their gate,
    measurements, resets, SPAM flips, batches) with every copy simulated by the
    oracle give the counts of the oracle's per-shot loop."""
    import ketsv_oracle as O
    import ket_amd.noise as N
    from helpers import build, shot_calls
    from ket_amd.circuit import make_prng
    monkeypatch.setattr(N, "State", _OracleCopies)
    monkeypatch.setattr(N, "MAX_COPIES", 8)
    n = 3
    calls = shot_calls(n, 20, seed)
    c = build(n, calls, cregs={"c": 2})
    params = N.resolve_noise(noise)
    cregs = {"c": [0, 0]}
    shots = {1: 41, 2: 43, 3: 40}[seed]          # last batch of 1, 3 (in 4 copies) and 8 shots
    counts, cr = N.run_batched(c, shots, make_prng(seed), params, 0, cregs)
    ref, rcr = O.run_shots(n, calls, shots, seed, params, cregs={"c": 2})
    assert list(counts.items()) == list(ref.items())
    assert cr == rcr


@pytest.mark.parametrize("edges", [[(0, 1), (0, 1)], [(2, 5), (5, 2)], [(0, 1), (1, 2), (3, 1), (2, 1)]])
def test_qaoa_batch_rejects_repeated_edges(edges):
    """qaoa() and maxCutHamiltonian() apply every listed edge (src/algorithms.ts:231-276)
    while the batch's cut counts each vertex pair once: a repeat, in either order, is a
    RangeError from create, before any device call."""
    from ket_amd._native import QaoaBatch, RangeError
    with pytest.raises(RangeError, match="repeated edge"):
        QaoaBatch([(14, [(0, 3)]), (14, edges)], p=2)


def test_creg_value_is_javascripts_32_bit_integer():
    """cregValue (src/circuit.ts:164-166) ORs 1 << i with JavaScript's 32-bit shift
    (count mod 32) and |: bit 31 gives a negative value and bit 32 aliases bit 0.
    The host driver and the oracle agree with that arithmetic, restated here from
    the ECMAScript definitions of << and |."""
    import ketsv_oracle as O
    from ket_amd.noise import _creg_value

    def js(bits):
        acc = 0
        for i, b in enumerate(bits):
            x = ((1 << (i % 32)) & 0xFFFFFFFF) if b else 0       # 1 << i, ToInt32 bit pattern
            acc = (acc | x) & 0xFFFFFFFF                          # ToInt32(acc) | ToInt32(x)
        return acc - (1 << 32) if acc >= 1 << 31 else acc

    cases = [[], [True], [False, True], [True, True, False, True], [False] * 31 + [True],
             [False] * 32 + [True], [True] + [False] * 31 + [True], [True] * 40]
    for bits in cases:
        assert _creg_value({"c": bits}, "c") == js(bits) == O.creg_value(bits)
    assert js([False] * 31 + [True]) == -(1 << 31)
    assert js([False] * 32 + [True]) == 1
    assert _creg_value({}, "missing") == 0


def test_vectorised_gate_matrices_equal_the_scalar_constructors():
    """gatevec (many angles at once) against gates.py entry by entry, matrices and
    derivatives of every parametrised gate method."""
    import math as m
    from ket_amd import gatevec as V
    from ket_amd.algorithms import _dmat
    from ket_amd.circuit import _resolve_parametric
    rng = np.random.default_rng(5)
    B = 64
    for name, (fn, k) in V.MATRIX.items():
        vals = [rng.uniform(-7, 7, B) for _ in range(k)]
        got = fn(*vals)
        qs = (0,) if name in ("rx", "ry", "rz", "vz", "u1", "p", "u2", "u3", "gpi", "gpi2") else (0, 1)
        for t in range(B):
            want = _resolve_parametric(name, [float(v[t]) for v in vals], qs)
            g = want[2] if want[0] == "single" else want[3]
            assert np.max(np.abs(np.array(g) - got[t])) <= 4e-16, name
        for which in range(k):
            dv = V.derivative(name, vals, which)
            for t in range(0, B, 7):
                d = _dmat(name, [float(v[t]) for v in vals], which)
                assert np.max(np.abs(np.array(d) - dv[t])) <= 4e-16, (name, which)


def test_batched_tapes_equal_per_set_packing():
    """The batched path traces an ansatz once and fills every set's matrices and
    derivatives vectorised: identical ops and records, and every set's matrix array
    equal to tracing and packing that set alone (efficientSU2 and a random ansatz
    with shared parameters and affine angles over every parametrised gate kind)."""
    from helpers import random_ansatz
    from ket_amd.algorithms import _batch_tapes, _packed_with_derivatives, efficientSU2, trace
    rng = np.random.default_rng(3)
    for ans, npar in ((efficientSU2(6, 2), 36), (random_ansatz(7, 80, 5, 11, wide=False), 5)):
        sets = [list(rng.uniform(-4, 4, npar)) for _ in range(6)]
        n, ops, mats_all, rec = _batch_tapes(ans, sets)
        for t, ps in enumerate(sets):
            flat, records, _, n1 = trace(ans, ps)
            o, m, r = _packed_with_derivatives(flat, records)
            assert n1 == n and np.array_equal(o, ops) and np.array_equal(r, rec)
            assert np.max(np.abs(m - mats_all[t])) <= 1e-15


def test_node_api_addon_error_classes():
    """The Node-API addon (ket_amd/napi/ketsv_napi.c, what ket's TypeScript host
    loads) under this image's Node: ket's error classes reach JavaScript
    (RangeError for sizes and qubit indices, TypeError for malformed tapes and
    arguments, Error without a device), the dry-run scheduler answers, and the
    library's draw stream equals makePrng written in JavaScript."""
    import shutil
    import subprocess
    from ket_amd.build import NAPI_OUT, NAPI_SRC
    if not shutil.which("node") or not os.path.exists(NAPI_OUT):
        pytest.skip("node or the built addon is not present")
    if ket_amd.device_count() > 0:
        pytest.skip("a GPU is present (the GPU suite runs the device half)")
    r = subprocess.run(["node", os.path.join(os.path.dirname(NAPI_SRC), "test_napi.js"), NAPI_OUT],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
