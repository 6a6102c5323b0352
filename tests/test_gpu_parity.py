"""GPU parity: libketsv (through the C ABI) against the CPU oracle.

Bars: under ket's per-gate pruning (the default mode) amplitudes and
probabilities equal the oracle's bit for bit: the kernels round every
product and sum in ket's order, and the oracle is built without FMA
contraction.  Exact (unpruned) mode: within 1e-12 (FP64).  Seeded shot
indices identical, gradients within 1e-8 of the parameter-shift rule.
"""
import math

import numpy as np
import pytest

import ket_amd
from ket_amd import Circuit, workloads
from ket_amd._native import State

import ketsv_oracle as O
from helpers import build, dense_ref, random_calls, shot_calls

pytestmark = pytest.mark.gpu

TOL = 1e-12


def gpu_state(n, calls, prune=True, tile=0, init=0, fuse=False, kernel="auto", rtile=0, bq=0, btile=0,
              precision="f64", bthreads=0, bpipe=None):
    c = build(n, calls)
    st = State(n, precision, prune=prune, tile=tile, fuse=fuse, kernel=kernel, rtile=rtile, bq=bq, btile=btile,
               bthreads=bthreads, bpipe=bpipe)
    if init:
        st.init_basis(init)
    st.apply(c.tape())
    a = st.amplitudes()
    st.free()
    return a


def test_config1_probs_and_shots():
    """12 qubits, 20 layers (830 gates): exactProbs and 10 000 seeded shots."""
    n = 12
    calls = workloads.random_layer_calls(n, 20, seed=42)
    assert len(calls) == 830
    ref = dense_ref(n, calls, prune=True)
    c = build(n, calls)
    got = c.statevectorDense()
    assert np.array_equal(got, ref)
    probs = c.exactProbs()
    idx, p = O.probabilities(n, ref)
    assert len(probs) == len(idx)
    for i, v in zip(idx.tolist(), p.tolist()):
        bs = format(i, "b").zfill(n)[::-1]
        assert probs[bs] == v
    dist = c.run(shots=10000, seed=42)
    r = O.prng(42, 10000)
    want = O.sample(n, ref, r)
    u, cnt = np.unique(want, return_counts=True)
    assert len(dist.probs) == len(u)
    for i, k in zip(u.tolist(), cnt.tolist()):
        assert dist.histogram[str(i)] == k / 10000


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12])
@pytest.mark.parametrize("prune", [True, False])
def test_random_circuits_every_op(n, prune):
    calls = random_calls(n, 150, seed=1000 + n)
    ref = dense_ref(n, calls, prune=prune)
    got = gpu_state(n, calls, prune=prune)
    if prune:
        assert np.array_equal(got, ref)
    else:
        assert np.max(np.abs(got - ref)) < TOL


@pytest.mark.parametrize("n,tile", [(14, 8), (15, 10), (16, 12), (17, 11), (18, 13), (18, 12)])
@pytest.mark.parametrize("prune", [False, True])
def test_multi_pass_tiles(n, tile, prune):
    """LDS-tile passes: global controls, diagonal gates on global qubits; under
    pruning bit-equal to the oracle."""
    calls = random_calls(n, 120, seed=7 * n + tile, with_unitary=prune)
    ref = dense_ref(n, calls, prune=prune)
    got = gpu_state(n, calls, prune=prune, tile=tile, kernel="lds")
    if prune:
        assert np.array_equal(got, ref)
    else:
        assert np.max(np.abs(got - ref)) < TOL


@pytest.mark.parametrize("n,rtile", [(11, 11), (12, 12), (13, 12), (13, 13), (14, 11), (16, 12), (17, 13), (18, 12)])
@pytest.mark.parametrize("prune", [True, False])
def test_register_round_passes(n, rtile, prune):
    """Register-round passes: LDS transposes, slot/thread/global controls, every op kind."""
    calls = random_calls(n, 160, seed=31 * n + rtile, with_unitary=True)
    ref = dense_ref(n, calls, prune=prune)
    got = gpu_state(n, calls, prune=prune, kernel="reg", rtile=rtile)
    if prune:
        assert np.array_equal(got, ref)
    else:
        assert np.max(np.abs(got - ref)) < TOL


@pytest.mark.parametrize("n,btile,bq,nt", [(12, 12, 4, 256), (14, 12, 4, 256), (15, 13, 4, 256), (15, 13, 5, 256),
                                           (17, 12, 4, 256), (11, 11, 4, 256), (15, 11, 4, 256), (12, 12, 4, 512),
                                           (16, 12, 4, 512)])
def test_fused_block_passes(n, btile, bq, nt):
    """Dense fused blocks on MFMA (FP64): every op kind fused, wide unitaries as their own steps."""
    calls = random_calls(n, 140, seed=7 * n + bq, with_unitary=True)
    ref = dense_ref(n, calls, prune=False)
    got = gpu_state(n, calls, prune=False, kernel="mfma", btile=btile, bq=bq, bthreads=nt)
    assert np.max(np.abs(got - ref)) < TOL


@pytest.mark.parametrize("n,kind,seed", [(16, "random", 1), (20, "random", 2), (22, "random", 3), (18, "qft", 4),
                                          (21, "calls", 5)])
def test_pipelined_block_passes(n, kind, seed):
    """FP64 passes of 3-4 dense blocks run as a pipelined tile walk (each workgroup
    loads its next tile while the current one's blocks run): amplitudes bit-identical
    to the one-tile-per-workgroup kernel (same per-tile arithmetic) and within 1e-12
    of the oracle, for random layers, QFT (diagonal blocks) and every op kind."""
    from ket_amd._native import plan
    if kind == "random":
        calls = workloads.random_layer_calls(n, 12, seed=seed)
    elif kind == "qft":
        calls = [("x", (q,)) for q in range(0, n, 3)]
        for j in range(n - 1, -1, -1):
            calls.append(("h", (j,)))
            calls += [("cu1", (math.pi / 2 ** (j - k), k, j)) for k in range(j - 1, -1, -1)]
        calls += [("swap", (i, n - 1 - i)) for i in range(n // 2)]
    else:
        calls = random_calls(n, 300, seed=seed)
    c = build(n, calls)
    ref = dense_ref(n, calls, prune=False)
    a = gpu_state(n, calls, prune=False, kernel="mfma", bpipe=False)
    b = gpu_state(n, calls, prune=False, kernel="mfma", bpipe=True)
    assert np.array_equal(a, b)
    assert np.max(np.abs(b - ref)) < TOL
    assert plan(n, c.tape(), kernel="mfma", prune=False)["blocks"] > 0


@pytest.mark.parametrize("btile,nt", [(11, 256), (12, 512)])
def test_fused_block_variants_f32(btile, nt):
    n = 16
    calls = workloads.random_layer_calls(n, 10, seed=9)
    ref = dense_ref(n, calls, prune=False)
    got = gpu_state(n, calls, prune=False, kernel="mfma", btile=btile, bq=4, bthreads=nt, precision="f32")
    assert np.max(np.abs(got - ref)) < 1e-6


F32_TOL = 1e-6      # complex64 state, FP32 arithmetic: max |amp error| against the FP64 oracle


@pytest.mark.parametrize("bq,btile", [(4, 13), (5, 13)])
def test_fused_block_passes_f32(bq, btile):
    n = 16
    calls = workloads.random_layer_calls(n, 10, seed=5)
    ref = dense_ref(n, calls, prune=False)
    got = gpu_state(n, calls, prune=False, kernel="mfma", btile=btile, bq=bq, precision="f32")
    err = np.max(np.abs(got - ref))
    print(f"f32 fused-block bq={bq}: max |d amp| = {err:.3e}")
    assert err < F32_TOL


@pytest.mark.parametrize("rtile", [11, 12, 13])
def test_register_round_random_layers(rtile):
    n = 16
    calls = workloads.random_layer_calls(n, 8, seed=17)
    ref = dense_ref(n, calls, prune=False)
    for fuse in (False, True):
        got = gpu_state(n, calls, prune=False, kernel="reg", rtile=rtile, fuse=fuse)
        assert np.max(np.abs(got - ref)) < TOL


def test_fused_one_qubit_runs_exact_mode():
    n = 14
    calls = workloads.random_layer_calls(n, 6, seed=3)
    ref = dense_ref(n, calls, prune=False)
    got = gpu_state(n, calls, prune=False, fuse=True, tile=10)
    assert np.max(np.abs(got - ref)) < TOL


@pytest.mark.parametrize("prune", [False, True])
def test_wide_unitary(prune):
    import random as _r
    from helpers import random_unitary
    rnd = _r.Random(5)
    n = 9
    calls = [("h", (q,)) for q in range(n)]
    calls.append(("unitary", (random_unitary(32, rnd), 0, 4, 2, 7, 8)))
    calls.append(("unitary", (random_unitary(64, rnd), 8, 1, 3, 5, 6, 0)))
    calls.append(("unitary", (random_unitary(16, rnd), 3, 1, 8, 2)))
    ref = dense_ref(n, calls, prune=prune)
    got = gpu_state(n, calls, prune=prune)
    if prune:
        assert np.array_equal(got, ref)
    else:
        assert np.max(np.abs(got - ref)) < TOL


def test_qft_closed_form():
    n = 10
    c, bs, x = workloads.qft_input(n, seed=7)
    got = c.statevectorDense(initialState=bs)
    k = np.arange(1 << n)
    want = np.exp(2j * np.pi * x * k / (1 << n)) / math.sqrt(1 << n)
    assert np.max(np.abs(got - want)) < 1e-12


def test_initial_state_and_amplitude():
    c = Circuit(3).x(0).cnot(0, 2)
    assert c.exactProbs() == {"101": 1.0}
    assert abs(c.amplitude("101") - 1) < TOL
    assert c.amplitude("000") == 0
    assert c.exactProbs(initialState="100") == {"000": 1.0}
    assert c.exactProbs(initialState="001") == {"100": 1.0}
    b = Circuit(2).h(0).cnot(0, 1)
    assert b.stateAsString() == "0.7071|00⟩ + 0.7071|11⟩"
    assert abs(b.probability("11") - 0.5) < TOL


def test_readouts_against_oracle():
    n = 9
    calls = random_calls(n, 120, seed=99)
    ref = dense_ref(n, calls, prune=True)
    c = build(n, calls)
    assert np.array_equal(c.statevectorDense(), ref)
    m = c.marginals()
    assert np.max(np.abs(np.array(m) - O.marginals(n, ref))) < TOL
    for q in (0, 4, 8):
        th, ph = O.bloch(n, ref, q)
        b = c.blochAngles(q)
        assert abs(b["theta"] - th) < 1e-9 and (abs(b["phi"] - ph) < 1e-9 or abs(abs(b["phi"] - ph) - 2 * math.pi) < 1e-9)
    for pauli in ("XIZYIIXZY", "ZZZZZZZZZ", "IIIIXIIII", "YYIIIIIII"):
        assert abs(c.expectation(pauli) - O.expectation(n, ref, pauli)) < TOL
    H = workloads.tfim_hamiltonian(n)
    assert abs(ket_amd.vqe(c, H) - O.vqe(n, ref, H)) < TOL


def test_exact_mode_expectation_direct():
    n = 8
    calls = random_calls(n, 100, seed=11)
    ket_amd.configure(prune=False)
    ref = dense_ref(n, calls, prune=False)
    c = build(n, calls)
    for pauli in ("XYZIXYZI", "ZIIIIIIZ", "YIIIIIII"):
        x = z = 0
        for q, ch in enumerate(pauli):
            if ch in "XY":
                x |= 1 << q
            if ch in "YZ":
                z |= 1 << q
        assert abs(c.expectation(pauli) - O.expect_direct(n, ref, x, z)) < TOL


def test_topk_and_norm():
    n = 16
    calls = workloads.random_layer_calls(n, 4, seed=5)
    ket_amd.configure(prune=False)
    ref = dense_ref(n, calls, prune=False)
    st = build(n, calls).simulate()
    assert abs(st.norm2() - 1) < 1e-12
    idx, p = st.topk(1024)
    pr = np.abs(ref) ** 2
    order = np.lexsort((np.arange(len(pr)), -pr))[:1024]
    assert np.array_equal(idx.astype(np.int64), order)
    assert np.max(np.abs(p - pr[order])) < TOL
    st.free()


def test_topk_ties_qft():
    """All 2^n probabilities equal: top-k by ascending index among ties."""
    n = 14
    c, bs, x = workloads.qft_input(n, seed=7)
    ket_amd.configure(prune=False)
    st = c.simulate(initialState=bs)
    idx, p = st.topk(100)
    assert np.max(np.abs(p * (1 << n) - 1)) < 1e-9
    st.free()


@pytest.mark.parametrize("n", [4, 13, 17])
def test_sampling_matches_oracle(n):
    calls = workloads.random_layer_calls(n, 3, seed=21 + n)
    ref = dense_ref(n, calls, prune=True)
    st = build(n, calls).simulate()
    r = O.prng(123, 4000)
    got = st.sample(r)
    want = O.sample(n, ref, r)
    st.free()
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [12, 16])
def test_sampling_boundary_draws_bit_exact(n):
    """Draws equal to ket's running sums (cum += p in ascending index, one rounding
    per addition) and to their neighbouring doubles: run()'s draw (ksv_sample) gives
    the oracle's index for every one, and the per-shot draw (ksv_shots_sample, one
    copy and 64 copies) gives sampleSV's."""
    import itertools
    calls = workloads.random_layer_calls(n, 4, seed=5 + n)
    ref = dense_ref(n, calls, prune=True)
    st = build(n, calls).simulate()
    assert np.array_equal(st.amplitudes(), ref)
    p = ref.real * ref.real + ref.imag * ref.imag
    keep = np.nonzero(p > 1e-14)[0]
    cum = np.array(list(itertools.accumulate(p[keep].tolist())))
    sel = np.random.default_rng(n).choice(len(cum) - 1, size=min(2000, len(cum) - 1), replace=False)
    rs = np.concatenate([cum[sel], np.nextafter(cum[sel], 0), np.nextafter(cum[sel], 2)])
    assert np.array_equal(st.sample(rs), O.sample(n, ref, rs))
    for r in rs[::97]:
        assert int(st.shots_sample(n, np.array([r]))[0]) == O.sample_sv(n, ref, r)
    st.free()
    b = 6
    cs = State(n + b, "f64", prune=True)
    cs.shots_init(n, 0)
    cs.apply(build(n, calls).tape())
    draws = rs[:1 << b]
    got = cs.shots_sample(n, draws)
    cs.free()
    assert got.tolist() == [O.sample_sv(n, ref, r) for r in draws.tolist()]


def test_sampling_sparse_state_last_entry():
    """r above the last CDF value lands on the last non-zero index (cdf[-1] := 1)."""
    c = Circuit(14).h(3).cnot(3, 13)
    st = c.simulate()
    # p = (1/sqrt 2)^2 = 0.4999999999999999 each; the CDF total stays below 1
    got = st.sample(np.array([0.2, 0.3, 0.7, 0.9999999999999999]))
    st.free()
    assert got.tolist() == [0, 0, 8 | 8192, 8 | 8192]


def test_collapse_and_per_shot_measure():
    c = Circuit(2).h(0).measure(0, "c", 0).if_("c", 1, lambda k: k.x(1))
    d = c.run(shots=200, seed=9)
    # q1 always equals q0 after the conditional flip
    assert set(d.probs) <= {"00", "11"}
    assert abs(d.cregs["c"][0] - d.probs.get("11", 0)) < 1e-12


def test_noise_run_reproducible():
    c = Circuit(3).h(0).cnot(0, 1).cnot(1, 2)
    a = c.run(shots=100, seed=5, noise="harmony")
    b = c.run(shots=100, seed=5, noise="harmony")
    assert a.probs == b.probs


def _expected_dist(n, shots, counts, cr):
    from ket_amd.circuit import Distribution
    return Distribution(n, shots, counts, cr)


@pytest.mark.parametrize("n,noise,seed", [(3, "harmony", 1), (5, {"p1": 0.02, "p2": 0.05, "pMeas": 0.03}, 2),
                                          (6, None, 3), (9, "aria-1", 4), (14, "harmony", 5),
                                          (15, {"p1": 0.01, "p2": 0.03, "pMeas": 0.02}, 6)])
def test_batched_shots_equal_oracle(n, noise, seed):
    """run() without `if`: all shots as copies of one device state (ksv_shots_*)
    give the oracle's per-shot counts, creg frequencies and first-seen order.  At
    14 and 15 qubits a copy exceeds one tile: Pauli events run as their own
    launches and every measurement or reset splits the gate segments."""
    from ket_amd.noise import resolve_noise
    calls = shot_calls(n, 40, seed)
    d = build(n, calls, cregs={"c": 2}).run(shots=300, seed=seed, noise=noise)
    counts, cr = O.run_shots(n, calls, 300, seed, resolve_noise(noise), cregs={"c": 2})
    e = _expected_dist(n, 300, counts, cr)
    assert list(d.probs.items()) == list(e.probs.items())
    assert d.cregs == e.cregs


def test_noisy_bench_shape_equals_oracle():
    """The noisy benchmark's shape (12 qubits, the 20-layer random circuit, every qubit
    measured into a register, 'harmony' noise) through the register-round copy path:
    counts, register frequencies and first-seen order equal the oracle's per-shot
    restatement."""
    from ket_amd.noise import resolve_noise
    n, shots = 12, 200
    calls = workloads.random_layer_calls(n, 20, seed=42) + [("measure", (q, "m", q)) for q in range(n)]
    d = build(n, calls, cregs={"m": n}).run(shots=shots, seed=42, noise="harmony")
    counts, cr = O.run_shots(n, calls, shots, 42, resolve_noise("harmony"), cregs={"m": n})
    e = _expected_dist(n, shots, counts, cr)
    assert list(d.probs.items()) == list(e.probs.items())
    assert d.cregs == e.cregs


@pytest.mark.parametrize("n,seed,noise", [(3, 1, None), (5, 2, "harmony"), (7, 3, {"p1": 0.02, "p2": 0.05, "pMeas": 0.03}),
                                          (12, 4, "aria-1"), (14, 5, "harmony")])
def test_if_circuits_equal_oracle(n, seed, noise):
    """run() with `if` (the shot-by-shot driver; each shot's draw count depends on its
    measured values): seeded circuits with mid-circuit measurements into multi-bit
    registers, resets, nested `if` blocks holding gates and measurements, and noise,
    give the oracle's per-shot counts, register frequencies and first-seen order."""
    from helpers import if_calls
    from ket_amd.noise import resolve_noise
    calls, regs = if_calls(n, 16, seed)
    assert any(name == "if" for name, _ in calls)
    shots = 150
    d = build(n, calls, cregs=regs).run(shots=shots, seed=seed, noise=noise)
    counts, cr = O.run_shots(n, calls, shots, seed, resolve_noise(noise), cregs=regs)
    e = _expected_dist(n, shots, counts, cr)
    assert list(d.probs.items()) == list(e.probs.items())
    assert d.cregs == e.cregs


def test_batched_shots_equal_per_shot_path():
    """The batched and the shot-by-shot drivers agree exactly (12 qubits, noise)."""
    n = 12
    calls = shot_calls(n, 60, 8, measures=2, resets=2)
    c = build(n, calls, cregs={"c": 2})
    a = c.run(shots=200, seed=77, noise="harmony")
    old = ket_amd.configure(batch_shots=False)
    try:
        b = c.run(shots=200, seed=77, noise="harmony")
    finally:
        ket_amd.configure(**old)
    assert list(a.probs.items()) == list(b.probs.items())
    assert a.cregs == b.cregs


@pytest.mark.parametrize("n,precision,wide", [(13, "f64", False), (13, "f32", False), (10, "f32", False),
                                               (9, "f64", True)])
def test_batched_shots_equal_per_shot_path_variants(n, precision, wide):
    """Batched and shot-by-shot drivers agree exactly: copies that fit one tile
    (per-copy passes: LDS passes up to 10 qubits, register rounds for 11..13, FP64 and
    FP32), and a 5-qubit unitary that splits a per-copy segment."""
    import random
    from helpers import random_unitary
    calls = shot_calls(n, 50, 11 + n, measures=2, resets=1)
    if wide:
        calls.insert(20, ("unitary", (random_unitary(32, random.Random(5)), 0, 2, 4, 6, 8)))
    ket_amd.configure(precision=precision)
    c = build(n, calls, cregs={"c": 2})
    a = c.run(shots=150, seed=31, noise="harmony")
    old = ket_amd.configure(batch_shots=False)
    try:
        b = c.run(shots=150, seed=31, noise="harmony")
    finally:
        ket_amd.configure(**old)
    assert list(a.probs.items()) == list(b.probs.items())
    assert a.cregs == b.cregs


def test_batched_shots_in_several_batches(monkeypatch):
    """More shots than copies per batch: the draw stream continues across batches,
    and the last batch (1 or 4 shots) runs on a state of 1 or 4 copies."""
    import ket_amd.noise as N
    monkeypatch.setattr(N, "MAX_COPIES", 16)
    n = 4
    calls = shot_calls(n, 25, 5)
    for shots in (100, 97):
        d = build(n, calls, cregs={"c": 2}).run(shots=shots, seed=5, noise={"p1": 0.05, "p2": 0.1, "pMeas": 0.02})
        counts, cr = O.run_shots(n, calls, shots, 5, {"p1": 0.05, "p2": 0.1, "pMeas": 0.02}, cregs={"c": 2})
        e = _expected_dist(n, shots, counts, cr)
        assert list(d.probs.items()) == list(e.probs.items())
        assert d.cregs == e.cregs


def test_batched_shots_entry_points():
    """ksv_shots_* directly: per-copy Paulis, measurement with reset, one draw per copy."""
    n, b = 3, 2
    st = State(n + b, "f64")
    st.shots_init(n, 0b001)
    st.shots_pauli(n, 2, [1, 3], [1, 2])          # X on copy 1, Y on copy 3 (qubit 2)
    amps = st.amplitudes().reshape(4, 8)
    assert amps[0, 1] == 1 and amps[2, 1] == 1
    assert amps[1, 0b101] == 1 and amps[3, 0b101] == 1j
    o = st.shots_measure(n, 2, np.array([0.5, 0.5, 0.5, 0.5]), reset=np.array([0, 1, 0, 0], dtype=np.uint8))
    assert o.tolist() == [0, 1, 0, 1]
    amps = st.amplitudes().reshape(4, 8)
    assert amps[1, 0b001] == 1                      # reset moved copy 1 back to q2 = 0
    assert abs(amps[3, 0b101] - 1j) < 1e-15
    idx = st.shots_sample(n, np.array([0.1, 0.2, 0.3, 0.4]))
    assert idx.tolist() == [1, 1, 1, 5]
    st.free()


def test_circuit_matrix_known_answers():
    """ket's circuitMatrix() cases (src/circuit.test.ts:6425-6511), all 2^n columns
    from one run on a 2n-qubit state whose copy t starts as |t>."""
    sq2 = 1 / math.sqrt(2)
    I = np.array(Circuit(2).circuitMatrix())
    assert np.allclose(I, np.eye(4), atol=1e-12)
    assert np.allclose(np.array(Circuit(1).h(0).circuitMatrix()), [[sq2, sq2], [sq2, -sq2]], atol=1e-12)
    assert np.allclose(np.array(Circuit(1).x(0).circuitMatrix()), [[0, 1], [1, 0]], atol=1e-12)
    U = np.array(Circuit(2).h(0).cnot(0, 1).circuitMatrix())
    assert np.allclose(U.conj().T @ U, np.eye(4), atol=1e-12)
    cnot = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]
    assert np.allclose(np.array(Circuit(2).unitary(cnot, 0, 1).circuitMatrix()), cnot, atol=1e-12)
    S = np.array(Circuit(1).s(0).circuitMatrix())
    assert np.allclose(S, [[1, 0], [0, 1j]], atol=1e-12)
    with pytest.raises(TypeError):
        Circuit(1).h(0).creg("c", 1).measure(0, "c", 0).circuitMatrix()
    with pytest.raises(ket_amd.RangeError):
        Circuit(13).circuitMatrix()


@pytest.mark.parametrize("n", [3, 6, 10])
def test_circuit_matrix_equals_oracle_columns(n):
    """Every column against the oracle simulating that basis state alone (q0 = MSB indices)."""
    calls = random_calls(n, 60, seed=40 + n)
    got = np.array(build(n, calls).circuitMatrix())
    dim = 1 << n
    rev = np.array([int(format(i, "b").zfill(n)[::-1], 2) for i in range(dim)])
    for col in range(0, dim, max(1, dim // 64)):
        psi = dense_ref(n, calls, init=int(rev[col]))
        psi = np.where(np.abs(psi) ** 2 >= 1e-14, psi, 0)
        assert np.max(np.abs(got[rev, col] - psi)) < TOL


def test_dm_known_answers():
    """ket's dm() cases (src/circuit.test.ts:5017-5250) on the device density matrix."""
    sq2 = 1 / math.sqrt(2)
    for c in (Circuit(1), Circuit(1).x(0), Circuit(1).h(0), Circuit(2).h(0).cnot(0, 1)):
        assert abs(c.dm().purity() - 1) < 1e-10
    assert Circuit(1).dm().probabilities() == {"0": 1.0}
    assert Circuit(1).x(0).dm().probabilities() == {"1": 1.0}
    p = Circuit(2).h(0).cnot(0, 1).dm().probabilities()
    assert set(p) == {"00", "11"} and abs(p["00"] - 0.5) < 1e-10
    c = Circuit(3).h(0).h(1).h(2)
    ex, pr = c.exactProbs(), c.dm().probabilities()
    assert all(abs(pr[k] - v) < 1e-10 for k, v in ex.items())
    for c in (Circuit(1), Circuit(1).h(0), Circuit(2).h(0).cnot(0, 1)):
        assert abs(c.dm().entropy()) < 1e-10
    assert abs(Circuit(1).x(0).dm().blochAngles(0)["theta"] - math.pi) < 1e-10
    b = Circuit(1).h(0).dm().blochAngles(0)
    assert abs(b["theta"] - math.pi / 2) < 1e-10 and abs(b["phi"]) < 1e-10
    c = Circuit(1).rx(math.pi / 3, 0)
    sv, d = c.blochAngles(0), c.dm().blochAngles(0)
    assert abs(sv["theta"] - d["theta"]) < 1e-8 and abs(sv["phi"] - d["phi"]) < 1e-8
    assert Circuit(1).dm().get(0, 1) == 0
    with pytest.raises(TypeError):
        Circuit(1).creg("c", 1).measure(0, "c", 0).dm()
    with pytest.raises(TypeError):
        Circuit(1).h(0).dm(noise="nonexistent-device")
    pv = [Circuit(1).h(0).dm(noise={"p1": x}).purity() for x in (0.01, 0.1, 0.3)]
    assert pv[0] > pv[1] > pv[2]
    assert abs(Circuit(1).h(0).dm(noise={"p1": 0.1}).get(0, 1).real - (1 - 0.4 / 3) * 0.5) < 1e-8
    c = Circuit(2).h(0).cnot(0, 1)
    assert c.dm(noise="forte-1").purity() > c.dm(noise="harmony").purity()
    assert c.dm(noise={"p2": 0.05}).purity() < 1
    assert abs(Circuit(1).h(0).dm(noise={"p1": 0.75}).entropy() - 1) < 1e-6
    assert 0 < Circuit(1).h(0).dm(noise={"p1": 0.2}).entropy() < 1
    assert abs(sum(c.dm(noise={"p1": 0.05, "p2": 0.05}).probabilities().values()) - 1) < 1e-8
    assert abs(Circuit(3).x(0).x(1).ccx(0, 1, 2).dm().probabilities()["111"] - 1) < 1e-10
    assert abs(Circuit(3).x(0).x(2).cswap(0, 1, 2).dm().probabilities()["110"] - 1) < 1e-10


@pytest.mark.parametrize("n,noise,seed", [(3, None, 1), (3, {"p1": 0.05, "p2": 0.1}, 2), (4, "harmony", 3),
                                          (5, {"p1": 0.02, "p2": 0.03}, 4)])
def test_dm_equals_oracle(n, noise, seed):
    """rho of seeded circuits over every op kind (csrswap, 2- and 4-qubit unitaries,
    toffoli, cswap) with depolarising noise, against the oracle's dense runDM, which
    forms depolarize2's sums whole: 1e-12 without two-qubit noise, else within 2e-7
    (ket drops running sums below 1e-7; the bit-exact comparison with a Map-order
    restatement of ket is in test_gpu_dm.py)."""
    from ket_amd.density import DM_DEVICE_NOISE
    calls = random_calls(n, 30, seed=seed, with_unitary=True)
    calls.append(("csrswap", (0, 1, 2)))
    d = build(n, calls).dm(noise=noise)
    params = DM_DEVICE_NOISE[noise] if isinstance(noise, str) else noise
    ref = O.run_dm(n, calls, params)
    tol = 2e-7 if (params or {}).get("p2") else TOL
    assert np.max(np.abs(d._rho - ref)) < tol
    assert abs(d.purity() - float(np.sum(np.abs(ref) ** 2))) < (1e-6 if tol > TOL else 1e-12)


@pytest.mark.parametrize("precision,tol", [("f64", TOL), ("f32", 1e-6)])
def test_qft_diagonal_blocks(precision, tol):
    """QFT on 16 qubits: blocks of controlled phases run as a phase per row;
    against the oracle and against set_matrix re-parameterising a phase."""
    n = 16
    calls = [("x", (q,)) for q in (0, 3, 7, 12)]
    for j in range(n):
        calls.append(("h", (j,)))
        for k in range(j + 1, n):
            calls.append(("cu1", (math.pi / (1 << (k - j)), k, j)))
    ref = dense_ref(n, calls, prune=False)
    got = gpu_state(n, calls, prune=False, kernel="mfma", btile=12, bq=4, precision=precision)
    assert np.max(np.abs(got - ref)) < tol


def test_errors_map_to_ket_classes():
    with pytest.raises(ket_amd.RangeError):
        Circuit(2).h(2)
    with pytest.raises(TypeError):
        Circuit(2).cnot(1, 1)
    with pytest.raises(TypeError):
        Circuit(3).unitary([[1, 0], [0, 1], [0, 0]], 0)
    with pytest.raises(TypeError):
        Circuit(3).xx(0.3, 1, 1).statevector()
    with pytest.raises(TypeError):
        Circuit(1).rx("theta", 0).statevector()


def _psr_oracle(ansatz, H, params, n):
    out = []
    for i in range(len(params)):
        e = []
        for s in (math.pi / 2, -math.pi / 2):
            p = list(params)
            p[i] += s
            calls = [(op[1][0], op[1][1]) for op in _calls_of(ansatz(p))]
            psi = dense_ref(n, calls, prune=True)
            e.append(O.vqe(n, psi, H))
        out.append(0.5 * (e[0] - e[1]))
    return out


def _calls_of(circ):
    """Recover (method, args) calls from a Circuit built by efficientSU2/realAmplitudes."""
    out = []
    for op in circ.ops:
        if op[0] == "single":
            name, ps = op[3]
            out.append(("c", (name, tuple(float(v) for v in ps) + (op[1],))))
        elif op[0] == "cnot":
            out.append(("c", ("cnot", (op[1], op[2]))))
        else:
            raise AssertionError(op[0])
    return out


@pytest.mark.parametrize("which", ["su2", "real"])
def test_adjoint_matches_parameter_shift(which):
    n = 6
    ans = ket_amd.efficientSU2(n, 2) if which == "su2" else ket_amd.realAmplitudes(n, 3)
    params = workloads.ansatz_params(ans.paramCount, seed=1234)
    H = workloads.tfim_hamiltonian(n)
    g = ket_amd.gradient(ans, H, params)
    want = _psr_oracle(ans, H, params, n)
    assert np.max(np.abs(np.array(g) - np.array(want))) < 1e-8
    gp = ket_amd.gradient(ans, H, params, method="psr")
    assert np.max(np.abs(np.array(gp) - np.array(want))) < 1e-10


def test_adjoint_qaoa_against_finite_differences():
    n = 6
    edges = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (0, 5), (1, 4)]
    H = ket_amd.maxCutHamiltonian(n, edges)

    def ans(p):
        return ket_amd.qaoa(n, edges, p[:2], p[2:])

    params = [0.7, 0.4, 0.3, 1.1]
    ket_amd.configure(prune=False)
    g = ket_amd.gradient(ans, H, params, method="adjoint")
    h = 1e-6
    fd = []
    for i in range(4):
        pp = list(params); pp[i] += h
        pm = list(params); pm[i] -= h
        fd.append((ket_amd.vqe(ans(pp), H) - ket_amd.vqe(ans(pm), H)) / (2 * h))
    assert np.max(np.abs(np.array(g) - np.array(fd))) < 1e-6


def test_program_reuse_and_stats():
    n = 16
    c = workloads.random_layers(n, 4, seed=1)
    st = State(n, "f64", prune=False, timing=True)
    prog = st.program(c.tape())
    info = prog.info()
    assert info["passes"] >= 1
    prog.run()
    a1 = st.amplitudes()
    st.init_basis(0)
    prog.run()
    a2 = st.amplitudes()
    assert np.array_equal(a1, a2)
    s = st.stats()
    assert s["pass_launches"] == 2 * info["passes"] and s["pass_ms"] > 0
    prog.free()
    st.free()


@pytest.mark.parametrize("kernel,kw", [("mfma", {"btile": 12, "bq": 4}), ("reg", {"rtile": 11}), ("lds", {"tile": 9})])
def test_program_set_matrix(kernel, kw):
    """A scheduled program re-run with new angles equals the rebuilt circuit."""
    import random
    n = 13
    rnd = random.Random(3)
    angles_a = [rnd.uniform(0.2, 3.0) for _ in range(3 * n)]
    angles_b = [rnd.uniform(0.2, 3.0) for _ in range(3 * n)]

    def calls(an):
        out = []
        for q in range(n):
            out.append(("ry", (an[q], q)))
            out.append(("rz", (an[n + q], q)))
        for q in range(n - 1):
            out.append(("cnot", (q, q + 1)))
        for q in range(n - 1):
            out.append(("xx", (an[2 * n + q], q, q + 1)))
        return out

    ta, tb = build(n, calls(angles_a)).tape(), build(n, calls(angles_b)).tape()
    st = State(n, "f64", prune=False, kernel=kernel, **kw)
    prog = st.program(ta)
    for i in range(len(ta[0])):
        o = int(ta[0][i]["mat"])
        if o < 0:
            continue
        dim = {0: 2, 3: 4, 4: 2}[int(ta[0][i]["kind"])]
        m = tb[1][o:o + 2 * dim * dim].view(np.complex128)
        prog.set_matrix(i, m)
    st.init_basis(0)
    prog.run()
    got = st.amplitudes()
    prog.free()
    st.free()
    assert np.max(np.abs(got - dense_ref(n, calls(angles_b), prune=False))) < TOL


def test_adjoint_vs_parameter_shift_vs_finite_differences():
    """efficientSU2(8, 2), transverse-field Ising H: adjoint = device shift rule to 1e-10,
    = central finite differences (h = 1e-6) to 1e-6."""
    from ket_amd.algorithms import efficientSU2, energy_and_gradient, parameter_shift
    n = 8
    ans = efficientSU2(n, 2)
    params = workloads.ansatz_params(2 * n * 3, seed=1234)
    H = workloads.tfim_hamiltonian(n)
    e, g = energy_and_gradient(ans, H, params)
    _, gp = parameter_shift(ans, H, params)
    assert max(abs(a - b) for a, b in zip(g, gp)) < 1e-10
    idx = list(range(0, len(params), 5))
    _, gf = parameter_shift(ans, H, params, idx, shift=1e-6, central_difference=True)
    assert max(abs(g[i] - v) for i, v in zip(idx, gf)) < 1e-6


def _qaoa_ansatz(n, edges, p):
    from ket_amd.algorithms import AnsatzFn, qaoa
    return AnsatzFn(2 * p, lambda ps: qaoa(n, edges, list(ps[:p]), list(ps[p:])))


@pytest.mark.parametrize("precision,tol", [("f64", 1e-10), ("f32", 2e-4)])
def test_qaoa_batch_energy_and_gradients(precision, tol):
    """Batched QAOA (uneven widths, one launch per pass) against ket's qaoa() circuit run
    through the general engine: <H_C> and adjoint gradients (FP64 to 1e-10, complex64 with
    FP64 accumulation to the stated bound)."""
    from ket_amd._native import QaoaBatch
    from ket_amd.algorithms import energy_and_gradient, maxCutHamiltonian
    ket_amd.configure(prune=False)
    p = 2
    inst = workloads.qaoa_instances(count=6, seed=2026, nmin=12, nmax=15, p=p)
    batch = QaoaBatch([(n, e) for n, e, _, _ in inst], p=p, precision=precision)
    G = np.array([g for _, _, g, _ in inst])
    B = np.array([b for _, _, _, b in inst])
    e, dg, db = batch.run(G, B, grad=True)
    e2 = batch.run(G, B, grad=False)           # graph replay of the forward schedule
    assert np.max(np.abs(e - e2)) < 1e-12
    for i, (n, edges, g, b) in enumerate(inst):
        ans = _qaoa_ansatz(n, edges, p)
        ref_e, ref_g = energy_and_gradient(ans, maxCutHamiltonian(n, edges), list(g) + list(b))
        assert abs(e[i] - ref_e) < tol
        assert np.max(np.abs(np.concatenate([dg[i], db[i]]) - np.array(ref_g))) < tol * 10
    batch.free()


def test_qaoa_batch_state_and_finite_differences():
    from ket_amd._native import QaoaBatch
    from ket_amd.algorithms import qaoa
    ket_amd.configure(prune=False)
    n, p = 13, 3
    inst = workloads.qaoa_instances(count=1, seed=11, nmin=n, nmax=n, p=p)
    _, edges, g, b = inst[0]
    batch = QaoaBatch([(n, edges)], p=p)
    e, dg, db = batch.run([g], [b], grad=True)
    batch.run([g], [b], grad=False)
    got = batch.amplitudes(0)
    c = qaoa(n, edges, g, b)
    ops, mats = c.tape()
    ref = O.zero_state(n)
    O.apply_tape(n, ref, ops, mats, prune=False)
    assert np.max(np.abs(got - ref)) < TOL
    h = 1e-6
    for l in range(p):
        gp, gm = list(g), list(g)
        gp[l] += h
        gm[l] -= h
        fd = (batch.run([gp], [b])[0] - batch.run([gm], [b])[0]) / (2 * h)
        assert abs(fd - dg[0][l]) < 1e-6
        bp, bm = list(b), list(b)
        bp[l] += h
        bm[l] -= h
        fd = (batch.run([g], [bp])[0] - batch.run([g], [bm])[0]) / (2 * h)
        assert abs(fd - db[0][l]) < 1e-6
    batch.free()


@pytest.mark.gpu
@pytest.mark.parametrize("precision,n,seed,tol", [("f64", 14, 11, 1e-7), ("f64", 13, 12, 1e-7), ("f32", 15, 13, 3e-3)])
def test_adjoint_tiled_random_ansatz(precision, n, seed, tol):
    """Device adjoint sweep (tile passes over phi and lambda, global controls, wide
    steps) on a random ansatz over every parametrised gate kind: against central
    finite differences of the oracle's energy (h = 1e-5), and bit-for-bit-close to
    the host emulation of the same plan in FP64."""
    from ket_amd.algorithms import energy_and_gradient
    from test_adjoint_emu import _fd, _energy, emu_gradient
    from helpers import random_ansatz
    nparam = 7
    ans = random_ansatz(n, 110, nparam, seed)
    params = list(np.random.default_rng(seed).uniform(-2, 2, nparam))
    H = workloads.tfim_hamiltonian(n)
    ket_amd.configure(precision=precision, prune=False)
    e, g = energy_and_gradient(ans, H, params)
    fd = _fd(ans, params, H, range(nparam))
    assert np.max(np.abs(np.array(g) - fd)) < tol, (g, fd)
    assert abs(e - _energy(ans, params, H)) < (1e-10 if precision == "f64" else 1e-4)
    if precision == "f64":
        ge, _, _ = emu_gradient(ans, params, H, K=11)
        assert np.max(np.abs(np.array(g) - ge)) < 1e-11


def test_adjoint_without_terms_large_state():
    """An empty Hamiltonian at 27 qubits: lambda is cleared (2 GiB, a byte count past
    2^31) and the energy and every gradient entry are exactly zero."""
    from ket_amd.algorithms import efficientSU2, energy_and_gradient
    n = 27
    ans = efficientSU2(n, 1)
    params = workloads.ansatz_params(4 * n, seed=3)
    ket_amd.configure(prune=False)
    e, g = energy_and_gradient(ans, [], params)
    assert e == 0.0 and all(v == 0.0 for v in g)


def test_node_api_addon_on_the_device():
    """The Node-API addon drives the device from JavaScript: a Bell pair through
    create / apply / probs / sample / expect (draws from makePrng written in JS),
    a 10-qubit uniform state through amplitudes, error classes on the device path."""
    import os
    import shutil
    import subprocess
    from ket_amd.build import NAPI_OUT, NAPI_SRC
    if not shutil.which("node"):
        pytest.skip("node is not installed")
    assert os.path.exists(NAPI_OUT), "ketsv.node is not built"
    r = subprocess.run(["node", os.path.join(os.path.dirname(NAPI_SRC), "test_napi.js"), NAPI_OUT, "gpu"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
