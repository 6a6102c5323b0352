"""The benchmark configurations at full size on the GPU.

Config 1 (12 qubits) is covered bit for bit by test_gpu_parity.  Here:
  2a  30 qubits FP64, random Rx/Ry/Rz + CNOT-brick circuit, 40 layers (4180 gates)
  2b  qft(30) on a seeded basis state
  3   32 qubits complex64 (FP32 arithmetic, FP64 accumulation), and complex64
      against FP64 at 30 qubits
  4   efficientSU2(24, 24), 1200 parameters, transverse-field Ising H
  5   512 Max-Cut QAOA instances, 14..22 qubits, p = 4 (FP64 and complex64)
At these sizes the oracle cannot hold or run the whole workload, so the checks
are size-independent: closed forms (product states, the QFT), norms, agreement
of independent kernel families (fused MFMA blocks, register rounds with ket's
rounding), the oracle on pieces it can run (shallower ansatz, small QAOA
instances), the parameter-shift rule and finite differences.
"""
import math

import numpy as np
import pytest

import ket_amd
from ket_amd import workloads
from ket_amd._native import State

import ketsv_oracle as O

pytestmark = pytest.mark.gpu

CHUNK = 1 << 26          # amplitudes per read-back chunk (1 GiB of complex128)


def compare(a, b, n):
    """max |a - b| over all amplitudes and the total-variation distance of |a|^2, |b|^2."""
    N, step = 1 << n, min(CHUNK, 1 << n)
    maxd = tvd = 0.0
    for off in range(0, N, step):
        x = a.amplitudes(off, step)
        y = b.amplitudes(off, step)
        maxd = max(maxd, float(np.max(np.abs(x - y))))
        tvd += float(np.sum(np.abs(np.abs(x) ** 2 - np.abs(y) ** 2)))
    return maxd, 0.5 * tvd


def run_state(n, tape, precision="f64", prune=False, **kw):
    st = State(n, precision, prune=prune, **kw)
    st.apply(tape)
    return st


def brick_product_amplitudes(n, calls, idx):
    """Amplitudes after the first layer of random_layer_calls (rotations, then the
    even CNOT brick), in closed form: each qubit holds Rz Ry Rx |0> (a product state)
    and the brick permutes indices, psi1(x) = psi0(pi(x)) with pi flipping bit
    2i+1 where bit 2i is set."""
    v = np.zeros((n, 2), dtype=np.complex128)
    for q in range(n):
        st = np.array([1, 0], dtype=np.complex128)
        for name, args in calls[3 * q:3 * q + 3]:
            assert args[-1] == q
            st = np.asarray(O.gate_matrix(name, args[:-1])).reshape(2, 2) @ st
        v[q] = st
    x = idx.copy()
    for a in range(0, n - 1, 2):
        x ^= ((x >> np.uint64(a)) & np.uint64(1)) << np.uint64(a + 1)
    amp = np.ones(len(idx), dtype=np.complex128)
    for q in range(n):
        amp *= v[q][((x >> np.uint64(q)) & np.uint64(1)).astype(np.int64)]
    return amp


def ket_pruned_layer0(n, calls, idx):
    """Amplitudes after layer 0 under ket's per-gate pruning, with ket's rounding.

    Before the brick the state is built qubit by qubit: the three rotations on
    qubit q only mix x with x ^ 2^q, and every higher qubit is still |0>, so the
    amplitude of x after qubit q's gates follows from the single amplitude of x's
    lower bits: the pair (s, 0) goes through ket's applySingle arithmetic,
    add(mul(a, x0), mul(b, x1)) rounded step by step (numpy float64 ufuncs do not
    fuse), and accumulate()'s prune at |a|^2 < 1e-14 after each gate
    (src/statevector.ts:19-27, :47-65).  The brick permutes without pruning
    (applyCNOT, :70-80)."""
    x = idx.copy()
    for a in range(0, n - 1, 2):
        x ^= ((x >> np.uint64(a)) & np.uint64(1)) << np.uint64(a + 1)
    sr = np.ones(len(idx))
    si = np.zeros(len(idx))

    def mul(ar, ai, br, bi):
        return ar * br - ai * bi, ar * bi + ai * br

    def prune(r, i):
        keep = (r * r + i * i) >= 1e-14
        return np.where(keep, r, 0.0), np.where(keep, i, 0.0)

    for q in range(n):
        v0r, v0i, v1r, v1i = sr, si, np.zeros_like(sr), np.zeros_like(si)
        for name, args in calls[3 * q:3 * q + 3]:
            m = np.asarray(O.gate_matrix(name, args[:-1])).reshape(2, 2)
            ar, ai = mul(m[0, 0].real, m[0, 0].imag, v0r, v0i)
            br, bi = mul(m[0, 1].real, m[0, 1].imag, v1r, v1i)
            cr, ci = mul(m[1, 0].real, m[1, 0].imag, v0r, v0i)
            dr, di = mul(m[1, 1].real, m[1, 1].imag, v1r, v1i)
            (v0r, v0i), (v1r, v1i) = prune(ar + br, ai + bi), prune(cr + dr, ci + di)
        bit = ((x >> np.uint64(q)) & np.uint64(1)).astype(bool)
        sr, si = np.where(bit, v1r, v0r), np.where(bit, v1i, v0i)
    return sr + 1j * si


def test_config2a_first_layer_closed_form():
    """30 qubits, layer 0 (90 rotations + 15 CNOTs) at 2048 sampled indices: the exact
    paths (fused MFMA blocks, register rounds) against the product-state closed form,
    and the default mode (register rounds under ket's pruning and rounding) bit for
    bit against ket's pruned evolution of the same amplitudes."""
    n = 30
    calls = workloads.random_layer_calls(n, 1, seed=42)
    tape = workloads.build(n, calls).tape()
    rng = np.random.default_rng(1)
    idx = np.unique(rng.integers(0, 1 << n, size=2048, dtype=np.uint64))
    want = brick_product_amplitudes(n, calls, idx)
    pruned = ket_pruned_layer0(n, calls, idx)
    assert np.count_nonzero(pruned == 0) > 0          # the sample holds pruned amplitudes
    for kw in ({"prune": False, "kernel": "mfma"}, {"prune": False, "kernel": "reg"},
               {"prune": True, "kernel": "reg"}):
        st = run_state(n, tape, **kw)
        got = np.array([st.amplitude(int(i)) for i in idx])
        norm = st.norm2()
        st.free()
        if kw["prune"]:
            nd = int(np.count_nonzero(got != pruned))
            print(f"config 2a layer 0 {kw}: {nd} of {len(idx)} amplitudes differ from ket's pruned "
                  f"evolution; norm^2 = {norm!r}")
            assert np.array_equal(got, pruned)
            assert 1 - 1e-5 < norm < 1
        else:
            err = float(np.max(np.abs(got - want)))
            print(f"config 2a layer 0 {kw}: max |d amp| = {err:.2e}, |norm - 1| = {abs(norm - 1):.2e}")
            assert err < 1e-13
            assert abs(norm - 1) < 1e-12


def test_config2a_full_circuit():
    """30 qubits, 40 layers, 4180 gates: the fused-block MFMA path (the benchmark's)
    against register rounds (an independent kernel family) over all 2^30 amplitudes
    to 1e-12, norm 1 to 1e-10, equal top-1024.  The default mode (ket's per-gate
    pruning) drops the amplitudes below 1e-7 that the first layers' product states
    hold (norm^2 ~ 2e-5 of them); the rest of the circuit spreads that loss, so the
    pruned state stays within TVD 1e-2 of the exact one, with norm^2 in
    (1 - 1e-4, 1].  Layer 0 pins the pruned arithmetic bit for bit (above)."""
    n = 30
    tape = workloads.random_layers(n, 40, seed=42).tape()
    assert len(tape[0]) == 4180
    a = run_state(n, tape, prune=False, kernel="mfma")
    b = run_state(n, tape, prune=False, kernel="reg")
    na, nb = a.norm2(), b.norm2()
    assert abs(na - 1) < 1e-10 and abs(nb - 1) < 1e-10
    ia, pa = a.topk(1024)
    ib, pb = b.topk(1024)
    maxd, tvd = compare(a, b, n)
    b.free()
    print(f"config 2a: norm^2 {na!r} / {nb!r}, max |d amp| = {maxd:.2e}, TVD = {tvd:.2e}")
    assert maxd < 1e-12
    assert np.max(np.abs(pa - pb)) < 1e-15
    assert np.array_equal(ia[:64], ib[:64])
    c = run_state(n, tape, prune=True)
    nc = c.norm2()
    maxd, tvd = compare(a, c, n)
    a.free()
    c.free()
    print(f"config 2a, ket's pruning: norm^2 {nc!r}, max |d amp| = {maxd:.2e}, TVD = {tvd:.2e} against exact")
    assert 1 - 1e-4 < nc <= 1 + 1e-12
    assert tvd < 1e-2


def test_config2b_qft30():
    """qft(30) on |x>, x = floor(2^30 r) (seed 7): amplitude k = e^{2 pi i x k / 2^30} / 2^15
    at 512 sampled indices; all top-1024 probabilities equal 2^-30."""
    n = 30
    c, bs, x = workloads.qft_input(n, seed=7)
    tape = c.tape()
    for kw in ({"prune": False}, {"prune": True, "kernel": "reg"}):
        st = State(n, "f64", **kw)
        st.init_basis(x)
        st.apply(tape)
        ks = np.random.default_rng(3).integers(0, 1 << n, size=512)
        err = 0.0
        for k in ks.tolist():
            ph = 2 * math.pi * ((x * k) % (1 << n)) / (1 << n)
            err = max(err, abs(st.amplitude(k) - complex(math.cos(ph), math.sin(ph)) / (1 << 15)))
        _, p = st.topk(1024)
        prel = float(np.max(np.abs(p * (1 << n) - 1)))
        norm = st.norm2()
        st.free()
        print(f"config 2b {kw}: max |d amp| = {err:.2e}, top-1024 max rel p err = {prel:.2e}")
        assert err < 1e-14
        assert prel < 1e-12
        assert abs(norm - 1) < 1e-12


def test_config3_complex64_32q_and_against_fp64_at_30q():
    """complex64 at 32 qubits (32 GiB): the FP64-accumulated norm stays 1 to 1e-5;
    complex64 against FP64 on the same seeded circuit at 30 qubits: max |d amp| < 1e-6,
    TVD < 1e-5 (the stated FP32 bound)."""
    n = 32
    tape = workloads.random_layers(n, 40, seed=42).tape()
    st = run_state(n, tape, precision="f32", kernel="mfma")
    norm = st.norm2()
    _, p = st.topk(16)
    st.free()
    print(f"config 3: 32 qubits complex64 norm^2 = {norm!r}, top p = {p[0]:.3e}")
    assert abs(norm - 1) < 1e-5
    n = 30
    tape = workloads.random_layers(n, 40, seed=42).tape()
    a = run_state(n, tape, precision="f64")
    b = run_state(n, tape, precision="f32", kernel="mfma")
    maxd, tvd = compare(a, b, n)
    a.free()
    b.free()
    print(f"config 3: complex64 vs FP64 at 30 qubits: max |d amp| = {maxd:.2e}, TVD = {tvd:.2e}")
    assert maxd < 1e-6
    assert tvd < 1e-5


def _oracle_energy(n, tape, H):
    from ket_amd.algorithms import _terms
    psi = O.zero_state(n)
    O.apply_tape(n, psi, tape[0], tape[1], prune=False)
    x, z, c = _terms(H, n)
    return sum(float(ci) * O.expect_direct(n, psi, int(xi), int(zi)) for xi, zi, ci in zip(x, z, c))


def test_config4_adjoint_24q_1200_parameters():
    """efficientSU2(24, 24), 1200 parameters, H = -sum Z_i Z_i+1 - 0.7 sum X_i:
    adjoint gradient against the device shift rule on 24 parameters (<= 1e-8) and
    central finite differences h = 1e-6 on 8 (<= 1e-6); energy against the device's
    own forward simulation.  The oracle pins energy and gradient at 24 qubits on
    efficientSU2(24, 1): energy to 1e-10, and its central differences on two
    parameters against the adjoint."""
    from ket_amd.algorithms import efficientSU2, energy_and_gradient, parameter_shift, vqe
    n, reps = 24, 24
    ket_amd.configure(prune=False)
    ans = efficientSU2(n, reps)
    npar = 2 * n * (reps + 1)
    assert npar == 1200
    params = workloads.ansatz_params(npar, seed=1234)
    H = workloads.tfim_hamiltonian(n)
    assert len(H) == 47
    e, g = energy_and_gradient(ans, H, params)
    assert abs(e - vqe(ans(params), H)) < 1e-10
    idx = [int(i) for i in np.linspace(0, npar - 1, 24)]
    _, gp = parameter_shift(ans, H, params, idx)
    err_psr = max(abs(g[i] - v) for i, v in zip(idx, gp))
    fidx = idx[::3]
    _, gf = parameter_shift(ans, H, params, fidx, shift=1e-6, central_difference=True)
    err_fd = max(abs(g[i] - v) for i, v in zip(fidx, gf))
    print(f"config 4: E = {e:.12f}, max |adj - psr| = {err_psr:.2e}, max |adj - fd| = {err_fd:.2e}")
    assert err_psr < 1e-8
    assert err_fd < 1e-6
    # the oracle at 24 qubits on a shallower ansatz of the same family
    small = efficientSU2(n, 1)
    ps = workloads.ansatz_params(4 * n, seed=1234)
    es, gs = energy_and_gradient(small, H, ps)
    eo = _oracle_energy(n, small(ps).tape(), H)
    assert abs(es - eo) < 1e-10
    h = 1e-5
    for i in (5, 70):
        pp, pm = list(ps), list(ps)
        pp[i] += h
        pm[i] -= h
        fd = (_oracle_energy(n, small(pp).tape(), H) - _oracle_energy(n, small(pm).tape(), H)) / (2 * h)
        print(f"config 4 (reps 1): param {i}: adjoint {gs[i]:.12f}, oracle central difference {fd:.12f}")
        assert abs(gs[i] - fd) < 1e-7


def test_config5_qaoa_512_instances():
    """512 seeded Erdos-Renyi Max-Cut instances, n uniform in 14..22, p = 4, one batch:
    energies and adjoint gradients of a spread of instances against ket's qaoa()
    circuit through the general engine (1e-10); two instance states against the
    oracle (1e-12); central finite differences on two instances; complex64 batch
    against FP64 within the stated bound (|d E| < 1e-3, |d grad| < 1e-2)."""
    from ket_amd._native import QaoaBatch
    from ket_amd.algorithms import AnsatzFn, energy_and_gradient, maxCutHamiltonian, qaoa
    ket_amd.configure(prune=False)
    P = 4
    inst = workloads.qaoa_instances(count=512, seed=2026, nmin=14, nmax=22, p=P)
    ns = np.array([n for n, _, _, _ in inst])
    assert ns.min() == 14 and ns.max() == 22
    G = np.array([g for _, _, g, _ in inst])
    B = np.array([b for _, _, _, b in inst])
    out = {}
    for prec in ("f64", "f32"):
        batch = QaoaBatch([(n, e) for n, e, _, _ in inst], p=P, precision=prec)
        out[prec] = batch.run(G, B, grad=True)
        if prec == "f64":
            e, dg, db = out[prec]
            # instances of every width: first of each n
            spread = [int(np.nonzero(ns == n)[0][0]) for n in range(14, 23)]
            for i in spread:
                n, edges, g, b = inst[i]
                ans = AnsatzFn(2 * P, lambda ps, n=n, edges=edges: qaoa(n, edges, list(ps[:P]), list(ps[P:])))
                re_, rg = energy_and_gradient(ans, maxCutHamiltonian(n, edges), list(g) + list(b))
                assert abs(e[i] - re_) < 1e-10, i
                assert np.max(np.abs(np.concatenate([dg[i], db[i]]) - np.array(rg))) < 1e-9, i
            # two 14-qubit instance states against the oracle (forward run leaves them in place)
            batch.run(G, B, grad=False)
            small = [int(i) for i in np.nonzero(ns == 14)[0][:2]]
            for i in small:
                n, edges, g, b = inst[i]
                ops, mats = qaoa(n, edges, g, b).tape()
                ref = O.zero_state(n)
                O.apply_tape(n, ref, ops, mats, prune=False)
                assert np.max(np.abs(batch.amplitudes(i) - ref)) < 1e-12, i
            # central differences of the batch energies on two instances (all instances move together)
            h = 1e-6
            for l in (0, P - 1):
                Gp, Gm = G.copy(), G.copy()
                Gp[:, l] += h
                Gm[:, l] -= h
                fd = (batch.run(Gp, B) - batch.run(Gm, B)) / (2 * h)
                Bp, Bm = B.copy(), B.copy()
                Bp[:, l] += h
                Bm[:, l] -= h
                fdb = (batch.run(G, Bp) - batch.run(G, Bm)) / (2 * h)
                for i in small:
                    assert abs(fd[i] - dg[i][l]) < 1e-6
                    assert abs(fdb[i] - db[i][l]) < 1e-6
        batch.free()
    e, dg, db = out["f64"]
    e32, dg32, db32 = out["f32"]
    de = float(np.max(np.abs(e32 - e)))
    dgr = float(max(np.max(np.abs(dg32 - dg)), np.max(np.abs(db32 - db))))
    print(f"config 5: complex64 vs FP64: max |d E| = {de:.2e}, max |d grad| = {dgr:.2e}")
    assert de < 1e-3
    assert dgr < 1e-2
