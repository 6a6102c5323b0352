"""Benchmark: gate-updates/s and % of HBM peak for a 30-qubit FP64 circuit.

Workload: 30 qubits in complex128 (16 GiB in HBM), 40 layers of Rx, Ry, Rz on
every qubit (seeded angles) plus a CNOT brick = 4180 gates.  One step =
reset to |0...0> and apply the whole circuit (the tape is already resident
on the device).  The dominant kernel is the fused LDS pass; its achieved
bandwidth is algorithmic bytes (2 x 2^30 x 16 B per launch) over the
average launch time measured with HIP events on the stream the passes are
launched on.  The CPU baseline is the oracle's scalar FP64 restatement on a
bounded prefix of the same circuit at the same 30 qubits.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TFS = 78.6     # MI355X dense FP64 matrix peak
FP32_MFMA_PEAK_TFS = 157.3    # MI355X dense FP32 matrix peak
METRIC = "gate-updates/s & % HBM peak at 30q FP64; adjoint vs param-shift grads/s"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _omp_threads():
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        return os.cpu_count() or 1


def cpu_baseline(n, ops, mats, budget_s):
    """The oracle on the first gates of the same circuit: the OpenMP build over
    every host core it is given (the reported value), then the scalar build on
    one core for a third of the budget (a sub-field).  Both are dense FP64 C
    with ket's gate arithmetic; the two builds give identical amplitudes."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ketsv_oracle as O

    def timed(apply, seconds):
        psi = O.zero_state(n)
        done, t0 = 0, time.perf_counter()
        while done < len(ops) and time.perf_counter() - t0 < seconds:
            apply(n, psi, ops[done:done + 1], mats, prune=False)
            done += 1
        return done, time.perf_counter() - t0

    cores = _omp_threads()
    k, dt = timed(O.apply_tape_omp, budget_s)
    k1, dt1 = timed(O.apply_tape, budget_s / 3)
    return {"value": round(k / dt, 4), "unit": "gate-updates/s", "cores": cores, "kind": "port",
            "sample": f"first {k} of {len(ops)} gates of the same {n}-qubit circuit, oracle dense FP64 C "
                      f"with OpenMP over {cores} threads, {dt:.1f} s",
            "single_core": {"value": round(k1 / dt1, 4), "cores": 1,
                            "sample": f"first {k1} gates, scalar build, {dt1:.1f} s"}}


# statevector() wall times read off ket's published charts (benchmark/charts/
# uniform.svg, qft.svg: GitHub Actions ubuntu-latest, 2 cores, Node 22, one
# timed run after warm-up), in ms
KET_CHART_MS = {"uniform_20": 1940.0, "qft_20": 6100.0, "uniform_16": 50.4, "qft_16": 162.0}


def ket_algorithm_baseline(scaling=False):
    """ket's own CPU algorithm (the sparse Map<bigint> statevector with per-gate
    pruning, restated in JavaScript and run under Node on one host core, as ket
    runs) timed on the same inputs as the GPU path in its default ket-pruning
    mode:
      - the 12-qubit 20-layer random circuit (830 gates): exactProbs plus 10 000
        seeded shots, one warm-up and the median of five, against the GPU path's
        Circuit.exactProbs() + Circuit.run(10000) end to end;
      - H on 20 qubits and qft(20): one timed statevector() each, beside ket's
        published chart points for the same circuits and the GPU's simulate()
        (tape packing, planning and the passes, until a read of the state);
      - the 40-layer random circuit at 16 qubits (and 20, 22 with `scaling`):
        gate-updates/s once the state has filled in (the first layer untimed).
    The JavaScript restatement's amplitudes and seeded counts equal the C
    oracle's bit for bit (tested on the CPU)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ket_sparse as K
    if K.node() is None:
        return {"skipped": "node is not installed on this host"}
    import ket_amd
    from ket_amd import workloads
    from ket_amd.algorithms import qft

    old = ket_amd.configure(prune=True, precision="f64")
    out = {"kind": "port", "cores": 1, "algorithm": "ket's sparse Map<bigint> statevector, per-gate 1e-14 prune, "
                                                   "JavaScript under Node, one thread",
           "gpu_mode": "ket pruning (the library default)"}
    try:
        n, seed, shots = 12, 42, 10000
        c = workloads.random_layers(n, 20, seed=seed)
        ops, mats = c.tape()
        log("bench: ket algorithm (Node) on the 12-qubit circuit")
        r = K.run(n, ops, mats, task="probs_run", shots=shots, seed=seed, warmup=1, runs=5)
        out["node"] = r["node"]
        ts = []
        for i in range(6):
            t0 = time.perf_counter()
            c.exactProbs()
            c.run(shots=shots, seed=seed)
            if i:
                ts.append(time.perf_counter() - t0)
        gpu_ms = sorted(ts)[len(ts) // 2] * 1e3
        out["random_12q_830g_probs_10k_shots"] = {
            "cpu_median_ms": round(r["median_ms"], 1), "gpu_median_ms": round(gpu_ms, 2),
            "speedup": round(r["median_ms"] / gpu_ms, 1), "gates": len(ops),
            "cpu_gate_updates_per_s": round(len(ops) / (r["median_ms"] * 1e-3), 1)}
        for name, circ in (("uniform_20", _uniform(20)), ("qft_20", qft(20))):
            o, m = circ.tape()
            log(f"bench: ket algorithm (Node) on {name}")
            r = K.run(20, o, m, warmup=0, runs=1)
            circ.simulate().free()
            t0 = time.perf_counter()
            st = circ.simulate()
            st.amplitude(0)     # a blocking read: the state is complete
            g = (time.perf_counter() - t0) * 1e3
            st.free()
            out[name] = {"cpu_ms": round(r["median_ms"], 1), "ket_published_ms": KET_CHART_MS[name],
                         "cpu_over_published": round(r["median_ms"] / KET_CHART_MS[name], 2),
                         "gpu_simulate_ms": round(g, 3), "speedup": round(r["median_ms"] / g, 1), "gates": len(o)}
        rates = {}
        for nq in ((16, 20, 22) if scaling else (16,)):
            circ = workloads.random_layers(nq, 40, seed=42)
            o, m = circ.tape()
            log(f"bench: ket algorithm (Node) rate at {nq} qubits")
            r = K.run(nq, o, m, task="rate", skip=3 * nq + nq // 2, budget_ms=4000.0)
            rates[str(nq)] = {"gate_updates_per_s": round(r["gates"] / (r["ms"] * 1e-3), 3), "gates_timed": r["gates"],
                              "ms": round(r["ms"], 1), "map_entries": r["entries"]}
        out["random_layers_rate"] = rates
    finally:
        ket_amd.configure(**old)
    return out


def _uniform(n):
    from ket_amd.circuit import Circuit
    c = Circuit(n)
    for q in range(n):
        c = c.h(q)
    return c


def traffic_from_profiles(kernel_tag):
    """Measured HBM bytes per launch of `kernel_tag` from the committed counter summary, or None."""
    import re
    path = os.path.join(ROOT, "profiles", "hbm_traffic_k_bpass_random30.md")
    try:
        with open(path) as f:
            for line in f:
                m = re.match(r"^\| `([^`]+)` \| (\d+) \| (\d+) \|", line)
                if m and m.group(1) == kernel_tag:
                    return int(m.group(3))
    except OSError:
        pass
    return None


def streaming_ceiling(st, reps=3):
    """Read + write rate of the library's own streaming copy kernel (ksv_copy_state)
    on a state of the same size: 2 x state bytes per copy over the best of `reps`
    synchronous copies (GB/s), the measured ceiling beside the 8 TB/s peak."""
    from ket_amd._native import State
    cp = State(st.n, st.precision, prune=False)
    cp.copy_from(st)                       # first touch
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        cp.copy_from(st)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    cp.free()
    return 2 * (16 if st.precision == "f64" else 8) * (1 << st.n) / best / 1e9


def prune_mode_section(n, ops, mats, steps=2):
    """The library's default mode on the same circuit: ket's per-gate pruning
    (accumulate's 1e-14 rule after every gate, every product and sum rounded in
    ket's order), which forbids fusing gates, so it runs as register-round
    passes (VALU).  Reported beside the exact-mode line, not as `value`."""
    from ket_amd._native import State
    st = State(n, "f64", prune=True)
    prog = st.program((ops, mats))
    info = prog.info()
    dt, stats = run_program_steps(st, prog, steps, 1)
    norm = st.norm2()
    prog.free()
    st.free()
    avg_ms = stats["pass_ms"] / max(1, stats["pass_launches"])
    gbs = stats["pass_bytes"] / max(1, stats["pass_launches"]) / (avg_ms * 1e-3) / 1e9
    return {"gate_updates_per_s": round(len(ops) * steps / dt, 2), "ms_per_step": round(dt / steps * 1e3, 2),
            "passes": info["passes"], "kernel": "ksv::k_rpass", "avg_launch_ms": round(avg_ms, 3),
            "hbm_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4), "norm2": norm}


def run_program_steps(st, prog, steps, warmup):
    """Reset to |0...0> and run the program `warmup` + `steps` times; time the last `steps`."""
    for _ in range(warmup):
        st.init_basis(0)
        prog.run()
    st.set_option(2, 1)     # KSV_OPT_TIMING: HIP events around every pass
    st.reset_stats()
    t0 = time.perf_counter()
    for s in range(steps):
        st.init_basis(0)
        prog.run()
        log(f"bench: step {s + 1}/{steps} done")
    return time.perf_counter() - t0, st.stats()


def compare_states(a, b, n, chunk_log=26):
    """max |a - b| over amplitudes and total-variation distance of |a|^2, |b|^2 (chunked read-back)."""
    import numpy as np
    N, step = 1 << n, 1 << min(chunk_log, n)
    maxd, tvd = 0.0, 0.0
    for off in range(0, N, step):
        x = a.amplitudes(off, step)
        y = b.amplitudes(off, step)
        maxd = max(maxd, float(np.max(np.abs(x - y))))
        tvd += float(np.sum(np.abs(np.abs(x) ** 2 - np.abs(y) ** 2)))
    return maxd, 0.5 * tvd


def workload_f32(args):
    """Complex64 state at 32 qubits (32 GiB), the 40-layer random circuit fused
    into dense blocks on FP32 MFMA (the library default: 4-qubit blocks on
    12-qubit tiles; the 5-qubit-block schedule is timed beside it); accuracy
    against FP64 at 30 qubits."""
    from ket_amd import workloads
    from ket_amd._native import State
    n = args.qubits if args.qubits != 30 else 32
    circ = workloads.random_layers(n, args.layers, seed=42)
    ops, mats = circ.tape()
    gates = len(ops)

    def timed(bq, btile):
        st = State(n, "f32", prune=False, kernel="mfma", bq=bq, btile=btile)
        prog = st.program((ops, mats))
        info = prog.info()
        log(f"bench f32: {n} qubits, {gates} gates -> {info['blocks']} {bq}-qubit blocks in {info['passes']} passes")
        dt, stats = run_program_steps(st, prog, args.steps, args.warmup)
        norm = st.norm2()
        prog.free()
        st.free()
        return dt, stats, info, norm

    bq, bt = args.bq or 4, args.btile or 12
    dt, stats, info, norm = timed(bq, bt)
    five = None
    if not (args.bq or args.btile):
        dt5, _, info5, _ = timed(5, 13)
        five = {"gate_updates_per_s": round(gates * args.steps / dt5, 2), "blocks": info5["blocks"],
                "passes": info5["passes"], "ms_per_step": round(dt5 / args.steps * 1e3, 3)}
    launches = stats["pass_launches"]
    avg_ms = stats["pass_ms"] / launches
    per_launch = stats["pass_bytes"] / launches
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    flops = info["blocks"] * 8 * (1 << bq) * (1 << n) / info["passes"]
    tflops = flops / (avg_ms * 1e-3) / 1e12
    hbm = {"achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4)}
    # counted as 8 real flops per complex multiply-add; executed: 6 (Gauss's three products)
    mfma = {"achieved": round(tflops, 2), "peak": FP32_MFMA_PEAK_TFS, "unit": "TFLOP/s",
            "frac": round(tflops / FP32_MFMA_PEAK_TFS, 4), "executed_tflops": round(tflops * 6 / 8, 2),
            "executed_frac": round(tflops * 6 / 8 / FP32_MFMA_PEAK_TFS, 4)}
    mfma["binding_at_published_peaks"] = flops * 6 / 8 / (FP32_MFMA_PEAK_TFS * 1e12) > per_launch / (HBM_PEAK_GBS * 1e9)
    roofline = {"bound": "hbm", **hbm, "traffic": None, "kernel": "ksv::k_bpass", "avg_launch_ms": round(avg_ms, 4),
                "mfma": mfma}
    # accuracy at 30 qubits: the same seeded circuit in FP64 and in FP32
    na = 30
    ca = workloads.random_layers(na, args.layers, seed=42).tape()
    s64 = State(na, "f64", prune=False)
    s64.apply(ca)
    s32 = State(na, "f32", prune=False, kernel="mfma", bq=bq, btile=bt)
    s32.apply(ca)
    maxd, tvd = compare_states(s64, s32, na)
    s64.free()
    s32.free()
    return {"metric": "gate-updates/s & roofline, complex64 statevector with FP32 MFMA blocks",
            "value": round(gates * args.steps / dt, 2), "unit": "gate-updates/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"random Rx/Ry/Rz + CNOT-brick circuit, {n} qubits, {args.layers} layers, "
                                   f"complex64, fused {bq}-qubit blocks on {bt}-qubit tiles",
                       "qubits": n, "gates": gates, "blocks": info["blocks"], "passes": info["passes"]},
            "roofline": roofline, "five_qubit_blocks": five,
            "norm2_fp64_accumulated": norm,
            "accuracy_vs_fp64_30q": {"max_abs_amp_err": maxd, "tvd": tvd}}


def workload_qft(args):
    """qft(n) on the seeded basis state x = floor(2^n r): closed form
    amplitude_k = exp(2 pi i x k / 2^n) / sqrt(2^n), every p = 2^-n."""
    import math
    import numpy as np
    from ket_amd import workloads
    from ket_amd._native import State
    n = args.qubits
    c, bs, x = workloads.qft_input(n, seed=7)
    ops, mats = c.tape()
    tuned = None
    if not args.no_autotune:
        from ket_amd.autotune import autotune, state_for
        best, cands = autotune(n, (ops, mats), precision="f64", prune=False, reps=2, verbose=log)
        tuned = {"chosen": best, "candidates_ms": [[c_, round(ms, 2)] for c_, ms in cands]}
        st = state_for(n, best, "f64", prune=False)
    else:
        st = State(n, "f64", prune=False)
    prog = st.program((ops, mats))
    info = prog.info()
    for _ in range(args.warmup):
        st.init_basis(x)
        prog.run()
    st.set_option(2, 1)
    st.reset_stats()
    t0 = time.perf_counter()
    for s in range(args.steps):
        st.init_basis(x)
        prog.run()
    dt = time.perf_counter() - t0
    stats = st.stats()
    rng = np.random.RandomState(1)
    ks = rng.randint(0, 1 << n, size=256, dtype=np.int64)
    err = 0.0
    for k in ks.tolist():
        want = complex(math.cos(2 * math.pi * ((x * k) % (1 << n)) / (1 << n)),
                       math.sin(2 * math.pi * ((x * k) % (1 << n)) / (1 << n))) / math.sqrt(1 << n)
        err = max(err, abs(st.amplitude(k) - want))
    idx, p = st.topk(1024)
    prel = float(np.max(np.abs(p * (1 << n) - 1)))
    norm = st.norm2()
    prog.free()
    st.free()
    launches = stats["pass_launches"] + stats["other_launches"]
    ms = (stats["pass_ms"] + stats["other_ms"]) / max(1, launches)
    achieved = 2 * 16 * 2 ** n / (ms * 1e-3) / 1e9
    return {"metric": "gate-updates/s, qft on a seeded basis state", "value": round(len(ops) * args.steps / dt, 2),
            "unit": "gate-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"qft({n}) on |x>, x = {x}", "qubits": n, "gates": len(ops),
                       "passes": info["passes"], "blocks": info["blocks"], "reg_passes": info["reg_passes"],
                       "autotune": tuned},
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "avg_launch_ms": round(ms, 4)},
            "checks": {"max_abs_amp_err_256_samples": err, "top1024_max_rel_p_err": prel, "norm2": norm}}


def adjoint_section(n=24, reps=24, psr_sample=24, fd_sample=8):
    """efficientSU2(n, reps) (2 n (reps + 1) parameters, seeded angles) with the
    transverse-field Ising H = -sum Z_i Z_i+1 - 0.7 sum X_i (open chain):
    adjoint gradient against the device parameter-shift rule (a sample of
    parameters, per-parameter cost extrapolated to all) and central finite
    differences (h = 1e-6)."""
    import numpy as np
    from ket_amd import workloads
    from ket_amd.algorithms import efficientSU2, energy_and_gradient, parameter_shift, realAmplitudes
    ans = efficientSU2(n, reps)
    npar = 2 * n * (reps + 1)
    params = workloads.ansatz_params(npar, seed=1234)
    H = workloads.tfim_hamiltonian(n)
    energy_and_gradient(efficientSU2(n, 1), H, params[:4 * n])       # warm-up
    t0 = time.perf_counter()
    e, g = energy_and_gradient(ans, H, params)
    t_adj = time.perf_counter() - t0
    log(f"bench: adjoint {n} qubits, {npar} params: {t_adj:.3f} s, E = {e:.12f}")
    idx = [int(i) for i in np.linspace(0, npar - 1, psr_sample)]
    tm = {}
    _, gp = parameter_shift(ans, H, params, idx, timings=tm)
    per_param = tm["loop_s"] / len(idx)
    err_psr = max(abs(g[i] - v) for i, v in zip(idx, gp))
    log(f"bench: parameter shift {len(idx)} params: {tm['loop_s']:.3f} s; max |adj - psr| = {err_psr:.2e}")
    fidx = idx[:: max(1, len(idx) // fd_sample)][:fd_sample]
    _, gf = parameter_shift(ans, H, params, fidx, shift=1e-6, central_difference=True)
    err_fd = max(abs(g[i] - v) for i, v in zip(fidx, gf))
    ra = realAmplitudes(n, reps)
    pr = workloads.ansatz_params(n * (reps + 1), seed=1234)
    t0 = time.perf_counter()
    energy_and_gradient(ra, H, pr)
    t_ra = time.perf_counter() - t0
    return {"workload": f"efficientSU2({n},{reps}), {npar} parameters, transverse-field Ising H with "
                        f"{len(H)} Pauli terms, FP64",
            "adjoint_s": round(t_adj, 4), "adjoint_grads_per_s": round(1 / t_adj, 3),
            "param_shift_s_per_param": round(per_param, 5),
            "param_shift_s_full_extrapolated": round(per_param * npar, 3),
            "param_shift_grads_per_s": round(1 / (per_param * npar), 5),
            "speedup_vs_param_shift": round(per_param * npar / t_adj, 1),
            "param_shift_sample": len(idx), "max_abs_err_vs_param_shift": err_psr,
            "fd_sample": len(fidx), "max_abs_err_vs_central_fd_h1e-6": err_fd,
            "realAmplitudes_adjoint_s": round(t_ra, 4), "realAmplitudes_params": n * (reps + 1)}


def workload_adjoint(args):
    """The 24-qubit adjoint section as its own line (gradient evaluations per second)."""
    sec = adjoint_section()
    return {"metric": "full adjoint gradients/s (energy + 1200-parameter gradient)", "value": sec["adjoint_grads_per_s"],
            "unit": "gradients/s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": round(sec["adjoint_s"] * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": sec["workload"]}, "adjoint": sec}


def workload_noisy(args):
    """run() with device noise: the 12-qubit, 20-layer random circuit (830 gates) plus
    a measurement of every qubit into a register, 'harmony' error rates, 10 000
    shots.  Batched trajectories (all shots as copies of one device state) against
    the shot-by-shot driver on a sample of shots; both give identical counts."""
    import ket_amd
    from ket_amd import workloads
    n, shots, seed = 12, args.shots or 10000, 42
    c = workloads.random_layers(n, 20, seed=seed).creg("m", n)
    for q in range(n):
        c = c.measure(q, "m", q)
    c.run(shots=64, seed=1, noise="harmony")                     # warm-up (library load, first launches)
    t0 = time.perf_counter()
    d = c.run(shots=shots, seed=seed, noise="harmony")
    t_b = time.perf_counter() - t0
    log(f"bench: batched {shots} shots in {t_b:.3f} s")
    sample = 40
    old = ket_amd.configure(batch_shots=False)
    try:
        t0 = time.perf_counter()
        d1 = c.run(shots=sample, seed=seed, noise="harmony")
        t_p = (time.perf_counter() - t0) / sample
    finally:
        ket_amd.configure(**old)
    old = ket_amd.configure(batch_shots=True)
    d2 = c.run(shots=sample, seed=seed, noise="harmony")
    ket_amd.configure(**old)
    same = list(d1.probs.items()) == list(d2.probs.items()) and d1.cregs == d2.cregs
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import ketsv_oracle as O
        from ket_amd.noise import resolve_noise
        calls = workloads.random_layer_calls(n, 20, seed) + [("measure", (q, "m", q)) for q in range(n)]
        k, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < min(args.cpu_seconds, 15.0):
            O.run_shots(n, calls, 1, seed + k, resolve_noise("harmony"), cregs={"m": n})
            k += 1
        dt = time.perf_counter() - t0
        cpu = {"value": round(k / dt, 3), "unit": "shots/s", "cores": 1, "kind": "port",
               "sample": f"{k} shots of the same circuit through the oracle's per-shot restatement, {dt:.1f} s"}
    return {"metric": "noisy run() shots/s, 12 qubits, 830 gates + 12 measurements, harmony noise",
            "value": round(shots / t_b, 1), "unit": "shots/s", "n_gpus": 1, "steps": 1, "warmup": 1,
            "ms_per_step": round(t_b * 1e3, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": "random_layers(12, 20, seed 42) + measure all, noise 'harmony', "
                                   f"{shots} shots", "qubits": n, "shots": shots},
            "per_shot_driver_s_per_shot": round(t_p, 5), "speedup_vs_per_shot_driver": round(t_p * shots / t_b, 1),
            "per_shot_sample": sample, "identical_counts_on_sample": same,
            "most_likely": d.most, "entropy_bits": round(d.entropy, 4), "cpu_baseline": cpu}


def workload_batch(args):
    """Many parameter sets of one ansatz per evaluation: 64 sets of
    efficientSU2(16, 8) (272 parameters, transverse-field Ising H) as copies of
    one 22-qubit device state, energies and adjoint gradients of all sets per
    graph replay, against the same sets evaluated one by one."""
    import numpy as np
    from ket_amd import workloads
    from ket_amd.algorithms import efficientSU2, energy_and_gradient, energy_and_gradient_batch
    import ket_amd
    ket_amd.configure(prune=False)
    n, reps, B = 16, 8, 64
    ans = efficientSU2(n, reps)
    H = workloads.tfim_hamiltonian(n)
    npar = 2 * n * (reps + 1)
    rng = np.random.default_rng(2026)
    sets = [list(rng.uniform(-np.pi, np.pi, npar)) for _ in range(B)]
    energy_and_gradient_batch(ans, H, sets)                 # plan, capture, warm-up
    t0 = time.perf_counter()
    for _ in range(args.steps):
        es, gs = energy_and_gradient_batch(ans, H, sets)
    dt = (time.perf_counter() - t0) / args.steps
    energy_and_gradient(ans, H, sets[0])
    sample = 8
    t0 = time.perf_counter()
    errs = []
    for i in range(sample):
        e1, g1 = energy_and_gradient(ans, H, sets[i])
        errs.append(max(abs(e1 - es[i]), float(np.max(np.abs(np.array(g1) - np.array(gs[i]))))))
    t1 = (time.perf_counter() - t0) / sample
    return {"metric": "parameter sets/s (energy + adjoint gradient), batched", "value": round(B / dt, 2),
            "unit": "parameter sets/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
            "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"{B} parameter sets of efficientSU2({n},{reps}) ({npar} parameters each), "
                                   f"transverse-field Ising H, one {n + 6}-qubit state of {B} copies"},
            "single_set_s": round(t1, 5), "speedup_vs_single_sets": round(t1 * B / dt, 1),
            "max_abs_err_vs_single_sets": max(errs)}


def workload_qaoa(args):
    """512 seeded Erdos-Renyi Max-Cut instances (n uniform in 14..22, edge
    probability 0.3, p = 4, gamma and beta uniform in [0, pi)) in one batch:
    <H_C> and adjoint gradients for every instance per evaluation (HIP graph
    replay); complex64 variant against FP64; a parity sample against the
    general engine."""
    import numpy as np
    from ket_amd import workloads
    from ket_amd._native import QaoaBatch
    from ket_amd.algorithms import AnsatzFn, energy_and_gradient, maxCutHamiltonian, qaoa
    import ket_amd
    ket_amd.configure(prune=False)
    P = 4
    inst = workloads.qaoa_instances(count=512, seed=2026, nmin=14, nmax=22, p=P)
    G = np.array([g for _, _, g, _ in inst])
    B = np.array([b for _, _, _, b in inst])
    out = {}
    for prec in ("f64", "f32"):
        batch = QaoaBatch([(n, e) for n, e, _, _ in inst], p=P, precision=prec)
        info = batch.info()
        batch.run(G, B, grad=True)              # capture + warm-up
        t0 = time.perf_counter()
        for _ in range(args.steps):
            e, dg, db = batch.run(G, B, grad=True)
        dt = (time.perf_counter() - t0) / args.steps
        batch.run(G, B, grad=False)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ef = batch.run(G, B, grad=False)
        dtf = (time.perf_counter() - t0) / args.steps
        out[prec] = (e, dg, db, dt, dtf, info)
        batch.free()
        log(f"bench qaoa {prec}: {dt * 1e3:.2f} ms per energy+gradient batch, {dtf * 1e3:.2f} ms energies only")
    e, dg, db, dt, dtf, info = out["f64"]
    e32, dg32, db32, dt32, dtf32, _ = out["f32"]
    amp_bytes = 16 * info["amplitudes"]
    # passes: forward 2P + init (write) + energy (read psi, write lambda); backward 2P passes of psi and lambda
    fwd_bytes = amp_bytes * (1 + 2 * 2 * P + 2)
    bwd_bytes = amp_bytes * (2 * P * 4)
    errs = []
    for i in (0, 1, 2):
        n, edges, g, b = inst[i]
        ans = AnsatzFn(2 * P, lambda ps, n=n, edges=edges: qaoa(n, edges, list(ps[:P]), list(ps[P:])))
        re_, rg = energy_and_gradient(ans, maxCutHamiltonian(n, edges), list(g) + list(b))
        errs.append(max(abs(e[i] - re_), float(np.max(np.abs(np.concatenate([dg[i], db[i]]) - np.array(rg))))))
    return {"metric": "QAOA instances/s (energy + adjoint gradient), batched",
            "value": round(len(inst) / dt, 2), "unit": "instances/s", "n_gpus": 1, "steps": args.steps,
            "warmup": 1, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"{len(inst)} Max-Cut QAOA instances, n uniform 14..22, ER p=0.3, depth {P}",
                       "amplitudes_total": info["amplitudes"], "work_items_per_launch": info["work_items"],
                       "state_bytes": amp_bytes},
            "energies_only_ms": round(dtf * 1e3, 3),
            "hbm_gbs_energy_and_gradient": round((fwd_bytes + bwd_bytes) / dt / 1e9, 1),
            "hbm_gbs_energies_only": round(fwd_bytes / dtf / 1e9, 1),
            "parity_vs_general_engine_3_instances": max(errs),
            "f32": {"ms_per_step": round(dt32 * 1e3, 3), "max_abs_energy_err_vs_f64": float(np.max(np.abs(e32 - e))),
                    "max_abs_grad_err_vs_f64": float(max(np.max(np.abs(dg32 - dg)), np.max(np.abs(db32 - db))))}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--layers", type=int, default=40)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--no-fuse", action="store_true")
    ap.add_argument("--kernel", default="auto", choices=["auto", "lds", "reg", "mfma"])
    ap.add_argument("--bq", type=int, default=0, help="fused-block width (4, 5)")
    ap.add_argument("--btile", type=int, default=0, help="fused-block tile qubits (11..13)")
    ap.add_argument("--bthreads", type=int, default=0, help="fused-block workgroup size (256, 512)")
    ap.add_argument("--rtile", type=int, default=0, help="register-round tile qubits (11..13)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ket-js", action="store_true", help="skip the timing of ket's own algorithm under Node")
    ap.add_argument("--ket-js-scaling", action="store_true",
                    help="add the 20- and 22-qubit points of ket's algorithm (about three minutes more)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the 24-qubit adjoint vs parameter-shift section")
    ap.add_argument("--no-prune-line", action="store_true", help="skip the ket-pruning-mode timing of the same circuit")
    ap.add_argument("--no-autotune", action="store_true", help="use the library's default schedule")
    ap.add_argument("--shots", type=int, default=0, help="noisy workload: shots (default 10000)")
    ap.add_argument("--workload", default="random", choices=["random", "f32", "qft", "qaoa", "noisy", "adjoint", "batch"],
                    help="random: 30-qubit FP64 random circuit (default line); f32: complex64 32-qubit run "
                         "with accuracy against FP64; qft: qft(n) on a seeded basis state; qaoa: 512 batched "
                         "Max-Cut QAOA instances with adjoint gradients; noisy: run() with device noise, "
                         "batched trajectories; adjoint: the 24-qubit adjoint-gradient section alone; batch: 64 "
                         "parameter sets of one ansatz per evaluation")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design, not a run)", flush=True)
        sys.exit(2)

    if args.workload != "random":
        line = {"f32": workload_f32, "qft": workload_qft, "qaoa": workload_qaoa,
                "noisy": workload_noisy, "adjoint": workload_adjoint, "batch": workload_batch}[args.workload](args)
        print(json.dumps(line), flush=True)
        return

    import numpy as np
    from ket_amd import workloads
    from ket_amd._native import State

    n, layers = args.qubits, args.layers
    log(f"bench: building {n}-qubit, {layers}-layer circuit")
    circ = workloads.random_layers(n, layers, seed=42)
    ops, mats = circ.tape()
    gates = len(ops)
    tuned = None
    if not args.no_autotune and args.kernel == "auto" and not (args.rtile or args.bq or args.btile or args.bthreads):
        from ket_amd.autotune import autotune
        log("bench: autotuning the schedule")
        best, cands = autotune(n, (ops, mats), precision="f64", prune=False, reps=2, verbose=log)
        tuned = {"chosen": best, "candidates_ms": [[c, round(ms, 2)] for c, ms in cands]}
        args.kernel, args.rtile = best["kernel"], best.get("rtile", 0)
        args.bq, args.btile = best.get("bq", 0), best.get("btile", 0)
        args.bthreads = best.get("bthreads", 0)
    st = State(n, "f64", prune=False, tile=args.tile, fuse=not args.no_fuse, kernel=args.kernel,
               rtile=args.rtile, bq=args.bq, btile=args.btile, bthreads=args.bthreads)
    prog = st.program((ops, mats))
    info = prog.info()
    log(f"bench: {gates} gates -> {info['passes']} passes ({info['reg_passes']} register-round with "
        f"{info['exchanges']} LDS transposes, {info['blk_passes']} fused-block with {info['blocks']} blocks) "
        f"of {info['tile_qubits']}-qubit tiles")
    for w in range(args.warmup):
        st.init_basis(0)
        prog.run()
        log(f"bench: warmup {w + 1}/{args.warmup} done")
    st.set_option(2, 1)     # KSV_OPT_TIMING: HIP events around every pass
    st.reset_stats()
    t0 = time.perf_counter()
    for s in range(args.steps):
        st.init_basis(0)
        prog.run()
        log(f"bench: step {s + 1}/{args.steps} done")
    dt = time.perf_counter() - t0
    stats = st.stats()
    norm = st.norm2()
    idx, p = st.topk(1024)
    log(f"bench: norm^2 = {norm:.15f}, top-1 p = {p[0]:.6e}")
    if abs(norm - 1) > 1e-10:
        raise SystemExit(f"bench: norm check failed ({norm})")
    prog.free()
    ceiling = streaming_ceiling(st, reps=3)
    log(f"bench: own streaming copy (read + write of the state): {ceiling:.1f} GB/s")
    st.free()

    ms_per_step = dt / args.steps * 1e3
    value = gates * args.steps / dt
    launches = stats["pass_launches"]
    avg_ms = stats["pass_ms"] / launches
    per_launch = stats["pass_bytes"] / launches
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    traffic = (traffic_from_profiles(f"k_bpass<double,{args.bq or 4},{args.btile or 12},{args.bthreads or 256}>@{n}q")
               if info["blk_passes"] == info["passes"] else None)
    # fused-block passes run as k_bpass (one tile per workgroup) or, for FP64 passes of
    # 3-4 dense blocks, as the pipelined tile walk k_bpass_pipe; the roofline covers
    # every pass launch of the schedule
    kname = ("ksv::k_bpass + ksv::k_bpass_pipe (every fused-block pass launch)" if info["blk_passes"]
             else "ksv::k_rpass" if info["reg_passes"] else "ksv::k_pass")
    hbm = {"achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4)}
    if info["blk_passes"] == info["passes"] and info["blocks"]:
        # a dense 2^b-qubit block is a complex 2^b x 2^b by 2^b x 2^(n-b) product:
        # 8 * 2^b real flops per amplitude counted the usual way (4 multiplies and 4 adds
        # per complex multiply-add); the kernel executes 6 * 2^b on MFMA (Gauss's three
        # real products)
        bq = args.bq or 4
        flops_per_launch = info["blocks"] * 8 * (1 << bq) * (1 << n) / info["passes"]
        executed = flops_per_launch * 6 / 8
        tflops = flops_per_launch / (avg_ms * 1e-3) / 1e12
        mfma = {"achieved": round(tflops, 2), "peak": FP64_MFMA_PEAK_TFS, "unit": "TFLOP/s",
                "frac": round(tflops / FP64_MFMA_PEAK_TFS, 4), "flops_per_launch": flops_per_launch,
                "executed_tflops": round(tflops * 6 / 8, 2), "executed_frac": round(tflops * 6 / 8 / FP64_MFMA_PEAK_TFS, 4)}
        # the bound is the resource whose time at its peak is the larger (executed MFMA flops)
        mfma_bound = executed / (FP64_MFMA_PEAK_TFS * 1e12) > per_launch / (HBM_PEAK_GBS * 1e9)
    else:
        mfma, mfma_bound = None, False
    if mfma:
        mfma.pop("flops_per_launch", None)
        mfma["binding_at_published_peaks"] = bool(mfma_bound)
    # the pass streams the state once each way: HBM is the roofline (algorithmic bytes
    # 2 x 2^n x 16 per launch over the HIP-event launch time); the fused blocks' MFMA
    # work is reported beside it
    roofline = {"bound": "hbm", **hbm, "traffic": traffic,
                "traffic_source": ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of these kernels and workload, "
                                   "calibrated by a plain copy (profiles/hbm_traffic_k_bpass_random30.md)")
                if traffic else None,
                "kernel": kname, "bytes_per_launch": per_launch, "avg_launch_ms": round(avg_ms, 4),
                "launches_per_step": launches // args.steps,
                "ceiling_own_copy_gbs": round(ceiling, 1), "frac_of_own_copy": round(achieved / ceiling, 4),
                "mfma": mfma}
    line = {"metric": METRIC, "value": round(value, 2), "unit": "gate-updates/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"random Rx/Ry/Rz + CNOT-brick circuit, {n} qubits, {layers} layers, complex128",
                       "qubits": n, "layers": layers, "gates": gates, "seed": 42, "passes": info["passes"],
                       "tile_qubits": info["tile_qubits"], "fused_1q": not args.no_fuse,
                       "unfused_bound_gates_per_s": round(HBM_PEAK_GBS * 1e9 / (2 * 16 * 2 ** n), 1),
                       "schedule": {"kernel": args.kernel, "bq": args.bq, "btile": args.btile, "rtile": args.rtile,
                                    "bthreads": args.bthreads},
                       "autotune": tuned, "parallelism": "single GPU",
                       "prune": False, "mode": "exact (no per-gate pruning; the fused-block schedule)"},
            "roofline": roofline}
    if not args.no_prune_line:
        log("bench: ket-pruning mode")
        line["ket_pruning_mode"] = prune_mode_section(n, ops, mats)
    if not args.no_adjoint:
        log("bench: adjoint section")
        line["adjoint"] = adjoint_section()
    if not args.no_cpu:
        log("bench: cpu baseline")
        line["cpu_baseline"] = cpu_baseline(n, ops, mats, args.cpu_seconds)
    if not args.no_cpu and not args.no_ket_js:
        line["ket_algorithm_cpu"] = ket_algorithm_baseline(args.ket_js_scaling)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
