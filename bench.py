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
METRIC = "gate-updates/s & % HBM peak at 30q FP64; adjoint vs param-shift grads/s"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(n, ops, mats, budget_s):
    """Oracle (1 core, scalar FP64) on the first gates of the same circuit."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ketsv_oracle as O
    psi = O.zero_state(n)
    done, t0 = 0, time.perf_counter()
    while done < len(ops) and time.perf_counter() - t0 < budget_s:
        O.apply_tape(n, psi, ops[done:done + 1], mats, prune=False)
        done += 1
    dt = time.perf_counter() - t0
    del psi
    return {"value": done / dt, "unit": "gate-updates/s", "cores": 1, "kind": "port",
            "sample": f"first {done} of {len(ops)} gates of the same {n}-qubit circuit, oracle scalar C, {dt:.1f} s"}


def traffic_from_profiles(kernel_tag):
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(kernel_tag, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


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
    ap.add_argument("--btile", type=int, default=0, help="fused-block tile qubits (12, 13)")
    ap.add_argument("--rtile", type=int, default=0, help="register-round tile qubits (11..13)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design, not a run)", flush=True)
        sys.exit(2)

    import numpy as np
    from ket_amd import workloads
    from ket_amd._native import State

    n, layers = args.qubits, args.layers
    log(f"bench: building {n}-qubit, {layers}-layer circuit")
    circ = workloads.random_layers(n, layers, seed=42)
    ops, mats = circ.tape()
    gates = len(ops)
    st = State(n, "f64", prune=False, tile=args.tile, fuse=not args.no_fuse, kernel=args.kernel,
               rtile=args.rtile, bq=args.bq, btile=args.btile)
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
    st.free()

    ms_per_step = dt / args.steps * 1e3
    value = gates * args.steps / dt
    launches = stats["pass_launches"]
    avg_ms = stats["pass_ms"] / launches
    per_launch = stats["pass_bytes"] / launches
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profiles(f"k_pass_f64_{n}q"),
                "kernel": ("ksv::k_bpass" if info["blk_passes"] else "ksv::k_rpass" if info["reg_passes"] else "ksv::k_pass"),
                "bytes_per_launch": per_launch,
                "avg_launch_ms": round(avg_ms, 4), "launches_per_step": launches // args.steps}
    line = {"metric": METRIC, "value": round(value, 2), "unit": "gate-updates/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"random Rx/Ry/Rz + CNOT-brick circuit, {n} qubits, {layers} layers, complex128",
                       "qubits": n, "layers": layers, "gates": gates, "seed": 42, "passes": info["passes"],
                       "tile_qubits": info["tile_qubits"], "fused_1q": not args.no_fuse,
                       "unfused_bound_gates_per_s": round(HBM_PEAK_GBS * 1e9 / (2 * 16 * 2 ** n), 1),
                       "parallelism": "single GPU"},
            "roofline": roofline}
    if not args.no_cpu:
        log("bench: cpu baseline")
        line["cpu_baseline"] = cpu_baseline(n, ops, mats, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
