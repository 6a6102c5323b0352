"""TEST INFRASTRUCTURE ONLY: drives oracle/ket_sparse.js (the JavaScript
restatement of ket's sparse Map<bigint> statevector algorithm) under Node on a
packed tape, for the CPU-baseline timings and for the check of that
restatement against the C oracle.  Nothing in the product imports this."""
import json
import os
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SCRIPT = os.path.join(HERE, "ket_sparse.js")


def node():
    """Path of the node executable, or None."""
    return shutil.which("node")


def run(n, ops, mats, task="statevector", shots=0, seed=1, warmup=1, runs=5, dump=False, skip=0,
        budget_ms=0.0, timeout=3600):
    """Run `task` ('statevector', 'exactProbs', 'run', 'probs_run' or 'rate')
    on the tape under Node.

    Returns the job's JSON result: per-run times in ms, their median, the final
    Map size, Node's version, 'counts' (first-seen order, index as a decimal
    string) for 'run' and 'probs_run', the gates done within budget_ms after
    the first `skip` for 'rate', and 'amplitudes' (dense complex128) when dump is set."""
    exe = node()
    if exe is None:
        raise RuntimeError("node is not on PATH")
    with tempfile.TemporaryDirectory(prefix="ket_sparse_") as tmp:
        po, pm, pj, pd = (os.path.join(tmp, f) for f in ("ops.bin", "mats.bin", "job.json", "amps.bin"))
        np.ascontiguousarray(ops).tofile(po)
        np.ascontiguousarray(mats, dtype=np.float64).tofile(pm)
        job = {"n": int(n), "ops": po, "mats": pm, "task": task, "shots": int(shots), "seed": int(seed),
               "warmup": int(warmup), "runs": int(runs), "skip": int(skip),
               "budget_ms": float(budget_ms)}
        if dump:
            job["dump"] = pd
        with open(pj, "w") as f:
            json.dump(job, f)
        proc = subprocess.run([exe, "--max-old-space-size=16384", SCRIPT, pj], capture_output=True, text=True,
                              timeout=timeout)
        if proc.returncode != 0:
            raise RuntimeError(f"ket_sparse.js failed ({proc.returncode}): {proc.stderr.strip()[-2000:]}")
        out = json.loads(proc.stdout.strip().splitlines()[-1])
        if dump:
            out["amplitudes"] = np.fromfile(pd, dtype=np.complex128)
        return out


def run_dm(n, ops, mats, p1=0.0, p2=0.0, timeout=3600):
    """ket's runDM restated over a Map (oracle/ket_dm.js) on the tape.

    Returns (rho, order, info): the dense 2^n x 2^n rho (row, col), the keys
    (row << n) | col of the final Map in its insertion order, and the job's
    JSON result (entries, dropped partial sums and their total magnitude, ms)."""
    exe = node()
    if exe is None:
        raise RuntimeError("node is not on PATH")
    with tempfile.TemporaryDirectory(prefix="ket_dm_") as tmp:
        po, pm, pj, pr, pk = (os.path.join(tmp, f) for f in ("ops.bin", "mats.bin", "job.json", "rho.bin",
                                                              "order.bin"))
        np.ascontiguousarray(ops).tofile(po)
        np.ascontiguousarray(mats, dtype=np.float64).tofile(pm)
        with open(pj, "w") as f:
            json.dump({"n": int(n), "ops": po, "mats": pm, "p1": float(p1), "p2": float(p2), "dump_rho": pr,
                       "dump_order": pk}, f)
        proc = subprocess.run([exe, "--max-old-space-size=16384", os.path.join(HERE, "ket_dm.js"), pj],
                              capture_output=True, text=True, timeout=timeout)
        if proc.returncode != 0:
            raise RuntimeError(f"ket_dm.js failed ({proc.returncode}): {proc.stderr.strip()[-2000:]}")
        info = json.loads(proc.stdout.strip().splitlines()[-1])
        rho = np.fromfile(pr, dtype=np.complex128).reshape(1 << n, 1 << n)
        order = np.fromfile(pk, dtype=np.uint64)
        return rho, order, info
