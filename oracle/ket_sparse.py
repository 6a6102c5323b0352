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
