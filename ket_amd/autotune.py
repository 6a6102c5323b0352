"""Schedule autotuner: picks the pass kernel, tile width and fusion width for
a circuit by timing the candidates on the device.

Candidates (exact mode): fused MFMA blocks of 4 qubits on 11-, 12- and
13-qubit tiles, of 5 qubits on 13-qubit tiles, and register-round passes
on 11-, 12- and 13-qubit tiles.  (512-thread workgroups,
KSV_OPT_BLOCK_THREADS, time within noise of 256 and are left out.)  Under ket's per-gate pruning only register-round
passes keep ket's gate order, so only their tile width is tuned.  Every
candidate is first scheduled without a device (ksv_plan_blocks) so invalid
or degenerate ones are skipped; each valid one is then built once and run
`reps` times on a scratch state with HIP-event timing of its passes.  The
winner is cached per (qubits, precision, pruning, circuit structure).
"""
import hashlib

import numpy as np

from ._native import State, plan

_CACHE = {}


def _candidates(prune):
    if prune:
        return [{"kernel": "reg", "rtile": r} for r in (11, 12, 13)]
    return ([{"kernel": "mfma", "bq": 4, "btile": 12},
             {"kernel": "mfma", "bq": 4, "btile": 11}, {"kernel": "mfma", "bq": 4, "btile": 13},
             {"kernel": "mfma", "bq": 5, "btile": 13}] +
            [{"kernel": "reg", "rtile": r} for r in (11, 12, 13)])


def structure_key(n, tape, precision, prune):
    """Hash of what the schedule depends on: kinds and qubits of every op (not the angles)."""
    ops, _ = tape
    h = hashlib.sha1()
    h.update(np.ascontiguousarray(ops["kind"]).tobytes())
    h.update(np.ascontiguousarray(ops["nq"]).tobytes())
    h.update(np.ascontiguousarray(ops["q"]).tobytes())
    return (int(n), precision, bool(prune), h.hexdigest())


def autotune(n, tape, precision="f64", prune=False, reps=2, state=None, verbose=None):
    """Return (best config dict, list of (config, ms per run)) for `tape` on n qubits.

    `state` (optional) is a State of the right shape used as scratch; it is
    reset to |0...0> before every timed run and left in an unspecified state.
    """
    key = structure_key(n, tape, precision, prune)
    if key in _CACHE:
        return _CACHE[key]
    results = []
    scratch = state
    for cand in _candidates(prune):
        try:
            info = plan(n, tape, precision=precision, kernel=cand["kernel"], prune=prune,
                        rtile=cand.get("rtile", 0), btile=cand.get("btile", 0), bq=cand.get("bq", 0))
        except (ValueError, TypeError):
            continue
        if info["passes"] == 0:
            continue
        st = scratch if scratch is not None else State(n, precision, prune=prune)
        scratch = st
        st.set_option(0, 1 if prune else 0)
        st.set_option(4, {"lds": 1, "reg": 2, "mfma": 3}[cand["kernel"]])
        st.set_option(5, cand.get("rtile", 0))
        st.set_option(6, cand.get("bq", 0))
        st.set_option(7, cand.get("btile", 0))
        st.set_option(8, cand.get("bthreads", 0))
        st.set_option(2, 1)
        prog = st.program(tape)
        best = None
        for _ in range(max(1, reps)):
            st.init_basis(0)
            st.reset_stats()
            prog.run()
            s = st.stats()
            ms = s["pass_ms"] + s["other_ms"]
            best = ms if best is None else min(best, ms)
        prog.free()
        results.append((dict(cand, passes=info["passes"]), best))
        if verbose:
            verbose(f"autotune: {cand} -> {info['passes']} passes, {best:.3f} ms")
    if scratch is not None and state is None:
        scratch.free()
    if not results:
        raise RuntimeError("autotune: no candidate schedule")
    best = min(results, key=lambda r: r[1])
    _CACHE[key] = (best[0], results)
    return _CACHE[key]


def state_for(n, config, precision="f64", prune=False, timing=False):
    """A State configured with an autotuned schedule."""
    return State(n, precision, prune=prune, timing=timing, kernel=config["kernel"], rtile=config.get("rtile", 0),
                 bq=config.get("bq", 0), btile=config.get("btile", 0), bthreads=config.get("bthreads", 0))
