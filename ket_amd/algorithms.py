"""Algorithm library on the hot path (src/algorithms.ts): QFT, QAOA / Max-Cut,
hardware-efficient ansätze, and the VQE energy, gradient and minimiser.

vqe() simulates the ansatz once on the GPU and reads every Pauli term off
the device state; ket re-simulates the circuit once per term
(src/algorithms.ts:301-336).  gradient() computes the adjoint gradient on the
GPU when ket's parameter-shift rule is exact for the ansatz (then the two
agree to rounding), and otherwise runs ket's parameter-shift formula on the
GPU so the numbers stay ket's; method='adjoint' / 'psr' force either.
"""
import math

import numpy as np

from . import gates as G
from .circuit import OPTIONS, Circuit, Sym, _TraceCtx, _flatten, pauli_masks
from ._native import PARAM_DTYPE, RangeError, State
from .tape import pack


# ---- QFT (src/algorithms.ts:16-39) --------------------------------------------
def qft(n):
    c = Circuit(n)
    for j in range(n - 1, -1, -1):
        c = c.h(j)
        for k in range(j - 1, -1, -1):
            c = c.cu1(math.pi / 2 ** (j - k), k, j)
    for i in range(n // 2):
        c = c.swap(i, n - 1 - i)
    return c


def iqft(n):
    c = Circuit(n)
    for i in range(n // 2):
        c = c.swap(i, n - 1 - i)
    for j in range(n):
        for k in range(j - 1, -1, -1):
            c = c.cu1(-math.pi / 2 ** (j - k), k, j)
        c = c.h(j)
    return c


# ---- QAOA / Max-Cut (src/algorithms.ts:231-276) ----------------------------------
def maxCutHamiltonian(n, edges):
    if len(edges) == 0:
        return []
    terms = [{"coeff": len(edges) / 2, "ops": "I" * n}]
    for u, v in edges:
        arr = ["I"] * n
        arr[n - 1 - u] = "Z"
        arr[n - 1 - v] = "Z"
        terms.append({"coeff": -0.5, "ops": "".join(arr)})
    return terms


def qaoa(n, edges, gamma, beta):
    if len(gamma) != len(beta):
        raise TypeError(f"gamma and beta must have equal length (got {len(gamma)} vs {len(beta)})")
    c = Circuit(n)
    for q in range(n):
        c = c.h(q)
    for l in range(len(gamma)):
        for u, v in edges:
            c = c.cnot(u, v).rz(-gamma[l], v).cnot(u, v)
        for q in range(n):
            c = c.rx(2 * beta[l], q)
    return c


# ---- ansätze (src/algorithms.ts:362-399) ----------------------------------------
class AnsatzFn:
    """An ansatz function with its parameter count (src/algorithms.ts:340-348)."""

    def __init__(self, param_count, fn):
        self.paramCount = param_count
        self._fn = fn

    def __call__(self, params):
        return self._fn(params)


def realAmplitudes(n, reps=3):
    count = n * (reps + 1)

    def build(params):
        if len(params) != count:
            raise RangeError(f"realAmplitudes({n}, {reps}) needs {count} params, got {len(params)}")
        c, p = Circuit(n), 0
        for _ in range(reps):
            for q in range(n):
                c = c.ry(params[p], q)
                p += 1
            for q in range(n - 1):
                c = c.cnot(q, q + 1)
        for q in range(n):
            c = c.ry(params[p], q)
            p += 1
        return c

    return AnsatzFn(count, build)


def efficientSU2(n, reps=3):
    count = 2 * n * (reps + 1)

    def build(params):
        if len(params) != count:
            raise RangeError(f"efficientSU2({n}, {reps}) needs {count} params, got {len(params)}")
        c, p = Circuit(n), 0
        for _ in range(reps):
            for q in range(n):
                c = c.ry(params[p], q)
                c = c.rz(params[p + 1], q)
                p += 2
            for q in range(n - 1):
                c = c.cnot(q, q + 1)
        for q in range(n):
            c = c.ry(params[p], q)
            c = c.rz(params[p + 1], q)
            p += 2
        return c

    return AnsatzFn(count, build)


# ---- Hamiltonians -----------------------------------------------------------------
def _terms(hamiltonian, n):
    """PauliTerms -> (xmask, zmask, coeff) arrays; ops[0] acts on qubit n-1."""
    xs, zs, cs = [], [], []
    for t in hamiltonian:
        coeff = t["coeff"] if isinstance(t, dict) else t.coeff
        ops = t["ops"] if isinstance(t, dict) else t.ops
        if len(ops) != n:
            raise TypeError(f"ops '{ops}' length must equal ansatz.qubits ({n})")
        up = ops.upper()
        if any(ch not in "IXYZ" for ch in up):
            raise TypeError(f"ops '{ops}' must contain only I, X, Y, Z")
        x, z = pauli_masks(up, q0_left=False)
        xs.append(x)
        zs.append(z)
        cs.append(float(coeff))
    return np.array(xs, dtype=np.uint64), np.array(zs, dtype=np.uint64), np.array(cs, dtype=np.float64)


def vqe(ansatz, hamiltonian):
    """<ansatz|H|ansatz> (src/algorithms.ts:301-336)."""
    n = ansatz.qubits
    x, z, c = _terms(hamiltonian, n)
    if len(c) == 0:
        return 0.0
    st = ansatz.simulate()
    _, tot = st.expect_pauli(x, z, c, mode=1 if OPTIONS["prune"] else 0)
    st.free()
    return tot


# ---- gradients --------------------------------------------------------------------
# Parametrised gates whose angles enter as exp(-i theta P / 2)-like factors with
# unit frequency: ket's two-point shift by pi/2 is exact for them.
_PSR_EXACT = {"rx", "ry", "rz", "vz", "u1", "p", "u2", "u3", "xx", "yy", "zz", "cu1"}


def _dmat(name, vals, which):
    """Derivative of a gate's 2x2 / 4x4 matrix w.r.t. its `which`-th angle."""
    if name in ("rx", "crx"):
        return G.dRx(vals[0])
    if name in ("ry", "cry"):
        return G.dRy(vals[0])
    if name in ("rz", "vz", "crz"):
        return G.dRz(vals[0])
    if name in ("u1", "p", "cu1"):
        return G.dU1(vals[0])
    if name in ("u2", "cu2"):
        return G.dU2(vals[0], vals[1], which)
    if name in ("u3", "cu3"):
        return G.dU3(vals[0], vals[1], vals[2], which)
    if name == "gpi":
        return G.dGpi(vals[0])
    if name == "gpi2":
        return G.dGpi2(vals[0])
    if name == "xx":
        return G.dXx(vals[0])
    if name == "yy":
        return G.dYy(vals[0])
    if name == "zz":
        return G.dZz(vals[0])
    if name == "xy":
        return G.dXy(vals[0])
    if name == "ms":
        return G.dMs(vals[0], vals[1], which)
    return None


def trace(ansatz, params):
    """Call `ansatz` with symbolic parameters.

    Returns (flat_ops, records, psr_exact) where records lists
    (op index, param index, coeff, derivative matrix), or None when the
    ansatz does something other than affine arithmetic with its parameters.
    """
    ctx = _TraceCtx()
    syms = [Sym({i: 1.0}, 0.0, float(v), ctx) for i, v in enumerate(params)]
    try:
        circ = ansatz(syms)
    except (TypeError, ValueError, AttributeError):
        return None
    if ctx.escaped or not isinstance(circ, Circuit):
        return None
    from . import gatevec as V
    flat = _flatten(circ._ops())
    pending = []            # (op, param, coeff, name, which, angle values)
    uses = {}
    exact = True
    for i, op in enumerate(flat):
        meta = op[-1] if op[0] in ("single", "two", "controlled") else None
        if not meta or not isinstance(meta, tuple):
            continue
        name, ps = meta
        vals = None
        for which, p in enumerate(ps):
            if not isinstance(p, Sym):
                continue
            for pi, coeff in p.terms.items():
                if coeff == 0:
                    continue
                if name not in V.MATRIX:
                    return None
                if vals is None:
                    vals = [p.value if isinstance(p, Sym) else float(p) for p in ps]
                pending.append((i, pi, coeff, name, which, vals))
                uses[pi] = uses.get(pi, 0) + 1
                if name not in _PSR_EXACT or abs(coeff) != 1.0:
                    exact = False
    if any(v != 1 for v in uses.values()):
        exact = False
    # every derivative matrix of one gate kind and angle in one vectorised call
    groups = {}
    for j, rec in enumerate(pending):
        groups.setdefault((rec[3], rec[4]), []).append(j)
    dmats = [None] * len(pending)
    for (name, which), js in groups.items():
        k = V.MATRIX[name][1]
        args = [np.array([pending[j][5][a] for j in js], dtype=np.float64) for a in range(k)]
        res = V.derivative(name, args, which)
        if res is None:
            return None
        for j, row in zip(js, res):
            dmats[j] = row
    records = [(i, pi, coeff, d) for (i, pi, coeff, _, _, _), d in zip(pending, dmats)]
    return flat, records, exact, circ.qubits


def _adjoint(flat, records, nparam, n, hamiltonian):
    ops, mats, rec = _packed_with_derivatives(flat, records)
    x, z, c = _terms(hamiltonian, n)
    st = State(n, OPTIONS["precision"], prune=False, tile=OPTIONS["tile"], adj_tile=OPTIONS["adj_tile"],
               adj_threads=OPTIONS["adj_threads"])
    grad, energy = st.adjoint((ops, mats), rec, nparam, x, z, c)
    st.free()
    return grad, energy


def gradient(ansatz, hamiltonian, params, method="auto"):
    """dE/dparams (src/algorithms.ts:536-547).

    method: 'auto' (adjoint when ket's shift rule is exact, else ket's shift
    rule), 'adjoint' (exact gradient by adjoint differentiation) or 'psr'
    (ket's parameter-shift formula, each energy simulated on the GPU).
    """
    params = [float(p) for p in params]
    if method not in ("auto", "adjoint", "psr"):
        raise TypeError(f"unknown gradient method '{method}'")
    if method != "psr":
        tr = trace(ansatz, params)
        if tr is not None:
            flat, records, exact, n = tr
            if method == "adjoint" or exact:
                grad, _ = _adjoint(flat, records, len(params), n, hamiltonian)
                return [float(g) for g in grad]
        if method == "adjoint":
            raise TypeError("adjoint gradient needs an ansatz built from affine functions of its parameters")
    shift = math.pi / 2
    out = []
    for i in range(len(params)):
        plus = [p + shift if j == i else p for j, p in enumerate(params)]
        minus = [p - shift if j == i else p for j, p in enumerate(params)]
        out.append(0.5 * (vqe(ansatz(plus), hamiltonian) - vqe(ansatz(minus), hamiltonian)))
    return out


def energy_and_gradient(ansatz, hamiltonian, params):
    """(E, dE/dparams) from one adjoint sweep (exact, no pruning)."""
    params = [float(p) for p in params]
    tr = trace(ansatz, params)
    if tr is None:
        raise TypeError("adjoint gradient needs an ansatz built from affine functions of its parameters")
    flat, records, _, n = tr
    grad, e = _adjoint(flat, records, len(params), n, hamiltonian)
    return float(e), [float(g) for g in grad]


# ---- many parameter sets at once ---------------------------------------------------
_BATCHES = {}            # structure key -> AdjointBatch (the few most recent)
_BATCH_KEEP = 4


def _packed_with_derivatives(flat, records):
    """(ops, mats with every record's dG appended, records as PARAM_DTYPE)."""
    ops, mats = pack(flat)
    rec = np.zeros(len(records), dtype=PARAM_DTYPE)
    if not records:
        return ops, mats, rec
    ds = [np.asarray(d, dtype=np.complex128).reshape(-1) for _, _, _, d in records]
    sizes = np.array([2 * len(d) for d in ds], dtype=np.int64)
    rec["op"] = [r[0] for r in records]
    rec["param"] = [r[1] for r in records]
    rec["coeff"] = [r[2] for r in records]
    rec["dmat"] = len(mats) + np.concatenate([[0], np.cumsum(sizes)[:-1]])
    extra = np.concatenate(ds).view(np.float64)
    return ops, np.concatenate([mats, extra]), rec


def _scatter(mats_all, offs, res, m, d):
    """res (m * B x d, op-major) into columns offs[o] .. of every row, interleaved (re, im)."""
    B = mats_all.shape[0]
    cols = (np.asarray(offs, dtype=np.int64)[:, None] + 2 * np.arange(d)[None, :]).reshape(-1)
    r = res.reshape(m, B, d).transpose(1, 0, 2).reshape(B, m * d)
    mats_all[:, cols] = r.real
    mats_all[:, cols + 1] = r.imag


def _batch_tapes(ansatz, params_list):
    """The tape of an ansatz for many parameter sets: traced once (the trace
    fails when the structure depends on parameter values), then every
    parametrised gate's matrix and derivative evaluated for all sets at once
    from its affine angles, one vectorised call per gate kind.  Returns
    (n, ops, mats_all [B x nmats], records)."""
    from . import gatevec as V
    theta = np.asarray(params_list, dtype=np.float64)
    tr = trace(ansatz, [float(p) for p in params_list[0]])
    if tr is None:
        raise TypeError("a batched gradient needs an ansatz built from affine functions of its parameters")
    flat, records, _, n = tr
    ops, mats0, rec = _packed_with_derivatives(flat, records)
    B = theta.shape[0]
    mats_all = np.tile(mats0, (B, 1))
    gates, derivs = {}, {}         # name -> [(offset, vals)]; (name, which) -> [(offset, vals)]
    ri = 0
    for i, op in enumerate(flat):
        meta = op[-1] if op[0] in ("single", "two", "controlled") else None
        if not meta or not isinstance(meta, tuple):
            continue
        name, ps = meta
        if not any(isinstance(p, Sym) for p in ps):
            continue
        vals = []
        for p in ps:
            if isinstance(p, Sym):
                v = np.full(B, float(p.const))
                for k, cf in p.terms.items():
                    v = v + cf * theta[:, k]
                vals.append(v)
            else:
                vals.append(np.full(B, float(p)))
        gates.setdefault(name, []).append((int(ops["mat"][i]), vals))
        for which, p in enumerate(ps):
            if not isinstance(p, Sym):
                continue
            for pi, coeff in p.terms.items():
                if coeff == 0:
                    continue
                if int(rec[ri]["op"]) != i or int(rec[ri]["param"]) != pi:
                    raise AssertionError("parameter records out of step with the trace")
                derivs.setdefault((name, which), []).append((int(rec[ri]["dmat"]), vals))
                ri += 1
    if ri != len(rec):
        raise AssertionError("parameter records out of step with the trace")
    for name, items in gates.items():
        fn, k = V.MATRIX[name]
        args = [np.concatenate([v[j] for _, v in items]) for j in range(k)]
        res = fn(*args)
        _scatter(mats_all, [o for o, _ in items], res, len(items), res.shape[1])
    for (name, which), items in derivs.items():
        k = V.MATRIX[name][1]
        args = [np.concatenate([v[j] for _, v in items]) for j in range(k)]
        res = V.derivative(name, args, which)
        _scatter(mats_all, [o for o, _ in items], res, len(items), res.shape[1])
    return n, ops, mats_all, rec


def _batch_run(ansatz, hamiltonian, params_list, grad):
    from ._native import AdjointBatch
    params_list = [list(p) for p in params_list]
    if not params_list:
        return ([], []) if grad else []
    count = len(params_list)
    b = max(0, (count - 1).bit_length())
    padded = params_list + [params_list[-1]] * ((1 << b) - count)
    n, ops, mats_all, rec = _batch_tapes(ansatz, padded)
    nparam = len(params_list[0])
    x, z, c = _terms(hamiltonian, n)
    key = (id(ansatz), n, b, OPTIONS["precision"], ops.tobytes(), rec.tobytes(), x.tobytes(), z.tobytes(),
           c.tobytes(), nparam, OPTIONS["adj_tile"], OPTIONS["adj_threads"])
    bt = _BATCHES.pop(key, None)
    for attempt in (0, 1):
        if bt is None:
            while len(_BATCHES) >= _BATCH_KEEP:
                _BATCHES.pop(next(iter(_BATCHES))).free()
            bt = AdjointBatch(n, b, OPTIONS["precision"], (ops, mats_all[0]), mats_all, rec, nparam, x, z, c,
                              adj_tile=OPTIONS["adj_tile"], adj_threads=OPTIONS["adj_threads"])
        try:
            out = bt.run(mats_all, grad=grad)
            break
        except TypeError as e:
            # an op that was diagonal for every set at creation is not any more
            if attempt or "diagonal" not in str(e):
                raise
            bt.free()
            bt = None
    _BATCHES[key] = bt
    if grad:
        e, g = out
        return [float(v) for v in e[:count]], [[float(v) for v in row] for row in g[:count]]
    return [float(v) for v in out[:count]]


def energy_and_gradient_batch(ansatz, hamiltonian, params_list):
    """(energies, gradients) for many parameter sets of one ansatz in one batched
    evaluation: each set runs on its own copy of one device state, both adjoint
    sweeps cover every copy per launch, and the launch sequence is a HIP graph
    replayed per call (no ket counterpart; ket evaluates sets one by one,
    src/algorithms.ts:301-336, :536-547).  Exact (no pruning)."""
    return _batch_run(ansatz, hamiltonian, params_list, True)


def vqe_batch(ansatz, hamiltonian, params_list):
    """<H> for many parameter sets of one ansatz in one batched evaluation (exact)."""
    return _batch_run(ansatz, hamiltonian, params_list, False)


def _shifted_matrix(flat_op, pi, delta):
    """Tape matrix of a traced op with parameter `pi` moved by `delta`."""
    from .circuit import _resolve_parametric
    kind = flat_op[0]
    name, ps = flat_op[-1]
    vals = [(p.value + p.terms.get(pi, 0.0) * delta) if isinstance(p, Sym) else float(p) for p in ps]
    qubits = (flat_op[1],) if kind == "single" else (flat_op[1], flat_op[2])
    r = _resolve_parametric(name, vals, qubits)
    m = r[2] if r[0] == "single" else r[3]
    return np.asarray(m, dtype=np.complex128).reshape(-1)


def parameter_shift(ansatz, hamiltonian, params, indices=None, shift=math.pi / 2, central_difference=False,
                    timings=None):
    """dE/dparams by the two-point shift rule on the device, for the parameters in
    `indices` (all by default): E(theta + s e_i) and E(theta - s e_i) from one
    scheduled program whose shifted gate matrices are rewritten in place
    (ksv_program_set_matrix), with exact expectations.  With
    central_difference=True the result is (E+ - E-) / (2 s) (finite differences);
    otherwise ket's (E+ - E-) / 2 (src/algorithms.ts:536-547).
    Returns (indices, gradient entries); `timings` (a dict) receives the setup
    and evaluation-loop wall times."""
    import time
    t_start = time.perf_counter()
    params = [float(p) for p in params]
    tr = trace(ansatz, params)
    if tr is None:
        raise TypeError("parameter_shift needs an ansatz built from affine functions of its parameters")
    flat, records, _, n = tr
    if indices is None:
        indices = list(range(len(params)))
    by_param = {}
    for oi, pi, coeff, d in records:
        by_param.setdefault(pi, []).append(oi)
    x, z, c = _terms(hamiltonian, n)
    ops, mats = pack(flat)
    st = State(n, OPTIONS["precision"], prune=False, tile=OPTIONS["tile"])
    prog = st.program((ops, mats))
    base = {oi: np.asarray(_shifted_matrix(flat[oi], -1, 0.0)) for oi in range(len(flat))
            if oi in {o for o, _, _, _ in records}}
    out = []
    t_loop = time.perf_counter()
    for pi in indices:
        es = []
        for sgn in (1.0, -1.0):
            for oi in by_param.get(pi, []):
                prog.set_matrix(oi, _shifted_matrix(flat[oi], pi, sgn * shift))
            st.init_basis(0)
            prog.run()
            es.append(st.expect_pauli(x, z, c, mode=0)[1])
        for oi in by_param.get(pi, []):
            prog.set_matrix(oi, base[oi])
        g = (es[0] - es[1]) / (2 * shift) if central_difference else 0.5 * (es[0] - es[1])
        out.append(g)
    if timings is not None:
        timings["setup_s"] = t_loop - t_start
        timings["loop_s"] = time.perf_counter() - t_loop
    prog.free()
    st.free()
    return list(indices), out


def minimize(ansatz, hamiltonian, initialParams, lr=0.1, steps=200, tol=1e-6):
    """Gradient descent (src/algorithms.ts:591-609)."""
    params = [float(p) for p in initialParams]
    for step in range(steps):
        grad = gradient(ansatz, hamiltonian, params)
        gnorm = math.sqrt(sum(g * g for g in grad))
        if gnorm < tol:
            return {"params": params, "energy": vqe(ansatz(params), hamiltonian), "steps": step,
                    "converged": True, "truncated": False}
        params = [p - lr * g for p, g in zip(params, grad)]
    return {"params": params, "energy": vqe(ansatz(params), hamiltonian), "steps": steps, "converged": False,
            "truncated": False}
