"""Circuit, Distribution and the seeded PRNG of ket, served by libketsv.

Mirrors the public surface of ket's Circuit (src/circuit.ts:1088-1763,
:3403-3482, :3694-3753, :4102-4129): an immutable builder whose gate
methods return new circuits, and terminals (`run`, `statevector`,
`exactProbs`, `amplitude`, `probability`, `marginals`, `expectation`,
`blochAngles`, `stateAsString`, `stateAsArray`, `circuitMatrix`) that
simulate on the GPU through the C ABI.  Bitstrings are q0-leftmost; state
indices have qubit q at bit q.  Errors are ket's: RangeError for qubit
indices, TypeError for malformed input.
"""
import math
import time

import numpy as np

from . import gates as G
from ._native import RangeError, State
from .tape import pack

# ---------------------------------------------------------------------------
# backend options (ket has none; these select the device mode)
# ---------------------------------------------------------------------------
OPTIONS = {
    "prune": True,       # ket's accumulate(): drop |a|^2 < 1e-14 after every gate
    "precision": "f64",  # 'f64' complex128 or 'f32' complex64
    "tile": 0,           # LDS tile qubits (0 = library default)
    "fuse": False,       # exact mode only: multiply runs of 1-qubit gates
    "batch_shots": True,  # run() without `if`: all shots as copies of one device state
    "adj_tile": 0,       # adjoint backward passes: tile qubits (0 = library default)
    "adj_threads": 0,    # ... workgroup size (0 = library default)
}


def configure(**kw):
    """Set backend options (prune, precision, tile, fuse); returns the old values."""
    old = dict(OPTIONS)
    for k, v in kw.items():
        if k not in OPTIONS:
            raise TypeError(f"unknown option '{k}'")
        OPTIONS[k] = v
    return old


def to_uint32(x):
    """ECMAScript ToUint32 (the `>>> 0` of src/circuit.ts:814)."""
    if isinstance(x, float):
        if not math.isfinite(x):
            return 0
        x = int(x)          # truncation toward zero
    return int(x) & 0xFFFFFFFF


def make_prng(seed=None):
    """xorshift32 of src/circuit.ts:813-819; returns a callable -> float in [0, 1)."""
    if seed is not None:
        s = to_uint32(seed) or 1
    else:
        s = (int(time.time() * 1000) & 0xFFFFFFFF) or 1
    state = [s]

    def rng():
        v = state[0]
        v ^= (v << 13) & 0xFFFFFFFF
        v ^= v >> 17
        v ^= (v << 5) & 0xFFFFFFFF
        state[0] = v
        return v / 4294967296.0

    def fill(count):
        """The next `count` draws as a float64 array (same stream as `count` calls)."""
        from ._native import xorshift_fill
        st = np.array([state[0]], dtype=np.uint32)
        out = xorshift_fill(st, count)
        state[0] = int(st[0])
        return out

    rng.fill = fill
    return rng


def prng_draws(seed, count):
    """`count` successive draws of make_prng(seed) as a float64 array."""
    rng = make_prng(seed)
    return np.fromiter((rng() for _ in range(count)), dtype=np.float64, count=count)


def _bitstring(idx, n):
    """State index -> q0-leftmost bitstring (src/circuit.ts:911)."""
    return format(int(idx), "b").zfill(n)[::-1]


def _index(bs):
    """q0-leftmost bitstring -> state index."""
    return int(bs[::-1], 2)


class Distribution:
    """Measurement result (src/circuit.ts:879-1042)."""

    def __init__(self, qubits, shots, counts, creg_counts=None, truncated=False):
        self.qubits = qubits
        self.shots = shots
        self.truncated = truncated
        probs, hist = {}, {}
        for idx, cnt in counts.items():
            p = cnt / shots
            probs[_bitstring(idx, qubits)] = p
            hist[str(int(idx))] = p
        self.probs = probs
        self.histogram = hist
        self.cregs = {name: tuple(c / shots for c in bits) for name, bits in (creg_counts or {}).items()}

    @property
    def most(self):
        best, bp = "", -1.0
        for bs, p in self.probs.items():
            if p > bp:
                best, bp = bs, p
        return best

    @property
    def entropy(self):
        h = 0.0
        for p in self.probs.values():
            if p > 0:
                h -= p * math.log2(p)
        return h

    def render(self):
        entries = sorted(self.probs.items())
        if not entries:
            return ""
        maxp = max(p for _, p in entries)
        lines = []
        for bs, p in entries:
            bar = "█" * round(p / maxp * 40)
            lines.append(f"|{bs}⟩ {p * 100:5.1f}%  {bar}")
        return "\n".join(lines)


# ---------------------------------------------------------------------------
# symbolic parameters (gradient tracing)
# ---------------------------------------------------------------------------
class _TraceCtx:
    def __init__(self):
        self.escaped = False


class Sym:
    """An affine expression const + sum_i c_i * theta_i, carrying its value.

    gradient() calls an ansatz with Sym parameters to recover, gate by gate,
    which parameters feed which angles.  Arithmetic with numbers stays affine;
    any other use (float(), math.cos, comparisons) marks the trace as escaped
    and gradient() falls back to parameter shifts.
    """
    __slots__ = ("terms", "const", "value", "ctx")

    def __init__(self, terms, const, value, ctx):
        self.terms = terms
        self.const = const
        self.value = value
        self.ctx = ctx

    def _lin(self, other, sign):
        if isinstance(other, Sym):
            t = dict(self.terms)
            for k, v in other.terms.items():
                t[k] = t.get(k, 0.0) + sign * v
            return Sym(t, self.const + sign * other.const, self.value + sign * other.value, self.ctx)
        if isinstance(other, (int, float)) and not isinstance(other, bool):
            return Sym(dict(self.terms), self.const + sign * other, self.value + sign * other, self.ctx)
        return NotImplemented

    def __add__(self, o):
        return self._lin(o, 1.0)

    __radd__ = __add__

    def __sub__(self, o):
        return self._lin(o, -1.0)

    def __rsub__(self, o):
        r = self._lin(o, -1.0)
        if r is NotImplemented:
            return r
        return -r

    def __mul__(self, o):
        if isinstance(o, (int, float)) and not isinstance(o, bool):
            return Sym({k: v * o for k, v in self.terms.items()}, self.const * o, self.value * o, self.ctx)
        self.ctx.escaped = True
        return NotImplemented

    __rmul__ = __mul__

    def __truediv__(self, o):
        if isinstance(o, (int, float)) and not isinstance(o, bool):
            return self * (1.0 / o)
        self.ctx.escaped = True
        return NotImplemented

    def __neg__(self):
        return self * -1.0

    def __pos__(self):
        return self

    def __float__(self):
        self.ctx.escaped = True
        return float(self.value)

    def __lt__(self, o):
        self.ctx.escaped = True
        return self.value < float(o)

    def __gt__(self, o):
        self.ctx.escaped = True
        return self.value > float(o)

    def __repr__(self):
        return f"Sym({self.value!r})"


def _num(x):
    if isinstance(x, Sym):
        return x.value
    return float(x)


def _is_str(x):
    return isinstance(x, str)


# ---------------------------------------------------------------------------
# op helpers
# ---------------------------------------------------------------------------
_PURE_BLOCKERS = ("measure", "reset", "if")


def _remap(op, qmap):
    """src/circuit.ts:84-107"""
    k = op[0]
    q = qmap.__getitem__
    if k == "single":
        return (k, q(op[1])) + op[2:]
    if k in ("cnot", "swap"):
        return (k, q(op[1]), q(op[2]))
    if k in ("two", "controlled"):
        return (k, q(op[1]), q(op[2])) + op[3:]
    if k in ("toffoli", "cswap", "csrswap"):
        return (k, q(op[1]), q(op[2]), q(op[3]))
    if k == "subcircuit":
        return (k, op[1], tuple(q(i) for i in op[2]), op[3])
    if k in ("unitary",):
        return (k, tuple(q(i) for i in op[1]), op[2])
    if k == "barrier":
        return (k, tuple(q(i) for i in op[1]))
    return op


def _collect_params(ops):
    """src/circuit.ts:1076-1086"""
    names = []
    seen = set()
    for op in ops:
        if op[0] == "parametric":
            for p in op[2]:
                if isinstance(p, str) and p not in seen:
                    seen.add(p)
                    names.append(p)
        elif op[0] == "if":
            for p in _collect_params(op[3]):
                if p not in seen:
                    seen.add(p)
                    names.append(p)
    return names


def _flatten(ops):
    """src/circuit.ts:113-133"""
    has_sub = False
    for op in ops:
        if op[0] == "parametric":
            names = ", ".join(_collect_params(ops))
            raise TypeError(f"Circuit has unbound parameters: [{names}]. Call bind({{ {names} }}) first.")
        if op[0] == "subcircuit":
            has_sub = True
    if not has_sub:
        return ops
    out = []
    for op in ops:
        if op[0] == "subcircuit":
            out.extend(_flatten([_remap(inner, op[2]) for inner in op[3]]))
        else:
            out.append(op)
    return out


_CTRL_GATES = {"crx": G.Rx, "cry": G.Ry, "crz": G.Rz, "cu1": G.U1, "cu2": G.U2, "cu3": G.U3}


def _resolve_parametric(name, ps, qubits):
    """src/circuit.ts:1047-1073"""
    meta = (name, tuple(ps))
    one = {"rx": G.Rx, "ry": G.Ry, "rz": G.Rz, "vz": G.Rz, "u1": G.U1, "p": G.U1, "u2": G.U2, "u3": G.U3,
           "gpi": G.Gpi, "gpi2": G.Gpi2}
    two = {"xx": G.Xx, "yy": G.Yy, "zz": G.Zz, "xy": G.Xy, "ms": G.Ms}
    vals = [_num(p) for p in ps]
    if name in one:
        return ("single", qubits[0], one[name](*vals), meta)
    if name in two:
        return ("two", qubits[0], qubits[1], two[name](*vals), meta)
    if name in _CTRL_GATES:
        return ("controlled", qubits[0], qubits[1], _CTRL_GATES[name](*vals), meta)
    raise TypeError(f"bind: unknown parametric gate '{name}'")


def _svfrom(s, qubits):
    """src/circuit.ts:136-140: validates and returns the basis index."""
    if not isinstance(s, str) or len(s) != qubits or not s or any(ch not in "01" for ch in s):
        raise TypeError(f"initialState '{s}' must be a {qubits}-character binary string")
    return _index(s)


class Circuit:
    """Immutable circuit builder (src/circuit.ts:1088)."""

    __slots__ = ("qubits", "_parent", "_op", "_len", "_ops_cache", "_cregs", "_gates")

    def __init__(self, qubits, ops=None, cregs=None, gates=None):
        self.qubits = int(qubits)
        self._parent = None
        self._op = None
        self._ops_cache = tuple(ops) if ops is not None else ()
        self._len = len(self._ops_cache)
        self._cregs = dict(cregs or {})
        self._gates = dict(gates or {})

    # ---- op storage: a shared-prefix chain, materialised on demand ---------
    def _ops(self):
        if self._ops_cache is None:
            chain = []
            c = self
            while c._ops_cache is None:
                chain.append(c._op)
                c = c._parent
            ops = list(c._ops_cache)
            ops.extend(reversed(chain))
            self._ops_cache = tuple(ops)
            self._parent = None
        return self._ops_cache

    def _child(self, op, cregs=None, gates=None):
        c = Circuit.__new__(Circuit)
        c.qubits = self.qubits
        c._parent = self
        c._op = op
        c._len = self._len + 1
        c._ops_cache = None
        c._cregs = self._cregs if cregs is None else cregs
        c._gates = self._gates if gates is None else gates
        return c

    @property
    def ops(self):
        return self._ops()

    def __len__(self):
        return self._len

    def _check_q(self, i):
        if not isinstance(i, int) or i < 0 or i >= self.qubits:
            raise RangeError(f"qubit index {i} is out of range for a {self.qubits}-qubit circuit")

    def _check_op(self, op):
        """src/circuit.ts:1113-1146"""
        k = op[0]
        q = self._check_q

        def diff(a, b):
            if a == b:
                raise TypeError(f"control and target qubits must differ (got {a})")

        if k == "single":
            q(op[1])
        elif k == "cnot" or k == "controlled":
            q(op[1]); q(op[2]); diff(op[1], op[2])
        elif k in ("swap", "two"):
            q(op[1]); q(op[2])
        elif k in ("toffoli", "cswap", "csrswap"):
            q(op[1]); q(op[2]); q(op[3])
        elif k in ("measure", "reset"):
            q(op[1])
        elif k == "barrier":
            for i in op[1]:
                q(i)
        elif k == "parametric":
            for i in op[3]:
                q(i)
        elif k == "unitary":
            qs, m = op[1], op[2]
            if len(qs) == 0:
                raise TypeError("unitary: qubits must be non-empty")
            exp = 1 << len(qs)
            if len(m) != exp or any(len(row) != exp for row in m):
                got1 = len(m[0]) if len(m) else 0
                raise TypeError(f"unitary: matrix must be {exp}×{exp} for {len(qs)} qubit(s), got {len(m)}×{got1}")
            for i in qs:
                q(i)
        elif k == "if":
            for inner in op[3]:
                self._check_op(inner)
        elif k == "subcircuit":
            for i in op[2]:
                q(i)

    def _add(self, op):
        self._check_op(op)
        return self._child(op)

    def _ctrl(self, control, target, gate, meta):
        return self._add(("controlled", control, target, gate, meta))

    # ---- single-qubit gates (src/circuit.ts:1150-1205) ---------------------
    def id(self, q): return self._add(("single", q, G.Id, ("id", ())))
    def h(self, q): return self._add(("single", q, G.H, ("h", ())))
    def x(self, q): return self._add(("single", q, G.X, ("x", ())))
    def y(self, q): return self._add(("single", q, G.Y, ("y", ())))
    def z(self, q): return self._add(("single", q, G.Z, ("z", ())))
    def s(self, q): return self._add(("single", q, G.S, ("s", ())))
    def si(self, q): return self._add(("single", q, G.Si, ("si", ())))
    def sdg(self, q): return self._add(("single", q, G.Si, ("sdg", ())))
    def t(self, q): return self._add(("single", q, G.T, ("t", ())))
    def ti(self, q): return self._add(("single", q, G.Ti, ("ti", ())))
    def tdg(self, q): return self._add(("single", q, G.Ti, ("tdg", ())))
    def v(self, q): return self._add(("single", q, G.V, ("v", ())))
    def vi(self, q): return self._add(("single", q, G.Vi, ("vi", ())))
    def srn(self, q): return self._add(("single", q, G.V, ("srn", ())))
    def srndg(self, q): return self._add(("single", q, G.Vi, ("srndg", ())))

    def rx(self, theta, q):
        if _is_str(theta):
            return self._add(("parametric", "rx", (theta,), (q,)))
        return self._add(("single", q, G.Rx(_num(theta)), ("rx", (theta,))))

    def ry(self, theta, q):
        if _is_str(theta):
            return self._add(("parametric", "ry", (theta,), (q,)))
        return self._add(("single", q, G.Ry(_num(theta)), ("ry", (theta,))))

    def rz(self, theta, q):
        if _is_str(theta):
            return self._add(("parametric", "rz", (theta,), (q,)))
        return self._add(("single", q, G.Rz(_num(theta)), ("rz", (theta,))))

    def vz(self, theta, q):
        if _is_str(theta):
            return self._add(("parametric", "vz", (theta,), (q,)))
        return self._add(("single", q, G.Rz(_num(theta)), ("vz", (theta,))))

    def r2(self, q): return self._add(("single", q, G.R2, ("r2", ())))
    def r4(self, q): return self._add(("single", q, G.R4, ("r4", ())))
    def r8(self, q): return self._add(("single", q, G.R8, ("r8", ())))

    def u1(self, lam, q):
        if _is_str(lam):
            return self._add(("parametric", "u1", (lam,), (q,)))
        return self._add(("single", q, G.U1(_num(lam)), ("u1", (lam,))))

    def p(self, lam, q):
        if _is_str(lam):
            return self._add(("parametric", "p", (lam,), (q,)))
        return self._add(("single", q, G.U1(_num(lam)), ("p", (lam,))))

    def u2(self, phi, lam, q):
        if _is_str(phi) or _is_str(lam):
            return self._add(("parametric", "u2", (phi, lam), (q,)))
        return self._add(("single", q, G.U2(_num(phi), _num(lam)), ("u2", (phi, lam))))

    def u3(self, theta, phi, lam, q):
        if _is_str(theta) or _is_str(phi) or _is_str(lam):
            return self._add(("parametric", "u3", (theta, phi, lam), (q,)))
        return self._add(("single", q, G.U3(_num(theta), _num(phi), _num(lam)), ("u3", (theta, phi, lam))))

    # ---- two-qubit gates ------------------------------------------------------
    def cnot(self, control, target): return self._add(("cnot", control, target))
    def cx(self, control, target): return self.cnot(control, target)
    def swap(self, a, b): return self._add(("swap", a, b))

    def _two(self, name, make, params, a, b):
        if any(_is_str(p) for p in params):
            return self._add(("parametric", name, tuple(params), (a, b)))
        return self._add(("two", a, b, make(*[_num(p) for p in params]), (name, tuple(params))))

    def xx(self, theta, a, b): return self._two("xx", G.Xx, (theta,), a, b)
    def yy(self, theta, a, b): return self._two("yy", G.Yy, (theta,), a, b)
    def zz(self, theta, a, b): return self._two("zz", G.Zz, (theta,), a, b)
    def xy(self, theta, a, b): return self._two("xy", G.Xy, (theta,), a, b)
    def iswap(self, a, b): return self._add(("two", a, b, G.ISwap, ("iswap", ())))
    def srswap(self, a, b): return self._add(("two", a, b, G.SrSwap, ("srswap", ())))
    def ms(self, phi0, phi1, a, b): return self._two("ms", G.Ms, (phi0, phi1), a, b)

    # ---- controlled gates (src/circuit.ts:1282-1352) ----------------------------
    def cy(self, c, t): return self._ctrl(c, t, G.Y, ("cy", ()))
    def cz(self, c, t): return self._ctrl(c, t, G.Z, ("cz", ()))
    def ch(self, c, t): return self._ctrl(c, t, G.H, ("ch", ()))

    def _cparam(self, name, params, c, t):
        if any(_is_str(p) for p in params):
            return self._add(("parametric", name, tuple(params), (c, t)))
        return self._ctrl(c, t, _CTRL_GATES[name](*[_num(p) for p in params]), (name, tuple(params)))

    def crx(self, theta, c, t): return self._cparam("crx", (theta,), c, t)
    def cry(self, theta, c, t): return self._cparam("cry", (theta,), c, t)
    def crz(self, theta, c, t): return self._cparam("crz", (theta,), c, t)
    def cr2(self, c, t): return self._ctrl(c, t, G.R2, ("cr2", ()))
    def cr4(self, c, t): return self._ctrl(c, t, G.R4, ("cr4", ()))
    def cr8(self, c, t): return self._ctrl(c, t, G.R8, ("cr8", ()))
    def cu1(self, lam, c, t): return self._cparam("cu1", (lam,), c, t)
    def cu2(self, phi, lam, c, t): return self._cparam("cu2", (phi, lam), c, t)
    def cu3(self, theta, phi, lam, c, t): return self._cparam("cu3", (theta, phi, lam), c, t)
    def cs(self, c, t): return self._ctrl(c, t, G.S, ("cs", ()))
    def ct(self, c, t): return self._ctrl(c, t, G.T, ("ct", ()))
    def csdg(self, c, t): return self._ctrl(c, t, G.Si, ("csdg", ()))
    def ctdg(self, c, t): return self._ctrl(c, t, G.Ti, ("ctdg", ()))
    def csrn(self, c, t): return self._ctrl(c, t, G.V, ("csrn", ()))

    # ---- IonQ native ----------------------------------------------------------
    def gpi(self, phi, q):
        if _is_str(phi):
            return self._add(("parametric", "gpi", (phi,), (q,)))
        return self._add(("single", q, G.Gpi(_num(phi)), ("gpi", (phi,))))

    def gpi2(self, phi, q):
        if _is_str(phi):
            return self._add(("parametric", "gpi2", (phi,), (q,)))
        return self._add(("single", q, G.Gpi2(_num(phi)), ("gpi2", (phi,))))

    # ---- three-qubit gates ------------------------------------------------------
    def ccx(self, c1, c2, target): return self._add(("toffoli", c1, c2, target))
    def cswap(self, control, a, b): return self._add(("cswap", control, a, b))
    def csrswap(self, control, a, b): return self._add(("csrswap", control, a, b))

    def barrier(self, *qubits):
        qs = tuple(qubits) if qubits else tuple(range(self.qubits))
        return self._add(("barrier", qs))

    def unitary(self, matrix, *qubits):
        """Custom N-qubit unitary; qubits[0] is the MSB (src/circuit.ts:1409-1421)."""
        norm = tuple(tuple(complex(v["re"], v["im"]) if isinstance(v, dict) else complex(v) for v in row)
                     for row in matrix)
        return self._add(("unitary", tuple(qubits), norm))

    # ---- classical registers, measurement, composition (src/circuit.ts:1595-1722) --
    def creg(self, name, size):
        cr = dict(self._cregs)
        cr[name] = size
        return Circuit(self.qubits, self._ops(), cr, self._gates)

    def measure(self, q, creg, bit):
        size = max(self._cregs.get(creg, 0), bit + 1)
        cr = dict(self._cregs)
        cr[creg] = size
        return self._child(("measure", q, creg, bit), cregs=cr)

    def reset(self, q, value=0):
        r = self._add(("reset", q))
        return r.x(q) if value == 1 else r

    def if_(self, creg, value, build):
        inner = build(Circuit(self.qubits, (), {}, self._gates))
        return self._add(("if", creg, value, inner._ops()))

    def defineGate(self, name, sub):
        if any(op[0] in _PURE_BLOCKERS for op in sub._ops()):
            raise TypeError(f"Gate '{name}' contains classical ops (measure/reset/if), which are not supported inside named gates")
        g = dict(self._gates)
        g[name] = sub
        return Circuit(self.qubits, self._ops(), self._cregs, g)

    def gate(self, name, *qubits):
        sub = self._gates.get(name)
        if sub is None:
            raise TypeError(f"Unknown gate '{name}'. Register it first with .defineGate(name, subcircuit).")
        if len(qubits) != sub.qubits:
            raise TypeError(f"Gate '{name}' expects {sub.qubits} qubit(s), got {len(qubits)}")
        return self._add(("subcircuit", name, tuple(qubits), sub._ops()))

    def decompose(self):
        return Circuit(self.qubits, _flatten(self._ops()), self._cregs, self._gates)

    def compose(self, other):
        if other.qubits != self.qubits:
            raise TypeError(f"compose: qubit count mismatch — this has {self.qubits}, other has {other.qubits}")
        cr = dict(self._cregs)
        for k, v in other._cregs.items():
            cr[k] = max(cr.get(k, 0), v)
        g = dict(other._gates)
        g.update(self._gates)
        return Circuit(self.qubits, tuple(self._ops()) + tuple(other._ops()), cr, g)

    @property
    def params(self):
        return sorted(_collect_params(self._ops()))

    def bind(self, values):
        def resolve(p):
            if not isinstance(p, str):
                return p
            if p in values:
                return values[p]
            raise TypeError(f"bind: unbound parameter '{p}'. Provide a value for it in the bind() call.")

        def rebuild(ops):
            out = []
            for op in ops:
                if op[0] == "parametric":
                    out.append(_resolve_parametric(op[1], [resolve(p) for p in op[2]], op[3]))
                elif op[0] == "if":
                    out.append(("if", op[1], op[2], tuple(rebuild(op[3]))))
                else:
                    out.append(op)
            return out

        return Circuit(self.qubits, rebuild(self._ops()), self._cregs, self._gates)

    # ---- wire formats (src/circuit.ts:1960-2090, :3768-3803) --------------------
    def toJSON(self):
        """CircuitJSON: {ket: 1, qubits, cregs, gates, ops}; gate matrices are
        rebuilt from each op's name and params on load."""
        from .wire import to_json
        return to_json(self)

    @staticmethod
    def fromJSON(j):
        """A circuit from a CircuitJSON object or its JSON text."""
        from .wire import from_json
        return from_json(j)

    def toQASM(self):
        """OpenQASM 2.0 text of this circuit."""
        from .wire import to_qasm
        return to_qasm(self)

    @staticmethod
    def fromQASM(source):
        """A circuit from OpenQASM 2.0 or 3.0 text."""
        from .wire import from_qasm
        return from_qasm(source)

    # ---- simulation -------------------------------------------------------------
    def _is_pure(self):
        return not any(op[0] in _PURE_BLOCKERS for op in self._ops())

    def tape(self):
        """The packed (ops, mats) tape of this pure circuit."""
        return pack(_flatten(self._ops()))

    def simulate(self, initialState=None, precision=None, prune=None):
        """Run the pure circuit on the GPU and return the device State."""
        if not self._is_pure():
            raise TypeError("statevector() requires a pure circuit — remove measure/reset/if ops")
        unbound = _collect_params(self._ops())
        if unbound:
            names = ", ".join(f"{p}: value" for p in unbound)
            raise TypeError(f"statevector() requires bound parameters. Call bind({{ {names} }}) first.")
        init = _svfrom(initialState, self.qubits) if initialState is not None else 0
        tape = self.tape()
        st = State(self.qubits, precision or OPTIONS["precision"],
                   prune=OPTIONS["prune"] if prune is None else prune, tile=OPTIONS["tile"], fuse=OPTIONS["fuse"])
        if init:
            st.init_basis(init)
        st.apply(tape)
        return st

    def statevector(self, initialState=None):
        """Sparse view {index: complex} of the amplitudes with |a|^2 >= 1e-14 (src/circuit.ts:1431)."""
        if self.qubits > 24:
            raise RangeError(f"statevector(): {self.qubits} qubits exceeds the 2^24-entry map limit; "
                             "use simulate().amplitudes() or exactProbsTop()")
        st = self.simulate(initialState)
        amps = st.amplitudes()
        st.free()
        p = amps.real * amps.real + amps.imag * amps.imag
        nz = np.nonzero(p >= 1e-14)[0]
        return {int(i): complex(amps[i]) for i in nz}

    def statevectorDense(self, initialState=None):
        """All 2^n amplitudes as a complex128 numpy array (index bit q = qubit q)."""
        st = self.simulate(initialState)
        a = st.amplitudes()
        st.free()
        return a

    def circuitMatrix(self):
        """src/circuit.ts:1463-1485: matrix[row][col], q0 = MSB of row/col indices."""
        if not self._is_pure():
            raise TypeError("circuitMatrix() requires a pure circuit — remove measure/reset/if ops")
        n = self.qubits
        dim = 1 << n
        if n > 12:
            raise RangeError(f"circuitMatrix(): circuit too large ({n} qubits = {dim}×{dim} matrix)")
        rev = np.array([int(format(i, "b").zfill(n)[::-1], 2) for i in range(dim)], dtype=np.int64)
        # all 2^n columns at once: a 2n-qubit state whose copy t starts as |t>
        st = State(2 * n, OPTIONS["precision"], prune=OPTIONS["prune"], tile=OPTIONS["tile"])
        try:
            st.shots_init_identity(n)
            st.apply(self.tape())
            a = st.amplitudes().reshape(dim, dim)           # a[t, i] = <i| U |t>
        finally:
            st.free()
        p = a.real ** 2 + a.imag ** 2
        a = np.where(p >= 1e-14, a, 0)
        mat = a.T[np.ix_(rev, rev)]                         # q0 = MSB on both indices
        return [[complex(v) for v in row] for row in mat]

    def dm(self, noise=None):
        """src/circuit.ts:4549-4579: exact density matrix with optional per-gate depolarising noise."""
        from .density import DM_DEVICE_NOISE, run_dm
        if not self._is_pure():
            raise TypeError("dm() requires a pure circuit — remove measure/reset/if ops")
        if isinstance(noise, str):
            params = DM_DEVICE_NOISE.get(noise)
            if params is None:
                raise TypeError(f"Unknown device profile '{noise}'. Known: {', '.join(DM_DEVICE_NOISE)}")
        else:
            params = noise
        return run_dm(self.qubits, _flatten(self._ops()), params)

    def amplitude(self, bitstring):
        if not isinstance(bitstring, str) or len(bitstring) != self.qubits or any(ch not in "01" for ch in bitstring) or not bitstring:
            raise TypeError(f"bitstring '{bitstring}' must be a {self.qubits}-character binary string")
        st = self.simulate()
        a = st.amplitude(_index(bitstring))
        st.free()
        if a.real * a.real + a.imag * a.imag < 1e-14:
            return complex(0, 0)
        return a

    def probability(self, bitstring):
        a = self.amplitude(bitstring)
        return a.real * a.real + a.imag * a.imag

    def marginals(self, initialState=None):
        st = self.simulate(initialState)
        m = st.marginals()
        st.free()
        return [float(v) for v in m]

    def exactProbs(self, initialState=None):
        """src/circuit.ts:3700-3711: {q0-leftmost bitstring: p} for p > 1e-14."""
        if not self._is_pure():
            raise TypeError("exactProbs() requires a pure circuit — no measure, reset, or if ops")
        if self.qubits > 24:
            raise RangeError(f"exactProbs(): {self.qubits} qubits exceeds the 2^24-entry limit; use exactProbsTop(k)")
        st = self.simulate(initialState)
        idx, p = st.probs_filtered(1e-14)
        st.free()
        n = self.qubits
        return {_bitstring(i, n): float(v) for i, v in zip(idx.tolist(), p.tolist())}

    def exactProbsTop(self, k, initialState=None):
        """The k most probable outcomes {bitstring: p}, p descending (any width)."""
        st = self.simulate(initialState)
        idx, p = st.topk(k)
        st.free()
        return {_bitstring(i, self.qubits): float(v) for i, v in zip(idx.tolist(), p.tolist())}

    def stateAsString(self, initialState=None):
        """src/circuit.ts:1528-1566"""
        sv = self.statevector(initialState)
        eps = 1e-10

        def fmt_num(x):
            return _js_num_str(float(f"{x:.4g}"))

        def fmt_amp(a):
            rz, iz = abs(a.real) < eps, abs(a.imag) < eps
            if rz and iz:
                return "0"
            if iz:
                return fmt_num(a.real)
            if rz:
                return f"{fmt_num(a.imag)}i"
            sign = "" if a.imag < 0 else "+"
            return f"({fmt_num(a.real)}{sign}{fmt_num(a.imag)}i)"

        entries = sorted((i, a) for i, a in sv.items() if a.real * a.real + a.imag * a.imag > eps * eps)
        if not entries:
            return "0"
        terms = [f"{fmt_amp(a)}|{_bitstring(i, self.qubits)}⟩" for i, a in entries]
        res = terms[0]
        for t in terms[1:]:
            res += f" - {t[1:]}" if t.startswith("-") else f" + {t}"
        return res

    def stateAsArray(self):
        """src/circuit.ts:1579-1593"""
        sv = self.statevector()
        out = []
        for i, a in sv.items():
            prob = a.real * a.real + a.imag * a.imag
            if prob < 1e-10:
                continue
            out.append({"bitstring": _bitstring(i, self.qubits), "re": a.real, "im": a.imag, "prob": prob,
                        "phase": math.atan2(a.imag, a.real)})
        out.sort(key=lambda e: -e["prob"])
        return out

    def blochAngles(self, q, initialState=None):
        """src/circuit.ts:4102-4129"""
        st = self.simulate(initialState)
        r00, r11, re, im = st.reduce_1q(q)
        st.free()
        rz = r00 - r11
        rx = 2 * re
        ry = -2 * im
        theta = math.acos(max(-1.0, min(1.0, rz)))
        phi = math.atan2(ry, rx)
        return {"theta": theta, "phi": phi}

    def expectation(self, pauli):
        """src/circuit.ts:3728-3753: pauli[q] acts on qubit q (q0-leftmost)."""
        n = self.qubits
        pauli = pauli.upper()
        if len(pauli) != n:
            raise TypeError(f"pauli '{pauli}' length must equal qubits ({n})")
        if not pauli or any(ch not in "IXYZ" for ch in pauli):
            raise TypeError("pauli must contain only I, X, Y, Z")
        if not any(ch in "XYZ" for ch in pauli):
            return 1
        x, z = pauli_masks(pauli, q0_left=True)
        st = self.simulate()
        _, tot = st.expect_pauli([x], [z], [1.0], mode=1 if OPTIONS["prune"] else 0)
        st.free()
        return tot

    # ---- run ------------------------------------------------------------------------
    def run(self, shots=1024, seed=None, noise=None, initialState=None):
        """src/circuit.ts:3406-3482"""
        from .noise import batchable, resolve_noise, run_batched, run_per_shot
        rng = make_prng(seed)
        init = _svfrom(initialState, self.qubits) if initialState is not None else 0
        params = resolve_noise(noise)
        creg_counts = {name: [0] * size for name, size in self._cregs.items()}
        if params is None and self._is_pure():
            st = self.simulate(initialState)
            r = np.fromiter((rng() for _ in range(shots)), dtype=np.float64, count=shots)
            idx = st.sample(r)
            st.free()
            u, cnt = np.unique(idx, return_counts=True)
            # first-occurrence order of the sampled indices, like ket's Map insertion
            first = {}
            for i in idx.tolist():
                if i not in first:
                    first[i] = 0
            cmap = dict(zip(u.tolist(), cnt.tolist()))
            counts = {i: cmap[i] for i in first}
            return Distribution(self.qubits, shots, counts, creg_counts)
        if OPTIONS["batch_shots"] and batchable(self):
            counts, creg_counts = run_batched(self, shots, rng, params, init, creg_counts)
        else:
            counts, creg_counts = run_per_shot(self, shots, rng, params, init, creg_counts)
        return Distribution(self.qubits, shots, counts, creg_counts)


# `if` is a Python keyword: Circuit.if_ is the method, getattr(c, "if") its alias
setattr(Circuit, "if", Circuit.if_)


def _js_num_str(x):
    """Number -> string as JavaScript's Number.prototype.toString for short decimals."""
    if x == int(x) and abs(x) < 1e21:
        return str(int(x))
    r = repr(x)
    if "e" in r:
        mant, exp = r.split("e")
        e = int(exp)
        return f"{mant}e{'+' if e > 0 else '-'}{abs(e)}"
    return r


def pauli_masks(ops, q0_left):
    """Pauli string -> (xmask, zmask) with bit q = qubit q.

    q0_left: pauli[q] acts on qubit q (Circuit.expectation); otherwise ops[0]
    acts on qubit n-1 (PauliTerm, src/algorithms.ts:281-289).
    """
    n = len(ops)
    x = z = 0
    for i, ch in enumerate(ops.upper()):
        q = i if q0_left else n - 1 - i
        if ch == "X":
            x |= 1 << q
        elif ch == "Y":
            x |= 1 << q
            z |= 1 << q
        elif ch == "Z":
            z |= 1 << q
    return x, z
