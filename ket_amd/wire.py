"""Wire formats: ket's CircuitJSON and OpenQASM 2.0 / 3.0.

toJSON / fromJSON follow src/circuit.ts:3768-3803 with the op encoding of
opsToJSON / opsFromJSON (:455-595): gate matrices are not stored, they are
rebuilt from each op's meta (name and params) on load.  toQASM / fromQASM
follow :1960-2090 with the helpers at :326-450 (angle formatting, the angle
expression grammar, the gate-name mapping and the gate dispatch).

Both loaders build the op list once through the Circuit builder (each op
checked as ket's #checkOp does, appended to a shared prefix, so a circuit of
G gates loads in O(G)), and `qasm_tape` / `json_tape` hand back the packed
ksv_op tape that the device takes, ready for ksv_apply or a program.
"""
import json
import math
import re

from . import gates as G
from ._native import RangeError
from .circuit import Circuit, _flatten, _js_num_str

# ---------------------------------------------------------------------------
# CircuitJSON
# ---------------------------------------------------------------------------
_SINGLE_META = {
    "h": G.H, "x": G.X, "y": G.Y, "z": G.Z, "s": G.S, "si": G.Si, "sdg": G.Si, "t": G.T, "ti": G.Ti, "tdg": G.Ti,
    "v": G.V, "vi": G.Vi, "srn": G.V, "srndg": G.Vi, "sx": G.V, "sxdg": G.Vi, "id": G.Id,
    "r2": G.R2, "r4": G.R4, "r8": G.R8,
}
_SINGLE_PARAM = {"rx": G.Rx, "ry": G.Ry, "rz": G.Rz, "vz": G.Rz, "u1": G.U1, "p": G.U1, "u2": G.U2, "u3": G.U3,
                 "gpi": G.Gpi, "gpi2": G.Gpi2}
_TWO_META = {"iswap": G.ISwap, "srswap": G.SrSwap}
_TWO_PARAM = {"xx": G.Xx, "yy": G.Yy, "zz": G.Zz, "xy": G.Xy, "ms": G.Ms}
_CTRL_META = {"cx": G.X, "cy": G.Y, "cz": G.Z, "ch": G.H, "cs": G.S, "ct": G.T, "csdg": G.Si, "ctdg": G.Ti,
              "cr2": G.R2, "cr4": G.R4, "cr8": G.R8, "csrn": G.V}
_CTRL_PARAM = {"crx": G.Rx, "cry": G.Ry, "crz": G.Rz, "cu1": G.U1, "cu2": G.U2, "cu3": G.U3}


def _nargs(fn):
    return {G.U2: 2, G.U3: 3, G.Ms: 2}.get(fn, 1)


def _gate_from_meta(fixed, param, name, params, what):
    """gate2x2FromMeta / gate4x4FromMeta / ctrlGate2x2FromMeta (src/circuit.ts:467-529)."""
    if name in fixed:
        return fixed[name]
    fn = param.get(name)
    if fn is None:
        raise TypeError(f"fromJSON: unknown {what} gate '{name}'")
    k = _nargs(fn)
    ps = [float(p) for p in (list(params) + [float("nan")] * k)[:k]]
    return fn(*ps)


def _meta_json(meta):
    name, params = meta
    out = {"name": name}
    if len(params):
        out["params"] = [p if isinstance(p, str) else float(p) for p in params]
    return out


def _meta_tuple(m):
    if not isinstance(m, dict) or "name" not in m:
        raise TypeError("fromJSON: op without gate meta")
    return m["name"], tuple(m.get("params") or ())


def ops_to_json(ops):
    """opsToJSON (src/circuit.ts:562-590)."""
    out = []
    for op in ops:
        k = op[0]
        if k == "single":
            out.append({"kind": "single", "q": op[1], "meta": _meta_json(op[3])})
        elif k == "cnot":
            out.append({"kind": "cnot", "control": op[1], "target": op[2]})
        elif k == "swap":
            out.append({"kind": "swap", "a": op[1], "b": op[2]})
        elif k == "two":
            out.append({"kind": "two", "a": op[1], "b": op[2], "meta": _meta_json(op[4])})
        elif k == "controlled":
            out.append({"kind": "controlled", "control": op[1], "target": op[2], "meta": _meta_json(op[4])})
        elif k == "toffoli":
            out.append({"kind": "toffoli", "c1": op[1], "c2": op[2], "target": op[3]})
        elif k in ("cswap", "csrswap"):
            out.append({"kind": k, "control": op[1], "a": op[2], "b": op[3]})
        elif k == "measure":
            out.append({"kind": "measure", "q": op[1], "creg": op[2], "bit": op[3]})
        elif k == "reset":
            out.append({"kind": "reset", "q": op[1]})
        elif k == "if":
            out.append({"kind": "if", "creg": op[1], "value": op[2], "ops": ops_to_json(op[3])})
        elif k == "subcircuit":
            out.append({"kind": "subcircuit", "name": op[1], "qubits": list(op[2]), "def": ops_to_json(op[3])})
        elif k == "barrier":
            out.append({"kind": "barrier", "qubits": list(op[1])})
        elif k == "unitary":
            out.append({"kind": "unitary", "qubits": list(op[1]),
                        "matrix": [[[complex(v).real, complex(v).imag] for v in row] for row in op[2]]})
        elif k == "parametric":
            out.append({"kind": "parametric", "name": op[1], "params": list(op[2]), "qubits": list(op[3])})
        else:
            raise TypeError(f"toJSON: unknown op kind '{k}'")
    return out


def ops_from_json(raw):
    """opsFromJSON (src/circuit.ts:532-559): op tuples with their gate matrices rebuilt."""
    out = []
    for o in raw:
        k = o.get("kind") if isinstance(o, dict) else None
        if k == "single":
            meta = _meta_tuple(o.get("meta"))
            out.append(("single", o["q"], _gate_from_meta(_SINGLE_META, _SINGLE_PARAM, *meta, "single-qubit"), meta))
        elif k == "cnot":
            out.append(("cnot", o["control"], o["target"]))
        elif k == "swap":
            out.append(("swap", o["a"], o["b"]))
        elif k == "two":
            meta = _meta_tuple(o.get("meta"))
            out.append(("two", o["a"], o["b"], _gate_from_meta(_TWO_META, _TWO_PARAM, *meta, "two-qubit"), meta))
        elif k == "controlled":
            meta = _meta_tuple(o.get("meta"))
            out.append(("controlled", o["control"], o["target"],
                        _gate_from_meta(_CTRL_META, _CTRL_PARAM, *meta, "controlled"), meta))
        elif k == "toffoli":
            out.append(("toffoli", o["c1"], o["c2"], o["target"]))
        elif k in ("cswap", "csrswap"):
            out.append((k, o["control"], o["a"], o["b"]))
        elif k == "measure":
            out.append(("measure", o["q"], o["creg"], o["bit"]))
        elif k == "reset":
            out.append(("reset", o["q"]))
        elif k == "if":
            out.append(("if", o["creg"], o["value"], tuple(ops_from_json(o["ops"]))))
        elif k == "subcircuit":
            out.append(("subcircuit", o["name"], tuple(o["qubits"]), tuple(ops_from_json(o["def"]))))
        elif k == "barrier":
            out.append(("barrier", tuple(o["qubits"])))
        elif k == "parametric":
            out.append(("parametric", o["name"], tuple(o["params"]), tuple(o["qubits"])))
        elif k == "unitary":
            m = tuple(tuple(complex(re, im) for re, im in row) for row in o["matrix"])
            out.append(("unitary", tuple(o["qubits"]), m))
        else:
            raise TypeError(f"fromJSON: unknown op kind '{k}'")
    return out


def to_json(circ):
    """Circuit.toJSON() (src/circuit.ts:3768-3778): {ket: 1, qubits, cregs, gates, ops}."""
    gates = {name: {"qubits": sub.qubits, "ops": ops_to_json(sub._ops())} for name, sub in circ._gates.items()}
    return {"ket": 1, "qubits": circ.qubits, "cregs": dict(circ._cregs), "gates": gates,
            "ops": ops_to_json(circ._ops())}


def from_json(j):
    """Circuit.fromJSON(json | string) (src/circuit.ts:3791-3803); TypeError on a
    schema version other than 1, an unknown op kind or an unknown gate name."""
    if isinstance(j, (str, bytes)):
        j = json.loads(j)
    if j.get("ket") != 1:
        raise TypeError(f"fromJSON: unsupported schema version {j.get('ket')}")
    gates = {name: Circuit(d["qubits"], ops_from_json(d["ops"])) for name, d in j.get("gates", {}).items()}
    return Circuit(j["qubits"], ops_from_json(j["ops"]), dict(j.get("cregs", {})), gates)


def json_tape(j):
    """(qubits, packed ksv_op tape) of a pure CircuitJSON circuit."""
    c = from_json(j)
    return c.qubits, c.tape()


# ---------------------------------------------------------------------------
# OpenQASM
# ---------------------------------------------------------------------------
def fmt_angle(r, pi_token="pi"):
    """fmtAngle (src/circuit.ts:330-344): rational multiples of pi up to /16."""
    if abs(r) < 1e-14:
        return "0"
    f = r / math.pi
    for d in (1, 2, 3, 4, 6, 8, 12, 16):
        for n in range(-16, 17):
            if n == 0:
                continue
            if abs(f - n / d) < 1e-12:
                sign = "-" if n < 0 else ""
                a = abs(n)
                if d == 1:
                    return f"{sign}{pi_token}" if a == 1 else f"{sign}{a}*{pi_token}"
                return f"{sign}{pi_token}/{d}" if a == 1 else f"{sign}{a}*{pi_token}/{d}"
    return _js_num_str(r)


def _parse_float_prefix(t):
    """JavaScript parseFloat on a string of digits and dots: the longest valid
    decimal prefix, NaN when there is none."""
    m = re.match(r"\d*(?:\.\d*)?", t)
    s = m.group(0) if m else ""
    if s in ("", "."):
        return float("nan")
    return float(s)


def parse_angle(expr):
    """parseAngle (src/circuit.ts:349-379): numbers, pi, unary +/-, * / + -, parentheses."""
    s = re.sub(r"\s", "", expr)
    pos = [0]

    def factor():
        i = pos[0]
        if i < len(s) and s[i] == "-":
            pos[0] += 1
            return -factor()
        if i < len(s) and s[i] == "+":
            pos[0] += 1
            return factor()
        if i < len(s) and s[i] == "(":
            pos[0] += 1
            v = expr_()
            if pos[0] < len(s) and s[pos[0]] == ")":
                pos[0] += 1
            return v
        if s.startswith("pi", i):
            pos[0] += 2
            return math.pi
        j = i
        while pos[0] < len(s) and s[pos[0]] in "0123456789.":
            pos[0] += 1
        return _parse_float_prefix(s[j:pos[0]])

    def term():
        v = factor()
        while pos[0] < len(s) and s[pos[0]] in "*/":
            op = s[pos[0]]
            pos[0] += 1
            rhs = factor()
            if op == "*":
                v = v * rhs
            else:
                v = _js_div(v, rhs)
        return v

    def expr_():
        v = term()
        while pos[0] < len(s) and s[pos[0]] in "+-":
            op = s[pos[0]]
            pos[0] += 1
            rhs = term()
            v = v + rhs if op == "+" else v - rhs
        return v

    result = expr_()
    if pos[0] != len(s) or not math.isfinite(result):
        raise TypeError(f'Invalid angle expression: "{expr}"')
    return result


def _js_div(a, b):
    """a / b with IEEE results for a zero divisor (JavaScript's /)."""
    if b == 0:
        if a == 0 or math.isnan(a):
            return float("nan")
        return math.copysign(math.inf, a) * math.copysign(1.0, b)
    return a / b


def _qasm_gate_name(meta):
    """qasmGateName (src/circuit.ts:385-406)."""
    name, params = meta
    fixed = {"si": ("sdg", []), "ti": ("tdg", []), "v": ("sx", []), "vi": ("sxdg", []),
             "r2": ("rz", [math.pi / 2]), "r4": ("rz", [math.pi / 4]), "r8": ("rz", [math.pi / 8]),
             "cr2": ("crz", [math.pi / 2]), "cr4": ("crz", [math.pi / 4]), "cr8": ("crz", [math.pi / 8]),
             "cs": ("cu1", [math.pi / 2]), "ct": ("cu1", [math.pi / 4]), "csdg": ("cu1", [-math.pi / 2]),
             "ctdg": ("cu1", [-math.pi / 4])}
    if name in fixed:
        return fixed[name]
    if name == "vz":
        return "rz", [float(p) for p in params]
    return name, [float(p) for p in params]


def to_qasm(circ):
    """Circuit.toQASM() (src/circuit.ts:1960-2003): OpenQASM 2.0 text; TypeError for
    gates with no QASM 2.0 form (gpi, gpi2, two-qubit interaction gates, csrswap,
    unitary) and for `if`."""
    lines = ["OPENQASM 2.0;", 'include "qelib1.inc";', "", f"qreg q[{circ.qubits}];"]
    for name, size in circ._cregs.items():
        lines.append(f"creg {name}[{size}];")
    if circ._cregs:
        lines.append("")
    for op in _flatten(circ._ops()):
        k = op[0]
        if k == "cnot":
            lines.append(f"cx q[{op[1]}],q[{op[2]}];")
        elif k == "swap":
            lines.append(f"swap q[{op[1]}],q[{op[2]}];")
        elif k == "toffoli":
            lines.append(f"ccx q[{op[1]}],q[{op[2]}],q[{op[3]}];")
        elif k == "cswap":
            lines.append(f"cswap q[{op[1]}],q[{op[2]}],q[{op[3]}];")
        elif k == "csrswap":
            raise TypeError("Gate 'csrswap' has no OpenQASM 2.0 representation")
        elif k == "measure":
            lines.append(f"measure q[{op[1]}] -> {op[2]}[{op[3]}];")
        elif k == "reset":
            lines.append(f"reset q[{op[1]}];")
        elif k == "barrier":
            lines.append("barrier " + ",".join(f"q[{q}]" for q in op[1]) + ";")
        elif k == "if":
            raise TypeError("if ops cannot be serialized to OpenQASM 2.0")
        elif k == "single":
            meta = op[3]
            if not meta:
                raise TypeError("Single-qubit op missing serialization meta")
            if meta[0] in ("gpi", "gpi2"):
                raise TypeError(f"Gate '{meta[0]}' has no OpenQASM 2.0 representation")
            qn, qp = _qasm_gate_name(meta)
            ps = f"({','.join(fmt_angle(p) for p in qp)})" if qp else ""
            lines.append(f"{qn}{ps} q[{op[1]}];")
        elif k == "controlled":
            meta = op[4]
            if not meta:
                raise TypeError("Controlled op missing serialization meta")
            qn, qp = _qasm_gate_name(meta)
            ps = f"({','.join(fmt_angle(p) for p in qp)})" if qp else ""
            lines.append(f"{qn}{ps} q[{op[1]}],q[{op[2]}];")
        elif k == "two":
            raise TypeError(f"Gate '{op[4][0] if op[4] else 'two'}' has no OpenQASM 2.0 representation")
        elif k == "unitary":
            raise TypeError("Gate 'unitary' has no OpenQASM 2.0 representation")
        elif k == "parametric":
            raise TypeError(f"Gate '{op[1]}' has unbound parameters")
    return "\n".join(lines)


def _apply_qasm_gate(c, name, params, qs):
    """applyQASMGate (src/circuit.ts:409-450)."""
    a = qs[0] if len(qs) > 0 else None
    b = qs[1] if len(qs) > 1 else None
    d = qs[2] if len(qs) > 2 else None
    p = (list(params) + [None, None, None])[:3]
    one = {"id": c.id, "h": c.h, "x": c.x, "y": c.y, "z": c.z, "s": c.s, "sdg": c.si, "t": c.t, "tdg": c.ti,
           "sx": c.v, "sxdg": c.vi, "srn": c.srn, "srndg": c.srndg}
    if name == "barrier":
        return c.barrier(*qs) if qs else c.barrier()
    if name in one:
        return one[name](a)
    if name in ("p", "rx", "ry", "rz", "u1"):
        return getattr(c, name)(p[0], a)
    if name == "u2":
        return c.u2(p[0], p[1], a)
    if name == "u3":
        return c.u3(p[0], p[1], p[2], a)
    if name == "cx":
        return c.cnot(a, b)
    if name in ("cy", "cz", "ch", "csrn"):
        return getattr(c, name)(a, b)
    if name in ("crx", "cry", "crz", "cu1"):
        return getattr(c, name)(p[0], a, b)
    if name == "cu2":
        return c.cu2(p[0], p[1], a, b)
    if name == "cu3":
        return c.cu3(p[0], p[1], p[2], a, b)
    if name == "swap":
        return c.swap(a, b)
    if name == "ccx":
        return c.ccx(a, b, d)
    if name == "cswap":
        return c.cswap(a, b, d)
    raise TypeError(f"Unknown QASM gate: '{name}'")


_A = re.ASCII
_QREG2 = re.compile(r"^qreg\s+\w+\[(\d+)\]$", _A)
_QREG3 = re.compile(r"^qubit(?:\[(\d+)\])?\s+\w+$", _A)
_CREG2 = re.compile(r"^creg\s+(\w+)\[(\d+)\]$", _A)
_CREG3 = re.compile(r"^bit(?:\[(\d+)\])?\s+(\w+)$", _A)
_SKIP = re.compile(r"^(OPENQASM|include|qreg|creg|qubit|bit)\b", _A)
_MEAS2 = re.compile(r"^measure\s+\w+\[(\d+)\]\s*->\s*(\w+)\[(\d+)\]$", _A)
_MEAS3 = re.compile(r"^(\w+)\[(\d+)\]\s*=\s*measure\s+\w+\[(\d+)\]$", _A)
_RESET = re.compile(r"^reset\s+\w+\[(\d+)\]$", _A)
_GATE = re.compile(r"^(\w+)(?:\(([^)]*)\))?\s+([\w\[\],\s]+)$", _A)
_INDEX = re.compile(r"\[(\d+)\]")


def from_qasm(source):
    """Circuit.fromQASM(source) (src/circuit.ts:2017-2090): OpenQASM 2.0 or 3.0.
    Statements that match no form are skipped, as in ket; an unknown gate name is
    a TypeError and an out-of-range qubit a RangeError."""
    text = re.sub(r"/\*[\s\S]*?\*/", "", source)
    text = re.sub(r"//[^\n]*", "", text)
    stmts = [t.strip() for t in text.split(";")]
    stmts = [t for t in stmts if t]
    qubits = 0
    cregs = {}
    for st in stmts:
        m = _QREG2.match(st)
        if m:
            qubits += int(m.group(1))
            continue
        m = _QREG3.match(st)
        if m:
            qubits += int(m.group(1) or 1)
            continue
        m = _CREG2.match(st)
        if m:
            cregs[m.group(1)] = int(m.group(2))
            continue
        m = _CREG3.match(st)
        if m:
            cregs[m.group(2)] = int(m.group(1) or 1)
    c = Circuit(qubits)
    for name, size in cregs.items():
        c = c.creg(name, size)
    for st in stmts:
        if _SKIP.match(st):
            continue
        m = _MEAS2.match(st)
        if m:
            c = c.measure(int(m.group(1)), m.group(2), int(m.group(3)))
            continue
        m = _MEAS3.match(st)
        if m:
            c = c.measure(int(m.group(3)), m.group(1), int(m.group(2)))
            continue
        m = _RESET.match(st)
        if m:
            c = c.reset(int(m.group(1)))
            continue
        m = _GATE.match(st)
        if m:
            params = [parse_angle(x) for x in m.group(2).split(",")] if m.group(2) else []
            qs = [int(x) for x in _INDEX.findall(m.group(3))]
            c = _apply_qasm_gate(c, m.group(1), params, qs)
    return c


def qasm_tape(source):
    """(qubits, packed ksv_op tape) of a pure OpenQASM circuit."""
    c = from_qasm(source)
    return c.qubits, c.tape()


__all__ = ["to_json", "from_json", "json_tape", "to_qasm", "from_qasm", "qasm_tape", "parse_angle", "fmt_angle",
           "ops_to_json", "ops_from_json", "RangeError"]
