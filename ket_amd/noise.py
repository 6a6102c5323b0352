"""run()'s per-shot path: noise, mid-circuit measurement, reset and `if`
(src/circuit.ts:169-207, :209-304, :3455-3481).

Two drivers.  `run_batched` (circuits without `if`) holds up to 8192 shots
as copies of one device state and walks the circuit once for all of them;
`run_per_shot` re-simulates shot by shot and serves circuits with `if`,
whose draw count depends on measured values.

Each shot re-simulates the circuit, consuming the seeded xorshift32 stream
in ket's order: one draw per 1-qubit gate when p1 > 0, one per cnot /
controlled / swap / two gate when p2 > 0, one per measure (a second when
pMeas > 0), one per reset, one for the final sample, then one per qubit
when pMeas > 0.  Gates between two draws that need the state (measure,
reset) are applied as one tape.
"""
import numpy as np

from . import gates as G
from ._native import KSV_SHOT_MEASURE, KSV_SHOT_PAULI, KSV_SHOT_RESET, SHOT_EVENT_DTYPE, State
from .tape import pack

# Published device error rates (src/circuit.ts:249-261).
DEVICES = {
    "aria-1": {"qubits": 25, "nativeGates": ("gpi", "gpi2", "ms", "vz"), "noise": {"p1": 0.0003, "p2": 0.005, "pMeas": 0.004}},
    "forte-1": {"qubits": 36, "nativeGates": ("gpi", "gpi2", "ms", "vz", "zz"), "noise": {"p1": 0.0001, "p2": 0.002, "pMeas": 0.002}},
    "harmony": {"qubits": 11, "nativeGates": ("gpi", "gpi2", "ms", "vz"), "noise": {"p1": 0.001, "p2": 0.015, "pMeas": 0.01}},
    "ibm_sherbrooke": {"qubits": 127, "noise": {"p1": 2.4e-4, "p2": 7.4e-3, "pMeas": 1.35e-2}},
    "ibm_brisbane": {"qubits": 127, "noise": {"p1": 2.4e-4, "p2": 7.6e-3, "pMeas": 1.35e-2}},
    "ibm_torino": {"qubits": 133, "noise": {"p1": 2.0e-4, "p2": 3.0e-3, "pMeas": 1.0e-2}},
    "h1-1": {"qubits": 20, "noise": {"p1": 1.8e-5, "p2": 9.7e-4, "pMeas": 2.3e-3}},
    "h2-1": {"qubits": 56, "noise": {"p1": 1.9e-5, "p2": 1.1e-3, "pMeas": 1.0e-3}},
}

# 15 non-identity two-qubit Paulis (src/circuit.ts:273-278)
_TWO_PAULI = [(None, G.X), (None, G.Y), (None, G.Z),
              (G.X, None), (G.X, G.X), (G.X, G.Y), (G.X, G.Z),
              (G.Y, None), (G.Y, G.X), (G.Y, G.Y), (G.Y, G.Z),
              (G.Z, None), (G.Z, G.X), (G.Z, G.Y), (G.Z, G.Z)]


def resolve_noise(noise):
    """Device name -> NoiseParams dict, dict as-is, None -> None (src/circuit.ts:3411-3417)."""
    if noise is None:
        return None
    if isinstance(noise, str):
        d = DEVICES.get(noise)
        if d is None:
            raise TypeError(f"Unknown device profile '{noise}'. Known: {', '.join(DEVICES)}")
        return dict(d["noise"])
    return dict(noise)


def _creg_value(shot_cregs, name):
    """cregValue (src/circuit.ts:164-166): acc | (1 << i) over the set bits, with
    JavaScript's 32-bit << (count mod 32) and | (a signed 32-bit result)."""
    v = 0
    for i, b in enumerate(shot_cregs.get(name, [])):
        if b:
            v |= 1 << (i & 31)
    return v - (1 << 32) if v & (1 << 31) else v


class _Shot:
    def __init__(self, st, rng, noise):
        self.st = st
        self.rng = rng
        self.p1 = (noise or {}).get("p1", 0) or 0
        self.p2 = (noise or {}).get("p2", 0) or 0
        self.pm = (noise or {}).get("pMeas", 0) or 0
        self.pending = []

    def flush(self):
        if self.pending:
            self.st.apply(pack(self.pending))
            self.pending = []

    def dep1(self, q):
        rand = self.rng()
        if rand >= self.p1:
            return
        r = rand / self.p1
        g = G.X if r < 1 / 3 else (G.Y if r < 2 / 3 else G.Z)
        self.pending.append(("single", q, g, None))

    def dep2(self, a, b):
        rand = self.rng()
        if rand >= self.p2:
            return
        pa, pb = _TWO_PAULI[min(int(rand / self.p2 * 15), 14)]
        if pa is not None:
            self.pending.append(("single", a, pa, None))
        if pb is not None:
            self.pending.append(("single", b, pb, None))

    def run_ops(self, ops, shot_cregs):
        from .circuit import _flatten
        for op in _flatten(ops):
            k = op[0]
            if k == "single":
                self.pending.append(op)
                if self.p1:
                    self.dep1(op[1])
            elif k in ("cnot", "controlled", "swap", "two"):
                self.pending.append(op)
                if self.p2:
                    self.dep2(op[1], op[2])
            elif k in ("toffoli", "cswap", "csrswap", "unitary"):
                self.pending.append(op)
            elif k == "measure":
                self.flush()
                outcome = self.st.collapse(op[1], self.rng())
                reported = (1 - outcome) if (self.pm and self.rng() < self.pm) else outcome
                reg = shot_cregs.get(op[2])
                if reg is not None:
                    reg[op[3]] = reported == 1
            elif k == "reset":
                self.flush()
                outcome = self.st.collapse(op[1], self.rng())
                if outcome == 1:
                    self.pending.append(("single", op[1], G.X, None))
            elif k == "if":
                self.flush()
                if _creg_value(shot_cregs, op[1]) == op[2]:
                    self.run_ops(op[3], shot_cregs)
            # barrier: nothing
        self.flush()


def run_per_shot(circ, shots, rng, noise, init, creg_counts):
    """Per-shot simulation; returns (counts {index: n}, creg_counts)."""
    from .circuit import OPTIONS
    # a state of at most 13 qubits runs on the kernel and tile the batched copies
    # use (ksv_shots_run: register rounds on an n-qubit tile for 11..13 qubits, LDS
    # passes below), so both drivers give bit-identical states
    n = circ.qubits
    kernel, rtile = ("lds", 0) if n <= 10 else ("reg", n) if n <= 13 else ("auto", 0)
    st = State(n, OPTIONS["precision"], prune=OPTIONS["prune"], tile=OPTIONS["tile"], kernel=kernel, rtile=rtile)
    counts = {}
    pm = (noise or {}).get("pMeas", 0) or 0
    ops = circ._ops()
    for _ in range(shots):
        shot_cregs = {name: [False] * size for name, size in circ._cregs.items()}
        st.init_basis(init)
        sh = _Shot(st, rng, noise)
        sh.run_ops(ops, shot_cregs)
        idx = int(st.shots_sample(n, np.array([rng()]))[0])     # sampleSV on the one state
        if pm:
            for q in range(circ.qubits):
                if rng() < pm:
                    idx ^= 1 << q
        counts[idx] = counts.get(idx, 0) + 1
        for name, bits in shot_cregs.items():
            acc = creg_counts[name]
            for j, b in enumerate(bits):
                if b:
                    acc[j] += 1
    st.free()
    return counts, creg_counts


# ---------------------------------------------------------------------------
# Batched trajectories
# ---------------------------------------------------------------------------
MAX_BATCH_AMPS = 1 << 28          # amplitudes per batched state (4 GiB FP64)
MAX_COPIES = 8192


def batchable(circ):
    """True when every shot consumes the same number of draws: no `if`."""
    from .circuit import _flatten
    return not any(op[0] == "if" for op in _flatten(circ._ops()))


def draws_per_shot(flat, n, p1, p2, pm):
    """Draws one shot consumes, in ket's order (src/circuit.ts:169-207, :3455-3481)."""
    d = 0
    for op in flat:
        k = op[0]
        if k == "single":
            d += 1 if p1 else 0
        elif k in ("cnot", "controlled", "swap", "two"):
            d += 1 if p2 else 0
        elif k == "measure":
            d += 2 if pm else 1
        elif k == "reset":
            d += 1
    return d + 1 + (n if pm else 0)


def run_batched(circ, shots, rng, noise, init, creg_counts):
    """All shots of a circuit without `if` as copies of one device state.

    Shot s consumes draws [s D, (s + 1) D) of the stream, so each batch's
    draws are the next cnt * D of it, drawn when the batch starts (host
    memory holds one batch of draws, not all shots); gates go to every copy at once, Pauli errors, measurements,
    resets and the final draw copy by copy (ksv_shots_*).  Gives the same
    counts as `run_per_shot` for the same seed.
    """
    from .circuit import OPTIONS, _flatten
    n = circ.qubits
    flat = _flatten(circ._ops())
    p1 = (noise or {}).get("p1", 0) or 0
    p2 = (noise or {}).get("p2", 0) or 0
    pm = (noise or {}).get("pMeas", 0) or 0
    D = draws_per_shot(flat, n, p1, p2, pm)

    def next_draws(cnt):
        if hasattr(rng, "fill"):
            return rng.fill(cnt * D).reshape(cnt, D)
        return np.fromiter((rng() for _ in range(cnt * D)), dtype=np.float64, count=cnt * D).reshape(cnt, D)

    b = 0
    while (1 << (b + 1)) <= max(1, shots) and (1 << (b + 1)) <= MAX_COPIES and (2 << b) << n <= MAX_BATCH_AMPS:
        b += 1
    B = 1 << b
    counts, memo, states = {}, {}, {}
    try:
        for s0 in range(0, shots, B):
            cnt = min(B, shots - s0)
            bb = b                                    # the last batch: the smallest power of two that holds it
            while bb > 0 and (1 << (bb - 1)) >= cnt:
                bb -= 1
            Bc = 1 << bb
            if bb not in states:
                states[bb] = State(n + bb, OPTIONS["precision"], prune=OPTIONS["prune"], tile=OPTIONS["tile"])
            draws = next_draws(cnt)                   # the stream is sequential: batch by batch
            if cnt == Bc:
                dr = draws
            else:                                     # the last copies of a short batch idle on 0.5s
                dr = np.full((Bc, D), 0.5)
                dr[:cnt] = draws
            regs = {name: np.zeros((Bc, size), dtype=bool) for name, size in circ._cregs.items()}
            idx = _run_copies(states[bb], n, flat, dr, p1, p2, pm, init, regs, memo)
            for t in range(cnt):
                i = int(idx[t])
                counts[i] = counts.get(i, 0) + 1
            for name, bits in regs.items():
                acc = creg_counts[name]
                for j in range(bits.shape[1]):
                    acc[j] += int(bits[:cnt, j].sum())
    finally:
        for st in states.values():
            st.free()
    return counts, creg_counts


# Pauli codes of the ksv_shots_pauli ABI (1 X, 2 Y, 3 Z; 0 none) for the 15
# two-qubit Paulis of _TWO_PAULI, first and second qubit
_PAIR_A = np.array([0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3], dtype=np.uint8)
_PAIR_B = np.array([1, 2, 3, 0, 1, 2, 3, 0, 1, 2, 3, 0, 1, 2, 3], dtype=np.uint8)


_TAPE_KINDS = ("single", "cnot", "controlled", "swap", "two", "toffoli", "cswap", "csrswap", "unitary")


def _hits(dr, thr):
    """Flat indices of dr < thr (row-major order), rows split over a few threads
    (numpy's comparison and nonzero release the GIL)."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    B, D = dr.shape
    nt = min(8, os.cpu_count() or 1, max(1, B // 1024))
    if nt <= 1:
        return _sparse_flatnonzero(dr < thr)
    cut = [B * i // nt for i in range(nt + 1)]

    def part(i):
        return cut[i] * D + _sparse_flatnonzero(dr[cut[i]:cut[i + 1]] < thr)

    with ThreadPoolExecutor(nt) as ex:
        return np.concatenate(list(ex.map(part, range(nt))))


def _sparse_flatnonzero(mask):
    """np.flatnonzero for a mask with few set entries: scans 8 bytes at a time."""
    flat = np.ascontiguousarray(mask).reshape(-1).view(np.uint8)
    body = len(flat) // 8 * 8
    words = np.flatnonzero(flat[:body].view(np.uint64))
    idx = (words[:, None] * 8 + np.arange(8)).reshape(-1)
    idx = idx[flat[idx] != 0]
    tail = body + np.flatnonzero(flat[body:])
    return np.concatenate([idx, tail]).astype(np.int64)


def shot_events(flat, dr, p1, p2, pm):
    """Host side of one batch: the gate tape and the event list of ksv_shots_run.

    `dr` holds one row of draws per copy.  Returns (gates, ev, copies, codes, r,
    meas, k): Pauli events are grouped per (gate, qubit) with their copies in
    ascending order, measure / reset events follow the noise of the gate before
    them, `r` holds the draws of every measure / reset event (B per event),
    `meas` lists (op, offset in r, flip mask or None) per measurement and `k` is
    the next unused draw column.  dep1 picks X, Y, Z by thirds of rand / p1 and
    dep2 one of the 15 Pauli pairs (src/circuit.ts:263-304); the columns are
    evaluated for all copies at once.
    """
    B = dr.shape[0]
    gates, ones, twos, ms = [], [], [], []
    k = 0
    for op in flat:                      # draw columns, in ket's order
        kind = op[0]
        if kind in _TAPE_KINDS:
            gates.append(op)
        at = len(gates)
        if kind == "single" and p1:
            ones.append((at, op[1], k))
            k += 1
        elif kind in ("cnot", "controlled", "swap", "two") and p2:
            twos.append((at, op[1], op[2], k))
            k += 1
        elif kind in ("measure", "reset"):
            fc = k + 1 if (kind == "measure" and pm) else -1
            ms.append((at, kind, op, k, fc))
            k += 1 if fc < 0 else 2
    # one entry per (Pauli, copy): sort keys (at, sub, copy), then qubit and code.
    # Every noise column is compared with its rate in one pass over the draws.
    parts = []
    if ones or twos:
        D = dr.shape[1]
        thr = np.full(D, -1.0)                 # draws are >= 0: never a hit
        at_of = np.zeros(D, np.int64)
        qa_of = np.zeros(D, np.int64)
        qb_of = np.zeros(D, np.int64)
        if ones:
            a = np.array(ones, dtype=np.int64)
            thr[a[:, 2]] = p1
            at_of[a[:, 2]], qa_of[a[:, 2]] = a[:, 0], a[:, 1]
        if twos:
            a = np.array(twos, dtype=np.int64)
            thr[a[:, 3]] = p2
            at_of[a[:, 3]], qa_of[a[:, 3]], qb_of[a[:, 3]] = a[:, 0], a[:, 1], a[:, 2]
        c, col = np.divmod(_hits(dr, thr), D)
        rv = dr[c, col]
        two = np.zeros(D, bool)
        if twos:
            two[np.array(twos, dtype=np.int64)[:, 3]] = True
        t2 = two[col]
        s = ~t2
        if s.any():
            cs, r = col[s], rv[s] / p1
            parts.append((at_of[cs], np.zeros(len(cs), np.int64), c[s], qa_of[cs],
                          (1 + (r >= 1 / 3) + (r >= 2 / 3)).astype(np.uint8)))
        if t2.any():
            cs, c2 = col[t2], c[t2]
            pick = np.minimum(np.floor(rv[t2] / p2 * 15).astype(np.int64), 14)
            for sub, q_of, table in ((0, qa_of, _PAIR_A), (1, qb_of, _PAIR_B)):
                code = table[pick]
                m = code > 0
                parts.append((at_of[cs[m]], np.full(int(m.sum()), sub, np.int64), c2[m], q_of[cs[m]], code[m]))
    if parts:
        at, sub, cp, q, code = (np.concatenate(x) for x in zip(*parts))
        o = np.lexsort((cp, sub, at))
        at, sub, cp, q, code = at[o], sub[o], cp[o], q[o], code[o]
        first = np.flatnonzero(np.r_[True, (at[1:] != at[:-1]) | (sub[1:] != sub[:-1])])
        cnt = np.diff(np.r_[first, len(at)])
        p_at, p_sub, p_q = at[first], sub[first], q[first]
    else:
        cp, code = np.zeros(0, np.int64), np.zeros(0, np.uint8)
        first = cnt = p_at = p_sub = p_q = np.zeros(0, np.int64)
    nm = len(ms)
    ev = np.zeros(len(first) + nm, dtype=SHOT_EVENT_DTYPE)
    key_at = np.r_[p_at, np.array([m[0] for m in ms], dtype=np.int64)]
    key_sub = np.r_[p_sub, 2 + np.arange(nm, dtype=np.int64)]   # noise first, then measures in order
    ev["at"] = key_at
    ev["kind"] = np.r_[np.full(len(first), KSV_SHOT_PAULI),
                       [KSV_SHOT_MEASURE if m[1] == "measure" else KSV_SHOT_RESET for m in ms]]
    ev["q"] = np.r_[p_q, np.array([m[2][1] for m in ms], dtype=np.int64)]
    ev["off"] = np.r_[first, np.arange(nm, dtype=np.int64) * B]
    ev["count"] = np.r_[cnt, np.zeros(nm, np.int64)]
    ev = ev[np.lexsort((key_sub, key_at))]
    r = np.ascontiguousarray(dr[:, [m[3] for m in ms]].T).ravel() if nm else np.zeros(0)
    meas = [(m[2], i * B, (dr[:, m[4]] < pm) if m[4] >= 0 else None)
            for i, m in enumerate(ms) if m[1] == "measure"]
    return gates, ev, cp.astype(np.uint32), code.astype(np.uint8), r, meas, k


def _run_copies(st, n, flat, dr, p1, p2, pm, init, regs, memo):
    """One batch: events are collected on the host, then the whole batch runs in
    one ksv_shots_run call; returns the drawn index of every copy.  `memo` keeps
    the packed gate tape, which is the same for every batch of a run."""
    st.shots_init(n, init)
    gates, ev, copies, codes, r, meas, k = shot_events(flat, dr, p1, p2, pm)
    if "tape" not in memo:
        memo["tape"] = pack(gates)
    out = st.shots_run(n, memo["tape"], ev, copies, codes, r)
    B = dr.shape[0]
    for op, off, flip in meas:
        o = out[off:off + B]
        if flip is not None:
            o = o ^ flip
        reg = regs.get(op[2])
        if reg is not None:
            reg[:, op[3]] = o == 1
    idx = st.shots_sample(n, dr[:, k]).astype(np.uint64)
    k += 1
    if pm:
        for q in range(n):
            idx ^= (dr[:, k] < pm).astype(np.uint64) << np.uint64(q)
            k += 1
    assert k == dr.shape[1], "draw layout mismatch"
    return idx
