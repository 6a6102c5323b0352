'use strict'
/*
 * TEST INFRASTRUCTURE ONLY (a checker; never the product).
 *
 * A JavaScript restatement of ket's density-matrix simulation, runDM, as ket
 * computes it: rho is a Map from the BigInt key (row << n) | col to {re, im},
 * every channel builds a new Map by walking the old one in insertion order,
 * every write prunes |v|^2 < 1e-14, and depolarize2 accumulates through
 * ket's dmAcc, which deletes a running sum the moment it is negligible, so
 * its results depend on the Map's order.  It records that order and how much
 * the partial-sum pruning dropped, for the device path that reproduces it.
 *
 * Followed, function by function (reference files, read not run):
 *   dmGet / dmSet / dmAcc      src/density.ts:19-35
 *   applySingle (U rho U^+)    src/density.ts:45-77 (rows by U, then columns by U^+)
 *   applyTwo                   src/density.ts:80-118
 *   applyPerm                  src/density.ts:121-129 (order kept)
 *   applyUnitaryN              src/density.ts:132-178
 *   depolarize1                src/density.ts:193-224 (closed form per context)
 *   PAULI15 / depolarize2      src/density.ts:226-272 (dmAcc in Map order)
 *   runDM                      src/density.ts:475-588 (noise after single /
 *                              two-qubit ops; csrswap as four control blocks
 *                              merged in the order 00, 01, 10, 11)
 *   controlledGate             src/mps.ts:53-60
 *
 *   node ket_dm.js job.json
 *   job = {n, ops, mats, p1, p2, dump_rho, dump_order}
 * Prints one JSON line {entries, dropped, dropped_abs, ms}: `dropped` counts
 * the running sums dmAcc deleted and `dropped_abs` adds their magnitudes (a
 * bound on how far ket's result lies from sums formed whole).  dump_rho gets
 * the dense rho (complex128 by key), dump_order the keys in Map order (uint64).
 */
const fs = require('fs')
const { KIND, readTape } = require('./ket_sparse.js')

const EPS = 1e-14
const Z = { re: 0, im: 0 }
const ONE = { re: 1, im: 0 }
const cmul = (a, b) => ({ re: a.re * b.re - a.im * b.im, im: a.re * b.im + a.im * b.re })
const cadd = (a, b) => ({ re: a.re + b.re, im: a.im + b.im })
const conj = (a) => ({ re: a.re, im: -a.im })
const small = (a) => a.re * a.re + a.im * a.im < EPS

const stats = { dropped: 0, dropped_abs: 0 }

class Rho {
  constructor(n) {
    this.n = n
    this.sh = BigInt(n)
    this.lo = (1n << this.sh) - 1n
  }
  key(r, c) { return (r << this.sh) | c }
  row(k) { return k >> this.sh }
  col(k) { return k & this.lo }
}

function get(R, m, r, c) { return m.get(R.key(r, c)) || Z }

function set(R, m, r, c, v) {
  const k = R.key(r, c)
  if (small(v)) m.delete(k)
  else m.set(k, v)
}

// dmAcc: add into the running entry; a negligible running sum is deleted
function acc(R, m, r, c, v) {
  const k = R.key(r, c)
  const had = m.get(k)
  const nx = had ? cadd(had, v) : v
  if (small(nx)) {
    m.delete(k)
    stats.dropped++
    stats.dropped_abs += Math.hypot(nx.re, nx.im)
  } else m.set(k, nx)
}

// local index i of the k qubits (qs[0] the most significant) -> mask
function expand(masks, i) {
  const k = masks.length
  let b = 0n
  for (let j = 0; j < k; j++) if ((i >> (k - 1 - j)) & 1) b |= masks[j]
  return b
}

function single(R, m, q, g) {
  const next = new Map()
  const qm = 1n << BigInt(q)
  const both = (qm << R.sh) | qm
  const done = new Set()
  const [a, b, c, d] = g
  const ca = conj(a), cb = conj(b), cc = conj(c), cd = conj(d)
  for (const k of m.keys()) {
    const ctx = k & ~both
    if (done.has(ctx)) continue
    done.add(ctx)
    const r0 = R.row(ctx), c0 = R.col(ctx), r1 = r0 | qm, c1 = c0 | qm
    const p00 = get(R, m, r0, c0), p01 = get(R, m, r0, c1), p10 = get(R, m, r1, c0), p11 = get(R, m, r1, c1)
    const t00 = cadd(cmul(a, p00), cmul(b, p10)), t01 = cadd(cmul(a, p01), cmul(b, p11))
    const t10 = cadd(cmul(c, p00), cmul(d, p10)), t11 = cadd(cmul(c, p01), cmul(d, p11))
    set(R, next, r0, c0, cadd(cmul(ca, t00), cmul(cb, t01)))
    set(R, next, r0, c1, cadd(cmul(cc, t00), cmul(cd, t01)))
    set(R, next, r1, c0, cadd(cmul(ca, t10), cmul(cb, t11)))
    set(R, next, r1, c1, cadd(cmul(cc, t10), cmul(cd, t11)))
  }
  return next
}

// U rho U^+ for a 2^k x 2^k U on qubits qs: rows first (t = U p), then
// columns (t U^+), each sum started from zero and taken in index order; the
// two-qubit case is applyTwo, wider ones applyUnitaryN (the same loops).
function block(R, m, qs, U) {
  const dim = 1 << qs.length
  const masks = qs.map((q) => 1n << BigInt(q))
  const all = masks.reduce((x, y) => x | y, 0n)
  const both = (all << R.sh) | all
  const next = new Map()
  const done = new Set()
  const rows = new Array(dim), cols = new Array(dim)
  for (const k of m.keys()) {
    const ctx = k & ~both
    if (done.has(ctx)) continue
    done.add(ctx)
    const rb = R.row(ctx), cb = R.col(ctx)
    for (let i = 0; i < dim; i++) { rows[i] = rb | expand(masks, i); cols[i] = cb | expand(masks, i) }
    const p = []
    for (let i = 0; i < dim; i++) {
      const line = new Array(dim)
      for (let j = 0; j < dim; j++) line[j] = get(R, m, rows[i], cols[j])
      p.push(line)
    }
    const t = []
    for (let i = 0; i < dim; i++) {
      const line = new Array(dim)
      for (let j = 0; j < dim; j++) {
        let v = Z
        for (let l = 0; l < dim; l++) v = cadd(v, cmul(U[i * dim + l], p[l][j]))
        line[j] = v
      }
      t.push(line)
    }
    for (let i = 0; i < dim; i++) {
      for (let j = 0; j < dim; j++) {
        let v = Z
        for (let l = 0; l < dim; l++) v = cadd(v, cmul(t[i][l], conj(U[j * dim + l])))
        set(R, next, rows[i], cols[j], v)
      }
    }
  }
  return next
}

function perm(R, m, f) {
  const next = new Map()
  for (const [k, v] of m) next.set(R.key(f(R.row(k)), f(R.col(k))), v)
  return next
}

function dep1(R, m, q, p) {
  if (p <= 0) return m
  const next = new Map()
  const qm = 1n << BigInt(q)
  const both = (qm << R.sh) | qm
  const done = new Set()
  const same = 1 - 2 * p / 3, flip = 2 * p / 3, cross = 1 - 4 * p / 3
  for (const k of m.keys()) {
    const ctx = k & ~both
    if (done.has(ctx)) continue
    done.add(ctx)
    const r0 = R.row(ctx), c0 = R.col(ctx), r1 = r0 | qm, c1 = c0 | qm
    const p00 = get(R, m, r0, c0), p11 = get(R, m, r1, c1)
    set(R, next, r0, c0, { re: same * p00.re + flip * p11.re, im: same * p00.im + flip * p11.im })
    set(R, next, r1, c1, { re: same * p11.re + flip * p00.re, im: same * p11.im + flip * p00.im })
    const p01 = get(R, m, r0, c1), p10 = get(R, m, r1, c0)
    set(R, next, r0, c1, { re: cross * p01.re, im: cross * p01.im })
    set(R, next, r1, c0, { re: cross * p10.re, im: cross * p10.im })
  }
  return next
}

// the 15 non-identity two-qubit Paulis P_a (x) P_b in ket's order (II skipped):
// b runs fastest over I, X, Y, Z.  flip: X or Y; phase: Y or Z.
const PAULIS = []
for (let pa = 0; pa < 4; pa++) {
  for (let pb = 0; pb < 4; pb++) {
    if (pa || pb) PAULIS.push([pa === 1 || pa === 2, pb === 1 || pb === 2, pa >= 2, pb >= 2])
  }
}

function dep2(R, m, a, b, p) {
  if (p <= 0) return m
  const next = new Map()
  const am = 1n << BigInt(a), bm = 1n << BigInt(b)
  const keep = 1 - p, w = p / 15
  for (const [k, v] of m) {
    const r = R.row(k), c = R.col(k)
    acc(R, next, r, c, { re: keep * v.re, im: keep * v.im })
    const da = ((r ^ c) & am) !== 0n, db = ((r ^ c) & bm) !== 0n
    for (const [fa, fb, za, zb] of PAULIS) {
      const x = (fa ? am : 0n) | (fb ? bm : 0n)
      const s = ((za && da) !== (zb && db)) ? -w : w
      acc(R, next, r ^ x, c ^ x, { re: s * v.re, im: s * v.im })
    }
  }
  return next
}

function csrswap(R, m, ctl, a, b, V) {
  const cm = 1n << BigInt(ctl)
  const parts = [new Map(), new Map(), new Map(), new Map()]
  for (const [k, v] of m) parts[((R.row(k) & cm) ? 2 : 0) | ((R.col(k) & cm) ? 1 : 0)].set(k, v)
  const tr = (x) => { const o = new Map(); for (const [k, v] of x) o.set(R.key(R.col(k), R.row(k)), v); return o }
  const b01 = tr(block(R, tr(parts[1]), [a, b], V))
  const b10 = block(R, parts[2], [a, b], V)
  const b11 = block(R, parts[3], [a, b], V)
  const out = new Map(parts[0])
  for (const part of [b01, b10, b11]) for (const [k, v] of part) acc(R, out, R.row(k), R.col(k), v)
  return out
}

function runDM(n, tape, p1, p2) {
  const R = new Rho(n)
  const h = 1 / Math.sqrt(2)
  const V = [ONE, Z, Z, Z, Z, { re: h, im: 0 }, { re: 0, im: h }, Z, Z, { re: 0, im: h }, { re: h, im: 0 }, Z,
    Z, Z, Z, ONE]
  const bit = (q) => 1n << BigInt(q)
  let m = new Map([[0n, { re: 1, im: 0 }]])
  for (const op of tape) {
    const q = op.q
    switch (op.kind) {
      case KIND.SINGLE:
        m = single(R, m, q[0], op.m)
        if (p1) m = dep1(R, m, q[0], p1)
        break
      case KIND.CNOT: {
        const cb = bit(q[0]), tb = bit(q[1])
        m = perm(R, m, (i) => ((i & cb) ? i ^ tb : i))
        if (p2) m = dep2(R, m, q[0], q[1], p2)
        break
      }
      case KIND.SWAP: {
        const ab = bit(q[0]), bb = bit(q[1])
        m = perm(R, m, (i) => (((i & ab) !== 0n) === ((i & bb) !== 0n) ? i : i ^ ab ^ bb))
        if (p2) m = dep2(R, m, q[0], q[1], p2)
        break
      }
      case KIND.TWO:
        m = block(R, m, [q[0], q[1]], op.m)
        if (p2) m = dep2(R, m, q[0], q[1], p2)
        break
      case KIND.CONTROLLED: {
        const g = op.m
        m = block(R, m, [q[0], q[1]], [ONE, Z, Z, Z, Z, ONE, Z, Z, Z, Z, g[0], g[1], Z, Z, g[2], g[3]])
        if (p2) m = dep2(R, m, q[0], q[1], p2)
        break
      }
      case KIND.TOFFOLI: {
        const c1 = bit(q[0]), c2 = bit(q[1]), tb = bit(q[2])
        m = perm(R, m, (i) => ((i & c1) && (i & c2) ? i ^ tb : i))
        break
      }
      case KIND.CSWAP: {
        const cb = bit(q[0]), ab = bit(q[1]), bb = bit(q[2])
        m = perm(R, m, (i) => (!(i & cb) || ((i & ab) !== 0n) === ((i & bb) !== 0n) ? i : i ^ ab ^ bb))
        break
      }
      case KIND.CSRSWAP:
        m = csrswap(R, m, q[0], q[1], q[2], V)
        break
      case KIND.UNITARY:
        m = block(R, m, q, op.m)
        break
      default:
        break       // barrier
    }
  }
  return m
}

function main() {
  const job = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'))
  const n = job.n
  const tape = readTape(job.ops, job.mats)
  const t0 = process.hrtime.bigint()
  const m = runDM(n, tape, job.p1 || 0, job.p2 || 0)
  const ms = Number(process.hrtime.bigint() - t0) / 1e6
  if (job.dump_rho) {
    const buf = Buffer.alloc(16 * 4 ** n)
    for (const [k, v] of m) {
      const i = Number(k)
      buf.writeDoubleLE(v.re, 16 * i)
      buf.writeDoubleLE(v.im, 16 * i + 8)
    }
    fs.writeFileSync(job.dump_rho, buf)
  }
  if (job.dump_order) {
    const buf = Buffer.alloc(8 * m.size)
    let i = 0
    for (const k of m.keys()) buf.writeBigUInt64LE(k, 8 * i++)
    fs.writeFileSync(job.dump_order, buf)
  }
  process.stdout.write(JSON.stringify({ entries: m.size, dropped: stats.dropped, dropped_abs: stats.dropped_abs,
    ms }) + '\n')
}

main()
