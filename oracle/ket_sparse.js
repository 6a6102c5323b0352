'use strict'
/*
 * TEST INFRASTRUCTURE ONLY (a CPU baseline and a checker; never the product).
 *
 * A JavaScript restatement of ket's own CPU statevector algorithm, run under
 * Node as ket runs: the state is a Map from BigInt basis index to {re, im},
 * every gate builds a new Map by walking the old one in insertion order and
 * pairing amplitudes through a `seen` set, and every write goes through ket's
 * accumulate prune (|a|^2 < 1e-14 deleted).  It exists to time ket's own
 * algorithm on the host cores of the machine the GPU path is measured on, and
 * to check that restatement against the C oracle bit for bit.
 *
 * Followed, function by function (reference files, read not run):
 *   complex arithmetic        src/complex.ts:11-24 (mul = re*re - im*im, ...)
 *   accumulate                src/statevector.ts:19-27
 *   applySingle               src/statevector.ts:47-66
 *   applyCNOT / applySWAP     src/statevector.ts:70-103
 *   applyTwo                  src/statevector.ts:110-133
 *   applyToffoli / applyCSwap src/statevector.ts:138-171
 *   applyCsrSwap              src/statevector.ts:177-207
 *   applyControlled           src/statevector.ts:212-235
 *   applyUnitary              src/statevector.ts:240-270
 *   probabilities             src/statevector.ts:273-280
 *   simulatePure              src/circuit.ts:143-161
 *   run() fast path           src/circuit.ts:3424-3450 (sorted CDF, last = 1,
 *                             binary search for the first cum >= r)
 *   exactProbs                src/circuit.ts:3700-3711 (q0-leftmost keys)
 *   makePrng (xorshift32)     src/circuit.ts:813-819
 *   timing                    benchmark/run.mjs:9-25 (one warm-up, median of N)
 *
 * Input is the packed tape of include/ketsv.h (64-byte ksv_op records and the
 * interleaved re/im matrix array) in two files, and a JSON job:
 *   node ket_sparse.js job.json
 *   job = {n, ops, mats, task, shots, seed, warmup, runs, dump, skip, budget_ms}
 * task: "statevector" (simulatePure), "exactProbs", "run" (simulate once and
 * draw `shots` seeded samples), "probs_run" (exactProbs and run's sampling of
 * one simulation), or "rate" (apply the first `skip` gates untimed, then the
 * next ones in order until budget_ms has passed: a bounded sample of a circuit
 * too long to finish, timed once its state has filled in).
 * Prints one JSON line: the per-run times (ms), their median, the final Map
 * size, for "run" the counts in first-seen order, for "rate" the gates done.
 * `dump` (optional) gets the final amplitudes as float64 (re, im) pairs in
 * ascending index, dense.
 */
const fs = require('fs')

const KIND = { SINGLE: 0, CNOT: 1, SWAP: 2, TWO: 3, CONTROLLED: 4, TOFFOLI: 5, CSWAP: 6, CSRSWAP: 7,
  UNITARY: 8, BARRIER: 9 }
const OP_BYTES = 64
const MAX_OPERANDS = 12
const EPS = 1e-14
const NONE = { re: 0, im: 0 }

// ---- complex arithmetic, each operation rounded as JavaScript rounds it ----
const cmul = (a, b) => ({ re: a.re * b.re - a.im * b.im, im: a.re * b.im + a.im * b.re })
const cadd = (a, b) => ({ re: a.re + b.re, im: a.im + b.im })
const small = (a) => a.re * a.re + a.im * a.im < EPS

function put(next, idx, amp) {
  const had = next.get(idx)
  const v = had === undefined ? amp : cadd(had, amp)
  if (small(v)) next.delete(idx)
  else next.set(idx, v)
}

// ---- the gate kinds ----
function single(sv, q, m) {
  const next = new Map()
  const bit = 1n << BigInt(q)
  const done = new Set()
  for (const k of sv.keys()) {
    const lo = k & ~bit
    if (done.has(lo)) continue
    done.add(lo)
    const hi = lo | bit
    const x0 = sv.get(lo) || NONE
    const x1 = sv.get(hi) || NONE
    put(next, lo, cadd(cmul(m[0], x0), cmul(m[1], x1)))
    put(next, hi, cadd(cmul(m[2], x0), cmul(m[3], x1)))
  }
  return next
}

function controlled(sv, c, t, m) {
  const next = new Map()
  const cb = 1n << BigInt(c)
  const tb = 1n << BigInt(t)
  const done = new Set()
  for (const [k, amp] of sv) {
    if ((k & cb) === 0n) { put(next, k, amp); continue }
    const lo = k & ~tb
    if (done.has(lo)) continue
    done.add(lo)
    const hi = lo | tb
    const x0 = sv.get(lo) || NONE
    const x1 = sv.get(hi) || NONE
    put(next, lo, cadd(cmul(m[0], x0), cmul(m[1], x1)))
    put(next, hi, cadd(cmul(m[2], x0), cmul(m[3], x1)))
  }
  return next
}

function permute(sv, f) {
  const next = new Map()
  for (const [k, amp] of sv) next.set(f(k), amp)
  return next
}

function cnot(sv, c, t) {
  const cb = 1n << BigInt(c), tb = 1n << BigInt(t)
  return permute(sv, (k) => ((k & cb) !== 0n ? k ^ tb : k))
}

function swap(sv, a, b) {
  const ab = 1n << BigInt(a), bb = 1n << BigInt(b)
  return permute(sv, (k) => (((k & ab) !== 0n) === ((k & bb) !== 0n) ? k : k ^ ab ^ bb))
}

function toffoli(sv, c1, c2, t) {
  const m1 = 1n << BigInt(c1), m2 = 1n << BigInt(c2), tb = 1n << BigInt(t)
  return permute(sv, (k) => ((k & m1) !== 0n && (k & m2) !== 0n ? k ^ tb : k))
}

function cswap(sv, c, a, b) {
  const cb = 1n << BigInt(c), ab = 1n << BigInt(a), bb = 1n << BigInt(b)
  return permute(sv, (k) => ((k & cb) === 0n || ((k & ab) !== 0n) === ((k & bb) !== 0n) ? k : k ^ ab ^ bb))
}

// Dense k-qubit block on qubits qs (qs[0] the most significant local bit):
// the row sums start from zero and add the columns in order, as applyTwo and
// applyUnitary do; applyTwo is the k = 2 case of the same loop.
function block(sv, qs, m) {
  const k = qs.length
  const dim = 1 << k
  const bits = qs.map((q) => 1n << BigInt(q))
  const all = bits.reduce((x, y) => x | y, 0n)
  const next = new Map()
  const done = new Set()
  const idx = new Array(dim)
  const x = new Array(dim)
  for (const key of sv.keys()) {
    const ctx = key & ~all
    if (done.has(ctx)) continue
    done.add(ctx)
    for (let i = 0; i < dim; i++) {
      let g = ctx
      for (let j = 0; j < k; j++) if ((i >> (k - 1 - j)) & 1) g |= bits[j]
      idx[i] = g
      x[i] = sv.get(g) || NONE
    }
    for (let r = 0; r < dim; r++) {
      let acc = NONE
      for (let c = 0; c < dim; c++) acc = cadd(acc, cmul(m[r * dim + c], x[c]))
      put(next, idx[r], acc)
    }
  }
  return next
}

function csrswap(sv, c, a, b) {
  const cb = 1n << BigInt(c), ab = 1n << BigInt(a), bb = 1n << BigInt(b)
  const h = 1 / Math.sqrt(2)
  const s = { re: h, im: 0 }, is = { re: 0, im: h }
  const next = new Map()
  const done = new Set()
  for (const key of sv.keys()) {
    const ctx = key & ~(cb | ab | bb)
    if (done.has(ctx)) continue
    done.add(ctx)
    for (const half of [0n, cb]) {
      const i00 = ctx | half, i01 = i00 | bb, i10 = i00 | ab, i11 = i00 | ab | bb
      const x00 = sv.get(i00) || NONE, x01 = sv.get(i01) || NONE
      const x10 = sv.get(i10) || NONE, x11 = sv.get(i11) || NONE
      if (half === 0n) {
        put(next, i00, x00); put(next, i01, x01); put(next, i10, x10); put(next, i11, x11)
      } else {
        put(next, i00, x00)
        put(next, i11, x11)
        put(next, i01, cadd(cmul(s, x01), cmul(is, x10)))
        put(next, i10, cadd(cmul(is, x01), cmul(s, x10)))
      }
    }
  }
  return next
}

// ---- the packed tape ----
function readTape(opsPath, matsPath) {
  const ob = fs.readFileSync(opsPath)
  const mb = matsPath ? fs.readFileSync(matsPath) : Buffer.alloc(0)
  const od = new DataView(ob.buffer, ob.byteOffset, ob.byteLength)
  const md = new DataView(mb.buffer, mb.byteOffset, mb.byteLength)
  const dbl = (i) => md.getFloat64(8 * i, true)
  const mat = (off, dim) => {
    const out = new Array(dim * dim)
    for (let i = 0; i < dim * dim; i++) out[i] = { re: dbl(off + 2 * i), im: dbl(off + 2 * i + 1) }
    return out
  }
  const tape = []
  for (let p = 0; p + OP_BYTES <= ob.byteLength; p += OP_BYTES) {
    const kind = od.getInt32(p, true)
    const nq = od.getInt32(p + 4, true)
    const q = []
    for (let j = 0; j < nq && j < MAX_OPERANDS; j++) q.push(od.getInt32(p + 8 + 4 * j, true))
    const off = Number(od.getBigInt64(p + 56, true))
    let m = null
    if (kind === KIND.SINGLE || kind === KIND.CONTROLLED) m = mat(off, 2)
    else if (kind === KIND.TWO) m = mat(off, 4)
    else if (kind === KIND.UNITARY) m = mat(off, 1 << nq)
    tape.push({ kind, q, m })
  }
  return tape
}

function apply(sv, op) {
  const q = op.q
  switch (op.kind) {
    case KIND.SINGLE: return single(sv, q[0], op.m)
    case KIND.CNOT: return cnot(sv, q[0], q[1])
    case KIND.SWAP: return swap(sv, q[0], q[1])
    case KIND.TWO: return block(sv, q, op.m)
    case KIND.CONTROLLED: return controlled(sv, q[0], q[1], op.m)
    case KIND.TOFFOLI: return toffoli(sv, q[0], q[1], q[2])
    case KIND.CSWAP: return cswap(sv, q[0], q[1], q[2])
    case KIND.CSRSWAP: return csrswap(sv, q[0], q[1], q[2])
    case KIND.UNITARY: return block(sv, q, op.m)
    case KIND.BARRIER: return sv
    default: throw new TypeError(`unknown op kind ${op.kind}`)
  }
}

function simulate(tape) {
  let sv = new Map([[0n, { re: 1, im: 0 }]])
  for (const op of tape) sv = apply(sv, op)
  return sv
}

// The first `skip` gates untimed (until the state has filled in), then gates
// in tape order until `budget` ms have passed.
function rate(tape, skip, budget) {
  let sv = new Map([[0n, { re: 1, im: 0 }]])
  for (let i = 0; i < skip && i < tape.length; i++) sv = apply(sv, tape[i])
  const t0 = process.hrtime.bigint()
  let done = Math.min(skip, tape.length)
  while (done < tape.length && Number(process.hrtime.bigint() - t0) / 1e6 < budget) sv = apply(sv, tape[done++])
  return { sv, done: done - Math.min(skip, tape.length), ms: Number(process.hrtime.bigint() - t0) / 1e6 }
}

function probabilities(sv) {
  const out = new Map()
  for (const [k, a] of sv) {
    const p = a.re * a.re + a.im * a.im
    if (p > EPS) out.set(k, p)
  }
  return out
}

function exactProbs(n, sv) {
  const out = {}
  for (const [k, p] of probabilities(sv)) out[k.toString(2).padStart(n, '0').split('').reverse().join('')] = p
  return Object.freeze(out)
}

function xorshift32(seed) {
  let s = ((seed >>> 0) || 1)
  return () => {
    s ^= s << 13; s ^= s >>> 17; s ^= s << 5
    return (s >>> 0) / 0x100000000
  }
}

function sampleRun(sv, shots, seed) {
  const rng = xorshift32(seed)
  const sorted = Array.from(probabilities(sv).entries()).sort(([a], [b]) => (a < b ? -1 : 1))
  const keys = sorted.map(([k]) => k)
  const cum = new Float64Array(sorted.length)
  let acc = 0
  for (let i = 0; i < sorted.length; i++) { acc += sorted[i][1]; cum[i] = acc }
  if (cum.length) cum[cum.length - 1] = 1.0
  const counts = new Map()
  for (let s = 0; s < shots; s++) {
    const r = rng()
    let lo = 0, hi = cum.length - 1
    while (lo < hi) {
      const mid = (lo + hi) >> 1
      if (cum[mid] < r) lo = mid + 1
      else hi = mid
    }
    const k = cum.length ? keys[lo] : 0n
    counts.set(k, (counts.get(k) || 0) + 1)
  }
  return counts
}

function main() {
  const job = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'))
  const tape = readTape(job.ops, job.mats)
  const n = job.n
  if (job.task === 'rate') {
    const r = rate(tape, job.skip || 0, job.budget_ms)
    process.stdout.write(JSON.stringify({ gates: r.done, of: tape.length, ms: r.ms, entries: r.sv.size,
      node: process.version }) + '\n')
    return
  }
  let last = null
  const work = () => {
    const sv = simulate(tape)
    last = { sv }
    if (job.task === 'exactProbs' || job.task === 'probs_run') last.probs = exactProbs(n, sv)
    if (job.task === 'run' || job.task === 'probs_run') last.counts = sampleRun(sv, job.shots, job.seed)
  }
  const times = []
  for (let i = 0; i < (job.warmup || 0); i++) work()
  for (let i = 0; i < (job.runs || 1); i++) {
    const t0 = process.hrtime.bigint()
    work()
    times.push(Number(process.hrtime.bigint() - t0) / 1e6)
  }
  const sorted = [...times].sort((a, b) => a - b)
  const out = { ms: times, median_ms: sorted[Math.floor(sorted.length / 2)], entries: last.sv.size,
    node: process.version }
  if (last.counts) out.counts = Array.from(last.counts, ([k, c]) => [k.toString(), c])
  if (job.dump) {
    const buf = Buffer.alloc(16 * (2 ** n))
    for (const [k, a] of last.sv) {
      const i = Number(k)
      buf.writeDoubleLE(a.re, 16 * i)
      buf.writeDoubleLE(a.im, 16 * i + 8)
    }
    fs.writeFileSync(job.dump, buf)
  }
  process.stdout.write(JSON.stringify(out) + '\n')
}

if (require.main === module) main()

module.exports = { KIND, readTape }
