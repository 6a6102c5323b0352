// Node-API addon checks.  Usage: node test_napi.js <path to ketsv.node> [gpu]
// Without "gpu": the error classes (RangeError / TypeError / Error) that reach
// JavaScript, the dry-run scheduler and makePrng's stream, no device needed.
// With "gpu": a Bell circuit and a 10-qubit circuit through create / apply /
// probs / sample / expect / amplitudes, against values computed here.
'use strict'
const assert = require('assert')
const addon = require(process.argv[2])
const gpu = process.argv[3] === 'gpu'

// ket's makePrng (src/circuit.ts:813-819), written out in JavaScript
function makePrng(seed) {
  let s = ((seed >>> 0) || 1)
  return () => { s ^= s << 13; s ^= s >>> 17; s ^= s << 5; return (s >>> 0) / 0x100000000 }
}

// one op of the ksv_op tape: kind, nq, q[12], mat (int64 as two int32)
function tape(list) {
  const t = new Int32Array(16 * list.length)
  list.forEach(([kind, qs, mat], i) => {
    t[16 * i] = kind; t[16 * i + 1] = qs.length
    qs.forEach((q, j) => { t[16 * i + 2 + j] = q })
    t[16 * i + 14] = mat; t[16 * i + 15] = mat < 0 ? -1 : 0
  })
  return t
}
const s2 = 1 / Math.sqrt(2)
const H = [s2, 0, s2, 0, s2, 0, -s2, 0]

// ---- no device -------------------------------------------------------------
assert.throws(() => addon.create(99), RangeError)
assert.throws(() => addon.create(0), RangeError)
assert.throws(() => addon.create('x'), TypeError)
assert.throws(() => addon.plan(3, tape([[42, [0], -1]]), new Float64Array(0)), TypeError)          // unknown kind
assert.throws(() => addon.plan(3, tape([[0, [7], 0]]), Float64Array.from(H)), RangeError)           // qubit out of range
assert.throws(() => addon.plan(3, tape([[1, [1, 1], -1]]), new Float64Array(0)), TypeError)         // control == target
assert.throws(() => addon.plan(3, tape([[0, [0], 4]]), Float64Array.from(H)), TypeError)            // matrix past the array
assert.throws(() => addon.plan(3, new Int32Array(15), new Float64Array(0)), TypeError)             // ragged tape
assert.throws(() => addon.plan(3, [0, 1], new Float64Array(0)), TypeError)                         // not a typed array
const p = addon.plan(12, tape([[0, [0], 0], [1, [0, 11], -1], [0, [5], 0]]), Float64Array.from(H))
assert.ok(p.passes === 1 && p.deviceOps >= 1)       // the three gates fit one pass
for (const seed of [42, 1, 0, 123456789, -5, 2 ** 32 + 7]) {
  const r = addon.xorshift(seed, 3000), js = makePrng(seed)
  for (let i = 0; i < r.length; i++) assert.strictEqual(r[i], js())
}
if (!gpu) {
  assert.throws(() => addon.create(3), (e) => e instanceof Error && !(e instanceof RangeError) && !(e instanceof TypeError))
  console.log('napi: error classes, plan and makePrng ok (no device)')
  process.exit(0)
}

// ---- on the device -----------------------------------------------------------
const s = addon.create(2)
addon.apply(s, tape([[0, [0], 0], [1, [0, 1], -1]]), Float64Array.from(H))          // h(0), cnot(0, 1)
const [idx, pr] = addon.probs(s)
assert.deepStrictEqual(Array.from(idx), [0n, 3n])
for (const v of pr) assert.ok(Math.abs(v - 0.5) < 1e-15)
const draws = Float64Array.from({ length: 1000 }, makePrng(7))
const got = addon.sample(s, draws)
draws.forEach((r, i) => assert.strictEqual(got[i], r <= pr[0] ? 0n : 3n))
const zz = addon.expect(s, BigUint64Array.from([0n]), BigUint64Array.from([3n]), Float64Array.from([1]), 0)
assert.ok(Math.abs(zz - 1) < 1e-15)
assert.throws(() => addon.apply(s, tape([[0, [4], 0]]), Float64Array.from(H)), RangeError)
addon.free(s)
assert.throws(() => addon.probs(s), Error)

// H on every qubit of 10, then |x> -> amplitudes all 2^-5, probabilities 2^-10
const n = 10, t = addon.create(n)
addon.initBasis(t, 0)
addon.apply(t, tape(Array.from({ length: n }, (_, q) => [0, [q], 0])), Float64Array.from(H))
const a = addon.amplitudes(t)
assert.strictEqual(a.length, 2 << n)
for (let i = 0; i < a.length; i += 2) assert.ok(Math.abs(a[i] - 1 / 32) < 1e-15 && a[i + 1] === 0)
assert.deepStrictEqual(addon.amplitude(t, 5n), [a[10], a[11]])
console.log('napi: device path ok')
