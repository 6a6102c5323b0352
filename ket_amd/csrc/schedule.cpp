// Host-side lowering and pass scheduling for libketsv.
//
// lower_tape() turns ket's FlatOp tape (src/circuit.ts:22-47, dispatched by
// simulatePure, src/circuit.ts:143-161) into logical ops of four classes:
// dense matrix on targets with controls, diagonal, controlled-X permutation
// and (controlled) swap.  schedule() then packs them into HBM passes: each
// pass holds a set of K "tile" qubits; a workgroup loads the 2^K amplitudes
// that share one value of the other n-K bits into LDS, applies every op of
// the pass there, and writes them back.  An op fits a pass when its
// non-diagonal targets are tile qubits; controls and diagonal qubits may lie
// anywhere (outside the tile they are constant per workgroup).
#include <algorithm>
#include <array>
#include <cstring>

#include "ksv_internal.h"

namespace ksv {

static bool mat_is_diag(const cd* m, int dim) {
    for (int r = 0; r < dim; r++)
        for (int c = 0; c < dim; c++)
            if (r != c && m[r * dim + c] != cd(0, 0)) return false;
    return true;
}

static bool mat_is_x(const cd* m) {
    return m[0] == cd(0, 0) && m[1] == cd(1, 0) && m[2] == cd(1, 0) && m[3] == cd(0, 0);
}

static uint64_t bits_of(const std::vector<int>& v) {
    uint64_t b = 0;
    for (int q : v) b |= 1ull << q;
    return b;
}

static void finish(LOp& o) {
    o.touch = bits_of(o.tg) | bits_of(o.ctl);
    if (o.cls == L_DIAG) o.need = 0;
    else o.need = bits_of(o.tg);
}

// Diagonal single-target op with controls; a pure |1..1> phase when d0 == 1.
static void push_diag1(std::vector<LOp>& out, int src, int t, std::vector<int> ctl, cd d0, cd d1) {
    LOp o;
    o.cls = L_DIAG;
    o.src = src;
    if (d0 == cd(1, 0)) {
        ctl.push_back(t);
        o.ctl = ctl;
        o.m = {d1};
    } else {
        o.tg = {t};
        o.ctl = ctl;
        o.m = {d0, d1};
    }
    finish(o);
    out.push_back(o);
}

int lower_tape(int n, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
               std::vector<LOp>& out, std::string& err) {
    out.clear();
    out.reserve(nops);
    auto getm = [&](const ksv_op& op, int dim, std::vector<cd>& m) -> bool {
        size_t need = (size_t)dim * dim * 2;
        if (op.mat < 0 || (size_t)op.mat + need > nmats || !mats) return false;
        m.resize((size_t)dim * dim);
        for (size_t i = 0; i < (size_t)dim * dim; i++) m[i] = cd(mats[op.mat + 2 * i], mats[op.mat + 2 * i + 1]);
        return true;
    };
    static const int expect_nq[] = {1, 2, 2, 2, 2, 3, 3, 3, -1, -1};
    for (size_t i = 0; i < nops; i++) {
        const ksv_op& op = ops[i];
        if (op.kind < 0 || op.kind > KSV_BARRIER) {
            err = "op " + std::to_string(i) + ": unknown kind " + std::to_string(op.kind);
            return KSV_ERR_TYPE;
        }
        if (op.nq < 0 || op.nq > KSV_MAX_OPERANDS ||
            (expect_nq[op.kind] >= 0 && op.nq != expect_nq[op.kind])) {
            err = "op " + std::to_string(i) + ": wrong operand count " + std::to_string(op.nq);
            return KSV_ERR_TYPE;
        }
        for (int j = 0; j < op.nq; j++)
            if (op.q[j] < 0 || op.q[j] >= n) {
                err = "qubit index " + std::to_string(op.q[j]) + " is out of range for a " +
                      std::to_string(n) + "-qubit circuit";
                return KSV_ERR_RANGE;
            }
        const int* q = op.q;
        std::vector<cd> m;
        int src = (int)i;
        switch (op.kind) {
        case KSV_SINGLE: {
            if (!getm(op, 2, m)) { err = "op " + std::to_string(i) + ": bad matrix offset"; return KSV_ERR_TYPE; }
            if (mat_is_diag(m.data(), 2)) { push_diag1(out, src, q[0], {}, m[0], m[3]); break; }
            LOp o;
            o.src = src;
            o.tg = {q[0]};
            if (mat_is_x(m.data())) o.cls = L_PERMX;
            else { o.cls = L_MAT; o.m = m; }
            finish(o);
            out.push_back(o);
            break;
        }
        case KSV_CNOT: {
            if (q[0] == q[1]) { err = "control and target qubits must differ (got " + std::to_string(q[0]) + ")"; return KSV_ERR_TYPE; }
            LOp o;
            o.src = src;
            o.cls = L_PERMX;
            o.tg = {q[1]};
            o.ctl = {q[0]};
            finish(o);
            out.push_back(o);
            break;
        }
        case KSV_SWAP: {
            if (q[0] == q[1]) break;      // applySWAP with a == b is the identity
            LOp o;
            o.src = src;
            o.cls = L_SWAP;
            o.tg = {q[0], q[1]};
            finish(o);
            out.push_back(o);
            break;
        }
        case KSV_TWO: {
            if (q[0] == q[1]) { err = "two-qubit gate needs distinct qubits (got " + std::to_string(q[0]) + ")"; return KSV_ERR_TYPE; }
            if (!getm(op, 4, m)) { err = "op " + std::to_string(i) + ": bad matrix offset"; return KSV_ERR_TYPE; }
            LOp o;
            o.src = src;
            o.tg = {q[0], q[1]};
            if (mat_is_diag(m.data(), 4)) {
                o.cls = L_DIAG;
                o.m = {m[0], m[5], m[10], m[15]};
            } else {
                o.cls = L_MAT;
                o.m = m;
            }
            finish(o);
            out.push_back(o);
            break;
        }
        case KSV_CONTROLLED: {
            if (q[0] == q[1]) { err = "control and target qubits must differ (got " + std::to_string(q[0]) + ")"; return KSV_ERR_TYPE; }
            if (!getm(op, 2, m)) { err = "op " + std::to_string(i) + ": bad matrix offset"; return KSV_ERR_TYPE; }
            if (mat_is_diag(m.data(), 2)) { push_diag1(out, src, q[1], {q[0]}, m[0], m[3]); break; }
            LOp o;
            o.src = src;
            o.tg = {q[1]};
            o.ctl = {q[0]};
            if (mat_is_x(m.data())) o.cls = L_PERMX;
            else { o.cls = L_MAT; o.m = m; }
            finish(o);
            out.push_back(o);
            break;
        }
        case KSV_TOFFOLI: {
            if (q[2] == q[0] || q[2] == q[1]) { err = "ccx target must differ from its controls"; return KSV_ERR_TYPE; }
            LOp o;
            o.src = src;
            o.cls = L_PERMX;
            o.tg = {q[2]};
            o.ctl = {q[0]};
            if (q[1] != q[0]) o.ctl.push_back(q[1]);
            finish(o);
            out.push_back(o);
            break;
        }
        case KSV_CSWAP: {
            if (q[1] == q[2]) break;      // swapping a qubit with itself is the identity
            if (q[0] == q[1] || q[0] == q[2]) { err = "cswap control must differ from the swapped qubits"; return KSV_ERR_TYPE; }
            LOp o;
            o.src = src;
            o.cls = L_SWAP;
            o.tg = {q[1], q[2]};
            o.ctl = {q[0]};
            finish(o);
            out.push_back(o);
            break;
        }
        case KSV_CSRSWAP: {
            if (q[0] == q[1] || q[0] == q[2] || q[1] == q[2]) { err = "csrswap needs three distinct qubits"; return KSV_ERR_TYPE; }
            // src/statevector.ts:177-207: |c01>,|c10> mixed by [[1, i],[i, 1]]/sqrt(2)
            const double s = 1 / std::sqrt(2.0);
            LOp o;
            o.src = src;
            o.cls = L_MAT;
            o.tg = {q[1], q[2]};
            o.ctl = {q[0]};
            o.m.assign(16, cd(0, 0));
            o.m[0] = 1;
            o.m[5] = cd(s, 0);
            o.m[6] = cd(0, s);
            o.m[9] = cd(0, s);
            o.m[10] = cd(s, 0);
            o.m[15] = 1;
            finish(o);
            out.push_back(o);
            break;
        }
        case KSV_UNITARY: {
            int k = op.nq;
            if (k < 1) { err = "unitary: qubits must be non-empty"; return KSV_ERR_TYPE; }
            for (int a = 0; a < k; a++)
                for (int b = a + 1; b < k; b++)
                    if (q[a] == q[b]) { err = "unitary: qubits must be distinct"; return KSV_ERR_TYPE; }
            int dim = 1 << k;
            if (!getm(op, dim, m)) { err = "op " + std::to_string(i) + ": bad matrix offset"; return KSV_ERR_TYPE; }
            LOp o;
            o.src = src;
            o.tg.assign(q, q + k);
            if (k <= DIAG_MAX && mat_is_diag(m.data(), dim)) {
                o.cls = L_DIAG;
                o.m.resize(dim);
                for (int r = 0; r < dim; r++) o.m[r] = m[(size_t)r * dim + r];
            } else {
                o.cls = L_MAT;
                o.m = m;
                o.big = k > MK_MAX;
            }
            finish(o);
            out.push_back(o);
            break;
        }
        case KSV_BARRIER:
            break;
        }
    }
    return KSV_OK;
}

// --------------------------------------------------------------------------
// Exact-mode fusion: a run of uncontrolled one-qubit gates on the same qubit
// (nothing else touching that qubit in between) becomes one 2x2 product.
// Forbidden under ket's per-gate pruning (each gate prunes its outputs).
// --------------------------------------------------------------------------
static void to_dense2(const LOp& o, cd* m) {
    if (o.cls == L_MAT) { for (int i = 0; i < 4; i++) m[i] = o.m[i]; return; }
    if (o.cls == L_PERMX) { m[0] = 0; m[1] = 1; m[2] = 1; m[3] = 0; return; }
    // diagonal (either {d0,d1} on tg, or a |1> phase {d1} as a control)
    if (o.tg.size() == 1) { m[0] = o.m[0]; m[1] = 0; m[2] = 0; m[3] = o.m[1]; }
    else { m[0] = 1; m[1] = 0; m[2] = 0; m[3] = o.m[0]; }
}

static bool is_plain_1q(const LOp& o, int& q) {
    if (o.big) return false;
    if (o.cls == L_MAT && o.tg.size() == 1 && o.ctl.empty()) { q = o.tg[0]; return true; }
    if (o.cls == L_PERMX && o.ctl.empty()) { q = o.tg[0]; return true; }
    if (o.cls == L_DIAG && o.tg.size() == 1 && o.ctl.empty()) { q = o.tg[0]; return true; }
    if (o.cls == L_DIAG && o.tg.empty() && o.ctl.size() == 1) { q = o.ctl[0]; return true; }
    return false;
}

static void fuse_1q(int n, std::vector<LOp>& lops) {
    std::vector<int> last(n, -1);     // index (in out) of the last op touching qubit q
    std::vector<LOp> out;
    out.reserve(lops.size());
    for (auto& o : lops) {
        int q;
        if (is_plain_1q(o, q) && last[q] >= 0) {
            LOp& p = out[last[q]];
            int pq;
            if (is_plain_1q(p, pq) && pq == q) {
                cd a[4], b[4], r[4];
                to_dense2(p, a);
                to_dense2(o, b);
                r[0] = b[0] * a[0] + b[1] * a[2];
                r[1] = b[0] * a[1] + b[1] * a[3];
                r[2] = b[2] * a[0] + b[3] * a[2];
                r[3] = b[2] * a[1] + b[3] * a[3];
                LOp f;
                f.src = o.src;
                if (r[1] == cd(0, 0) && r[2] == cd(0, 0)) {
                    f.cls = L_DIAG;
                    f.tg = {q};
                    f.m = {r[0], r[3]};
                } else {
                    f.cls = L_MAT;
                    f.tg = {q};
                    f.m.assign(r, r + 4);
                }
                finish(f);
                p = f;
                continue;
            }
        }
        out.push_back(o);
        for (int b = 0; b < n; b++)
            if (o.touch >> b & 1) last[b] = (int)out.size() - 1;
    }
    lops.swap(out);
}

// --------------------------------------------------------------------------
// Pass packing.
// --------------------------------------------------------------------------
static void emit_dop(const LOp& o, const int* pos, Program& prog) {
    DOp d;
    std::memset(&d, 0, sizeof d);
    int ins[16], nins = 0;
    uint32_t lctl = 0;
    uint64_t gctl = 0;
    for (int c : o.ctl) {
        if (pos[c] >= 0) { lctl |= 1u << pos[c]; ins[nins++] = pos[c]; }
        else gctl |= 1ull << c;
    }
    d.k = (int)o.tg.size();
    switch (o.cls) {
    case L_MAT:
        d.kind = d.k == 1 ? DK_M1 : (d.k == 2 ? DK_M2 : DK_MK);
        for (int i = 0; i < d.k; i++) { d.t[i] = (int8_t)pos[o.tg[i]]; ins[nins++] = pos[o.tg[i]]; }
        break;
    case L_DIAG:
        d.kind = DK_DIAG;
        for (int i = 0; i < d.k; i++) {
            int q = o.tg[i];
            d.t[i] = pos[q] >= 0 ? (int8_t)pos[q] : (int8_t)(-(q + 1));
        }
        break;
    case L_PERMX:
        d.kind = DK_PERMX;
        d.t[0] = (int8_t)pos[o.tg[0]];
        ins[nins++] = pos[o.tg[0]];
        break;
    case L_SWAP:
        d.kind = DK_SWAP;
        d.t[0] = (int8_t)pos[o.tg[0]];
        d.t[1] = (int8_t)pos[o.tg[1]];
        ins[nins++] = pos[o.tg[0]];
        ins[nins++] = pos[o.tg[1]];
        break;
    }
    std::sort(ins, ins + nins);
    d.nins = nins;
    for (int i = 0; i < nins; i++) d.ins[i] = (int8_t)ins[i];
    d.lctl = lctl;
    d.gctl = gctl;
    d.cond = (uint32_t)o.cond;
    d.mat = (int32_t)prog.mats.size();
    prog.mats.insert(prog.mats.end(), o.m.begin(), o.m.end());
    prog.dops.push_back(d);
}

static int popc(uint64_t x) { return __builtin_popcountll(x); }

// Choose the ops of one pass: scan the remaining ops in order and take every
// op whose qubits no earlier skipped op touches (so program order between
// dependent ops is kept) and whose must-be-local qubits fit the tile.
// With `strict` (ket's per-gate pruning) the pass ends at the first op that
// does not fit: pruning after every gate makes even commuting gates
// order-sensitive, so no op may move ahead of another.
// Returns the tile mask (forced bits included, not yet padded).
template <typename Allowed>
static uint64_t pick_pass(const std::vector<LOp>& lops, std::vector<int>& rem, int K, uint64_t forced,
                          uint64_t all, Allowed allowed, std::vector<int>& chosen, std::vector<int>& rest,
                          bool strict) {
    uint64_t T = forced, blocked = 0;
    chosen.clear();
    rest.clear();
    size_t i = 0;
    for (; i < rem.size(); i++) {
        const LOp& o = lops[rem[i]];
        if (!allowed(o) || (o.touch & blocked)) { blocked |= o.touch; rest.push_back(rem[i]); }
        else {
            uint64_t nt = T | o.need;
            if (popc(nt) <= K) { T = nt; chosen.push_back(rem[i]); }
            else { blocked |= o.touch; rest.push_back(rem[i]); }
        }
        if (strict && blocked) blocked = all;
        if ((blocked & all) == all) { i++; break; }
    }
    for (; i < rem.size(); i++) rest.push_back(rem[i]);
    return T;
}

static void emit_big(const LOp& o, int id, Program& prog) {
    if (id >= 0 && (size_t)id < prog.mat_of_lop.size()) prog.mat_of_lop[id] = (int64_t)prog.mats.size();
    Step s;
    s.big = true;
    s.big_op = (int)prog.big_ops.size();
    BigOp b;
    b.tg = o.tg;
    b.mat = (int)prog.mats.size();
    prog.mats.insert(prog.mats.end(), o.m.begin(), o.m.end());
    prog.big_ops.push_back(b);
    prog.steps.push_back(s);
}

// An LDS pass over tile T holding the ops `idx`.
static void emit_lds_pass(int n, int K, uint64_t T, const std::vector<LOp>& lops, const std::vector<int>& idx,
                          Program& prog) {
    for (int q = 0; q < n && popc(T) < K; q++) T |= 1ull << q;
    Step s;
    s.k = K;
    int pos[64];
    for (int q = 0; q < 64; q++) pos[q] = -1;
    int j = 0;
    for (int q = 0; q < n; q++)
        if (T >> q & 1) { s.tq[j] = (int8_t)q; pos[q] = j; j++; }
    s.op_begin = (int)prog.dops.size();
    for (int id : idx) {
        if ((size_t)id < prog.mat_of_lop.size() && lops[id].cls != L_PERMX && lops[id].cls != L_SWAP)
            prog.mat_of_lop[id] = (int64_t)prog.mats.size();
        emit_dop(lops[id], pos, prog);
    }
    s.op_count = (int)prog.dops.size() - s.op_begin;
    prog.steps.push_back(s);
}

static void schedule_lds(int n, int tile, std::vector<LOp>& lops, Program& prog) {
    const bool strict = prog.prune != 0;
    const int K = std::min(n, tile);
    const int c = std::min(4, K);
    const uint64_t all = n >= 64 ? ~0ull : ((1ull << n) - 1);
    const uint64_t forced = (1ull << c) - 1;
    std::vector<int> rem(lops.size());
    for (size_t i = 0; i < lops.size(); i++) rem[i] = (int)i;
    std::vector<int> chosen, rest;
    while (!rem.empty()) {
        const LOp& first = lops[rem[0]];
        if (first.big) {
            emit_big(first, rem[0], prog);
            rem.erase(rem.begin());
            continue;
        }
        uint64_t T = pick_pass(lops, rem, K, forced, all, [](const LOp& o) { return !o.big; }, chosen, rest,
                               strict);
        emit_lds_pass(n, K, T, lops, chosen, prog);
        rem.swap(rest);
    }
}

int plan_copies(Program& p, int ntot, int nsys, int prec, const ksv_op* ops, size_t nops, const double* mats,
                size_t nmats, const std::vector<CopyPauli>& paulis, int prune, std::string& err) {
    if (nsys < 1 || nsys > KMAX_TILE || nsys > ntot) {
        err = "a copy must fit one LDS tile";
        return KSV_ERR_RANGE;
    }
    std::vector<LOp> gl;
    int rc = lower_tape(nsys, ops, nops, mats, nmats, gl, err);
    if (rc) return rc;
    // the Pauli events in order, each before the first lowered op of tape op `at`
    std::vector<LOp> lops;
    lops.reserve(gl.size() + paulis.size());
    size_t e = 0;
    auto flush = [&](size_t upto) {
        for (; e < paulis.size() && paulis[e].at <= upto; e++) {
            const CopyPauli& c = paulis[e];
            LOp o;
            o.tg = {c.q};
            if (c.pauli == 1) {
                o.cls = L_PERMX;
            } else if (c.pauli == 3) {
                o.cls = L_DIAG;
                o.m = {cd(1, 0), cd(-1, 0)};
            } else {
                o.cls = L_MAT;
                o.m = {cd(0, 0), cd(0, -1), cd(0, 1), cd(0, 0)};
            }
            finish(o);
            o.cond = (int)p.bits.size() + 1;
            p.bits.insert(p.bits.end(), c.words.begin(), c.words.end());
            lops.push_back(o);
        }
    };
    for (LOp& o : gl) {
        flush((size_t)o.src);
        lops.push_back(std::move(o));
        lops.back().src = -1;
    }
    flush(nops);
    p.n = ntot;
    p.nsys = nsys;
    p.precision = prec;
    p.prune = prune;
    p.tile = nsys;
    p.tape_ops = nops;
    // every op lies inside the tile of one copy: in order, one pass (plus a
    // step of its own for each unitary too wide for a pass) of register rounds
    // for copies of 11..13 qubits, an LDS pass below
    if (nsys >= 11 && nsys <= 13) {
        for (const LOp& o : lops)
            if (o.cond && !(o.cls == L_PERMX || o.cls == L_MAT || (o.cls == L_DIAG && o.tg.size() <= 2))) {
                err = "copy-conditional op of an unexpected kind";
                return KSV_ERR_TYPE;
            }
        schedule_rounds(ntot, nsys, false, lops, p);
        const size_t nw = ((size_t(1) << (ntot - nsys)) + 31) / 32, nr = p.rops.size();
        bool any = false;
        for (const ROp& r : p.rops) any = any || (r.flags & ROP_COND);
        if (any) {
            p.rbits.assign(nw * nr, ~0u);       // unconditional: every copy
            for (size_t g = 0; g < nr; g++)
                if (p.rops[g].flags & ROP_COND)
                    for (size_t w = 0; w < nw; w++) p.rbits[w * nr + g] = p.bits[(size_t)p.rops[g].mat + w];
        }
    } else {
        schedule_lds(ntot, nsys, lops, p);
    }
    return KSV_OK;
}

void schedule(int n, int tile, bool fuse, std::vector<LOp>& lops, Program& prog) {
    if (fuse) fuse_1q(n, lops);
    schedule_lds(n, tile, lops, prog);
}

// --------------------------------------------------------------------------
// Register-round passes.
//
// A pass over K = TB + 4 tile qubits is split into rounds.  In each round
// four tile qubits are register slots and TB are thread bits; every op of
// the round needs its non-diagonal targets in slots.  Rounds are chosen
// greedily: scan the remaining ops in order, take each op whose slot needs
// still fit in four slots and that no skipped op blocks.  Between rounds an
// LDS transpose moves the tile into the next assignment.  Tile bits 0..2
// (physical qubits 0..2) are thread bits 0..2 at load and store, so every 8
// lanes move one 128-byte run of HBM.
// --------------------------------------------------------------------------
static bool round_ok(const LOp& o) {
    if (o.big) return false;
    if (o.cls == L_MAT) return o.tg.size() <= 2 || (o.tg.size() <= (size_t)RSLOTS && o.ctl.empty());
    return true;
}

struct RLayout {
    int slot[RSLOTS];   // tile bit of slot s
    int thr[12];        // tile bit of thread bit i
};

static RLayout make_layout(int K, uint32_t S) {
    RLayout L;
    int ns = 0, nt = 0;
    for (int p = 0; p < K; p++) {
        if (S >> p & 1) L.slot[ns++] = p;
        else L.thr[nt++] = p;
    }
    for (; nt < 12; nt++) L.thr[nt] = 0;
    return L;
}

// Add the highest tile bits above 2 until S has RSLOTS members.
static uint32_t fill_slots(int K, uint32_t S) {
    for (int p = K - 1; p >= 3 && popc(S) < RSLOTS; p--) S |= 1u << p;
    return S;
}

// Rank over GF(2) of the first m (<= 4) words of v (v is a copy).
static int gf2_rank(std::array<uint32_t, 4> v, int m) {
    int r = 0;
    for (int bit = 31; bit >= 0; bit--) {
        int piv = -1;
        for (int i = r; i < m; i++)
            if (v[i] >> bit & 1) { piv = i; break; }
        if (piv < 0) continue;
        std::swap(v[r], v[piv]);
        for (int i = 0; i < m; i++)
            if (i != r && (v[i] >> bit & 1)) v[i] ^= v[r];
        r++;
    }
    return r;
}

// LDS swizzle of one exchange: address(x) = x ^ XOR_{p >= 4, x_p = 1} v[p]
// (v[p] < 16).  Chosen so that lanes 0..3 of both the writing and the reading
// layout reach four distinct 16-byte columns of every 256-byte LDS row
// (and lanes 0..2 distinct ones of every 128-byte half), i.e. conflict-free
// 16-byte accesses on both sides.
static void emit_exch(int TB, const RLayout& W, const RLayout& R, Program& prog) {
    uint32_t v[16] = {0};
    auto addr = [&](int p) -> uint32_t { return (1u << p) ^ (p >= 4 ? v[p] : 0u); };
    auto good = [&](const RLayout& L) {
        std::array<uint32_t, 4> a3{}, a4{};
        int m3 = 0, m4 = 0;
        for (int i = 0; i < 4 && i < TB; i++) {
            uint32_t a = addr(L.thr[i]);
            if (i < 3) a3[m3++] = a & 7;
            a4[m4++] = a & 15;
        }
        return gf2_rank(a3, m3) == m3 && gf2_rank(a4, m4) == m4;
    };
    std::vector<int> free_p;
    for (int i = 0; i < 4 && i < TB; i++) {
        if (W.thr[i] >= 4 && std::find(free_p.begin(), free_p.end(), W.thr[i]) == free_p.end()) free_p.push_back(W.thr[i]);
        if (R.thr[i] >= 4 && std::find(free_p.begin(), free_p.end(), R.thr[i]) == free_p.end()) free_p.push_back(R.thr[i]);
    }
    // The search cannot succeed when a layout has tile bit 3 among its first three
    // thread bits (its address has no low three bits, whatever the swizzle), or
    // when no thread bit can be swizzled; it then ends with v = 0, as below.
    auto possible = [&](const RLayout& L) {
        for (int i = 0; i < 3 && i < TB; i++)
            if (L.thr[i] == 3) return false;
        return true;
    };
    const bool search = !free_p.empty() && possible(W) && possible(R);
    uint32_t rng = 0x9E3779B9u;
    for (int attempt = 0; search && attempt < 4096 && !(good(W) && good(R)); attempt++) {
        for (int p : free_p) {
            rng ^= rng << 13;
            rng ^= rng >> 17;
            rng ^= rng << 5;
            v[p] = rng & 15;
        }
    }
    if (!(good(W) && good(R))) for (int p : free_p) v[p] = 0;   // unreachable in practice; only costs LDS cycles
    XRec x;
    std::memset(&x, 0, sizeof x);
    for (int i = 0; i < TB; i++) { x.wt[i] = (uint16_t)addr(W.thr[i]); x.rt[i] = (uint16_t)addr(R.thr[i]); }
    for (int s = 0; s < RSLOTS; s++) { x.wr[s] = (uint16_t)addr(W.slot[s]); x.rr[s] = (uint16_t)addr(R.slot[s]); }
    ROp r;
    std::memset(&r, 0, sizeof r);
    r.kind = R_EXCH;
    std::memcpy(r.pay, &x, sizeof x);
    prog.rops.push_back(r);
}

// Up to four complex entries in the program's precision, inline in the record.
static void put_inline(ROp& r, const cd* v, int count, int prec) {
    if (prec == KSV_F64) {
        double* d = reinterpret_cast<double*>(r.pay);
        for (int i = 0; i < count; i++) { d[2 * i] = v[i].real(); d[2 * i + 1] = v[i].imag(); }
    } else {
        float* f = reinterpret_cast<float*>(r.pay);
        for (int i = 0; i < count; i++) { f[2 * i] = (float)v[i].real(); f[2 * i + 1] = (float)v[i].imag(); }
    }
}

// Structure of a 2x2 matrix for the register-round kernel (ROP_REAL / ROP_RXF).
static uint16_t m1_form(const cd* m) {
    auto real = [](const cd& v) { return v.imag() == 0.0; };
    auto imag = [](const cd& v) { return v.real() == 0.0; };
    if (real(m[0]) && real(m[1]) && real(m[2]) && real(m[3])) return ROP_REAL;
    if (real(m[0]) && imag(m[1]) && imag(m[2]) && real(m[3])) return ROP_RXF;
    return 0;
}

// One op in the slot / thread-bit / global form of layout L.
static void emit_rop(const LOp& o, const RLayout& L, const int* pos, Program& prog) {
    auto slot_of = [&](int q) {
        for (int s = 0; s < RSLOTS; s++)
            if (L.slot[s] == pos[q]) return s;
        return -1;
    };
    auto thr_of = [&](int q) {
        for (int i = 0; i < 12; i++)
            if (L.thr[i] == pos[q]) return i;
        return -1;
    };
    ROp r;
    std::memset(&r, 0, sizeof r);
    for (int c : o.ctl) {
        if (pos[c] < 0) r.gctl |= 1ull << c;
        else if (slot_of(c) >= 0) r.rctl |= 1u << slot_of(c);
        else r.tctl |= 1u << thr_of(c);
    }
    r.mat = (int32_t)prog.mats.size();
    switch (o.cls) {
    case L_MAT: {
        const int k = (int)o.tg.size();
        if (k == 1) {
            r.kind = R_M1;
            r.a = (uint8_t)slot_of(o.tg[0]);
            put_inline(r, o.m.data(), 4, prog.precision);
            r.flags = m1_form(o.m.data());
        } else {
            // dense 2^k x 2^k (k = 2..4) on slots loc[0] (MSB) .. loc[k-1]
            r.kind = R_DENSE;
            r.k = (uint8_t)k;
            for (int i = 0; i < k; i++) r.loc[i] = (uint8_t)slot_of(o.tg[i]);
            prog.mats.insert(prog.mats.end(), o.m.begin(), o.m.end());
        }
        break;
    }
    case L_DIAG:
        r.kind = R_DIAG;
        r.k = (uint8_t)o.tg.size();
        for (size_t i = 0; i < o.tg.size(); i++) {
            int q = o.tg[i];
            if (pos[q] < 0) r.loc[i] = (uint8_t)(32 + q);
            else if (slot_of(q) >= 0) r.loc[i] = (uint8_t)(16 + slot_of(q));
            else r.loc[i] = (uint8_t)thr_of(q);
        }
        if (o.tg.size() <= 2) put_inline(r, o.m.data(), (int)o.m.size(), prog.precision);
        else prog.mats.insert(prog.mats.end(), o.m.begin(), o.m.end());
        break;
    case L_PERMX:
        r.kind = R_PERMX;
        r.a = (uint8_t)slot_of(o.tg[0]);
        break;
    case L_SWAP: {
        r.kind = R_SWAP;
        int sa = slot_of(o.tg[0]), sb = slot_of(o.tg[1]);
        r.a = (uint8_t)std::max(sa, sb);
        r.b = (uint8_t)std::min(sa, sb);
        break;
    }
    }
    if (o.cond) {   // copy-conditional (Pauli events): kinds whose matrix is inline
        r.flags |= ROP_COND;
        r.mat = o.cond - 1;
    }
    prog.rops.push_back(r);
}

static void round_pass(int n, int K, uint64_t T, const std::vector<LOp>& lops, const std::vector<int>& idx,
                       bool strict, Program& prog) {
    const int TB = K - RSLOTS;
    for (int q = 0; q < n && popc(T) < K; q++) T |= 1ull << q;
    Step st;
    st.reg = true;
    st.k = K;
    int pos[64];
    for (int q = 0; q < 64; q++) pos[q] = -1;
    int j = 0;
    for (int q = 0; q < n; q++)
        if (T >> q & 1) { st.tq[j] = (int8_t)q; pos[q] = j; j++; }
    const uint32_t low = 7;   // tile bits 0..2 = physical qubits 0..2

    std::vector<uint32_t> sets;
    std::vector<std::vector<int>> takes;
    std::vector<int> rem = idx, rest, take;
    bool first = true;
    while (!rem.empty()) {
        uint32_t S = 0;
        uint64_t blocked = 0;
        take.clear();
        rest.clear();
        for (size_t i = 0; i < rem.size(); i++) {
            const LOp& o = lops[rem[i]];
            bool ok = !(o.touch & blocked);
            uint32_t need = 0;
            if (ok && o.cls != L_DIAG)
                for (int q : o.tg) need |= 1u << pos[q];
            if (ok && first && (need & low)) ok = false;
            if (ok && popc(S | need) > RSLOTS) ok = false;
            if (ok) {
                S |= need;
                take.push_back(rem[i]);
            } else {
                blocked |= o.touch;
                rest.push_back(rem[i]);
                if (strict) {
                    rest.insert(rest.end(), rem.begin() + i + 1, rem.end());
                    break;
                }
            }
        }
        sets.push_back(fill_slots(K, S));
        takes.push_back(take);
        rem.swap(rest);
        first = false;
    }
    if (sets.empty()) { sets.push_back(fill_slots(K, 0)); takes.emplace_back(); }
    if (sets.back() & low) { sets.push_back(fill_slots(K, sets.back() & ~low)); takes.emplace_back(); }

    const RLayout L0 = make_layout(K, sets.front()), LN = make_layout(K, sets.back());
    for (int i = 0; i < TB; i++) { st.ld_t[i] = st.tq[L0.thr[i]]; st.st_t[i] = st.tq[LN.thr[i]]; }
    for (int s = 0; s < RSLOTS; s++) { st.ld_r[s] = st.tq[L0.slot[s]]; st.st_r[s] = st.tq[LN.slot[s]]; }
    st.op_begin = (int)prog.rops.size();
    for (size_t r = 0; r < sets.size(); r++) {
        const RLayout L = make_layout(K, sets[r]);
        if (r > 0) {
            emit_exch(TB, make_layout(K, sets[r - 1]), L, prog);
            st.exchanges++;
        }
        for (int id : takes[r]) {
            if ((size_t)id < prog.rop_of_lop.size()) prog.rop_of_lop[id] = (int)prog.rops.size();
            emit_rop(lops[id], L, pos, prog);
        }
    }
    st.op_count = (int)prog.rops.size() - st.op_begin;
    prog.steps.push_back(st);
}

void schedule_rounds(int n, int rtile, bool fuse, std::vector<LOp>& lops, Program& prog) {
    if (fuse) fuse_1q(n, lops);
    const int K = rtile;
    if (n < K) {
        schedule_lds(n, prog.tile, lops, prog);
        return;
    }
    prog.rtile = K;
    const bool strict = prog.prune != 0;
    const uint64_t all = n >= 64 ? ~0ull : ((1ull << n) - 1);
    const uint64_t forced = 7;
    const int KL = std::min(n, prog.tile);
    std::vector<int> rem(lops.size());
    for (size_t i = 0; i < lops.size(); i++) rem[i] = (int)i;
    std::vector<int> chosen, rest;
    while (!rem.empty()) {
        const LOp& first = lops[rem[0]];
        if (first.big) {
            emit_big(first, rem[0], prog);
            rem.erase(rem.begin());
            continue;
        }
        if (!round_ok(first)) {
            // a controlled dense 3- or 4-qubit matrix: one LDS pass of its own
            uint64_t T = ((1ull << std::min(4, KL)) - 1) | first.need;
            emit_lds_pass(n, KL, T, lops, {rem[0]}, prog);
            rem.erase(rem.begin());
            continue;
        }
        uint64_t T = pick_pass(lops, rem, K, forced, all, round_ok, chosen, rest, strict);
        round_pass(n, K, T, lops, chosen, strict, prog);
        rem.swap(rest);
    }
}

// --------------------------------------------------------------------------
// Fused-block passes (exact mode).
//
// Fusion: every op joins the best existing block it may join, else opens a
// new one.  With b the latest block touching any qubit of the op, the op
// may join any block c >= b whose qubit set stays within bq qubits: blocks
// after b touch none of its qubits, so moving the op to block c keeps every
// dependence.  Blocks are then packed into passes over K-qubit tiles.
// --------------------------------------------------------------------------
struct FBlock {
    std::vector<int> q;       // row bit i = qubit q[i]
    std::vector<cd> u;        // 2^|q| x 2^|q|, row-major
    int lop = -1;             // >= 0: a standalone op too wide for a block
    uint64_t touch = 0;
    std::vector<int> ops;     // lowered ops absorbed, in application order
    bool diag = true;         // every absorbed op is diagonal (by kind, whatever its angles)
    bool mono = true;         // every absorbed op is diagonal or a permutation
};

// Apply a logical op to a small dense vector over qubits Q (bit i = Q[i]).
static void apply_small(const LOp& o, const std::vector<int>& Q, cd* v) {
    auto bit = [&](int q) {
        for (size_t i = 0; i < Q.size(); i++)
            if (Q[i] == q) return (int)i;
        return -1;
    };
    const uint32_t dim = 1u << Q.size();
    uint32_t cm = 0;
    for (int c : o.ctl) cm |= 1u << bit(c);
    switch (o.cls) {
    case L_MAT: {
        const int k = (int)o.tg.size(), kd = 1 << k;
        uint32_t tm = 0;
        int tb[16];
        for (int i = 0; i < k; i++) { tb[i] = bit(o.tg[i]); tm |= 1u << tb[i]; }
        std::vector<cd> x(kd), y(kd);
        std::vector<uint32_t> id(kd);
        for (uint32_t b0 = 0; b0 < dim; b0++) {
            if ((b0 & tm) || (b0 & cm) != cm) continue;
            for (int r = 0; r < kd; r++) {
                uint32_t g = b0;
                for (int i = 0; i < k; i++)
                    if ((r >> (k - 1 - i)) & 1) g |= 1u << tb[i];
                id[r] = g;
                x[r] = v[g];
            }
            for (int r = 0; r < kd; r++) {
                cd acc = 0;
                for (int c = 0; c < kd; c++) acc += o.m[(size_t)r * kd + c] * x[c];
                y[r] = acc;
            }
            for (int r = 0; r < kd; r++) v[id[r]] = y[r];
        }
        break;
    }
    case L_DIAG: {
        const int k = (int)o.tg.size();
        for (uint32_t i = 0; i < dim; i++) {
            if ((i & cm) != cm) continue;
            uint32_t row = 0;
            for (int t = 0; t < k; t++) row |= ((i >> bit(o.tg[t])) & 1) << (k - 1 - t);
            v[i] *= o.m[row];
        }
        break;
    }
    case L_PERMX: {
        const int t = bit(o.tg[0]);
        for (uint32_t i = 0; i < dim; i++)
            if (!((i >> t) & 1) && (i & cm) == cm) std::swap(v[i], v[i | (1u << t)]);
        break;
    }
    case L_SWAP: {
        const int a = bit(o.tg[0]), b = bit(o.tg[1]);
        for (uint32_t i = 0; i < dim; i++)
            if (((i >> a) & 1) && !((i >> b) & 1) && (i & cm) == cm) std::swap(v[i], v[i ^ (1u << a) ^ (1u << b)]);
        break;
    }
    }
}

// Grow block f to also cover qubits `add` (appended as higher row bits).
static void grow_block(FBlock& f, const std::vector<int>& add) {
    if (add.empty()) return;
    const uint32_t d0 = 1u << f.q.size();
    for (int q : add) f.q.push_back(q);
    const uint32_t d1 = 1u << f.q.size();
    std::vector<cd> u((size_t)d1 * d1, cd(0, 0));
    for (uint32_t r = 0; r < d1; r++)
        for (uint32_t c = 0; c < d1; c++)
            if ((r / d0) == (c / d0)) u[(size_t)r * d1 + c] = f.u[(size_t)(r % d0) * d0 + (c % d0)];
    f.u.swap(u);
    for (int q : add) f.touch |= 1ull << q;
}

static void absorb(FBlock& f, const LOp& o) {
    const uint32_t d = 1u << f.q.size();
    std::vector<cd> col(d);
    for (uint32_t c = 0; c < d; c++) {
        for (uint32_t r = 0; r < d; r++) col[r] = f.u[(size_t)r * d + c];
        apply_small(o, f.q, col.data());
        for (uint32_t r = 0; r < d; r++) f.u[(size_t)r * d + c] = col[r];
    }
}

// The dense matrix of a fused block from its ops: qubits join the block in
// the order the ops first touch them (ascending within an op), row bit i =
// qubit q[i].  The packings below record only which ops each block takes;
// matrices are built for the chosen packing alone.
static void block_matrix(int n, FBlock& f, const std::vector<LOp>& lops) {
    f.q.clear();
    f.u.assign(1, cd(1, 0));
    uint64_t have = 0;
    for (int i : f.ops) {
        const LOp& o = lops[i];
        std::vector<int> add;
        for (int q = 0; q < n; q++)
            if ((o.touch >> q & 1) && !(have >> q & 1)) add.push_back(q);
        grow_block(f, add);
        have = f.touch;
        absorb(f, o);
    }
}

static void fuse_blocks(int n, int bq, const std::vector<LOp>& lops, std::vector<FBlock>& blocks) {
    std::vector<int> last(n, -1);
    const int window = 96;       // candidate blocks searched per op
    for (size_t i = 0; i < lops.size(); i++) {
        const LOp& o = lops[i];
        std::vector<int> S;
        for (int q = 0; q < n; q++)
            if (o.touch >> q & 1) S.push_back(q);
        int b = -1;
        for (int q : S) b = std::max(b, last[q]);
        if (o.big || (int)S.size() > bq) {
            FBlock f;
            f.lop = (int)i;
            f.q = S;
            f.touch = o.touch;
            blocks.push_back(f);
            for (int q : S) last[q] = (int)blocks.size() - 1;
            continue;
        }
        int best = -1, bestov = -1;
        const int from = std::max(std::max(b, 0), (int)blocks.size() - window);
        for (int c = from; c < (int)blocks.size(); c++) {
            const FBlock& f = blocks[c];
            if (f.lop >= 0) continue;
            const int uni = popc(f.touch | o.touch);
            if (uni > bq) continue;
            const int ov = popc(f.touch & o.touch);
            if (ov >= bestov) { best = c; bestov = ov; }
        }
        if (best < 0) {
            FBlock f;
            f.u.assign(1, cd(1, 0));
            blocks.push_back(f);
            best = (int)blocks.size() - 1;
        }
        FBlock& f = blocks[best];
        f.touch |= o.touch;
        f.ops.push_back((int)i);
        f.diag = f.diag && o.cls == L_DIAG;
        f.mono = f.mono && (o.cls == L_DIAG || o.cls == L_PERMX || o.cls == L_SWAP);
        for (int q : S) last[q] = best;
    }
}

// Front-driven fusion.  Repeatedly take the first op not yet placed and grow
// its qubit set to bq qubits, each added qubit chosen to maximise how many of
// the next `look` unplaced ops the block could absorb; then absorb every op that acts
// inside the set and is not preceded, on one of its qubits, by an op left
// out (so moving the block ahead of everything still unplaced keeps every
// dependence).  On brick-wall circuits this finds the staggered windows that
// op-order greedy fusion misses (148 instead of 220 blocks for 30 qubits x 40
// layers, best lookahead).  Blocks come out in a valid execution order.
static void fuse_blocks_front(int n, int bq, const std::vector<LOp>& lops, std::vector<FBlock>& blocks, int look) {
    const uint64_t all = n >= 64 ? ~0ull : ((1ull << n) - 1);
    std::vector<int> rem(lops.size());
    for (size_t i = 0; i < lops.size(); i++) rem[i] = (int)i;
    std::vector<char> done(lops.size(), 0);
    size_t head = 0;
    auto fits = [&](const LOp& o) { return !o.big && popc(o.touch) <= bq; };
    auto scan = [&](uint64_t Q, int limit, std::vector<int>* out) {
        uint64_t blocked = 0;
        int cnt = 0, seen = 0;
        for (size_t j = head; j < rem.size(); j++) {
            const int i = rem[j];
            if (done[i]) continue;
            if (limit >= 0 && seen++ >= limit) break;
            const LOp& o = lops[i];
            if (fits(o) && (o.touch & ~Q) == 0 && (o.touch & blocked) == 0) {
                cnt++;
                if (out) out->push_back(i);
            } else {
                blocked |= o.touch & Q;
                if (blocked == Q) break;
            }
        }
        return cnt;
    };
    for (;;) {
        while (head < rem.size() && done[rem[head]]) head++;
        if (head == rem.size()) break;
        if (head > 4096 && head * 2 > rem.size()) {       // drop the placed prefix
            rem.erase(rem.begin(), rem.begin() + (long)head);
            head = 0;
        }
        const int first = rem[head];
        const LOp& f0 = lops[first];
        if (!fits(f0)) {
            FBlock f;
            f.lop = first;
            for (int q = 0; q < n; q++)
                if (f0.touch >> q & 1) f.q.push_back(q);
            f.touch = f0.touch;
            blocks.push_back(f);
            done[first] = 1;
            continue;
        }
        uint64_t Q = f0.touch;
        while (popc(Q) < bq && popc(Q) < n) {
            uint64_t cand = 0;
            int seen = 0;
            for (size_t j = head; j < rem.size() && seen < look; j++) {
                const int i = rem[j];
                if (done[i]) continue;
                seen++;
                if (lops[i].touch & Q) cand |= lops[i].touch;
            }
            cand &= all & ~Q;
            if (!cand) cand = all & ~Q;
            int bestq = -1, bestc = -1;
            for (int q = 0; q < n; q++) {
                if (!(cand >> q & 1)) continue;
                const int c = scan(Q | (1ull << q), look, nullptr);
                if (c > bestc) { bestc = c; bestq = q; }
            }
            Q |= 1ull << bestq;
        }
        std::vector<int> got;
        scan(Q, -1, &got);
        FBlock f;
        for (int i : got) {
            f.touch |= lops[i].touch;
            f.ops.push_back(i);
            f.diag = f.diag && lops[i].cls == L_DIAG;
            f.mono = f.mono && (lops[i].cls == L_DIAG || lops[i].cls == L_PERMX || lops[i].cls == L_SWAP);
            done[i] = 1;
        }
        blocks.push_back(f);
    }
}

// Row of the one entry of column c of a monomial block, or -1.
static int mono_row(const cd* u, int R, int c) {
    int row = -1;
    for (int r = 0; r < R; r++)
        if (u[(size_t)r * R + c] != cd(0, 0)) {
            if (row >= 0) return -1;
            row = r;
        }
    return row;
}

// pi of a monomial 4-qubit block as 16 nibbles (nibble r = pi(r)), or false.
static bool mono_perm(const cd* u, uint64_t& word) {
    word = 0;
    uint32_t seen = 0;
    for (int c = 0; c < 16; c++) {
        const int r = mono_row(u, 16, c);
        if (r < 0 || ((seen >> r) & 1)) return false;
        seen |= 1u << r;
        word |= (uint64_t)r << (4 * c);
    }
    return true;
}

void block_fragments(const cd* u, int bq, int64_t flags, double* out) {
    const int R = 1 << bq, MB = 2 * R / 16, KS = 2 * R / 4;
    if (flags & BREC_MONO) {
        // fragment 2t / 2t + 1, lane l: Re / Im of U[pi(r)][r], r = 4t + l / 16 (the
        // rows a lane reads in the kernel), the phase of row r on its way to pi(r)
        std::fill(out, out + (size_t)MB * KS * 64, 0.0);
        for (int t = 0; t < R / 4; t++)
            for (int l = 0; l < 64; l++) {
                const int r = 4 * t + l / 16;
                const int pr = std::max(mono_row(u, R, r), 0);
                out[(size_t)(2 * t) * 64 + l] = u[(size_t)pr * R + r].real();
                out[(size_t)(2 * t + 1) * 64 + l] = u[(size_t)pr * R + r].imag();
            }
        return;
    }
    if (flags & BREC_DIAG) {
        // fragment 2t / 2t + 1, lane l: Re / Im of U[r][r], r = 4t + l / 16 (the rows a
        // lane reads in the kernel); the remaining fragments are unused
        std::fill(out, out + (size_t)MB * KS * 64, 0.0);
        for (int t = 0; t < R / 4; t++)
            for (int l = 0; l < 64; l++) {
                const int r = 4 * t + l / 16;
                out[(size_t)(2 * t) * 64 + l] = u[(size_t)r * R + r].real();
                out[(size_t)(2 * t + 1) * 64 + l] = u[(size_t)r * R + r].imag();
            }
        return;
    }
    // Gauss's three products: fragments of Ur, Ui and Ur + Ui (which = 0, 1, 2), each
    // [mb][kk][lane] with lane l holding M[16 mb + l % 16][4 kk + l / 16]; the
    // kernel forms Yr = Ur Xr - Ui Xi, Yi = (Ur + Ui)(Xr + Xi) - Ur Xr - Ui Xi
    const int MB3 = R / 16, KS3 = R / 4;
    std::fill(out, out + (size_t)MB * KS * 64, 0.0);
    for (int which = 0; which < 3; which++)
        for (int mb = 0; mb < MB3; mb++)
            for (int kk = 0; kk < KS3; kk++)
                for (int l = 0; l < 64; l++) {
                    const cd v = u[(size_t)(16 * mb + l % 16) * R + 4 * kk + l / 16];
                    out[(((size_t)which * MB3 + mb) * KS3 + kk) * 64 + l] =
                        which == 0 ? v.real() : which == 1 ? v.imag() : v.real() + v.imag();
                }
}

// Pass swizzle: tile bit p >= 4 XORs v[p] into the low four address bits.
// Chosen so that, for every block of the pass, the lowest four column bits
// land in four independent low-bit directions (conflict-free 16-lane LDS
// accesses).
static uint64_t choose_swizzle(int K, const std::vector<std::vector<int>>& colbits) {
    uint32_t v[16] = {0};
    auto good = [&]() {
        for (const auto& cb : colbits) {
            std::array<uint32_t, 4> a{};
            int m = 0;
            for (size_t i = 0; i < cb.size() && i < 4; i++) {
                const int p = cb[i];
                a[m++] = p < 4 ? (1u << p) : v[p];
            }
            if (gf2_rank(a, m) != m) return false;
        }
        return true;
    };
    std::vector<int> free_p;
    for (const auto& cb : colbits)
        for (size_t i = 0; i < cb.size() && i < 4; i++)
            if (cb[i] >= 4 && std::find(free_p.begin(), free_p.end(), cb[i]) == free_p.end()) free_p.push_back(cb[i]);
    uint32_t rng = 0x2545F491u;
    bool ok = good();
    for (int attempt = 0; attempt < 4096 && !ok && !free_p.empty(); attempt++) {
        for (int p : free_p) {
            rng ^= rng << 13;
            rng ^= rng >> 17;
            rng ^= rng << 5;
            v[p] = rng & 15;
        }
        ok = good();
    }
    if (!ok)
        for (int p : free_p) v[p] = (uint32_t)(1u << (p & 3));   // fold by nibble: conflicts, never wrong
    uint64_t w = 0;
    for (int p = 4; p < K && p < 20; p++) w |= (uint64_t)(v[p] & 15) << (4 * (p - 4));
    return w;
}

static uint32_t swz_addr(uint64_t w, uint32_t x) {
    uint32_t g = 0;
    for (int p = 4; p < 20; p++)
        if ((x >> p) & 1) g ^= (uint32_t)((w >> (4 * (p - 4))) & 15);
    return x ^ g;
}

static void block_pass(int n, int K, int bq, uint64_t T, std::vector<FBlock>& blocks, const std::vector<int>& idx,
                       Program& prog) {
    for (int q = 0; q < n && popc(T) < K; q++) T |= 1ull << q;
    Step st;
    st.blk = true;
    st.k = K;
    st.bq = bq;
    int pos[64];
    for (int q = 0; q < 64; q++) pos[q] = -1;
    int j = 0;
    for (int q = 0; q < n; q++)
        if (T >> q & 1) { st.tq[j] = (int8_t)q; pos[q] = j; j++; }
    // pad every block to bq qubits with tile qubits it does not use
    std::vector<std::vector<int>> rowpos(idx.size()), colpos(idx.size());
    for (size_t b = 0; b < idx.size(); b++) {
        FBlock& f = blocks[idx[b]];
        std::vector<int> add;
        for (int p = 0; p < K && (int)(f.q.size() + add.size()) < bq; p++) {
            const int q = st.tq[p];
            if (!(f.touch >> q & 1)) add.push_back(q);
        }
        grow_block(f, add);
        for (int q : f.q) rowpos[b].push_back(pos[q]);
        for (int p = 0; p < K; p++)
            if (std::find(rowpos[b].begin(), rowpos[b].end(), p) == rowpos[b].end()) colpos[b].push_back(p);
    }
    st.swz = choose_swizzle(K, colpos);
    st.op_begin = (int)prog.brecs.size();
    const int64_t nfrag = (int64_t)(2 * (1 << bq) / 16) * (2 * (1 << bq) / 4) * 64;
    for (size_t b = 0; b < idx.size(); b++) {
        const FBlock& f = blocks[idx[b]];
        BRec r;
        std::memset(&r, 0, sizeof r);
        for (int i = 0; i < bq; i++) r.ar[i] = (uint16_t)swz_addr(st.swz, 1u << rowpos[b][i]);
        for (int i = 0; i < K - bq; i++) r.ac[i] = (uint16_t)swz_addr(st.swz, 1u << colpos[b][i]);
        r.frag = (int64_t)prog.brecs.size() * nfrag;
        r.flags = f.diag ? BREC_DIAG : 0;
        uint64_t perm = 0;
        if (!f.diag && f.mono && bq == 4 && mono_perm(f.u.data(), perm)) {
            r.flags = BREC_MONO;
            for (int i = 0; i < 4; i++) r.ar[4 + i] = (uint16_t)(perm >> (16 * i));
        }
        prog.bmat_off.push_back((int64_t)prog.bmats.size());
        prog.bmats.insert(prog.bmats.end(), f.u.begin(), f.u.end());
        for (int id : f.ops)
            if ((size_t)id < prog.blk_of_lop.size()) prog.blk_of_lop[id] = (int)prog.brecs.size();
        prog.blk_lops.push_back(f.ops);
        prog.blk_q.push_back(f.q);
        prog.blk_bq.push_back(bq);
        prog.brecs.push_back(r);
    }
    st.op_count = (int)prog.brecs.size() - st.op_begin;
    prog.steps.push_back(st);
}

void schedule_blocks(int n, int btile, int bq, std::vector<LOp>& lops, Program& prog) {
    const int K = btile;
    if (n < K || prog.prune) {
        schedule_rounds(n, prog.precision == KSV_F64 ? 11 : 12, false, lops, prog);
        return;
    }
    // several packings; the one with fewest blocks does the least MFMA work
    // (the best lookahead depends on the circuit's width and layering)
    std::vector<FBlock> blocks;
    fuse_blocks(n, bq, lops, blocks);
    for (int look : {48, 64, 96, 400}) {
        std::vector<FBlock> front;
        fuse_blocks_front(n, bq, lops, front, look);
        if (front.size() < blocks.size()) blocks.swap(front);
    }
    for (FBlock& f : blocks)
        if (f.lop < 0) {
            f.touch = 0;
            block_matrix(n, f, lops);
        }
    // pack blocks into passes (a standalone op is a step of its own)
    std::vector<LOp> proxies(blocks.size());
    for (size_t i = 0; i < blocks.size(); i++) {
        proxies[i].touch = blocks[i].touch;
        proxies[i].need = blocks[i].touch;
        proxies[i].big = blocks[i].lop >= 0;
    }
    const uint64_t all = n >= 64 ? ~0ull : ((1ull << n) - 1);
    const int KL = std::min(n, prog.tile);
    std::vector<int> rem(blocks.size()), chosen, rest;
    for (size_t i = 0; i < blocks.size(); i++) rem[i] = (int)i;
    while (!rem.empty()) {
        const FBlock& first = blocks[rem[0]];
        if (first.lop >= 0) {
            const LOp& o = lops[first.lop];
            if (o.big) emit_big(o, first.lop, prog);
            else {
                uint64_t T = ((1ull << std::min(4, KL)) - 1) | o.need;
                emit_lds_pass(n, KL, T, lops, {first.lop}, prog);
            }
            rem.erase(rem.begin());
            continue;
        }
        uint64_t T = pick_pass(proxies, rem, K, 7, all, [](const LOp& o) { return !o.big; }, chosen, rest, false);
        block_pass(n, K, bq, T, blocks, chosen, prog);
        rem.swap(rest);
    }
}

int plan_program(Program& p, int n, int prec, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                 int kernel, int tile, int rtile, bool fuse, int prune, std::string& err, int btile, int bq) {
    if (nops && !ops) {
        err = "null op tape";
        return KSV_ERR_TYPE;
    }
    std::vector<LOp> lops;
    int rc = lower_tape(n, ops, nops, mats, nmats, lops, err);
    if (rc) return rc;
    p.n = n;
    p.precision = prec;
    p.prune = prune;
    p.tile = tile > 0 ? tile : (prec == KSV_F64 ? 12 : 13);
    p.tape_ops = nops;
    p.tape.assign(ops, ops + nops);
    p.rop_of_lop.assign(lops.size(), -1);
    p.blk_of_lop.assign(lops.size(), -1);
    p.mat_of_lop.assign(lops.size(), -1);
    const int rt = rtile > 0 ? rtile : (prec == KSV_F64 ? 11 : 12);
    // auto: fused MFMA blocks in exact mode, ordered register rounds under ket's pruning
    const int bt = btile > 0 ? btile : 12;
    if (kernel == 0) kernel = prune ? 2 : 3;
    if (kernel == 1) schedule(n, p.tile, fuse && !prune, lops, p);
    else if (kernel == 3) schedule_blocks(n, bt, bq > 0 ? bq : 4, lops, p);
    else schedule_rounds(n, rt, fuse && !prune, lops, p);
    // one-qubit fusion (register rounds with `fuse`) merges lowered ops: such
    // programs keep no op map and cannot be re-parameterised
    if (p.rop_of_lop.size() == lops.size()) {
        p.lops = lops;
        p.lop_of_src.assign(nops, {});
        for (size_t i = 0; i < lops.size(); i++)
            if (lops[i].src >= 0 && (size_t)lops[i].src < nops) p.lop_of_src[lops[i].src].push_back((int)i);
    }
    return KSV_OK;
}

// Payload of an already emitted register-round record after its op's matrix changed.
static void refresh_rop(const LOp& o, ROp& r, Program& p, std::vector<std::pair<size_t, size_t>>& ranges) {
    switch (r.kind) {
    case R_M1:
        put_inline(r, o.m.data(), 4, p.precision);
        r.flags = (uint16_t)((r.flags & ROP_COND) | m1_form(o.m.data()));
        break;
    case R_DIAG:
        if (o.tg.size() <= 2) put_inline(r, o.m.data(), (int)o.m.size(), p.precision);
        else {
            std::copy(o.m.begin(), o.m.end(), p.mats.begin() + r.mat);
            ranges.push_back({(size_t)r.mat, o.m.size()});
        }
        break;
    case R_DENSE:
        std::copy(o.m.begin(), o.m.end(), p.mats.begin() + r.mat);
        ranges.push_back({(size_t)r.mat, o.m.size()});
        break;
    default: break;       // permutations carry no matrix
    }
}

static bool same_shape(const LOp& a, const LOp& b) {
    return a.cls == b.cls && a.tg == b.tg && a.ctl == b.ctl && a.m.size() == b.m.size() && a.big == b.big;
}

int update_program(Program& p, size_t op, const double* mat, size_t nmat, std::vector<int>& rops,
                   std::vector<int>& blocks, std::vector<std::pair<size_t, size_t>>& ranges, std::string& err) {
    rops.clear();
    blocks.clear();
    ranges.clear();
    if (p.lops.empty() && !p.tape.empty()) {
        err = "this program fused one-qubit gates and cannot be re-parameterised";
        return KSV_ERR_TYPE;
    }
    if (op >= p.tape.size()) {
        err = "op index out of range";
        return KSV_ERR_RANGE;
    }
    ksv_op o = p.tape[op];
    o.mat = 0;
    std::vector<LOp> nl;
    int rc = lower_tape(p.n, &o, 1, mat, nmat, nl, err);
    if (rc) return rc;
    const std::vector<int>& old = p.lop_of_src[op];
    if (nl.size() != old.size()) {
        err = "the new matrix changes the op's structure (diagonal / permutation / dense)";
        return KSV_ERR_TYPE;
    }
    for (size_t i = 0; i < old.size(); i++)
        if (!same_shape(nl[i], p.lops[old[i]])) {
            err = "the new matrix changes the op's structure (diagonal / permutation / dense)";
            return KSV_ERR_TYPE;
        }
    for (size_t i = 0; i < old.size(); i++) {
        const int id = old[i];
        p.lops[id].m = nl[i].m;
        if (p.rop_of_lop[id] >= 0) {
            refresh_rop(p.lops[id], p.rops[p.rop_of_lop[id]], p, ranges);
            rops.push_back(p.rop_of_lop[id]);
        } else if (p.blk_of_lop[id] >= 0) {
            if (std::find(blocks.begin(), blocks.end(), p.blk_of_lop[id]) == blocks.end())
                blocks.push_back(p.blk_of_lop[id]);
        } else if (p.mat_of_lop[id] >= 0) {
            std::copy(p.lops[id].m.begin(), p.lops[id].m.end(), p.mats.begin() + p.mat_of_lop[id]);
            ranges.push_back({(size_t)p.mat_of_lop[id], p.lops[id].m.size()});
        } else if (p.lops[id].cls == L_PERMX || p.lops[id].cls == L_SWAP) {
            // permutations carry no matrix
        } else {
            err = "the op was scheduled into an LDS or wide-unitary step, which cannot be re-parameterised";
            return KSV_ERR_TYPE;
        }
    }
    for (int b : blocks) {
        FBlock f;
        f.q = p.blk_q[b];
        const size_t d = (size_t)1 << f.q.size();
        f.u.assign(d * d, cd(0, 0));
        for (size_t i = 0; i < d; i++) f.u[i * d + i] = 1;
        for (int id : p.blk_lops[b]) absorb(f, p.lops[id]);
        std::copy(f.u.begin(), f.u.end(), p.bmats.begin() + p.bmat_off[b]);
    }
    return KSV_OK;
}

}  // namespace ksv
