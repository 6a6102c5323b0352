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

static void emit_big(const LOp& o, Program& prog) {
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
    for (int id : idx) emit_dop(lops[id], pos, prog);
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
            emit_big(first, prog);
            rem.erase(rem.begin());
            continue;
        }
        uint64_t T = pick_pass(lops, rem, K, forced, all, [](const LOp& o) { return !o.big; }, chosen, rest,
                               strict);
        emit_lds_pass(n, K, T, lops, chosen, prog);
        rem.swap(rest);
    }
}

void schedule(int n, int tile, bool fuse, std::vector<LOp>& lops, Program& prog) {
    if (fuse) fuse_1q(n, lops);
    schedule_lds(n, tile, lops, prog);
}

// --------------------------------------------------------------------------
// Register-tile passes.
// --------------------------------------------------------------------------
static bool wave_ok(const LOp& o) {
    if (o.big) return false;
    if (o.cls == L_MAT && o.tg.size() > 2) return false;
    if (o.cls == L_DIAG && o.tg.size() > 2) return false;
    return true;
}

// where[] codes: -1 outside the tile, 0..5 lane bit, 8+s register slot
static void wave_pass(int n, int K, int rlog, uint64_t T, const std::vector<LOp>& lops,
                      const std::vector<int>& idx, Program& prog) {
    for (int q = 0; q < n && popc(T) < K; q++) T |= 1ull << q;
    Step s;
    s.wave = true;
    s.k = K;
    s.rlog = rlog;
    int where[64];
    for (int q = 0; q < 64; q++) where[q] = -1;
    int j = 0;
    for (int q = 0; q < n; q++)
        if (T >> q & 1) s.tq[j++] = (int8_t)q;
    // targets that must sit in register slots, per op
    std::vector<std::vector<int>> tgs(idx.size());
    for (size_t i = 0; i < idx.size(); i++) {
        const LOp& o = lops[idx[i]];
        if (o.cls != L_DIAG) tgs[i] = o.tg;
    }
    // initial placement: state bits 0..2 on lane bits 0..2; the earliest
    // targets in the slots; the rest on lane bits 3..5
    for (int b = 0; b < 3; b++) where[b] = b;
    std::vector<int> order;
    for (size_t i = 0; i < idx.size(); i++)
        for (int q : tgs[i])
            if (q > 2 && std::find(order.begin(), order.end(), q) == order.end()) order.push_back(q);
    for (int q = 3; q < n; q++)
        if ((T >> q & 1) && std::find(order.begin(), order.end(), q) == order.end()) order.push_back(q);
    int slot_q[8], lane_q[6] = {0, 1, 2, -1, -1, -1};
    for (int i = 0; i < rlog; i++) { slot_q[i] = order[i]; where[order[i]] = 8 + i; }
    for (int i = 0; i < 3; i++) { lane_q[3 + i] = order[rlog + i]; where[order[rlog + i]] = 3 + i; }
    for (int b = 0; b < 6; b++) s.lane_in[b] = (int8_t)lane_q[b];
    for (int i = 0; i < rlog; i++) s.reg_in[i] = (int8_t)slot_q[i];

    s.op_begin = (int)prog.wops.size();
    auto next_use = [&](int q, size_t from) -> size_t {
        for (size_t i = from; i < idx.size(); i++)
            for (int t : tgs[i])
                if (t == q) return i;
        return (size_t)-1;
    };
    auto lswap = [&](int lanebit, int slot) {
        WOp w;
        std::memset(&w, 0, sizeof w);
        w.kind = W_LSWAP;
        w.a = (uint8_t)lanebit;
        w.b = (uint8_t)slot;
        prog.wops.push_back(w);
        int ql = lane_q[lanebit], qs = slot_q[slot];
        lane_q[lanebit] = qs;
        slot_q[slot] = ql;
        where[qs] = lanebit;
        where[ql] = 8 + slot;
    };
    for (size_t i = 0; i < idx.size(); i++) {
        const LOp& o = lops[idx[i]];
        for (int q : tgs[i]) {
            if (where[q] >= 8) continue;
            // evict the slot whose qubit is needed furthest ahead
            int best = -1;
            size_t bestd = 0;
            for (int sl = 0; sl < rlog; sl++) {
                int qq = slot_q[sl];
                if (std::find(tgs[i].begin(), tgs[i].end(), qq) != tgs[i].end()) continue;
                size_t d = next_use(qq, i);
                if (best < 0 || d > bestd) { best = sl; bestd = d; }
            }
            lswap(where[q], best);
        }
        WOp w;
        std::memset(&w, 0, sizeof w);
        for (int c : o.ctl) {
            int wc = where[c];
            if (wc < 0) w.gctl |= 1ull << c;
            else if (wc < 8) w.lctl |= 1u << wc;
            else w.rctl |= 1u << (wc - 8);
        }
        w.mat = (int32_t)prog.mats.size();
        switch (o.cls) {
        case L_MAT:
            if (o.tg.size() == 1) {
                w.kind = W_M1;
                w.a = (uint8_t)(where[o.tg[0]] - 8);
                prog.mats.insert(prog.mats.end(), o.m.begin(), o.m.end());
            } else {
                w.kind = W_M2;
                int sa = where[o.tg[0]] - 8, sb = where[o.tg[1]] - 8;
                if (sa > sb) {
                    prog.mats.insert(prog.mats.end(), o.m.begin(), o.m.end());
                    w.a = (uint8_t)sa;
                    w.b = (uint8_t)sb;
                } else {
                    static const int perm[4] = {0, 2, 1, 3};
                    for (int r = 0; r < 4; r++)
                        for (int c = 0; c < 4; c++) prog.mats.push_back(o.m[perm[r] * 4 + perm[c]]);
                    w.a = (uint8_t)sb;
                    w.b = (uint8_t)sa;
                }
            }
            break;
        case L_DIAG:
            if (o.tg.empty()) {
                w.kind = W_PHASE;
            } else {
                w.kind = W_DIAG;
                w.k = (uint8_t)o.tg.size();
                for (size_t t = 0; t < o.tg.size(); t++) {
                    int q = o.tg[t], wq = where[q];
                    w.loc[t] = (uint8_t)(wq < 0 ? 16 + q : wq);
                }
            }
            prog.mats.insert(prog.mats.end(), o.m.begin(), o.m.end());
            break;
        case L_PERMX:
            w.kind = W_PERMX;
            w.a = (uint8_t)(where[o.tg[0]] - 8);
            break;
        case L_SWAP: {
            w.kind = W_SWAPR;
            int sa = where[o.tg[0]] - 8, sb = where[o.tg[1]] - 8;
            w.a = (uint8_t)std::max(sa, sb);
            w.b = (uint8_t)std::min(sa, sb);
            break;
        }
        }
        prog.wops.push_back(w);
    }
    // bring state bits 0..2 back to lane bits 0..2 for a coalesced store
    for (int p = 0; p < 3; p++)
        if (where[p] >= 8) lswap(p, where[p] - 8);
    for (int b = 0; b < 6; b++) s.lane_out[b] = (int8_t)lane_q[b];
    for (int i = 0; i < rlog; i++) s.reg_out[i] = (int8_t)slot_q[i];
    s.op_count = (int)prog.wops.size() - s.op_begin;
    prog.steps.push_back(s);
}

void schedule_wave(int n, int rlog, bool fuse, std::vector<LOp>& lops, Program& prog) {
    if (fuse) fuse_1q(n, lops);
    const int K = WLANE_BITS + rlog;
    if (n < K + 2) {
        schedule_lds(n, prog.tile, lops, prog);
        return;
    }
    prog.rlog = rlog;
    const uint64_t all = n >= 64 ? ~0ull : ((1ull << n) - 1);
    const uint64_t forced = 7;
    const int KL = std::min(n, prog.tile);
    std::vector<int> rem(lops.size());
    for (size_t i = 0; i < lops.size(); i++) rem[i] = (int)i;
    std::vector<int> chosen, rest;
    while (!rem.empty()) {
        const LOp& first = lops[rem[0]];
        if (first.big) {
            emit_big(first, prog);
            rem.erase(rem.begin());
            continue;
        }
        if (!wave_ok(first)) {
            // a dense 3- or 4-qubit matrix or a wide diagonal: one LDS pass of its own
            uint64_t T = ((1ull << std::min(4, KL)) - 1) | first.need;
            emit_lds_pass(n, KL, T, lops, {rem[0]}, prog);
            rem.erase(rem.begin());
            continue;
        }
        uint64_t T = pick_pass(lops, rem, K, forced, all, wave_ok, chosen, rest, prog.prune != 0);
        wave_pass(n, K, rlog, T, lops, chosen, prog);
        rem.swap(rest);
    }
}

}  // namespace ksv
