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

void schedule(int n, int tile, bool fuse, std::vector<LOp>& lops, Program& prog) {
    if (fuse) fuse_1q(n, lops);
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
            Step s;
            s.big = true;
            s.big_op = (int)prog.big_ops.size();
            BigOp b;
            b.tg = first.tg;
            b.mat = (int)prog.mats.size();
            prog.mats.insert(prog.mats.end(), first.m.begin(), first.m.end());
            prog.big_ops.push_back(b);
            prog.steps.push_back(s);
            rem.erase(rem.begin());
            continue;
        }
        uint64_t T = forced, blocked = 0;
        chosen.clear();
        rest.clear();
        size_t i = 0;
        for (; i < rem.size(); i++) {
            const LOp& o = lops[rem[i]];
            if (o.big || (o.touch & blocked)) { blocked |= o.touch; rest.push_back(rem[i]); }
            else {
                uint64_t nt = T | o.need;
                if (popc(nt) <= K) { T = nt; chosen.push_back(rem[i]); }
                else { blocked |= o.touch; rest.push_back(rem[i]); }
            }
            if ((blocked & all) == all) { i++; break; }
        }
        for (; i < rem.size(); i++) rest.push_back(rem[i]);
        // pad the tile with the lowest unused qubits (longer contiguous runs)
        for (int q = 0; q < n && popc(T) < K; q++) T |= 1ull << q;
        Step s;
        s.k = K;
        int pos[64];
        for (int q = 0; q < 64; q++) pos[q] = -1;
        int j = 0;
        for (int q = 0; q < n; q++)
            if (T >> q & 1) { s.tq[j] = (int8_t)q; pos[q] = j; j++; }
        s.op_begin = (int)prog.dops.size();
        for (int idx : chosen) emit_dop(lops[idx], pos, prog);
        s.op_count = (int)prog.dops.size() - s.op_begin;
        prog.steps.push_back(s);
        rem.swap(rest);
    }
}

}  // namespace ksv
