// Host planner of the adjoint backward sweep (ksv_adjoint).
//
// The sweep applies the inverse tape G_N^dag, ..., G_1^dag to phi and to
// lambda and, at every op with parameter records, adds
// Re <lambda| dG |G^dag phi> (lambda taken before its own G^dag).  This file
// splits that inverse tape into passes over K-qubit tiles: a pass takes, in
// order, every op whose targets lie in the tile and whose qubits no earlier
// skipped op touches.  Taking an op ahead of skipped ops on other qubits is
// exact for the derivative terms too: such ops U act on phi and lambda alike
// and commute with dG, so <U lambda| dG |U phi> = <lambda| dG |phi>.
// Controls may lie outside the tile (the kernel tests them on the tile's base
// index).  Unitaries wider than three qubits run alone through the general
// engine, on phi and then on lambda.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>

#include "ksv_internal.h"

namespace ksv {

namespace {

struct AItem {
    int kind = AK_DENSE;        // AKind, or -1: wide unitary for the general engine
    std::vector<int> tg, ctl;   // targets (MSB first), controls
    std::vector<cd> m;          // G^dag (AK_DENSE and wide), row-major
    std::vector<int> recs;      // parameter records of the op
    int src = -1;               // tape index of the op
    uint64_t need = 0, touch = 0;
    bool diag = false;          // AK_DENSE: G^dag and every record's dG diagonal
};

int popc(uint64_t x) { return __builtin_popcountll(x); }

uint64_t mask_of(const std::vector<int>& v) {
    uint64_t m = 0;
    for (int q : v) m |= 1ull << q;
    return m;
}

std::vector<cd> dagger(const std::vector<cd>& m, int dim) {
    std::vector<cd> a((size_t)dim * dim);
    for (int r = 0; r < dim; r++)
        for (int c = 0; c < dim; c++) a[(size_t)r * dim + c] = std::conj(m[(size_t)c * dim + r]);
    return a;
}

bool read_mat(const double* mats, size_t nmats, int64_t off, int dim, std::vector<cd>& m) {
    const size_t cnt = 2 * (size_t)dim * dim;
    if (off < 0 || (size_t)off + cnt > nmats) return false;
    m.resize((size_t)dim * dim);
    for (size_t j = 0; j < m.size(); j++) m[j] = cd(mats[off + 2 * j], mats[off + 2 * j + 1]);
    return true;
}

// A one-qubit gate that is a rotation exp(-i theta P / 2) about a Pauli axis P
// has dG/dtheta = (-i/2) P G, so the backward sweep's derivative term
// Re <lambda| dG G^dag |phi> is Re <lambda| (-i/2) P |phi>.  Returns P (0 = X,
// 1 = Y, 2 = Z) when the record's derivative `d` (2x2, interleaved complex)
// times the op's G^dag `gd` equals (-i/2) P to rounding, else -1.
int pauli_rotation(const std::vector<cd>& gd, const double* d) {
    cd m[4];
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) {
            cd acc = 0;
            for (int k = 0; k < 2; k++) acc += cd(d[2 * (2 * r + k)], d[2 * (2 * r + k) + 1]) * gd[2 * k + c];
            m[2 * r + c] = acc;
        }
    const cd h(0, -0.5), z(0, 0);
    const cd want[3][4] = {{z, h, h, z}, {z, cd(-0.5, 0), cd(0.5, 0), z}, {h, z, z, -h}};
    for (int pk = 0; pk < 3; pk++) {
        bool ok = true;
        for (int i = 0; i < 4 && ok; i++) ok = std::abs(m[i] - want[pk][i]) <= 1e-14;
        if (ok) return pk;
    }
    return -1;
}

bool offdiag_zero(const std::vector<cd>& m, int dim) {
    for (int r = 0; r < dim; r++)
        for (int c = 0; c < dim; c++)
            if (r != c && m[(size_t)r * dim + c] != cd(0, 0)) return false;
    return true;
}

// Same element enumeration: dense ops on the same targets in the same order
// under the same controls.
bool same_support(const AItem& a, const AItem& b) {
    if (a.kind != AK_DENSE || b.kind != AK_DENSE || a.tg != b.tg) return false;
    return mask_of(a.ctl) == mask_of(b.ctl);
}

// An X (cnot, Toffoli) whose controls all lie in tile T can head a chain of
// uncontrolled one-qubit dense ops on its target: the kernel swaps each pair
// whose control bits are set, then runs the chain on it, in one LDS round trip.
bool x_head(const AItem& a, uint64_t T) {
    return a.kind == AK_X && a.tg.size() == 1 && !a.ctl.empty() && (mask_of(a.ctl) & ~T) == 0;
}

bool plain_1q(const AItem& b, int t) {
    return b.kind == AK_DENSE && b.tg.size() == 1 && b.tg[0] == t && b.ctl.empty();
}

// Order the ops of a pass into chains: op j joins the chain of an earlier op
// i with the same support when no op left between them touches those
// qubits (j then commutes with everything it moves past).  An X with local
// controls heads a chain of the uncontrolled one-qubit ops on its target that
// follow it, under the same rule on the target (ops between them may touch
// the controls: the X reads them at its own place).  run[x] = chain length at
// a chain head, 0 inside a chain.
void make_chains(const std::vector<AItem>& it, const std::vector<int>& chosen, uint64_t T, std::vector<int>& order,
                 std::vector<int>& run) {
    order.clear();
    run.clear();
    std::vector<char> used(chosen.size(), 0);
    for (size_t i = 0; i < chosen.size(); i++) {
        if (used[i]) continue;
        used[i] = 1;
        const AItem& a = it[chosen[i]];
        const size_t head = order.size();
        order.push_back(chosen[i]);
        run.push_back(1);
        if (x_head(a, T)) {
            const int t = a.tg[0];
            const uint64_t tm = 1ull << t;
            const AItem* sup = nullptr;
            uint64_t blk = 0;
            for (size_t j = i + 1; j < chosen.size() && run[head] < 1 + ACHAIN_1Q; j++) {
                if (used[j]) continue;
                const AItem& b = it[chosen[j]];
                const bool fits = sup ? same_support(*sup, b) : plain_1q(b, t);
                if (fits && !(b.touch & blk)) {
                    used[j] = 1;
                    order.push_back(chosen[j]);
                    run.push_back(0);
                    run[head]++;
                    if (!sup) sup = &b;
                } else {
                    blk |= b.touch;
                    if (blk & tm) break;
                }
            }
            continue;
        }
        if (a.kind != AK_DENSE || a.tg.size() > 2) continue;
        const int cap = a.tg.size() == 1 ? ACHAIN_1Q : ACHAIN_2Q;
        uint64_t blk = 0;
        for (size_t j = i + 1; j < chosen.size() && run[head] < cap; j++) {
            if (used[j]) continue;
            const AItem& b = it[chosen[j]];
            if (!(b.touch & blk) && same_support(a, b)) {
                used[j] = 1;
                order.push_back(chosen[j]);
                run.push_back(0);
                run[head]++;
            } else {
                blk |= b.touch;
                if (blk & a.touch) break;
            }
        }
    }
}

// Chains that may share one LDS round trip: adjacent chains of `order` whose
// ops all act on one qubit each (an X whose controls lie in the tile, then
// uncontrolled one-qubit ops on its target, or uncontrolled one-qubit ops
// alone), on distinct targets, at most AGRP_MAX of them.  Running them one
// after the other on the 2^k elements of each group over their targets is the
// same arithmetic as running them in turn over the whole tile.  grp[x] = chains
// of the group a chain head x starts (0 elsewhere, 1 for a chain left alone).
void make_groups(const std::vector<AItem>& it, const std::vector<int>& order, const std::vector<int>& run,
                 uint64_t T, std::vector<int>& grp, std::vector<int>& tgt) {
    grp.assign(order.size(), 0);
    tgt.assign(order.size(), -1);
    auto one_qubit = [&](size_t x) {   // target of the chain at head x, or -1
        const AItem& h = it[order[x]];
        if (h.tg.size() != 1) return -1;
        if (h.kind == AK_X) return x_head(h, T) ? h.tg[0] : -1;
        if (h.kind != AK_DENSE || !h.ctl.empty()) return -1;
        for (int c = 1; c < run[x]; c++)
            if (!plain_1q(it[order[x + c]], h.tg[0])) return -1;
        return h.tg[0];
    };
    size_t x = 0;
    while (x < order.size()) {
        const int t0 = one_qubit(x);
        tgt[x] = t0;
        if (t0 < 0) {
            grp[x] = 1;
            x += (size_t)std::max(run[x], 1);
            continue;
        }
        uint64_t used = 1ull << t0;
        size_t y = x + (size_t)std::max(run[x], 1);
        int nch = 1;
        while (nch < AGRP_MAX && y < order.size()) {
            const int t = one_qubit(y);
            if (t < 0 || ((used >> t) & 1)) break;
            tgt[y] = t;
            used |= 1ull << t;
            nch++;
            y += (size_t)std::max(run[y], 1);
        }
        grp[x] = nch;
        x = y;
    }
}

// Ops of `rem` that a pass over tile T takes (see the file comment), within
// ADJ_SLOTS parameter records.  Scans at most `window` ops when chosen == null.
int take(const std::vector<AItem>& it, const std::vector<int>& rem, uint64_t T, uint64_t all, size_t window,
         std::vector<int>* chosen, std::vector<int>* rest) {
    uint64_t blocked = 0;
    int nslot = 0, count = 0;
    size_t i = 0;
    const size_t lim = chosen ? rem.size() : std::min(rem.size(), window);
    for (; i < lim; i++) {
        const AItem& o = it[rem[i]];
        const bool ok = o.kind >= 0 && !(o.touch & blocked) && !(o.need & ~T) &&
                        nslot + (int)o.recs.size() <= ADJ_SLOTS;
        if (ok) {
            count++;
            nslot += (int)o.recs.size();
            if (chosen) chosen->push_back(rem[i]);
        } else {
            blocked |= o.touch;
            if (rest) rest->push_back(rem[i]);
        }
        if ((blocked & all) == all) {
            i++;
            break;
        }
    }
    if (rest)
        for (; i < rem.size(); i++) rest->push_back(rem[i]);
    return count;
}

// Grow T first-come: every op in order that is not blocked and still fits.
uint64_t fill_first_come(const std::vector<AItem>& it, const std::vector<int>& rem, uint64_t T, int K,
                         uint64_t all) {
    uint64_t blocked = 0;
    for (int id : rem) {
        const AItem& o = it[id];
        if (o.kind < 0 || (o.touch & blocked)) blocked |= o.touch;
        else if (popc(T | o.need) <= K) T |= o.need;
        else blocked |= o.touch;
        if ((blocked & all) == all || popc(T) == K) break;
    }
    return T;
}

// Tile of the next pass: greedy forward selection (add the qubit that lets
// the pass take the most ops) from the first op's targets, then first-come
// fill; kept if it takes more ops than plain first-come growth.
uint64_t choose_tile(const std::vector<AItem>& it, const std::vector<int>& rem, int n, int K, uint64_t forced,
                     uint64_t all) {
    constexpr size_t WINDOW = 2048;
    uint64_t T = forced | it[rem[0]].need;
    uint64_t cand = 0;
    for (size_t i = 0; i < std::min(rem.size(), WINDOW); i++)
        if (it[rem[i]].kind >= 0) cand |= it[rem[i]].need;
    int cur = take(it, rem, T, all, WINDOW, nullptr, nullptr);
    while (popc(T) < K) {
        int best = cur, bq = -1;
        for (int q = 0; q < n; q++) {
            if (!((cand >> q) & 1) || ((T >> q) & 1)) continue;
            const int s = take(it, rem, T | (1ull << q), all, WINDOW, nullptr, nullptr);
            if (s > best) { best = s; bq = q; }
        }
        if (bq < 0) break;
        T |= 1ull << bq;
        cur = best;
    }
    T = fill_first_come(it, rem, T, K, all);
    const uint64_t F = fill_first_come(it, rem, forced, K, all);
    if (take(it, rem, F, all, WINDOW, nullptr, nullptr) > take(it, rem, T, all, WINDOW, nullptr, nullptr)) T = F;
    for (int q = 0; q < n && popc(T) < K; q++) T |= 1ull << q;
    return T;
}

}  // namespace

int plan_adjoint(int n, int K, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                 const ksv_param* params, size_t nrec, AdjPlan& plan, std::string& err, bool pauli) {
    plan = AdjPlan();
    K = std::min(n, std::max(K, 8));
    plan.k = K;
    std::vector<std::vector<int>> recs_of(nops);
    for (size_t r = 0; r < nrec; r++) recs_of[(size_t)params[r].op].push_back((int)r);
    // inverse tape
    std::vector<AItem> items;
    items.reserve(nops);
    for (size_t i = nops; i-- > 0;) {
        const ksv_op& o = ops[i];
        const int32_t* q = o.q;
        AItem a;
        std::vector<cd> m;
        int dim = 0;
        switch (o.kind) {
        case KSV_BARRIER: continue;
        case KSV_SINGLE: a.tg = {q[0]}; dim = 2; break;
        case KSV_CONTROLLED: a.tg = {q[1]}; a.ctl = {q[0]}; dim = 2; break;
        case KSV_TWO: a.tg = {q[0], q[1]}; dim = 4; break;
        case KSV_CNOT: a.kind = AK_X; a.tg = {q[1]}; a.ctl = {q[0]}; break;
        case KSV_TOFFOLI: a.kind = AK_X; a.tg = {q[2]}; a.ctl = {q[0], q[1]}; break;
        case KSV_SWAP: a.kind = AK_SWAP; a.tg = {q[0], q[1]}; break;
        case KSV_CSWAP: a.kind = AK_SWAP; a.tg = {q[1], q[2]}; a.ctl = {q[0]}; break;
        case KSV_CSRSWAP: {
            // src/statevector.ts:177-207: |c01>, |c10> mixed by [[1, i], [i, 1]] / sqrt(2)
            const double s = 1 / std::sqrt(2.0);
            a.tg = {q[1], q[2]};
            a.ctl = {q[0]};
            m.assign(16, cd(0, 0));
            m[0] = m[15] = 1;
            m[5] = m[10] = cd(s, 0);
            m[6] = m[9] = cd(0, s);
            a.m = dagger(m, 4);
            break;
        }
        case KSV_UNITARY:
            if (o.nq < 1 || o.nq > KSV_MAX_OPERANDS) { err = "unitary: bad operand count"; return KSV_ERR_TYPE; }
            a.tg.assign(q, q + o.nq);
            dim = 1 << o.nq;
            if (o.nq > ADJ_DENSE_MAX) a.kind = -1;
            break;
        default: err = "op " + std::to_string(i) + ": unknown kind"; return KSV_ERR_TYPE;
        }
        for (int t : a.tg)
            if (t < 0 || t >= n) { err = "op " + std::to_string(i) + ": qubit out of range"; return KSV_ERR_RANGE; }
        for (int c : a.ctl)
            if (c < 0 || c >= n) { err = "op " + std::to_string(i) + ": qubit out of range"; return KSV_ERR_RANGE; }
        if (dim) {
            if (!read_mat(mats, nmats, o.mat, dim, m)) { err = "op " + std::to_string(i) + ": bad matrix offset"; return KSV_ERR_TYPE; }
            a.m = dagger(m, dim);
        }
        a.recs = recs_of[i];
        a.src = (int)i;
        if (a.kind == AK_DENSE) {
            const int d = 1 << a.tg.size();
            a.diag = offdiag_zero(a.m, d);
            for (int r : a.recs) {
                std::vector<cd> dm;
                if (a.diag && read_mat(mats, nmats, params[r].dmat, d, dm)) a.diag = offdiag_zero(dm, d);
            }
        }
        a.need = mask_of(a.tg);
        a.touch = a.need | mask_of(a.ctl);
        if (popc(a.touch) != (int)(a.tg.size() + a.ctl.size())) {
            err = "op " + std::to_string(i) + ": operands must be distinct";
            return KSV_ERR_TYPE;
        }
        items.push_back(std::move(a));
    }

    const uint64_t all = n >= 64 ? ~0ull : ((1ull << n) - 1);
    // the tile always holds the lowest qubits: 128-byte runs of amplitudes
    const uint64_t forced = (1ull << std::min(K, 3)) - 1;
    std::vector<int> rem(items.size()), chosen, rest, order, run, grp, tgt;
    std::iota(rem.begin(), rem.end(), 0);
    while (!rem.empty()) {
        const AItem& first = items[rem[0]];
        if (first.kind < 0) {
            AdjStep s;
            s.big = true;
            s.bop = (int)plan.bops.size();
            ksv_op w;
            std::memset(&w, 0, sizeof w);
            w.kind = KSV_UNITARY;
            w.nq = (int32_t)first.tg.size();
            for (size_t j = 0; j < first.tg.size(); j++) w.q[j] = first.tg[j];
            w.mat = (int64_t)plan.bmats.size();
            for (const cd& z : first.m) {
                plan.bmats.push_back(z.real());
                plan.bmats.push_back(z.imag());
            }
            plan.bops.push_back(w);
            plan.steps.push_back(s);
            rem.erase(rem.begin());
            continue;
        }
        const uint64_t T = choose_tile(items, rem, n, K, forced, all);
        chosen.clear();
        rest.clear();
        take(items, rem, T, all, 0, &chosen, &rest);
        make_chains(items, chosen, T, order, run);
        make_groups(items, order, run, T, grp, tgt);
        AdjStep s;
        s.k = K;
        int pos[64];
        for (int& p : pos) p = -1;
        int j = 0;
        for (int qb = 0; qb < n; qb++)
            if ((T >> qb) & 1) {
                s.tq[j] = (int8_t)qb;
                pos[qb] = j++;
            }
        s.op_begin = (int)plan.aops.size();
        s.slot_begin = (int)plan.slot_rec.size();
        for (size_t x = 0; x < order.size(); x++) {
            if (grp[x] > 1) {
                // group header: the targets of its chains, in order, and how many
                // records follow
                AOp g;
                std::memset(&g, 0, sizeof g);
                g.kind = AK_GRP;
                g.k = grp[x];
                std::vector<int> ins;
                int members = 0;
                size_t y = x;
                for (int c = 0; c < grp[x]; c++) {
                    g.t[c] = (int8_t)pos[tgt[y]];
                    ins.push_back(pos[tgt[y]]);
                    members += std::max(run[y], 1);
                    y += (size_t)std::max(run[y], 1);
                }
                std::sort(ins.begin(), ins.end());
                g.nins = (int32_t)ins.size();
                for (size_t b = 0; b < ins.size(); b++) g.ins[b] = (int8_t)ins[b];
                g.run = (uint8_t)members;
                plan.aops.push_back(g);
                plan.aop_src.push_back(-1);
                plan.aop_rec.push_back({-1, -1, -1});
            }
            const AItem& it = items[order[x]];
            AOp a;
            std::memset(&a, 0, sizeof a);
            a.kind = it.kind;
            a.run = (uint8_t)run[x];
            a.flags = it.diag ? AOP_DIAG : 0;
            a.k = (int32_t)it.tg.size();
            std::vector<int> ins;
            for (size_t b = 0; b < it.tg.size(); b++) {
                a.t[b] = (int8_t)pos[it.tg[b]];
                ins.push_back(pos[it.tg[b]]);
            }
            for (int c : it.ctl) {
                if (pos[c] >= 0) {
                    a.lctl |= 1u << pos[c];
                    ins.push_back(pos[c]);
                } else {
                    a.gctl |= 1ull << c;
                }
            }
            std::sort(ins.begin(), ins.end());
            a.nins = (int32_t)ins.size();
            for (size_t b = 0; b < ins.size(); b++) a.ins[b] = (int8_t)ins[b];
            if (it.kind == AK_DENSE) {
                a.mat = (int32_t)(plan.mats.size() / 2);
                for (const cd& z : it.m) {
                    plan.mats.push_back(z.real());
                    plan.mats.push_back(z.imag());
                }
            }
            a.nrec = (int32_t)it.recs.size();
            plan.aop_src.push_back(it.src);
            std::array<int32_t, 3> ar = {-1, -1, -1};
            for (size_t r = 0; r < it.recs.size() && r < 3; r++) ar[r] = it.recs[r];
            plan.aop_rec.push_back(ar);
            const int dim = 1 << a.k;
            for (size_t r = 0; r < it.recs.size(); r++) {
                const ksv_param& p = params[it.recs[r]];
                if (p.dmat < 0 || (size_t)p.dmat + 2 * (size_t)dim * dim > nmats) {
                    err = "bad derivative matrix offset";
                    return KSV_ERR_TYPE;
                }
                const int pk = pauli && it.kind == AK_DENSE && dim == 2 ? pauli_rotation(it.m, mats + p.dmat) : -1;
                if (pk >= 0) {   // dG G^dag = (-i/2) P: the kernel needs no derivative matrix
                    a.dmat[r] = -1 - pk;
                } else {
                    a.dmat[r] = (int32_t)(plan.mats.size() / 2);
                    plan.mats.insert(plan.mats.end(), mats + p.dmat, mats + p.dmat + 2 * (size_t)dim * dim);
                }
                a.slot[r] = (int16_t)((int)plan.slot_rec.size() - s.slot_begin);
                plan.slot_rec.push_back(it.recs[r]);
            }
            plan.aops.push_back(a);
        }
        s.op_count = (int)plan.aops.size() - s.op_begin;
        s.slot_count = (int)plan.slot_rec.size() - s.slot_begin;
        plan.steps.push_back(s);
        rem.swap(rest);
    }
    if (plan.mats.size() / 2 > (size_t)INT32_MAX) {
        err = "adjoint tape too large";
        return KSV_ERR_RANGE;
    }
    return KSV_OK;
}

// csrswap's 4x4 on (a, b) where the control is 1 (src/statevector.ts:177-207)
static std::vector<cd> csrswap_block() {
    const double s = 1 / std::sqrt(2.0);
    std::vector<cd> m(16, cd(0, 0));
    m[0] = m[15] = 1;
    m[5] = m[10] = cd(s, 0);
    m[6] = m[9] = cd(0, s);
    return m;
}

int adjoint_plan_mats(const AdjPlan& plan, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                      const ksv_param* params, std::vector<double>& out, std::string& err) {
    out.clear();
    out.reserve(plan.mats.size());
    std::vector<cd> m, dm;
    for (size_t x = 0; x < plan.aops.size(); x++) {
        const AOp& a = plan.aops[x];
        if (a.kind == AK_GRP) continue;   // a group header carries no matrices
        const int src = plan.aop_src[x];
        if (src < 0 || (size_t)src >= nops) {
            err = "adjoint plan and tape differ";
            return KSV_ERR_TYPE;
        }
        const ksv_op& o = ops[src];
        const int dim = 1 << a.k;
        if (a.kind == AK_DENSE) {
            if (a.mat != (int32_t)(out.size() / 2)) {
                err = "adjoint plan and tape differ";
                return KSV_ERR_TYPE;
            }
            if (o.kind == KSV_CSRSWAP) {
                m = csrswap_block();
            } else if (!read_mat(mats, nmats, o.mat, dim, m)) {
                err = "op " + std::to_string(src) + ": bad matrix offset";
                return KSV_ERR_TYPE;
            }
            m = dagger(m, dim);
            if ((a.flags & AOP_DIAG) && !offdiag_zero(m, dim)) {
                err = "op " + std::to_string(src) + " is no longer diagonal: the batch must be created again";
                return KSV_ERR_TYPE;
            }
            for (const cd& z : m) {
                out.push_back(z.real());
                out.push_back(z.imag());
            }
        }
        for (int r = 0; r < a.nrec; r++) {
            const ksv_param& p = params[plan.aop_rec[x][r]];
            if (a.dmat[r] < 0) {
                err = "adjoint plan holds Pauli-rotation records (made for one parameter set)";
                return KSV_ERR_TYPE;
            }
            if (!read_mat(mats, nmats, p.dmat, dim, dm)) {
                err = "bad derivative matrix offset";
                return KSV_ERR_TYPE;
            }
            if ((a.flags & AOP_DIAG) && !offdiag_zero(dm, dim)) {
                err = "op " + std::to_string(src) + " is no longer diagonal: the batch must be created again";
                return KSV_ERR_TYPE;
            }
            for (const cd& z : dm) {
                out.push_back(z.real());
                out.push_back(z.imag());
            }
        }
    }
    if (out.size() != plan.mats.size()) {
        err = "adjoint plan and tape differ";
        return KSV_ERR_TYPE;
    }
    return KSV_OK;
}

void adjoint_plan_undiag(AdjPlan& plan, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                         const ksv_param* params) {
    std::vector<cd> m;
    for (size_t x = 0; x < plan.aops.size(); x++) {
        AOp& a = plan.aops[x];
        if (!(a.flags & AOP_DIAG)) continue;
        const int src = plan.aop_src[x];
        const int dim = 1 << a.k;
        bool diag = src >= 0 && (size_t)src < nops;
        if (diag && ops[src].kind != KSV_CSRSWAP) diag = read_mat(mats, nmats, ops[src].mat, dim, m) && offdiag_zero(m, dim);
        for (int r = 0; diag && r < a.nrec; r++)
            diag = read_mat(mats, nmats, params[plan.aop_rec[x][r]].dmat, dim, m) && offdiag_zero(m, dim);
        if (!diag) a.flags &= (uint8_t)~AOP_DIAG;
    }
}

int invert_tape(const ksv_op* ops, size_t nops, const double* mats, size_t nmats, std::vector<ksv_op>& iops,
                std::vector<double>& imats, std::string& err) {
    iops.clear();
    imats.clear();
    std::vector<cd> m;
    for (size_t i = nops; i-- > 0;) {
        ksv_op o = ops[i];
        int dim = 0;
        switch (o.kind) {
        case KSV_SINGLE: case KSV_CONTROLLED: dim = 2; break;
        case KSV_TWO: dim = 4; break;
        case KSV_UNITARY:
            if (o.nq < 1 || o.nq > KSV_MAX_OPERANDS) { err = "unitary: bad operand count"; return KSV_ERR_TYPE; }
            dim = 1 << o.nq;
            break;
        case KSV_CSRSWAP: {
            // the inverse as a 3-qubit unitary on (control, a, b): identity where the
            // control is 0, the block's conjugate transpose where it is 1
            const std::vector<cd> v = dagger(csrswap_block(), 4);
            ksv_op u;
            std::memset(&u, 0, sizeof u);
            u.kind = KSV_UNITARY;
            u.nq = 3;
            u.q[0] = o.q[0];
            u.q[1] = o.q[1];
            u.q[2] = o.q[2];
            u.mat = (int64_t)imats.size();
            for (int r = 0; r < 8; r++)
                for (int c = 0; c < 8; c++) {
                    cd z = r < 4 ? (r == c ? cd(1, 0) : cd(0, 0)) : (c < 4 ? cd(0, 0) : v[(size_t)(r - 4) * 4 + (c - 4)]);
                    imats.push_back(z.real());
                    imats.push_back(z.imag());
                }
            iops.push_back(u);
            continue;
        }
        default: break;     // permutations and barriers are their own inverses
        }
        if (dim) {
            if (!read_mat(mats, nmats, o.mat, dim, m)) { err = "op " + std::to_string(i) + ": bad matrix offset"; return KSV_ERR_TYPE; }
            m = dagger(m, dim);
            o.mat = (int64_t)imats.size();
            for (const cd& z : m) {
                imats.push_back(z.real());
                imats.push_back(z.imag());
            }
        } else {
            o.mat = -1;
        }
        iops.push_back(o);
    }
    return KSV_OK;
}

}  // namespace ksv
