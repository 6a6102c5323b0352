// Host emulation of libketsv's scheduled programs (test infrastructure).
//
// Runs the exact program that ksv_program_create would upload (the same
// plan_program() from ket_amd/csrc/schedule.cpp) on a host FP64 state,
// executing every step the way its kernel does: register-round passes thread
// by thread with 16 registers each and the LDS transposes through their
// swizzled addresses, LDS-tile passes element by element, wide unitaries
// group by group.  Every HBM index and every LDS address is checked: in
// range, each tile loaded and stored exactly once, every LDS slot of an
// exchange written exactly once.  The CPU tests compare the result with the
// oracle, so scheduler and addressing mistakes show up without a GPU.
#include <algorithm>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "ksv_internal.h"

using namespace ksv;

static std::string g_msg;
static int g_bthreads = 256;     // fused-block workgroup size emulated

static uint64_t ins0(uint64_t p, int q) { return ((p >> q) << (q + 1)) | (p & ((1ull << q) - 1)); }

static bool prune_ok(cd& v, int prune) {
    if (prune && (v.real() * v.real() + v.imag() * v.imag()) < 1e-14) v = 0;
    return true;
}

#define CHECK(cond, msg)                     \
    do {                                     \
        if (!(cond)) {                       \
            g_msg = std::string(msg);        \
            return -1;                       \
        }                                    \
    } while (0)

static int emu_reg(const Program& p, const Step& s, cd* psi, int n) {
    const int K = s.k, TB = K - RSLOTS, NT = 1 << TB;
    const uint64_t N = 1ull << n;
    const ROp* ops = p.rops.data() + s.op_begin;
    const int nops = s.op_count;
    std::vector<cd> regs((size_t)NT * RREGS), lds((size_t)1 << K);
    std::vector<int> mark((size_t)1 << K);
    std::vector<uint64_t> ld_idx, st_idx;
    uint64_t tmask = 0;
    for (int i = 0; i < K; i++) tmask |= 1ull << s.tq[i];
    for (uint64_t b = 0; b < (N >> K); b++) {
        uint64_t base = b;
        for (int i = 0; i < K; i++) base = ins0(base, s.tq[i]);
        CHECK((base & tmask) == 0, "base overlaps the tile");
        ld_idx.clear();
        st_idx.clear();
        for (int t = 0; t < NT; t++) {
            uint64_t lo = base;
            for (int i = 0; i < TB; i++) lo |= (uint64_t)((t >> i) & 1) << s.ld_t[i];
            for (int j = 0; j < RREGS; j++) {
                uint64_t idx = lo;
                for (int q = 0; q < RSLOTS; q++)
                    if ((j >> q) & 1) idx |= 1ull << s.ld_r[q];
                CHECK(idx < N, "load index out of range");
                CHECK((idx & ~tmask) == base, "load index leaves the tile");
                ld_idx.push_back(idx);
                regs[(size_t)t * RREGS + j] = psi[idx];
            }
        }
        for (int o = 0; o < nops; o++) {
            const ROp& op = ops[o];
            if (op.kind == R_EXCH) {
                XRec x;
                std::memcpy(&x, op.pay, sizeof x);
                std::fill(mark.begin(), mark.end(), 0);
                for (int t = 0; t < NT; t++) {
                    uint32_t wb = 0;
                    for (int i = 0; i < TB; i++)
                        if ((t >> i) & 1) wb ^= x.wt[i];
                    for (int j = 0; j < RREGS; j++) {
                        uint32_t a = wb;
                        for (int q = 0; q < RSLOTS; q++)
                            if ((j >> q) & 1) a ^= x.wr[q];
                        CHECK(a < (1u << K), "LDS write address out of range");
                        mark[a]++;
                        lds[a] = regs[(size_t)t * RREGS + j];
                    }
                }
                for (int a = 0; a < (1 << K); a++) CHECK(mark[a] == 1, "LDS slot not written exactly once");
                for (int t = 0; t < NT; t++) {
                    uint32_t rb = 0;
                    for (int i = 0; i < TB; i++)
                        if ((t >> i) & 1) rb ^= x.rt[i];
                    for (int j = 0; j < RREGS; j++) {
                        uint32_t a = rb;
                        for (int q = 0; q < RSLOTS; q++)
                            if ((j >> q) & 1) a ^= x.rr[q];
                        CHECK(a < (1u << K), "LDS read address out of range");
                        regs[(size_t)t * RREGS + j] = lds[a];
                    }
                }
                continue;
            }
            if (!p.rbits.empty() && !(op.flags & ROP_COND)) {
                const uint64_t copy = base >> p.nsys;
                const size_t at = (size_t)(copy >> 5) * p.rops.size() + s.op_begin + o;
                CHECK(at < p.rbits.size() && p.rbits[at] == ~0u, "unconditional op without an all-ones by-op word");
            }
            if ((base & op.gctl) != op.gctl) continue;
            if (op.flags & ROP_COND) {      // `mat` is the copy bitmap's word offset
                CHECK(op.kind == R_M1 || op.kind == R_PERMX || (op.kind == R_DIAG && op.k <= 2), "conditional op of a kind with a matrix offset");
                const uint64_t copy = base >> p.nsys;
                CHECK(op.mat >= 0 && (size_t)op.mat + (copy >> 5) < p.bits.size(), "copy bitmap out of range");
                // the kernel reads the by-op copy of the bitmaps
                const size_t g = (size_t)s.op_begin + o, at = (size_t)(copy >> 5) * p.rops.size() + g;
                CHECK(at < p.rbits.size(), "by-op copy bitmap out of range");
                CHECK(p.rbits[at] == p.bits[(size_t)op.mat + (copy >> 5)], "by-op copy bitmap differs");
                if (!((p.rbits[at] >> (copy & 31)) & 1u)) continue;
            }
            CHECK((op.flags & ROP_COND) || (op.mat >= 0 && (size_t)op.mat <= p.mats.size()), "matrix offset out of range");
            const cd* wide = p.mats.data() + ((op.flags & ROP_COND) ? 0 : op.mat);
            cd inl[4];
            {
                const double* d = reinterpret_cast<const double*>(op.pay);
                for (int i = 0; i < 4; i++) inl[i] = cd(d[2 * i], d[2 * i + 1]);
            }
            for (int t = 0; t < NT; t++) {
                if ((t & op.tctl) != op.tctl) continue;
                cd* r = regs.data() + (size_t)t * RREGS;
                switch (op.kind) {
                case R_M1: {
                    CHECK(op.a < RSLOTS, "bad slot");
                    const cd* m = inl;
                    for (int j = 0; j < RREGS; j++) {
                        if ((j >> op.a) & 1) continue;
                        if (((uint32_t)j & op.rctl) != op.rctl) continue;
                        const int j1 = j | (1 << op.a);
                        cd x0 = r[j], x1 = r[j1];
                        cd y0 = m[0] * x0 + m[1] * x1, y1 = m[2] * x0 + m[3] * x1;
                        prune_ok(y0, p.prune);
                        prune_ok(y1, p.prune);
                        r[j] = y0;
                        r[j1] = y1;
                    }
                    break;
                }
                case R_DENSE: {
                    const int k = op.k, dim = 1 << k;
                    CHECK(k >= 2 && k <= RSLOTS, "bad dense width");
                    uint32_t tm = 0;
                    for (int i = 0; i < k; i++) { CHECK(op.loc[i] < RSLOTS, "bad dense slot"); tm |= 1u << op.loc[i]; }
                    CHECK(__builtin_popcount(tm) == k, "dense slots repeat");
                    CHECK((size_t)op.mat + (size_t)dim * dim <= p.mats.size(), "dense matrix out of range");
                    cd in[RREGS];
                    for (int j = 0; j < RREGS; j++) in[j] = r[j];
                    for (int j = 0; j < RREGS; j++) {
                        if (((uint32_t)j & op.rctl) != op.rctl) continue;
                        int row = 0;
                        for (int i = 0; i < k; i++) row |= ((j >> op.loc[i]) & 1) << (k - 1 - i);
                        cd acc = 0;
                        for (int c = 0; c < dim; c++) {
                            int jc = j & ~(int)tm;
                            for (int i = 0; i < k; i++) jc |= ((c >> (k - 1 - i)) & 1) << op.loc[i];
                            acc += wide[row * dim + c] * in[jc];
                        }
                        prune_ok(acc, p.prune);
                        r[j] = acc;
                    }
                    break;
                }
                case R_DIAG: {
                    const int k = op.k;
                    for (int j = 0; j < RREGS; j++) {
                        if (((uint32_t)j & op.rctl) != op.rctl) continue;
                        uint32_t row = 0;
                        for (int i = 0; i < k; i++) {
                            const int L = op.loc[i];
                            int bit;
                            if (L < 16) { CHECK(L < TB, "bad thread bit"); bit = (t >> L) & 1; }
                            else if (L < 32) { CHECK(L - 16 < RSLOTS, "bad slot"); bit = (j >> (L - 16)) & 1; }
                            else { CHECK(L - 32 < n, "bad global bit"); bit = (int)((base >> (L - 32)) & 1); }
                            row |= (uint32_t)bit << (k - 1 - i);
                        }
                        cd y = (k <= 2 ? inl[row] : wide[row]) * r[j];
                        prune_ok(y, p.prune);
                        r[j] = y;
                    }
                    break;
                }
                case R_PERMX: {
                    CHECK(op.a < RSLOTS, "bad slot");
                    for (int j = 0; j < RREGS; j++) {
                        if ((j >> op.a) & 1) continue;
                        if (((uint32_t)j & op.rctl) != op.rctl) continue;
                        std::swap(r[j], r[j | (1 << op.a)]);
                    }
                    break;
                }
                case R_SWAP: {
                    CHECK(op.a < RSLOTS && op.b < op.a, "bad slot pair");
                    for (int j = 0; j < RREGS; j++) {
                        if (!((j >> op.a) & 1) || ((j >> op.b) & 1)) continue;
                        if (((uint32_t)j & op.rctl) != op.rctl) continue;
                        std::swap(r[j], r[j ^ (1 << op.a) ^ (1 << op.b)]);
                    }
                    break;
                }
                default:
                    CHECK(false, "unknown register-round op");
                }
            }
        }
        for (int t = 0; t < NT; t++) {
            uint64_t lo = base;
            for (int i = 0; i < TB; i++) lo |= (uint64_t)((t >> i) & 1) << s.st_t[i];
            for (int j = 0; j < RREGS; j++) {
                uint64_t idx = lo;
                for (int q = 0; q < RSLOTS; q++)
                    if ((j >> q) & 1) idx |= 1ull << s.st_r[q];
                CHECK(idx < N, "store index out of range");
                st_idx.push_back(idx);
                psi[idx] = regs[(size_t)t * RREGS + j];
            }
        }
        std::sort(ld_idx.begin(), ld_idx.end());
        std::sort(st_idx.begin(), st_idx.end());
        CHECK(std::adjacent_find(ld_idx.begin(), ld_idx.end()) == ld_idx.end(), "an amplitude is loaded twice");
        CHECK(ld_idx == st_idx, "store set differs from load set");
        CHECK(ld_idx.size() == ((size_t)1 << K), "tile size mismatch");
    }
    return 0;
}

static int emu_lds(const Program& p, const Step& s, cd* psi, int n) {
    const int K = s.k;
    const uint64_t N = 1ull << n;
    const uint32_t tsize = 1u << K;
    std::vector<cd> T(tsize);
    for (uint64_t b = 0; b < (N >> K); b++) {
        uint64_t base = b;
        for (int i = 0; i < K; i++) base = ins0(base, s.tq[i]);
        auto scatter = [&](uint32_t l) {
            uint64_t off = 0;
            for (int i = 0; i < K; i++) off |= (uint64_t)((l >> i) & 1) << s.tq[i];
            return off;
        };
        for (uint32_t l = 0; l < tsize; l++) {
            CHECK((base | scatter(l)) < N, "LDS pass index out of range");
            T[l] = psi[base | scatter(l)];
        }
        for (int o = 0; o < s.op_count; o++) {
            const DOp& op = p.dops[s.op_begin + o];
            if ((base & op.gctl) != op.gctl) continue;
            if (op.cond) {
                const uint64_t copy = base >> p.nsys;
                CHECK(op.cond - 1 + (copy >> 5) < p.bits.size(), "copy bitmap out of range");
                if (!((p.bits[op.cond - 1 + (copy >> 5)] >> (copy & 31)) & 1u)) continue;
            }
            const cd* m = p.mats.data() + op.mat;
            const uint32_t ne = tsize >> op.nins;
            for (uint32_t e = 0; e < ne; e++) {
                uint32_t l = e;
                for (int i = 0; i < op.nins; i++) l = (uint32_t)ins0(l, op.ins[i]);
                l |= op.lctl;
                CHECK(l < tsize, "LDS element out of range");
                switch (op.kind) {
                case DK_M1: {
                    uint32_t l1 = l | (1u << op.t[0]);
                    cd x0 = T[l], x1 = T[l1];
                    cd y0 = m[0] * x0 + m[1] * x1, y1 = m[2] * x0 + m[3] * x1;
                    prune_ok(y0, p.prune);
                    prune_ok(y1, p.prune);
                    T[l] = y0;
                    T[l1] = y1;
                    break;
                }
                case DK_M2:
                case DK_MK: {
                    const int kk = op.k, dim = 1 << kk;
                    std::vector<uint32_t> id(dim);
                    std::vector<cd> x(dim);
                    for (int r = 0; r < dim; r++) {
                        uint32_t g = l;
                        for (int bb = 0; bb < kk; bb++)
                            if ((r >> (kk - 1 - bb)) & 1) g |= 1u << op.t[bb];
                        id[r] = g;
                        x[r] = T[g];
                    }
                    for (int r = 0; r < dim; r++) {
                        cd acc = 0;
                        for (int c = 0; c < dim; c++) acc += m[r * dim + c] * x[c];
                        prune_ok(acc, p.prune);
                        T[id[r]] = acc;
                    }
                    break;
                }
                case DK_DIAG: {
                    uint32_t r = 0;
                    for (int bb = 0; bb < op.k; bb++) {
                        int bit;
                        if (op.t[bb] >= 0) bit = (l >> op.t[bb]) & 1;
                        else bit = (int)((base >> (-op.t[bb] - 1)) & 1);
                        r |= (uint32_t)bit << (op.k - 1 - bb);
                    }
                    cd y = m[r] * T[l];
                    prune_ok(y, p.prune);
                    T[l] = y;
                    break;
                }
                case DK_PERMX:
                    std::swap(T[l], T[l | (1u << op.t[0])]);
                    break;
                case DK_SWAP:
                    std::swap(T[l | (1u << op.t[0])], T[l | (1u << op.t[1])]);
                    break;
                }
            }
        }
        for (uint32_t l = 0; l < tsize; l++) psi[base | scatter(l)] = T[l];
    }
    return 0;
}

static uint32_t emu_swz(uint64_t w, uint32_t x) {
    uint32_t g = 0;
    for (int p = 4; p < 20; p++)
        if ((x >> p) & 1) g ^= (uint32_t)((w >> (4 * (p - 4))) & 15);
    return x ^ g;
}

// Fused-block pass, lane by lane as k_bpass runs it; MFMA 16x16x4 f64
// operand layout (measured on gfx950): A[i][k] in lane i + 16k, B[k][j] in
// lane j + 16k, D[i][j] in lane j + 16(i % 4), element i / 4.
static int emu_blk(const Program& p, const Step& s, cd* psi, int n) {
    const int K = s.k, bq = s.bq, R = 1 << bq, MB = 2 * R / 16, KS = 2 * R / 4;
    const int NT = g_bthreads, NW = NT / 64, NC = 1 << (K - bq), NBW = NC / 16 / NW, PER = (1 << K) / NT;
    const uint64_t N = 1ull << n;
    CHECK(NBW >= 1, "tile too small for the block width");
    std::vector<cd> lds((size_t)1 << K);
    std::vector<int> rd((size_t)1 << K), wr((size_t)1 << K);
    std::vector<std::vector<double>> fr(s.op_count);
    const size_t nf = (size_t)MB * KS * 64;
    for (int i = 0; i < s.op_count; i++) {
        fr[i].resize(nf);
        block_fragments(p.bmats.data() + p.bmat_off[s.op_begin + i], bq, p.brecs[s.op_begin + i].flags,
                        fr[i].data());
    }
    for (uint64_t b = 0; b < (N >> K); b++) {
        uint64_t base = b;
        for (int i = 0; i < K; i++) base = ins0(base, s.tq[i]);
        std::fill(wr.begin(), wr.end(), 0);
        for (int t = 0; t < NT; t++)
            for (int i = 0; i < PER; i++) {
                const uint32_t x = (uint32_t)t + (uint32_t)i * NT;
                uint64_t g = base;
                for (int q = 0; q < K; q++)
                    if ((x >> q) & 1) g |= 1ull << s.tq[q];
                const uint32_t la = emu_swz(s.swz, x);
                CHECK(g < N && la < (1u << K), "block pass load out of range");
                wr[la]++;
                lds[la] = psi[g];
            }
        for (uint32_t a = 0; a < (1u << K); a++) CHECK(wr[a] == 1, "block pass LDS slot not loaded exactly once");
        for (int bi = 0; bi < s.op_count; bi++) {
            const BRec& r = p.brecs[s.op_begin + bi];
            const double* af = fr[bi].data();
            std::fill(rd.begin(), rd.end(), 0);
            std::fill(wr.begin(), wr.end(), 0);
            std::vector<cd> out(lds);
            if (r.flags & BREC_DIAG) {
                // phase per row: lane l of wave w, slice m reads rows 4t + l/16 of column
                // (l % 16, slice bits) and writes them back times the row's phase
                for (int wv = 0; wv < NW; wv++)
                    for (int m = 0; m < NBW; m++) {
                        const uint32_t nb = (uint32_t)wv + (uint32_t)NW * (uint32_t)m;
                        for (int l = 0; l < 64; l++) {
                            const uint32_t rl = l >> 4, cl = l & 15;
                            uint32_t col = 0;
                            for (int bb = 0; bb < 4; bb++)
                                if ((cl >> bb) & 1) col ^= r.ac[bb];
                            for (int bb = 0; bb < K - bq - 4; bb++)
                                if ((nb >> bb) & 1) col ^= r.ac[4 + bb];
                            for (int t = 0; t < R / 4; t++) {
                                const uint32_t row = 4u * (uint32_t)t + rl;
                                uint32_t ra = 0;
                                for (int bb = 0; bb < bq; bb++)
                                    if ((row >> bb) & 1) ra ^= r.ar[bb];
                                const uint32_t a = ra ^ col;
                                CHECK(a < (1u << K), "diagonal block address out of range");
                                rd[a]++;
                                wr[a]++;
                                const cd ph(af[(size_t)(2 * t) * 64 + l], af[(size_t)(2 * t + 1) * 64 + l]);
                                out[a] = lds[a] * ph;
                            }
                        }
                    }
                for (uint32_t a = 0; a < (1u << K); a++)
                    CHECK(rd[a] == 1 && wr[a] == 1, "diagonal block LDS slot not read and written exactly once");
                lds.swap(out);
                continue;
            }
            if (r.flags & BREC_MONO) {
                // one entry per column: lane l reads rows 4t + l/16 of its column and
                // writes row pi(row) (nibble row of ar[4..7]) times U[pi(row)][row]
                CHECK(bq == 4, "monomial flag on a 5-qubit block");
                uint64_t perm = 0;
                for (int i = 0; i < 4; i++) perm |= (uint64_t)r.ar[4 + i] << (16 * i);
                const cd* u = p.bmats.data() + p.bmat_off[s.op_begin + bi];
                for (int wv = 0; wv < NW; wv++)
                    for (int m = 0; m < NBW; m++) {
                        const uint32_t nb = (uint32_t)wv + (uint32_t)NW * (uint32_t)m;
                        for (int l = 0; l < 64; l++) {
                            const uint32_t rl = l >> 4, cl = l & 15;
                            uint32_t col = 0;
                            for (int bb = 0; bb < 4; bb++)
                                if ((cl >> bb) & 1) col ^= r.ac[bb];
                            for (int bb = 0; bb < K - bq - 4; bb++)
                                if ((nb >> bb) & 1) col ^= r.ac[4 + bb];
                            for (int t = 0; t < R / 4; t++) {
                                const uint32_t row = 4u * (uint32_t)t + rl;
                                const uint32_t pr = (uint32_t)(perm >> (4 * row)) & 15u;
                                uint32_t ra = 0, wa = 0;
                                for (int bb = 0; bb < bq; bb++) {
                                    if ((row >> bb) & 1) ra ^= r.ar[bb];
                                    if ((pr >> bb) & 1) wa ^= r.ar[bb];
                                }
                                const cd ph(af[(size_t)(2 * t) * 64 + l], af[(size_t)(2 * t + 1) * 64 + l]);
                                CHECK(ph == u[(size_t)pr * R + row], "monomial phase is not U[pi(r)][r]");
                                CHECK((ra ^ col) < (1u << K) && (wa ^ col) < (1u << K), "monomial block address out of range");
                                rd[ra ^ col]++;
                                wr[wa ^ col]++;
                                out[wa ^ col] = lds[ra ^ col] * ph;
                            }
                        }
                    }
                for (uint32_t a = 0; a < (1u << K); a++)
                    CHECK(rd[a] == 1 && wr[a] == 1, "monomial block LDS slot not read and written exactly once");
                lds.swap(out);
                continue;
            }
            for (int wv = 0; wv < NW; wv++)
                for (int m = 0; m < NBW; m++) {
                    const uint32_t nb = (uint32_t)wv + (uint32_t)NW * (uint32_t)m;
                    // per-lane gathered B values and addresses
                    std::vector<std::vector<cd>> x(64, std::vector<cd>(R / 4));
                    std::vector<uint32_t> colv(64);
                    for (int l = 0; l < 64; l++) {
                        const uint32_t rl = l >> 4, cl = l & 15;
                        uint32_t col = 0;
                        for (int bb = 0; bb < 4; bb++)
                            if ((cl >> bb) & 1) col ^= r.ac[bb];
                        for (int bb = 0; bb < K - bq - 4; bb++)
                            if ((nb >> bb) & 1) col ^= r.ac[4 + bb];
                        colv[l] = col;
                        for (int t = 0; t < R / 4; t++) {
                            const uint32_t row = 4u * (uint32_t)t + rl;
                            uint32_t ra = 0;
                            for (int bb = 0; bb < bq; bb++)
                                if ((row >> bb) & 1) ra ^= r.ar[bb];
                            const uint32_t a = ra ^ col;
                            CHECK(a < (1u << K), "block read address out of range");
                            rd[a]++;
                            x[l][t] = lds[a];
                        }
                    }
                    // MFMA accumulate: Gauss's three products T1 = Ur Xr, T2 = Ui Xi,
                    // T3 = (Ur + Ui)(Xr + Xi), fragments [which][mb][kk][lane]
                    const int MB3 = R / 16, KS3 = R / 4;
                    std::vector<std::vector<std::vector<double>>> T(
                        3, std::vector<std::vector<double>>(MB3, std::vector<double>(256, 0.0)));   // T[w][mb][i*16+j]
                    for (int w = 0; w < 3; w++)
                        for (int kk = 0; kk < KS3; kk++)
                            for (int mb = 0; mb < MB3; mb++)
                                for (int i = 0; i < 16; i++)
                                    for (int j = 0; j < 16; j++) {
                                        double acc = 0;
                                        for (int k = 0; k < 4; k++) {
                                            const double av = af[(((size_t)w * MB3 + mb) * KS3 + kk) * 64 + (i + 16 * k)];
                                            const cd xv = x[j + 16 * k][kk];
                                            const double bv = w == 0 ? xv.real() : w == 1 ? xv.imag() : xv.real() + xv.imag();
                                            acc += av * bv;
                                        }
                                        T[w][mb][i * 16 + j] += acc;
                                    }
                    for (int l = 0; l < 64; l++) {
                        const uint32_t rl = l >> 4;
                        for (int h = 0; h < R / 16; h++)
                            for (int v = 0; v < 4; v++) {
                                const uint32_t row = 16u * (uint32_t)h + 4u * (uint32_t)v + rl;
                                uint32_t wa = 0;
                                for (int bb = 0; bb < bq; bb++)
                                    if ((row >> bb) & 1) wa ^= r.ar[bb];
                                const uint32_t a = wa ^ colv[l];
                                CHECK(a < (1u << K), "block write address out of range");
                                wr[a]++;
                                const int i = 4 * v + (int)rl, j = l & 15;
                                const double t1 = T[0][h][i * 16 + j], t2 = T[1][h][i * 16 + j], t3 = T[2][h][i * 16 + j];
                                out[a] = cd(t1 - t2, t3 - t1 - t2);
                            }
                    }
                }
            for (uint32_t a = 0; a < (1u << K); a++)
                CHECK(rd[a] == 1 && wr[a] == 1, "block LDS slot not read and written exactly once");
            lds.swap(out);
        }
        for (int t = 0; t < NT; t++)
            for (int i = 0; i < PER; i++) {
                const uint32_t x = (uint32_t)t + (uint32_t)i * NT;
                uint64_t g = base;
                for (int q = 0; q < K; q++)
                    if ((x >> q) & 1) g |= 1ull << s.tq[q];
                psi[g] = lds[emu_swz(s.swz, x)];
            }
    }
    return 0;
}

static int emu_big(const Program& p, const Step& s, cd* psi, int n) {
    const BigOp& b = p.big_ops[s.big_op];
    const int k = (int)b.tg.size(), dim = 1 << k;
    std::vector<int> srt = b.tg;
    std::sort(srt.begin(), srt.end());
    const cd* m = p.mats.data() + b.mat;
    std::vector<cd> x(dim);
    std::vector<uint64_t> gi(dim);
    for (uint64_t g = 0; g < (1ull << (n - k)); g++) {
        uint64_t base = g;
        for (int q : srt) base = ins0(base, q);
        for (int r = 0; r < dim; r++) {
            uint64_t idx = base;
            for (int bb = 0; bb < k; bb++)
                if ((r >> (k - 1 - bb)) & 1) idx |= 1ull << b.tg[bb];
            gi[r] = idx;
            x[r] = psi[idx];
        }
        for (int r = 0; r < dim; r++) {
            cd acc = 0;
            for (int c = 0; c < dim; c++) acc += m[(size_t)r * dim + c] * x[c];
            prune_ok(acc, p.prune);
            psi[gi[r]] = acc;
        }
    }
    return 0;
}

extern "C" {

void emu_set_block_threads(int nt) { g_bthreads = nt; }

int emu_last_error(char* buf, size_t len) {
    if (buf && len) {
        size_t m = std::min(len - 1, g_msg.size());
        std::memcpy(buf, g_msg.data(), m);
        buf[m] = 0;
    }
    return (int)g_msg.size();
}

// stats[0] passes, [1] register-round passes, [2] LDS transposes, [3] device ops, [4] fused-block passes,
// [5] fused blocks
// emu_run_updated: after scheduling, tape op upd_ops[u] gets the matrix
// upd_mats[u * stride ...] through update_program (ksv_program_set_matrix's host side).
int emu_run_updated(int n, int prune, const ksv_op* ops, size_t nops, const double* mats, size_t nmats, int kernel,
                    int tile, int rtile, int fuse, double* psi, int64_t* stats, int btile, int bq,
                    const int64_t* upd_ops, const double* upd_mats, size_t nupd, size_t stride, size_t upd_nmat) {
    Program p;
    std::string err;
    int rc = plan_program(p, n, KSV_F64, ops, nops, mats, nmats, kernel, tile, rtile, fuse != 0, prune, err, btile,
                          bq);
    if (rc) {
        g_msg = err;
        return rc;
    }
    for (size_t u = 0; u < nupd; u++) {
        std::vector<int> r, b;
        std::vector<std::pair<size_t, size_t>> rg;
        rc = update_program(p, (size_t)upd_ops[u], upd_mats + u * stride, upd_nmat, r, b, rg, err);
        if (rc) {
            g_msg = err;
            return rc;
        }
    }
    cd* st = reinterpret_cast<cd*>(psi);
    int64_t regs = 0, exch = 0, blks = 0;
    for (const Step& s : p.steps) {
        int r;
        if (s.big) r = emu_big(p, s, st, n);
        else if (s.reg) { r = emu_reg(p, s, st, n); regs++; exch += s.exchanges; }
        else if (s.blk) { r = emu_blk(p, s, st, n); blks++; }
        else r = emu_lds(p, s, st, n);
        if (r) return r;
    }
    if (stats) {
        stats[0] = (int64_t)p.steps.size();
        stats[1] = regs;
        stats[2] = exch;
        int64_t dev = (int64_t)(p.dops.size() + p.big_ops.size());
        for (const ROp& r : p.rops)
            if (r.kind != R_EXCH) dev++;
        stats[3] = dev;
        stats[4] = blks;
        stats[5] = (int64_t)p.brecs.size();
    }
    return 0;
}

// Fused-block plan of a tape in exact mode: kinds[0] dense, [1] diagonal, [2] monomial blocks.
int emu_block_kinds(int n, const ksv_op* ops, size_t nops, const double* mats, size_t nmats, int btile, int bq,
                    int64_t* kinds) {
    Program p;
    std::string err;
    int rc = plan_program(p, n, KSV_F64, ops, nops, mats, nmats, 3, 0, 0, true, false, err, btile, bq);
    if (rc) {
        g_msg = err;
        return rc;
    }
    kinds[0] = kinds[1] = kinds[2] = 0;
    for (const BRec& r : p.brecs) kinds[(r.flags & BREC_DIAG) ? 1 : ((r.flags & BREC_MONO) ? 2 : 0)]++;
    return 0;
}

int emu_run(int n, int prune, const ksv_op* ops, size_t nops, const double* mats, size_t nmats, int kernel, int tile,
            int rtile, int fuse, double* psi, int64_t* stats, int btile, int bq) {
    return emu_run_updated(n, prune, ops, nops, mats, nmats, kernel, tile, rtile, fuse, psi, stats, btile, bq, nullptr,
                           nullptr, 0, 0, 0);
}

}  // extern "C"

// ---- adjoint backward sweep -------------------------------------------------

static void emu_apply_dense(int n, const std::vector<int>& tg, const cd* m, cd* v) {
    const int k = (int)tg.size(), dim = 1 << k;
    uint64_t tmask = 0;
    for (int q : tg) tmask |= 1ull << q;
    std::vector<cd> x(dim);
    std::vector<uint64_t> id(dim);
    for (uint64_t b = 0; b < (1ull << n); b++) {
        if (b & tmask) continue;
        for (int r = 0; r < dim; r++) {
            uint64_t g = b;
            for (int i = 0; i < k; i++)
                if ((r >> (k - 1 - i)) & 1) g |= 1ull << tg[i];
            id[r] = g;
            x[r] = v[g];
        }
        for (int r = 0; r < dim; r++) {
            cd y = 0;
            for (int c = 0; c < dim; c++) y += m[(size_t)r * dim + c] * x[c];
            v[id[r]] = y;
        }
    }
}

// The backward sweep of ksv_adjoint as k_adjpass runs it: the same plan
// (plan_adjoint), every pass tile by tile and every op element group by
// element group on host copies of phi and lambda, wide steps applied
// directly.  Every index is checked; every amplitude is loaded and stored
// once per pass; every record is owned by exactly one pass.  sums[r] gets
// Re <lambda| dG_r |G^dag phi> of record r.  stats: passes, wide steps,
// ops in the largest pass, ops inside chains (not heads), diagonal ops.
extern "C" int emu_adjoint(int n, int K, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                           const ksv_param* params, size_t nrec, double* phi_io, double* lam_io, double* sums,
                           int* stats) {
    AdjPlan plan;
    std::string err;
    int rc = plan_adjoint(n, K, ops, nops, mats, nmats, params, nrec, plan, err, true);
    if (rc) {
        g_msg = err;
        return rc;
    }
    cd* phi = reinterpret_cast<cd*>(phi_io);
    cd* lam = reinterpret_cast<cd*>(lam_io);
    const uint64_t N = 1ull << n;
    std::vector<int> owner(nrec, 0);
    for (int32_t r : plan.slot_rec) {
        CHECK(r >= 0 && (size_t)r < nrec, "slot names a missing record");
        owner[r]++;
    }
    for (size_t r = 0; r < nrec; r++) CHECK(owner[r] == 1, "record not owned by exactly one pass");
    for (size_t r = 0; r < nrec; r++) sums[r] = 0;
    int passes = 0, bigs = 0, maxops = 0;
    std::vector<char> seen(N);
    for (const AdjStep& s : plan.steps) {
        if (s.big) {
            bigs++;
            const ksv_op& w = plan.bops[s.bop];
            CHECK(w.kind == KSV_UNITARY && w.nq > ADJ_DENSE_MAX && w.nq <= KSV_MAX_OPERANDS, "bad wide step");
            CHECK(w.mat >= 0 && (size_t)w.mat + (2ull << (2 * w.nq)) <= plan.bmats.size(), "wide matrix out of range");
            std::vector<int> tg(w.q, w.q + w.nq);
            const cd* m = reinterpret_cast<const cd*>(plan.bmats.data() + w.mat);
            emu_apply_dense(n, tg, m, phi);
            emu_apply_dense(n, tg, m, lam);
            continue;
        }
        passes++;
        maxops = std::max(maxops, s.op_count);
        const int k = s.k;
        CHECK(k == plan.k && k >= 1 && k <= n && k <= 12, "bad tile width");
        CHECK(s.slot_count <= ADJ_SLOTS && s.slot_begin + s.slot_count <= (int)plan.slot_rec.size(), "bad slots");
        CHECK(s.op_begin + s.op_count <= (int)plan.aops.size(), "bad op range");
        const uint32_t ts = 1u << k;
        uint64_t tmask = 0;
        for (int i = 0; i < k; i++) {
            CHECK(s.tq[i] >= 0 && s.tq[i] < n, "tile bit out of range");
            CHECK(i == 0 || s.tq[i] > s.tq[i - 1], "tile bits not ascending");
            tmask |= 1ull << s.tq[i];
        }
        std::fill(seen.begin(), seen.end(), 0);
        std::vector<cd> P(ts), L(ts);
        std::vector<double> red(s.slot_count, 0.0);
        for (uint64_t tile = 0; tile < (N >> k); tile++) {
            uint64_t base = tile;
            for (int i = 0; i < k; i++) base = ins0(base, s.tq[i]);
            CHECK((base & tmask) == 0, "base overlaps the tile");
            for (uint32_t l = 0; l < ts; l++) {
                uint64_t off = base;
                for (int i = 0; i < k; i++)
                    if ((l >> i) & 1) off |= 1ull << s.tq[i];
                CHECK(off < N && !seen[off], "tile load out of range or repeated");
                seen[off] = 1;
                P[l] = phi[off];
                L[l] = lam[off];
            }
            for (int o = 0; o < s.op_count; o++) {
                // an X with local controls may head a chain of one-qubit ops on its
                // target: each pair is swapped where the controls are set, then the
                // chain runs on it
                uint32_t xctl = 0;
                int xpos = -1;
                const AOp& h = plan.aops[s.op_begin + o];
                if (h.kind == AK_GRP) {
                    // one-qubit chains on distinct targets, run in turn on the 2^M
                    // elements of each group over those targets
                    const int M = h.k;
                    CHECK((M == 2 || M == 3) && h.nins == M && h.gctl == 0 && h.lctl == 0 && h.run >= M &&
                              o + h.run < s.op_count,
                          "bad group header");
                    for (int i = 0; i < M; i++) {
                        CHECK(h.ins[i] >= 0 && h.ins[i] < k && (i == 0 || h.ins[i] > h.ins[i - 1]), "bad group bits");
                        CHECK(h.t[i] >= 0 && h.t[i] < k, "group target outside the tile");
                    }
                    const int D = 1 << M;
                    for (uint32_t e = 0; e < (ts >> M); e++) {
                        uint32_t l = e;
                        for (int i = 0; i < M; i++) l = (uint32_t)ins0(l, h.ins[i]);
                        uint32_t id[8];
                        cd x[8], y[8], a[8], b[8];
                        for (int r = 0; r < D; r++) {
                            uint32_t g = l;
                            for (int bb = 0; bb < M; bb++)
                                if ((r >> (M - 1 - bb)) & 1) g |= 1u << h.t[bb];
                            CHECK(g < ts, "group element outside the tile");
                            id[r] = g;
                            x[r] = P[g];
                            y[r] = L[g];
                        }
                        for (int c = 1; c <= h.run; c++) {
                            const AOp& m = plan.aops[s.op_begin + o + c];
                            int sub = -1;
                            for (int i = 0; i < M; i++)
                                if (h.t[i] == m.t[0]) sub = i;
                            CHECK(sub >= 0 && m.k == 1 && m.gctl == 0, "group member off the group's targets");
                            const int B = 1 << (M - 1 - sub);
                            if (m.kind == AK_X) {
                                CHECK((m.lctl & (1u << m.t[0])) == 0, "X control on its target");
                                for (int r = 0; r < D; r++)
                                    if (!(r & B) && (id[r] & m.lctl) == m.lctl) {
                                        std::swap(x[r], x[r | B]);
                                        std::swap(y[r], y[r | B]);
                                    }
                                continue;
                            }
                            CHECK(m.kind == AK_DENSE && m.lctl == 0, "group member kind");
                            const bool diag = (m.flags & AOP_DIAG) != 0;
                            CHECK(m.mat >= 0 && 2 * ((size_t)m.mat + 4) <= plan.mats.size(), "matrix out of range");
                            const cd* u = reinterpret_cast<const cd*>(plan.mats.data()) + m.mat;
                            if (diag) CHECK(u[1] == cd(0, 0) && u[2] == cd(0, 0), "diagonal flag on a dense matrix");
                            for (int r = 0; r < D; r++) {
                                if (r & B) continue;
                                const int r1 = r | B;
                                a[r] = u[0] * x[r] + u[1] * x[r1];
                                a[r1] = u[2] * x[r] + u[3] * x[r1];
                                b[r] = u[0] * y[r] + u[1] * y[r1];
                                b[r1] = u[2] * y[r] + u[3] * y[r1];
                            }
                            for (int j = 0; j < m.nrec; j++) {
                                CHECK(m.slot[j] >= 0 && m.slot[j] < s.slot_count, "bad slot");
                                double acc = 0;
                                if (m.dmat[j] < 0) {
                                    CHECK(m.dmat[j] >= -3, "Pauli record on a wrong op");
                                    const int pk = -1 - m.dmat[j];
                                    const cd hh(0, -0.5);
                                    for (int r = 0; r < D; r++) {
                                        if (r & B) continue;
                                        const int r1 = r | B;
                                        cd px0, px1;
                                        if (pk == 0) { px0 = x[r1]; px1 = x[r]; }
                                        else if (pk == 1) { px0 = cd(0, -1) * x[r1]; px1 = cd(0, 1) * x[r]; }
                                        else { px0 = x[r]; px1 = -x[r1]; }
                                        acc += (std::conj(y[r]) * hh * px0 + std::conj(y[r1]) * hh * px1).real();
                                    }
                                } else {
                                    CHECK(2 * ((size_t)m.dmat[j] + 4) <= plan.mats.size(), "derivative out of range");
                                    const cd* d = reinterpret_cast<const cd*>(plan.mats.data()) + m.dmat[j];
                                    if (diag) CHECK(d[1] == cd(0, 0) && d[2] == cd(0, 0), "diagonal flag on a dense derivative");
                                    for (int r = 0; r < D; r++) {
                                        if (r & B) continue;
                                        const int r1 = r | B;
                                        const cd t0 = d[0] * a[r] + d[1] * a[r1], t1 = d[2] * a[r] + d[3] * a[r1];
                                        acc += (std::conj(y[r]) * t0 + std::conj(y[r1]) * t1).real();
                                    }
                                }
                                red[m.slot[j]] += acc;
                            }
                            for (int r = 0; r < D; r++) {
                                x[r] = a[r];
                                y[r] = b[r];
                            }
                        }
                        for (int r = 0; r < D; r++) {
                            P[id[r]] = x[r];
                            L[id[r]] = y[r];
                        }
                    }
                    o += h.run;
                    continue;
                }
                if (h.kind == AK_X && h.run > 1) {
                    CHECK(h.k == 1 && h.gctl == 0 && h.lctl != 0 && h.run <= 1 + ACHAIN_1Q && o + h.run <= s.op_count,
                          "bad X-headed chain");
                    const AOp& m0 = plan.aops[s.op_begin + o + 1];
                    CHECK(m0.kind == AK_DENSE && m0.k == 1 && m0.t[0] == h.t[0] && m0.lctl == 0 && m0.gctl == 0 &&
                              m0.run == 0,
                          "X-headed chain member");
                    xctl = h.lctl;
                    xpos = h.t[0];
                    o++;
                }
                const AOp& op = plan.aops[s.op_begin + o];
                CHECK((op.gctl & tmask) == 0, "global control inside the tile");
                int run = 1;
                if (xpos >= 0) {
                    run = h.run - 1;
                    for (int c = 1; c < run; c++) {
                        const AOp& m = plan.aops[s.op_begin + o + c];
                        CHECK(m.kind == AK_DENSE && m.run == 0 && m.k == op.k && m.nins == op.nins &&
                                  m.lctl == op.lctl && m.gctl == op.gctl &&
                                  std::memcmp(m.t, op.t, sizeof op.t) == 0 && std::memcmp(m.ins, op.ins, sizeof op.ins) == 0,
                              "chain member with another support");
                    }
                } else if (op.kind == AK_DENSE) {
                    run = op.run;
                    const int cap = op.k == 1 ? ACHAIN_1Q : (op.k == 2 ? ACHAIN_2Q : 1);
                    CHECK(run >= 1 && run <= cap && o + run <= s.op_count, "bad chain length");
                    for (int c = 1; c < run; c++) {
                        const AOp& m = plan.aops[s.op_begin + o + c];
                        CHECK(m.kind == AK_DENSE && m.run == 0 && m.k == op.k && m.nins == op.nins &&
                                  m.lctl == op.lctl && m.gctl == op.gctl &&
                                  std::memcmp(m.t, op.t, sizeof op.t) == 0 && std::memcmp(m.ins, op.ins, sizeof op.ins) == 0,
                              "chain member with another support");
                    }
                } else {
                    CHECK(op.run == 0 || op.run == 1, "chain on a permutation");
                }
                if ((base & op.gctl) != op.gctl) {
                    o += run - 1;
                    continue;
                }
                CHECK(op.nins >= op.k && op.nins <= 5, "bad insertion count");
                for (int i = 0; i < op.nins; i++) CHECK(op.ins[i] >= 0 && op.ins[i] < k && (i == 0 || op.ins[i] > op.ins[i - 1]), "bad insertion bits");
                for (int b = 0; b < op.k; b++) CHECK(op.t[b] >= 0 && op.t[b] < k, "target outside the tile");
                const uint32_t ne = ts >> op.nins;
                for (uint32_t e = 0; e < ne; e++) {
                    uint32_t l = (uint32_t)e;
                    for (int i = 0; i < op.nins; i++) l = (uint32_t)ins0(l, op.ins[i]);
                    l |= op.lctl;
                    CHECK(l < ts, "element outside the tile");
                    if (op.kind == AK_X) {
                        const uint32_t l1 = l | (1u << op.t[0]);
                        CHECK(l1 != l, "X target not inserted");
                        std::swap(P[l], P[l1]);
                        std::swap(L[l], L[l1]);
                        continue;
                    }
                    if (op.kind == AK_SWAP) {
                        const uint32_t la = l | (1u << op.t[0]), lb = l | (1u << op.t[1]);
                        CHECK(la != l && lb != l, "swap targets not inserted");
                        std::swap(P[la], P[lb]);
                        std::swap(L[la], L[lb]);
                        continue;
                    }
                    CHECK(op.kind == AK_DENSE && op.k <= ADJ_DENSE_MAX, "bad op kind");
                    const int D = 1 << op.k;
                    uint32_t id[8];
                    cd x[8], y[8], a[8], b[8];
                    for (int r = 0; r < D; r++) {
                        uint32_t g = l;
                        for (int bb = 0; bb < op.k; bb++)
                            if ((r >> (op.k - 1 - bb)) & 1) g |= 1u << op.t[bb];
                        id[r] = g;
                        x[r] = P[g];
                        y[r] = L[g];
                    }
                    if (xpos >= 0) {
                        CHECK((xctl & (1u << xpos)) == 0 && (l & (1u << xpos)) == 0, "X control on its target");
                        if ((l & xctl) == xctl) {
                            std::swap(x[0], x[1]);
                            std::swap(y[0], y[1]);
                        }
                    }
                    // the chain, in registers
                    for (int c = 0; c < run; c++) {
                        const AOp& m = plan.aops[s.op_begin + o + c];
                        const bool diag = (m.flags & AOP_DIAG) != 0;
                        CHECK(m.mat >= 0 && 2 * ((size_t)m.mat + (size_t)D * D) <= plan.mats.size(), "matrix out of range");
                        const cd* u = reinterpret_cast<const cd*>(plan.mats.data()) + m.mat;
                        for (int r = 0; r < D; r++) {
                            a[r] = 0;
                            b[r] = 0;
                            for (int cc = 0; cc < D; cc++) {
                                if (diag && cc != r) {
                                    CHECK(u[r * D + cc] == cd(0, 0), "diagonal flag on a dense matrix");
                                    continue;
                                }
                                a[r] += u[r * D + cc] * x[cc];
                                b[r] += u[r * D + cc] * y[cc];
                            }
                        }
                        CHECK(m.nrec == 0 || m.k <= 2, "records on a wide op");
                        for (int j = 0; j < m.nrec; j++) {
                            CHECK(m.slot[j] >= 0 && m.slot[j] < s.slot_count, "bad slot");
                            if (m.dmat[j] < 0) {
                                // rotation about Pauli P: Re <y| (-i/2) P |x>, x before G^dag
                                CHECK(D == 2 && m.dmat[j] >= -3, "Pauli record on a wrong op");
                                const int pk = -1 - m.dmat[j];
                                const cd h(0, -0.5);
                                cd px[2];
                                if (pk == 0) { px[0] = x[1]; px[1] = x[0]; }
                                else if (pk == 1) { px[0] = cd(0, -1) * x[1]; px[1] = cd(0, 1) * x[0]; }
                                else { px[0] = x[0]; px[1] = -x[1]; }
                                red[m.slot[j]] += (std::conj(y[0]) * h * px[0] + std::conj(y[1]) * h * px[1]).real();
                                continue;
                            }
                            CHECK(m.dmat[j] >= 0 && 2 * ((size_t)m.dmat[j] + (size_t)D * D) <= plan.mats.size(), "derivative out of range");
                            const cd* d = reinterpret_cast<const cd*>(plan.mats.data()) + m.dmat[j];
                            double acc = 0;
                            for (int r = 0; r < D; r++) {
                                cd t = 0;
                                for (int cc = 0; cc < D; cc++) {
                                    if (diag && cc != r) {
                                        CHECK(d[r * D + cc] == cd(0, 0), "diagonal flag on a dense derivative");
                                        continue;
                                    }
                                    t += d[r * D + cc] * a[cc];
                                }
                                acc += (std::conj(y[r]) * t).real();
                            }
                            red[m.slot[j]] += acc;
                        }
                        for (int r = 0; r < D; r++) {
                            x[r] = a[r];
                            y[r] = b[r];
                        }
                    }
                    for (int r = 0; r < D; r++) {
                        P[id[r]] = x[r];
                        L[id[r]] = y[r];
                    }
                }
                o += run - 1;
            }
            for (uint32_t l = 0; l < ts; l++) {
                uint64_t off = base;
                for (int i = 0; i < k; i++)
                    if ((l >> i) & 1) off |= 1ull << s.tq[i];
                phi[off] = P[l];
                lam[off] = L[l];
            }
        }
        for (uint64_t i = 0; i < N; i++) CHECK(seen[i], "amplitude not covered by the pass");
        for (int i = 0; i < s.slot_count; i++) sums[plan.slot_rec[s.slot_begin + i]] += red[i];
    }
    if (stats) {
        stats[0] = passes;
        stats[1] = bigs;
        stats[2] = maxops;
        stats[3] = stats[4] = 0;
        for (const AOp& op : plan.aops) {
            if (op.kind == AK_DENSE && op.run == 0) stats[3]++;
            if (op.flags & AOP_DIAG) stats[4]++;
        }
    }
    return 0;
}

// Shape of the adjoint plan only: passes, wide steps, ops in the largest pass,
// kernel steps over all passes.
extern "C" int emu_adjoint_plan(int n, int K, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                                const ksv_param* params, size_t nrec, int* stats) {
    AdjPlan plan;
    std::string err;
    int rc = plan_adjoint(n, K, ops, nops, mats, nmats, params, nrec, plan, err, true);
    if (rc) {
        g_msg = err;
        return rc;
    }
    stats[0] = stats[1] = stats[2] = stats[3] = 0;
    for (const AdjStep& s : plan.steps) {
        if (s.big) stats[1]++;
        else {
            stats[0]++;
            stats[2] = std::max(stats[2], s.op_count);
            // kernel steps (one LDS round trip and barrier each): chain heads, permutations
            for (int o = 0; o < s.op_count; o++) {
                const AOp& op = plan.aops[s.op_begin + o];
                stats[3]++;
                // a group or a chain (dense or X-headed) is one step
                o += op.kind == AK_GRP ? op.run : (op.run > 1 ? op.run - 1 : 0);
            }
        }
    }
    return 0;
}

// Batched copies as ksv_shots_run plans them in per-copy mode: plan_copies
// for one segment (Pauli events given as (at, q, pauli, copy) entries, one
// copy per entry, grouped into bitmaps here exactly as the library does),
// then every step emulated on psi (2^ntot amplitudes).  stats: passes, LDS
// tile qubits of the first pass.
extern "C" int emu_copies(int ntot, int nsys, int prune, const ksv_op* ops, size_t nops, const double* mats,
                          size_t nmats, const uint64_t* ev_at, const int32_t* ev_q, const int32_t* ev_pauli,
                          const uint32_t* ev_copy, size_t nent, double* psi_io, int* stats) {
    const uint64_t copies = 1ull << (ntot - nsys);
    const size_t nwords = (size_t)((copies + 31) / 32);
    std::vector<CopyPauli> pend;
    // entries with equal (at, q) form one event; one bitmap per Pauli kind, X, Y, Z
    for (size_t i = 0; i < nent;) {
        size_t j = i;
        while (j < nent && ev_at[j] == ev_at[i] && ev_q[j] == ev_q[i]) j++;
        for (int pk = 1; pk <= 3; pk++) {
            CopyPauli c;
            c.at = ev_at[i];
            c.q = ev_q[i];
            c.pauli = pk;
            c.words.assign(nwords, 0u);
            bool any = false;
            for (size_t k = i; k < j; k++)
                if (ev_pauli[k] == pk) {
                    CHECK(ev_copy[k] < copies, "copy out of range");
                    c.words[ev_copy[k] >> 5] |= 1u << (ev_copy[k] & 31);
                    any = true;
                }
            if (any) pend.push_back(std::move(c));
        }
        i = j;
    }
    Program p;
    std::string err;
    int rc = plan_copies(p, ntot, nsys, KSV_F64, ops, nops, mats, nmats, pend, prune, err);
    if (rc) {
        g_msg = err;
        return rc;
    }
    cd* psi = reinterpret_cast<cd*>(psi_io);
    int passes = 0;
    for (const Step& s : p.steps) {
        int r;
        if (s.big) r = emu_big(p, s, psi, ntot);
        else {
            CHECK(!s.blk, "per-copy plan with a fused-block pass");
            if (passes == 0 && stats) stats[1] = s.k;
            passes++;
            r = s.reg ? emu_reg(p, s, psi, ntot) : emu_lds(p, s, psi, ntot);
        }
        if (r) return r;
    }
    if (stats) stats[0] = passes;
    return 0;
}
