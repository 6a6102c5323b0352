/*
 * ketsv_oracle.c — CPU restatement of ket's statevector path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the ket_amd package loads this file;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, and only as the checker (or the timed CPU baseline), never as the
 * product path.
 *
 * It restates, densely and in plain C, the arithmetic of
 *   src/complex.ts:1-24       complex add/mul/norm2 (no FMA: built with -ffp-contract=off)
 *   src/gates.ts:1-208        gate matrices
 *   src/statevector.ts:1-280  apply* kernels, accumulate's 1e-14 prune, probabilities
 *   src/circuit.ts:55-70      collapseQubit
 *   src/circuit.ts:813-819    makePrng (xorshift32)
 *   src/circuit.ts:3425-3450  run() fast path: sorted CDF, last := 1.0, binary search
 *   src/circuit.ts:1508-1518  marginals
 *   src/circuit.ts:4102-4129  blochAngles
 *   src/circuit.ts:3728-3753  expectation (rotate, exactProbs, parity)
 *   src/algorithms.ts:301-336 vqe (ops[0] acts on qubit n-1)
 *   src/circuit.ts:1150-1420  the Circuit gate methods -> (kind, matrix) mapping
 * A dense array replaces ket's sparse Map: an absent key is a zero
 * amplitude.  With prune = 1 every written amplitude with |a|^2 < 1e-14 is
 * zeroed, which is exactly the set of keys ket's accumulate() deletes.
 *
 * State layout: interleaved (re, im) doubles, index bit q = qubit q.
 *
 * Built twice (oracle/Makefile): plainly, and with -fopenmp as the
 * all-core dense CPU baseline.  The OpenMP pragmas split only loops whose
 * iterations own disjoint amplitudes, so both builds give the same bits.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/ketsv.h"

typedef struct { double re, im; } cx;

/* src/complex.ts:10-21 */
static inline cx cadd(cx a, cx b) { cx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cx cmul(cx a, cx b) {
    cx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline double cnorm2(cx a) { return a.re * a.re + a.im * a.im; }
static const cx CZERO = {0.0, 0.0};

/* src/complex.ts:24 + src/statevector.ts:19-27 */
static inline cx keep(cx v, int prune) {
    if (prune && cnorm2(v) < 1e-14) return CZERO;
    return v;
}

static inline cx ld(const double* psi, uint64_t i) { cx r = {psi[2 * i], psi[2 * i + 1]}; return r; }
static inline void st(double* psi, uint64_t i, cx v) { psi[2 * i] = v.re; psi[2 * i + 1] = v.im; }
static inline void swp(double* psi, uint64_t i, uint64_t j) {
    cx a = ld(psi, i), b = ld(psi, j);
    st(psi, i, b);
    st(psi, j, a);
}

/* ------------------------------------------------------------------------- */
/* gate matrices, src/gates.ts                                                */
/* ------------------------------------------------------------------------- */

static void set2(double* o, cx a, cx b, cx c, cx d) {
    o[0] = a.re; o[1] = a.im; o[2] = b.re; o[3] = b.im;
    o[4] = c.re; o[5] = c.im; o[6] = d.re; o[7] = d.im;
}
static cx C(double re, double im) { cx r = {re, im}; return r; }

static void rz_(double t, double* o) {            /* src/gates.ts:40-44 */
    double co = cos(t / 2), si = sin(t / 2);
    set2(o, C(co, -si), CZERO, CZERO, C(co, si));
}
static void u3_(double th, double ph, double la, double* o) {   /* src/gates.ts:71-78 */
    double co = cos(th / 2), si = sin(th / 2);
    set2(o, C(co, 0), C(-si * cos(la), -si * sin(la)),
         C(si * cos(ph), si * sin(ph)), C(co * cos(ph + la), co * sin(ph + la)));
}
static void xy_(double t, double* o) {             /* src/gates.ts:145-153 */
    memset(o, 0, 32 * sizeof(double));
    double co = cos(t / 2), is = sin(t / 2);
    o[0] = 1;                       /* [0][0] */
    o[2 * 5] = co;                  /* [1][1] */
    o[2 * 6 + 1] = is;              /* [1][2] */
    o[2 * 9 + 1] = is;              /* [2][1] */
    o[2 * 10] = co;                 /* [2][2] */
    o[2 * 15] = 1;                  /* [3][3] */
}

/* Returns the matrix dimension (2 or 4) and fills `o` (row-major,
 * interleaved); -1 if the name is not a ket gate matrix. */
int orc_gate_matrix(const char* name, const double* p, double* o) {
    const double sq2 = 1 / sqrt(2.0);
    if (!strcmp(name, "id")) { set2(o, C(1, 0), CZERO, CZERO, C(1, 0)); return 2; }
    if (!strcmp(name, "h")) { set2(o, C(sq2, 0), C(sq2, 0), C(sq2, 0), C(-sq2, 0)); return 2; }
    if (!strcmp(name, "x")) { set2(o, CZERO, C(1, 0), C(1, 0), CZERO); return 2; }
    if (!strcmp(name, "y")) { set2(o, CZERO, C(0, -1), C(0, 1), CZERO); return 2; }
    if (!strcmp(name, "z")) { set2(o, C(1, 0), CZERO, CZERO, C(-1, 0)); return 2; }
    if (!strcmp(name, "s")) { set2(o, C(1, 0), CZERO, CZERO, C(0, 1)); return 2; }
    if (!strcmp(name, "si")) { set2(o, C(1, 0), CZERO, CZERO, C(0, -1)); return 2; }
    if (!strcmp(name, "t")) { set2(o, C(1, 0), CZERO, CZERO, C(sq2, sq2)); return 2; }
    if (!strcmp(name, "ti")) { set2(o, C(1, 0), CZERO, CZERO, C(sq2, -sq2)); return 2; }
    if (!strcmp(name, "v")) { set2(o, C(.5, .5), C(.5, -.5), C(.5, -.5), C(.5, .5)); return 2; }
    if (!strcmp(name, "vi")) { set2(o, C(.5, -.5), C(.5, .5), C(.5, .5), C(.5, -.5)); return 2; }
    if (!strcmp(name, "rx")) {
        double co = cos(p[0] / 2), si = sin(p[0] / 2);
        set2(o, C(co, 0), C(0, -si), C(0, -si), C(co, 0));
        return 2;
    }
    if (!strcmp(name, "ry")) {
        double co = cos(p[0] / 2), si = sin(p[0] / 2);
        set2(o, C(co, 0), C(-si, 0), C(si, 0), C(co, 0));
        return 2;
    }
    if (!strcmp(name, "rz")) { rz_(p[0], o); return 2; }
    if (!strcmp(name, "r2")) { rz_(M_PI / 2, o); return 2; }
    if (!strcmp(name, "r4")) { rz_(M_PI / 4, o); return 2; }
    if (!strcmp(name, "r8")) { rz_(M_PI / 8, o); return 2; }
    if (!strcmp(name, "u1")) { set2(o, C(1, 0), CZERO, CZERO, C(cos(p[0]), sin(p[0]))); return 2; }
    if (!strcmp(name, "u2")) { u3_(M_PI / 2, p[0], p[1], o); return 2; }
    if (!strcmp(name, "u3")) { u3_(p[0], p[1], p[2], o); return 2; }
    if (!strcmp(name, "gpi")) {
        double co = cos(p[0]), si = sin(p[0]);
        set2(o, CZERO, C(co, -si), C(co, si), CZERO);
        return 2;
    }
    if (!strcmp(name, "gpi2")) {
        double co = cos(p[0]), si = sin(p[0]);
        set2(o, C(sq2, 0), C(-sq2 * si, -sq2 * co), C(sq2 * si, -sq2 * co), C(sq2, 0));
        return 2;
    }
    if (!strcmp(name, "xx") || !strcmp(name, "yy")) {   /* src/gates.ts:99-122 */
        memset(o, 0, 32 * sizeof(double));
        double co = cos(p[0] / 2), s = sin(p[0] / 2);
        double corner = !strcmp(name, "xx") ? -s : s;   /* XX: -i s on the anti-diagonal corners; YY: +i s */
        o[0] = co; o[2 * 3 + 1] = corner;
        o[2 * 5] = co; o[2 * 6 + 1] = -s;
        o[2 * 9 + 1] = -s; o[2 * 10] = co;
        o[2 * 12 + 1] = corner; o[2 * 15] = co;
        return 4;
    }
    if (!strcmp(name, "zz")) {                           /* src/gates.ts:129-138 */
        memset(o, 0, 32 * sizeof(double));
        double co = cos(p[0] / 2), s = sin(p[0] / 2);
        o[0] = co; o[1] = -s;
        o[2 * 5] = co; o[2 * 5 + 1] = s;
        o[2 * 10] = co; o[2 * 10 + 1] = s;
        o[2 * 15] = co; o[2 * 15 + 1] = -s;
        return 4;
    }
    if (!strcmp(name, "xy")) { xy_(p[0], o); return 4; }
    if (!strcmp(name, "iswap")) { xy_(M_PI, o); return 4; }
    if (!strcmp(name, "srswap")) { xy_(M_PI / 2, o); return 4; }
    if (!strcmp(name, "ms")) {                           /* src/gates.ts:196-208 */
        memset(o, 0, 32 * sizeof(double));
        double sp = sin(p[0] + p[1]), cp = cos(p[0] + p[1]);
        double sd = sin(p[0] - p[1]), cd = cos(p[0] - p[1]);
        o[0] = sq2;
        o[2 * 3] = -sp * sq2; o[2 * 3 + 1] = -cp * sq2;
        o[2 * 5] = sq2;
        o[2 * 6] = -sd * sq2; o[2 * 6 + 1] = -cd * sq2;
        o[2 * 9] = sd * sq2; o[2 * 9 + 1] = -cd * sq2;
        o[2 * 10] = sq2;
        o[2 * 12] = sp * sq2; o[2 * 12 + 1] = -cp * sq2;
        o[2 * 15] = sq2;
        return 4;
    }
    return -1;
}

/* ------------------------------------------------------------------------- */
/* apply kernels, src/statevector.ts                                          */
/* ------------------------------------------------------------------------- */

static inline uint64_t ins0(uint64_t p, int q) {   /* insert a 0 bit at position q */
    uint64_t lo = p & ((1ull << q) - 1);
    return ((p >> q) << (q + 1)) | lo;
}

/* applySingle, src/statevector.ts:47-65 */
static void apply_single(int n, double* psi, int q, const double* m, int prune) {
    cx a = {m[0], m[1]}, b = {m[2], m[3]}, c = {m[4], m[5]}, d = {m[6], m[7]};
    uint64_t half = 1ull << (n - 1), mk = 1ull << q;
#pragma omp parallel for schedule(static)
    for (uint64_t p = 0; p < half; p++) {
        uint64_t i0 = ins0(p, q), i1 = i0 | mk;
        cx x0 = ld(psi, i0), x1 = ld(psi, i1);
        st(psi, i0, keep(cadd(cmul(a, x0), cmul(b, x1)), prune));
        st(psi, i1, keep(cadd(cmul(c, x0), cmul(d, x1)), prune));
    }
}

/* applyControlled, src/statevector.ts:212-234 (control = 0 half is a
 * pass-through whose amplitudes already satisfy the prune invariant) */
static void apply_controlled(int n, double* psi, int ctl, int q, const double* m, int prune) {
    cx a = {m[0], m[1]}, b = {m[2], m[3]}, c = {m[4], m[5]}, d = {m[6], m[7]};
    uint64_t half = 1ull << (n - 1), mk = 1ull << q, cm = 1ull << ctl;
#pragma omp parallel for schedule(static)
    for (uint64_t p = 0; p < half; p++) {
        uint64_t i0 = ins0(p, q), i1 = i0 | mk;
        if (!(i0 & cm)) continue;
        cx x0 = ld(psi, i0), x1 = ld(psi, i1);
        st(psi, i0, keep(cadd(cmul(a, x0), cmul(b, x1)), prune));
        st(psi, i1, keep(cadd(cmul(c, x0), cmul(d, x1)), prune));
    }
}

/* applyTwo, src/statevector.ts:110-132 and applyUnitary :240-270:
 * out_r = ((0 + M[r][0] a_0) + M[r][1] a_1) + ... in column order. */
static void apply_k(int n, double* psi, int k, const int* qs, const double* m, int prune) {
    uint64_t dim = 1ull << k, all = 0;
    uint64_t masks[KSV_MAX_OPERANDS];
    for (int b = 0; b < k; b++) { masks[b] = 1ull << qs[b]; all |= masks[b]; }
    uint64_t N = 1ull << n;
#pragma omp parallel
    {
        uint64_t* base = (uint64_t*)malloc(dim * sizeof(uint64_t));
        cx* amp = (cx*)malloc(dim * sizeof(cx));
        cx* out = (cx*)malloc(dim * sizeof(cx));
#pragma omp for schedule(static)
        for (uint64_t ctx = 0; ctx < N; ctx++) {
            if (ctx & all) continue;
            for (uint64_t i = 0; i < dim; i++) {
                uint64_t g = ctx;
                for (int b = 0; b < k; b++)
                    if ((i >> (k - 1 - b)) & 1) g |= masks[b];
                base[i] = g;
                amp[i] = ld(psi, g);
            }
            for (uint64_t r = 0; r < dim; r++) {
                cx acc = CZERO;
                for (uint64_t c = 0; c < dim; c++) {
                    const double* e = m + 2 * (r * dim + c);
                    cx me = {e[0], e[1]};
                    acc = cadd(acc, cmul(me, amp[c]));
                }
                out[r] = acc;
            }
            for (uint64_t r = 0; r < dim; r++) st(psi, base[r], keep(out[r], prune));
        }
        free(base);
        free(amp);
        free(out);
    }
}

/* applyCsrSwap, src/statevector.ts:177-207 */
static void apply_csrswap(int n, double* psi, int ctl, int a, int b, int prune) {
    const double sq2 = 1 / sqrt(2.0);
    cx r2 = {sq2, 0}, i2 = {0, sq2};
    uint64_t N = 1ull << n, cm = 1ull << ctl, am = 1ull << a, bm = 1ull << b;
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < N; i++) {
        if (!(i & cm) || (i & am) || (i & bm)) continue;
        uint64_t i01 = i | bm, i10 = i | am;
        cx a01 = ld(psi, i01), a10 = ld(psi, i10);
        st(psi, i01, keep(cadd(cmul(r2, a01), cmul(i2, a10)), prune));
        st(psi, i10, keep(cadd(cmul(i2, a01), cmul(r2, a10)), prune));
    }
}

static int distinct(const int* q, int k) {
    for (int i = 0; i < k; i++)
        for (int j = i + 1; j < k; j++)
            if (q[i] == q[j]) return 0;
    return 1;
}

/* One tape op with its own matrix (the engine-level check). */
static int apply_op(int n, double* psi, const ksv_op* op, const double* mats, int prune) {
    const int* q = op->q;
    const double* m = op->mat >= 0 ? mats + op->mat : NULL;
    uint64_t N = 1ull << n;
    for (int i = 0; i < op->nq; i++)
        if (q[i] < 0 || q[i] >= n) return KSV_ERR_RANGE;
    switch (op->kind) {
    case KSV_SINGLE: apply_single(n, psi, q[0], m, prune); return 0;
    case KSV_CNOT: {                                   /* src/statevector.ts:70-80 */
        if (q[0] == q[1]) return KSV_ERR_TYPE;
        uint64_t cm = 1ull << q[0], tm = 1ull << q[1];
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < N; i++)
            if ((i & cm) && !(i & tm)) swp(psi, i, i | tm);
        return 0;
    }
    case KSV_SWAP: {                                   /* src/statevector.ts:85-102 */
        if (q[0] == q[1]) return 0;
        uint64_t am = 1ull << q[0], bm = 1ull << q[1];
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < N; i++)
            if ((i & am) && !(i & bm)) swp(psi, i, i ^ am ^ bm);
        return 0;
    }
    case KSV_TWO:
        if (q[0] == q[1]) return KSV_ERR_TYPE;
        apply_k(n, psi, 2, q, m, prune);
        return 0;
    case KSV_CONTROLLED:
        if (q[0] == q[1]) return KSV_ERR_TYPE;
        apply_controlled(n, psi, q[0], q[1], m, prune);
        return 0;
    case KSV_TOFFOLI: {                                /* src/statevector.ts:138-147 */
        if (q[2] == q[0] || q[2] == q[1]) return KSV_ERR_TYPE;
        uint64_t c1 = 1ull << q[0], c2 = 1ull << q[1], tm = 1ull << q[2];
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < N; i++)
            if ((i & c1) && (i & c2) && !(i & tm)) swp(psi, i, i | tm);
        return 0;
    }
    case KSV_CSWAP: {                                  /* src/statevector.ts:153-168 */
        if (q[1] == q[2]) return 0;
        if (q[0] == q[1] || q[0] == q[2]) return KSV_ERR_TYPE;
        uint64_t cm = 1ull << q[0], am = 1ull << q[1], bm = 1ull << q[2];
#pragma omp parallel for schedule(static)
        for (uint64_t i = 0; i < N; i++)
            if ((i & cm) && (i & am) && !(i & bm)) swp(psi, i, i ^ am ^ bm);
        return 0;
    }
    case KSV_CSRSWAP:
        if (!distinct(q, 3)) return KSV_ERR_TYPE;
        apply_csrswap(n, psi, q[0], q[1], q[2], prune);
        return 0;
    case KSV_UNITARY:
        if (op->nq < 1 || op->nq > KSV_MAX_OPERANDS || !distinct(q, op->nq)) return KSV_ERR_TYPE;
        apply_k(n, psi, op->nq, q, m, prune);
        return 0;
    case KSV_BARRIER: return 0;
    default: return KSV_ERR_TYPE;
    }
}

void orc_init_basis(int n, double* psi, uint64_t index) {
    memset(psi, 0, (size_t)2 * sizeof(double) << n);
    psi[2 * index] = 1.0;
}

/* Apply a tape of ksv_op records (each with its own matrix). */
int orc_apply(int n, double* psi, const ksv_op* ops, size_t nops, const double* mats, int prune) {
    for (size_t i = 0; i < nops; i++) {
        int rc = apply_op(n, psi, &ops[i], mats, prune);
        if (rc) return rc;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* named gates: the Circuit methods of src/circuit.ts:1150-1420               */
/* ------------------------------------------------------------------------- */

typedef struct orc_gate {
    char    name[12];
    int32_t nq;
    int32_t q[KSV_MAX_OPERANDS];
    double  p[3];
    int64_t mat;        /* 'unitary' only: offset (doubles) of its matrix */
} orc_gate;

/* Controlled methods and the 2x2 each applies (src/circuit.ts:1282-1352). */
static const char* ctrl_gate(const char* name) {
    static const char* tab[][2] = {
        {"cy", "y"}, {"cz", "z"}, {"ch", "h"}, {"crx", "rx"}, {"cry", "ry"}, {"crz", "rz"},
        {"cr2", "r2"}, {"cr4", "r4"}, {"cr8", "r8"}, {"cu1", "u1"}, {"cu2", "u2"}, {"cu3", "u3"},
        {"cs", "s"}, {"ct", "t"}, {"csdg", "si"}, {"ctdg", "ti"}, {"csrn", "v"},
    };
    for (size_t i = 0; i < sizeof(tab) / sizeof(tab[0]); i++)
        if (!strcmp(name, tab[i][0])) return tab[i][1];
    return NULL;
}

/* Aliases of single-qubit methods (src/circuit.ts:1152-1205). */
static const char* alias1(const char* name) {
    if (!strcmp(name, "sdg")) return "si";
    if (!strcmp(name, "tdg")) return "ti";
    if (!strcmp(name, "srn")) return "v";
    if (!strcmp(name, "srndg")) return "vi";
    if (!strcmp(name, "vz")) return "rz";
    if (!strcmp(name, "p")) return "u1";
    return name;
}

int orc_apply_named(int n, double* psi, const orc_gate* g, size_t ngates, const double* mats,
                    int prune) {
    double m[32];
    for (size_t i = 0; i < ngates; i++) {
        const char* nm = g[i].name;
        const int* q = g[i].q;
        for (int j = 0; j < g[i].nq; j++)
            if (q[j] < 0 || q[j] >= n) return KSV_ERR_RANGE;
        ksv_op op;
        memset(&op, 0, sizeof op);
        op.nq = g[i].nq;
        memcpy(op.q, q, sizeof(op.q));
        op.mat = 0;
        const char* cg = ctrl_gate(nm);
        int rc;
        if (!strcmp(nm, "cnot") || !strcmp(nm, "cx")) { op.kind = KSV_CNOT; rc = apply_op(n, psi, &op, m, prune); }
        else if (!strcmp(nm, "swap")) { op.kind = KSV_SWAP; rc = apply_op(n, psi, &op, m, prune); }
        else if (!strcmp(nm, "ccx")) { op.kind = KSV_TOFFOLI; rc = apply_op(n, psi, &op, m, prune); }
        else if (!strcmp(nm, "cswap")) { op.kind = KSV_CSWAP; rc = apply_op(n, psi, &op, m, prune); }
        else if (!strcmp(nm, "csrswap")) { op.kind = KSV_CSRSWAP; rc = apply_op(n, psi, &op, m, prune); }
        else if (!strcmp(nm, "barrier")) { rc = 0; }
        else if (!strcmp(nm, "unitary")) {
            op.kind = KSV_UNITARY; op.mat = g[i].mat;
            rc = apply_op(n, psi, &op, mats, prune);
        } else if (cg) {
            if (orc_gate_matrix(cg, g[i].p, m) != 2) return KSV_ERR_TYPE;
            op.kind = KSV_CONTROLLED;
            rc = apply_op(n, psi, &op, m, prune);
        } else {
            int dim = orc_gate_matrix(alias1(nm), g[i].p, m);
            if (dim == 2) op.kind = KSV_SINGLE;
            else if (dim == 4) op.kind = KSV_TWO;
            else return KSV_ERR_TYPE;
            rc = apply_op(n, psi, &op, m, prune);
        }
        if (rc) return rc;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* read-out                                                                   */
/* ------------------------------------------------------------------------- */

/* probabilities(), src/statevector.ts:273-280, ascending index */
size_t orc_probabilities(int n, const double* psi, double thresh, uint64_t* idx, double* p) {
    size_t k = 0;
    for (uint64_t i = 0; i < (1ull << n); i++) {
        double v = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        if (v > thresh) {
            if (idx) idx[k] = i;
            if (p) p[k] = v;
            k++;
        }
    }
    return k;
}

/* makePrng, src/circuit.ts:813-819: `seed` is already (seed >>> 0) || 1. */
void orc_prng(uint32_t seed, size_t count, double* out) {
    uint32_t s = seed ? seed : 1;
    for (size_t i = 0; i < count; i++) {
        s ^= s << 13;
        s ^= s >> 17;
        s ^= s << 5;
        out[i] = (double)s / 4294967296.0;
    }
}

/* run() fast path, src/circuit.ts:3425-3450, for caller-supplied draws r. */
int orc_sample(int n, const double* psi, const double* r, size_t shots, uint64_t* out) {
    uint64_t N = 1ull << n;
    uint64_t* idx = (uint64_t*)malloc(N * sizeof(uint64_t));
    double* cum = (double*)malloc(N * sizeof(double));
    if (!idx || !cum) { free(idx); free(cum); return KSV_ERR_DEVICE; }
    size_t k = orc_probabilities(n, psi, 1e-14, idx, cum);
    double acc = 0;
    for (size_t i = 0; i < k; i++) { acc += cum[i]; cum[i] = acc; }
    if (k) cum[k - 1] = 1.0;
    for (size_t s = 0; s < shots; s++) {
        if (!k) { out[s] = 0; continue; }
        size_t lo = 0, hi = k - 1;
        while (lo < hi) {
            size_t mid = (lo + hi) >> 1;
            if (cum[mid] < r[s]) lo = mid + 1;
            else hi = mid;
        }
        out[s] = idx[lo];
    }
    free(idx);
    free(cum);
    return 0;
}

/* sampleSV, src/circuit.ts:74-81 (the per-shot path's final draw): cum +=
 * |a|^2 over the map's entries (the non-zero amplitudes) in ascending index;
 * the first entry with rand <= cum, else the last entry, else index 0. */
uint64_t orc_sample_sv(int n, const double* psi, double rand) {
    double cum = 0;
    uint64_t last = 0;
    for (uint64_t i = 0; i < (1ull << n); i++) {
        double p = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        if (!(p > 0)) continue;
        cum += p;
        last = i;
        if (rand <= cum) return i;
    }
    return last;
}

/* marginals(), src/circuit.ts:1508-1518 */
void orc_marginals(int n, const double* psi, double* out) {
    for (int q = 0; q < n; q++) out[q] = 0;
    for (uint64_t i = 0; i < (1ull << n); i++) {
        double p = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        for (int q = 0; q < n; q++)
            if ((i >> q) & 1) out[q] += p;
    }
}

/* blochAngles(), src/circuit.ts:4102-4129 */
void orc_bloch(int n, const double* psi, int q, double* theta, double* phi) {
    uint64_t mk = 1ull << q;
    double r00 = 0, r11 = 0, re = 0, im = 0;
    for (uint64_t i = 0; i < (1ull << n); i++) {
        cx a = ld(psi, i);
        double p = cnorm2(a);
        if (i & mk) { r11 += p; continue; }
        r00 += p;
        cx b = ld(psi, i | mk);
        re += a.re * b.re + a.im * b.im;
        im += a.im * b.re - a.re * b.im;
    }
    double rz = r00 - r11, rx = 2 * re, ry = -2 * im;
    double c = rz < -1 ? -1 : (rz > 1 ? 1 : rz);
    *theta = acos(c);
    *phi = atan2(ry, rx);
}

/* <P> by ket's method on a copy of the state: X -> h(q), Y -> rx(pi/2, q),
 * then probabilities (p > 1e-14) and +-p by parity of the non-I bits.
 * `pq[q]` is the Pauli on qubit q (one of I X Y Z). */
static double ket_parity_expect(int n, const double* psi, const char* pq) {
    uint64_t N = 1ull << n;
    double* w = (double*)malloc((size_t)2 * sizeof(double) * N);
    memcpy(w, psi, (size_t)2 * sizeof(double) * N);
    double m[8], pa = M_PI / 2;
    uint64_t mask = 0;
    for (int q = 0; q < n; q++) {
        if (pq[q] == 'X') { orc_gate_matrix("h", NULL, m); apply_single(n, w, q, m, 1); }
        else if (pq[q] == 'Y') { orc_gate_matrix("rx", &pa, m); apply_single(n, w, q, m, 1); }
        if (pq[q] != 'I') mask |= 1ull << q;
    }
    double e = 0;
    for (uint64_t i = 0; i < N; i++) {
        double p = w[2 * i] * w[2 * i] + w[2 * i + 1] * w[2 * i + 1];
        if (!(p > 1e-14)) continue;
        e += (__builtin_popcountll(i & mask) & 1 ? -1 : 1) * p;
    }
    free(w);
    return e;
}

/* Circuit.expectation(pauli), pauli[q] acts on qubit q (q0-leftmost).
 * The state passed in is the circuit's simulated state. */
double orc_expectation(int n, const double* psi, const char* pauli) {
    int any = 0;
    for (int q = 0; q < n; q++) any |= pauli[q] != 'I';
    if (!any) return 1;
    return ket_parity_expect(n, psi, pauli);
}

/* vqe(): sum over terms; ops[t] has ops[t][0] acting on qubit n-1. */
double orc_vqe(int n, const double* psi, const char* const* ops, const double* coeff, size_t T) {
    double energy = 0;
    char* pq = (char*)malloc((size_t)n + 1);
    for (size_t t = 0; t < T; t++) {
        if (fabs(coeff[t]) < 1e-15) continue;
        int any = 0;
        for (int q = 0; q < n; q++) {
            pq[q] = ops[t][n - 1 - q];
            any |= pq[q] != 'I';
        }
        if (!any) { energy += coeff[t]; continue; }
        energy += coeff[t] * ket_parity_expect(n, psi, pq);
    }
    free(pq);
    return energy;
}

/* Direct <psi|P|psi> = sum_j conj(psi_{j^x}) i^{nY} (-1)^{|j & z|} psi_j. */
double orc_expect_direct(int n, const double* psi, uint64_t xmask, uint64_t zmask) {
    int ny = __builtin_popcountll(xmask & zmask) & 3;
    double sre = 0, sim = 0;
    for (uint64_t j = 0; j < (1ull << n); j++) {
        cx a = ld(psi, j), b = ld(psi, j ^ xmask);
        double sg = (__builtin_popcountll(j & zmask) & 1) ? -1 : 1;
        /* conj(b) * a */
        sre += sg * (b.re * a.re + b.im * a.im);
        sim += sg * (b.re * a.im - b.im * a.re);
    }
    /* multiply by i^ny and keep the real part */
    switch (ny) {
    case 0: return sre;
    case 1: return -sim;
    case 2: return -sre;
    default: return sim;
    }
}

/* collapseQubit, src/circuit.ts:55-70. Returns the outcome. */
int orc_collapse(int n, double* psi, int q, double r) {
    uint64_t mk = 1ull << q, N = 1ull << n;
    double p1 = 0;
    for (uint64_t i = 0; i < N; i++)
        if (i & mk) p1 += psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
    int outcome = r < p1 ? 1 : 0;
    double inv = 1 / sqrt(outcome ? p1 : 1 - p1);
    for (uint64_t i = 0; i < N; i++) {
        int bit = (i & mk) != 0;
        if (bit == outcome) { psi[2 * i] *= inv; psi[2 * i + 1] *= inv; }
        else { psi[2 * i] = 0; psi[2 * i + 1] = 0; }
    }
    return outcome;
}

size_t orc_gate_struct_size(void) { return sizeof(orc_gate); }
size_t orc_op_struct_size(void) { return sizeof(ksv_op); }
