// Adjoint differentiation of E(theta) = <psi|H|psi>, psi = G_N ... G_1 |init>,
// H = sum_t c_t P_t (replaces gradient()'s 2N parameter-shift evaluations,
// src/algorithms.ts:536-547).
//
//   phi <- psi, lambda <- H psi, E = Re<psi|lambda>
//   for k = N..1:  phi <- G_k^dag phi
//                  dE/dtheta += 2 Re <lambda| dG_k |phi>  (per parameter record)
//                  lambda <- G_k^dag lambda
//
// Runs of gates without parameters are applied to phi and lambda as fused
// LDS passes of the inverse gates; each parametrised gate is one kernel that
// reads and writes both states once and reduces its inner products in FP64.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>

#include "kcommon.h"

namespace ksv {

// Local index of element e of an op: zero bits inserted at the op's
// ascending positions, then its local control bits set.
// One and two insertions (plain one-qubit chains, X heads with one control)
// take a uniform branch of their own: the unrolled general form costs a select
// per possible insertion.
__device__ __forceinline__ uint32_t aop_elem(uint32_t e, const AOp& op) {
    uint32_t l = e;
    if (op.nins == 1) {
        l = ins0(l, op.ins[0]);
    } else if (op.nins == 2) {
        l = ins0(ins0(l, op.ins[0]), op.ins[1]);
    } else {
#pragma unroll
        for (int i = 0; i < 5; i++)
            if (i < op.nins) l = ins0(l, op.ins[i]);
    }
    return l | op.lctl;
}

struct AdjPassArgs {
    void* phi;
    void* lam;
    const AOp* ops;
    const double* mats;        // interleaved complex
    const int32_t* slot_rec;   // global record of each slot
    double* part;              // [record][gridDim.x]
    int nops;
    int nslot;
    int k;
    int single;                // 1: phi only (lambda is neither read nor written)
    uint64_t ntiles;
    uint64_t tq_lo, tq_hi;     // physical bit of tile bit i, ascending
    uint64_t mstride;          // doubles between the matrix arrays of consecutive copies
    int cb;                    // 2^cb copies: the workgroups are split evenly among them, each
                               // taking only its copy's tiles and reading its copy's matrices
    int pad;
};

// Fields of chain member c, read when its step starts (scalar loads), so
// only one member's fields and matrices occupy SGPRs at a time.
struct ChainOp {
    int32_t mat, nrec, diag;
    int32_t dmat[3];
};
__device__ __forceinline__ ChainOp chain_op(const KSV_CONST uint64_t* opw, int c) {
    AOp m;
    uint64_t w[sizeof(AOp) / 8];
#pragma unroll
    for (int i = 0; i < (int)(sizeof(AOp) / 8); i++) w[i] = opw[(sizeof(AOp) / 8) * c + i];
    __builtin_memcpy(&m, w, sizeof m);
    ChainOp q;
    q.mat = m.mat;
    q.nrec = m.nrec;
    q.diag = m.flags & AOP_DIAG;
#pragma unroll
    for (int j = 0; j < 3; j++) q.dmat[j] = m.dmat[j];
    return q;
}

// A chain of `run` dense ops on the same 2^KK-element groups: G groups per
// thread and iteration (all their LDS reads issued before any arithmetic),
// each op applies G^dag to phi and lambda in registers and, for each
// record j, acc[c][j] += Re sum_r conj(lambda_r) (dG_j G^dag phi)_r with
// lambda taken before its own G^dag.  Diagonal ops cost one complex
// product per amplitude and state.
template <typename V, int KK, int G, int CMAX, int NT>
__device__ __forceinline__ void adj_chain(V* P, V* L, const AOp& op, const KSV_CONST uint64_t* opw, int run,
                                          const KSV_CONST double* mats, uint32_t ne, uint32_t tid,
                                          double (&acc)[ACHAIN_1Q][3], uint32_t xctl) {
    constexpr int D = 1 << KK;
    for (uint32_t e0 = tid; e0 < ne; e0 += G * NT) {
        uint32_t id[G][D];
        double xr[G][D], xi[G][D], yr[G][D], yi[G][D];
#pragma unroll
        for (int g = 0; g < G; g++) {
            const uint32_t l = aop_elem(e0 + g * NT, op);
#pragma unroll
            for (int r = 0; r < D; r++) {
                uint32_t q = l;
#pragma unroll
                for (int b = 0; b < KK; b++)
                    if ((r >> (KK - 1 - b)) & 1) q |= 1u << op.t[b];
                id[g][r] = q;
            }
        }
#pragma unroll
        for (int g = 0; g < G; g++)
#pragma unroll
            for (int r = 0; r < D; r++) {
                const V x = P[id[g][r]], y = L[id[g][r]];
                xr[g][r] = x.x; xi[g][r] = x.y;
                yr[g][r] = y.x; yi[g][r] = y.y;
            }
        if (KK == 1 && xctl) {
            // the heading X: swap the pair where its control bits are set
#pragma unroll
            for (int g = 0; g < G; g++) {
                if ((id[g][0] & xctl) == xctl) {
                    double t;
                    t = xr[g][0]; xr[g][0] = xr[g][1]; xr[g][1] = t;
                    t = xi[g][0]; xi[g][0] = xi[g][1]; xi[g][1] = t;
                    t = yr[g][0]; yr[g][0] = yr[g][1]; yr[g][1] = t;
                    t = yi[g][0]; yi[g][0] = yi[g][1]; yi[g][1] = t;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < CMAX; c++) {
            if (c >= run) break;
            // keeps the scalar loads of this op's matrices from being hoisted
            // above the step: one op's matrices live in SGPRs at a time
            asm volatile("" ::: "memory");
            const ChainOp q = chain_op(opw, c);
            const KSV_CONST double* u = mats + 2 * (int64_t)q.mat;
            double ar[G][D], ai[G][D], br[G][D], bi[G][D];
            if (q.diag) {
#pragma unroll
                for (int r = 0; r < D; r++) {
                    const double mr = u[2 * (r * D + r)], mi = u[2 * (r * D + r) + 1];
#pragma unroll
                    for (int g = 0; g < G; g++) {
                        ar[g][r] = fma(mr, xr[g][r], -mi * xi[g][r]);
                        ai[g][r] = fma(mr, xi[g][r], mi * xr[g][r]);
                        br[g][r] = fma(mr, yr[g][r], -mi * yi[g][r]);
                        bi[g][r] = fma(mr, yi[g][r], mi * yr[g][r]);
                    }
                }
            } else {
#pragma unroll
                for (int r = 0; r < D; r++) {
#pragma unroll
                    for (int g = 0; g < G; g++) {
                        double sr = 0, si = 0, tr = 0, ti = 0;
#pragma unroll
                        for (int cc = 0; cc < D; cc++) {
                            const double mr = u[2 * (r * D + cc)], mi = u[2 * (r * D + cc) + 1];
                            sr = fma(-mi, xi[g][cc], fma(mr, xr[g][cc], sr));
                            si = fma(mi, xr[g][cc], fma(mr, xi[g][cc], si));
                            tr = fma(-mi, yi[g][cc], fma(mr, yr[g][cc], tr));
                            ti = fma(mi, yr[g][cc], fma(mr, yi[g][cc], ti));
                        }
                        ar[g][r] = sr; ai[g][r] = si;
                        br[g][r] = tr; bi[g][r] = ti;
                    }
                }
            }
            if (KK <= 2) {   // parameters sit on 1- and 2-qubit gates only
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    if (j >= q.nrec) break;
                    asm volatile("" ::: "memory");
                    if (KK == 1 && q.dmat[j] < 0) {
                        // rotation about Pauli P: Re <lambda| (-i/2) P |phi>, phi before G^dag
                        const int pk = -1 - q.dmat[j];
                        double s = 0;
#pragma unroll
                        for (int g = 0; g < G; g++) {
                            const double x0r = xr[g][0], x0i = xi[g][0], x1r = xr[g][1], x1i = xi[g][1];
                            const double y0r = yr[g][0], y0i = yi[g][0], y1r = yr[g][1], y1i = yi[g][1];
                            if (pk == 0)        // X
                                s += (y0r * x1i - y0i * x1r) + (y1r * x0i - y1i * x0r);
                            else if (pk == 1)   // Y
                                s += (y1r * x0r + y1i * x0i) - (y0r * x1r + y0i * x1i);
                            else                // Z
                                s += (y0r * x0i - y0i * x0r) - (y1r * x1i - y1i * x1r);
                        }
                        acc[c][j] += 0.5 * s;
                        continue;
                    }
                    const KSV_CONST double* d = mats + 2 * (int64_t)q.dmat[j];
                    double s = 0;
#pragma unroll
                    for (int r = 0; r < D; r++) {
                        if (q.diag) {
                            const double mr = d[2 * (r * D + r)], mi = d[2 * (r * D + r) + 1];
#pragma unroll
                            for (int g = 0; g < G; g++) {
                                const double tr = fma(mr, ar[g][r], -mi * ai[g][r]);
                                const double ti = fma(mr, ai[g][r], mi * ar[g][r]);
                                s = fma(yi[g][r], ti, fma(yr[g][r], tr, s));
                            }
                        } else {
#pragma unroll
                            for (int g = 0; g < G; g++) {
                                double tr = 0, ti = 0;
#pragma unroll
                                for (int cc = 0; cc < D; cc++) {
                                    const double mr = d[2 * (r * D + cc)], mi = d[2 * (r * D + cc) + 1];
                                    tr = fma(-mi, ai[g][cc], fma(mr, ar[g][cc], tr));
                                    ti = fma(mi, ar[g][cc], fma(mr, ai[g][cc], ti));
                                }
                                s = fma(yi[g][r], ti, fma(yr[g][r], tr, s));
                            }
                        }
                    }
                    acc[c][j] += s;
                }
            }
#pragma unroll
            for (int g = 0; g < G; g++)
#pragma unroll
                for (int r = 0; r < D; r++) {
                    xr[g][r] = ar[g][r]; xi[g][r] = ai[g][r];
                    yr[g][r] = br[g][r]; yi[g][r] = bi[g][r];
                }
        }
#pragma unroll
        for (int g = 0; g < G; g++)
#pragma unroll
            for (int r = 0; r < D; r++) {
                V x, y;
                x.x = xr[g][r]; x.y = xi[g][r];
                y.x = yr[g][r]; y.y = yi[g][r];
                P[id[g][r]] = x;
                L[id[g][r]] = y;
            }
    }
}

// The widest group batch that divides the op's element count: G * NT
// elements per iteration.
template <typename V, int KK, int GMAX, int CMAX, int NT>
__device__ __forceinline__ void adj_chain_g(V* P, V* L, const AOp& op, const KSV_CONST uint64_t* opw, int run,
                                            const KSV_CONST double* mats, uint32_t ne, uint32_t tid,
                                            double (&acc)[ACHAIN_1Q][3], uint32_t xctl) {
    if constexpr (GMAX >= 4) {
        if (ne >= 4 * NT) return adj_chain<V, KK, 4, CMAX, NT>(P, L, op, opw, run, mats, ne, tid, acc, xctl);
    }
    if constexpr (GMAX >= 2) {
        if (ne >= 2 * NT) return adj_chain<V, KK, 2, CMAX, NT>(P, L, op, opw, run, mats, ne, tid, acc, xctl);
    }
    adj_chain<V, KK, 1, CMAX, NT>(P, L, op, opw, run, mats, ne, tid, acc, xctl);
}

// One member of a group (AK_GRP) on the group's sub-target S: an X swaps each
// pair whose control bits are set; a one-qubit op applies G^dag to phi and
// lambda and adds its records' terms to red (one wave sum per record).
template <int M, int S, int G>
__device__ __forceinline__ void grp_member(double (&xr)[G][1 << M], double (&xi)[G][1 << M],
                                           double (&yr)[G][1 << M], double (&yi)[G][1 << M],
                                           const uint32_t (&id)[G][1 << M], const AOp& m,
                                           const KSV_CONST double* mats, double* red, int nw, uint32_t lane,
                                           uint32_t wave) {
    constexpr int D = 1 << M;
    constexpr int B = 1 << (M - 1 - S);   // the pair partner's element bit
    if (m.kind == AK_X) {
#pragma unroll
        for (int g = 0; g < G; g++)
#pragma unroll
            for (int r = 0; r < D; r++) {
                if (r & B) continue;
                if ((id[g][r] & m.lctl) == m.lctl) {
                    double t;
                    t = xr[g][r]; xr[g][r] = xr[g][r | B]; xr[g][r | B] = t;
                    t = xi[g][r]; xi[g][r] = xi[g][r | B]; xi[g][r | B] = t;
                    t = yr[g][r]; yr[g][r] = yr[g][r | B]; yr[g][r | B] = t;
                    t = yi[g][r]; yi[g][r] = yi[g][r | B]; yi[g][r | B] = t;
                }
            }
        return;
    }
    const KSV_CONST double* u = mats + 2 * (int64_t)m.mat;
    const bool diag = (m.flags & AOP_DIAG) != 0;
    double ar[G][D], ai[G][D], br[G][D], bi[G][D];
    if (diag) {
        const double d0r = u[0], d0i = u[1], d1r = u[6], d1i = u[7];
#pragma unroll
        for (int g = 0; g < G; g++)
#pragma unroll
            for (int r = 0; r < D; r++) {
                const double mr = (r & B) ? d1r : d0r, mi = (r & B) ? d1i : d0i;
                ar[g][r] = fma(mr, xr[g][r], -mi * xi[g][r]);
                ai[g][r] = fma(mr, xi[g][r], mi * xr[g][r]);
                br[g][r] = fma(mr, yr[g][r], -mi * yi[g][r]);
                bi[g][r] = fma(mr, yi[g][r], mi * yr[g][r]);
            }
    } else {
        const double u00r = u[0], u00i = u[1], u01r = u[2], u01i = u[3];
        const double u10r = u[4], u10i = u[5], u11r = u[6], u11i = u[7];
#pragma unroll
        for (int g = 0; g < G; g++)
#pragma unroll
            for (int r = 0; r < D; r++) {
                if (r & B) continue;
                const int r1 = r | B;
                const double x0r = xr[g][r], x0i = xi[g][r], x1r = xr[g][r1], x1i = xi[g][r1];
                const double y0r = yr[g][r], y0i = yi[g][r], y1r = yr[g][r1], y1i = yi[g][r1];
                ar[g][r] = fma(-u01i, x1i, fma(u01r, x1r, fma(-u00i, x0i, fma(u00r, x0r, 0.0))));
                ai[g][r] = fma(u01i, x1r, fma(u01r, x1i, fma(u00i, x0r, fma(u00r, x0i, 0.0))));
                ar[g][r1] = fma(-u11i, x1i, fma(u11r, x1r, fma(-u10i, x0i, fma(u10r, x0r, 0.0))));
                ai[g][r1] = fma(u11i, x1r, fma(u11r, x1i, fma(u10i, x0r, fma(u10r, x0i, 0.0))));
                br[g][r] = fma(-u01i, y1i, fma(u01r, y1r, fma(-u00i, y0i, fma(u00r, y0r, 0.0))));
                bi[g][r] = fma(u01i, y1r, fma(u01r, y1i, fma(u00i, y0r, fma(u00r, y0i, 0.0))));
                br[g][r1] = fma(-u11i, y1i, fma(u11r, y1r, fma(-u10i, y0i, fma(u10r, y0r, 0.0))));
                bi[g][r1] = fma(u11i, y1r, fma(u11r, y1i, fma(u10i, y0r, fma(u10r, y0i, 0.0))));
            }
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        if (j >= m.nrec) break;
        asm volatile("" ::: "memory");
        double sacc = 0;
        if (m.dmat[j] < 0) {
            // rotation about Pauli P: Re <lambda| (-i/2) P |phi>, phi before G^dag
            const int pk = -1 - m.dmat[j];
#pragma unroll
            for (int g = 0; g < G; g++)
#pragma unroll
                for (int r = 0; r < D; r++) {
                    if (r & B) continue;
                    const int r1 = r | B;
                    const double x0r = xr[g][r], x0i = xi[g][r], x1r = xr[g][r1], x1i = xi[g][r1];
                    const double y0r = yr[g][r], y0i = yi[g][r], y1r = yr[g][r1], y1i = yi[g][r1];
                    if (pk == 0)        // X
                        sacc += (y0r * x1i - y0i * x1r) + (y1r * x0i - y1i * x0r);
                    else if (pk == 1)   // Y
                        sacc += (y1r * x0r + y1i * x0i) - (y0r * x1r + y0i * x1i);
                    else                // Z
                        sacc += (y0r * x0i - y0i * x0r) - (y1r * x1i - y1i * x1r);
                }
            sacc *= 0.5;
        } else {
            // Re sum_r conj(lambda_r) (dG G^dag phi)_r, lambda before its own G^dag
            const KSV_CONST double* d = mats + 2 * (int64_t)m.dmat[j];
            const double d00r = d[0], d00i = d[1], d01r = d[2], d01i = d[3];
            const double d10r = d[4], d10i = d[5], d11r = d[6], d11i = d[7];
#pragma unroll
            for (int g = 0; g < G; g++)
#pragma unroll
                for (int r = 0; r < D; r++) {
                    if (r & B) continue;
                    const int r1 = r | B;
                    const double a0r = ar[g][r], a0i = ai[g][r], a1r = ar[g][r1], a1i = ai[g][r1];
                    double t0r, t0i, t1r, t1i;
                    if (diag) {
                        t0r = fma(d00r, a0r, -d00i * a0i);
                        t0i = fma(d00r, a0i, d00i * a0r);
                        t1r = fma(d11r, a1r, -d11i * a1i);
                        t1i = fma(d11r, a1i, d11i * a1r);
                    } else {
                        t0r = fma(-d01i, a1i, fma(d01r, a1r, fma(-d00i, a0i, fma(d00r, a0r, 0.0))));
                        t0i = fma(d01i, a1r, fma(d01r, a1i, fma(d00i, a0r, fma(d00r, a0i, 0.0))));
                        t1r = fma(-d11i, a1i, fma(d11r, a1r, fma(-d10i, a0i, fma(d10r, a0r, 0.0))));
                        t1i = fma(d11i, a1r, fma(d11r, a1i, fma(d10i, a0r, fma(d10r, a0i, 0.0))));
                    }
                    sacc = fma(yi[g][r], t0i, fma(yr[g][r], t0r, sacc));
                    sacc = fma(yi[g][r1], t1i, fma(yr[g][r1], t1r, sacc));
                }
        }
        const double v = wave_sum_last(sacc);
        if (lane == 63) red[m.slot[j] * nw + wave] += v;
    }
#pragma unroll
    for (int g = 0; g < G; g++)
#pragma unroll
        for (int r = 0; r < D; r++) {
            xr[g][r] = ar[g][r]; xi[g][r] = ai[g][r];
            yr[g][r] = br[g][r]; yi[g][r] = bi[g][r];
        }
}

// A group (AK_GRP): each thread takes G groups of the 2^M elements over the
// group's targets, runs the group's `nmem` member records on them in order in
// registers, and writes them back: one LDS round trip for up to AGRP_MAX chains.
template <typename V, int M, int G, int NT>
__device__ __forceinline__ void adj_group(V* P, V* L, const AOp& grp, const KSV_CONST uint64_t* memw, int nmem,
                                          const KSV_CONST double* mats, uint32_t ne, uint32_t tid, double* red,
                                          uint32_t lane, uint32_t wave) {
    constexpr int D = 1 << M;
    // every lane runs every iteration (the members' wave sums need the whole
    // wave); lanes past the tile's groups hold zeros and store nothing
    const uint32_t iters = (ne + G * NT - 1) / (G * NT);
    for (uint32_t it = 0; it < iters; it++) {
        const uint32_t e0 = tid + it * G * NT;
        uint32_t id[G][D];
        bool ok[G];
        double xr[G][D], xi[G][D], yr[G][D], yi[G][D];
#pragma unroll
        for (int g = 0; g < G; g++) {
            ok[g] = e0 + g * NT < ne;
            const uint32_t l = aop_elem(e0 + g * NT, grp);
#pragma unroll
            for (int r = 0; r < D; r++) {
                uint32_t q = l;
#pragma unroll
                for (int b = 0; b < M; b++)
                    if ((r >> (M - 1 - b)) & 1) q |= 1u << grp.t[b];
                id[g][r] = q;
            }
        }
#pragma unroll
        for (int g = 0; g < G; g++)
#pragma unroll
            for (int r = 0; r < D; r++) {
                V x, y;
                x.x = x.y = y.x = y.y = 0.0;
                if (ok[g]) {
                    x = P[id[g][r]];
                    y = L[id[g][r]];
                }
                xr[g][r] = x.x; xi[g][r] = x.y;
                yr[g][r] = y.x; yi[g][r] = y.y;
            }
#pragma unroll 2
        for (int c = 0; c < nmem; c++) {
            AOp m;
            {
                uint64_t w[sizeof(AOp) / 8];
#pragma unroll
                for (int i = 0; i < (int)(sizeof(AOp) / 8); i++) w[i] = memw[(sizeof(AOp) / 8) * c + i];
                __builtin_memcpy(&m, w, sizeof m);
            }
            const int sub = m.t[0] == grp.t[0] ? 0 : (m.t[0] == grp.t[1] ? 1 : 2);
            if (sub == 0) grp_member<M, 0, G>(xr, xi, yr, yi, id, m, mats, red, NT / 64, lane, wave);
            else if (sub == 1) grp_member<M, 1, G>(xr, xi, yr, yi, id, m, mats, red, NT / 64, lane, wave);
            else if constexpr (M > 2) grp_member<M, (M > 2 ? 2 : 0), G>(xr, xi, yr, yi, id, m, mats, red, NT / 64, lane, wave);
        }
#pragma unroll
        for (int g = 0; g < G; g++)
#pragma unroll
            for (int r = 0; r < D; r++) {
                V x, y;
                x.x = xr[g][r]; x.y = xi[g][r];
                y.x = yr[g][r]; y.y = yi[g][r];
                if (ok[g]) {
                    P[id[g][r]] = x;
                    L[id[g][r]] = y;
                }
            }
    }
}

// One backward pass: the workgroup walks tiles blockIdx.x, +gridDim.x, ...;
// for each it stages phi and lambda in LDS, applies the pass's ops to both
// and accumulates the records' inner products in red[slot][wave]; at the
// end it writes one partial per record of the pass.
template <typename V, int KT, int NT>
__global__ __launch_bounds__(NT) void k_adjpass(AdjPassArgs a) {
    constexpr int NW = NT / 64;
    constexpr int LNT = NT == 512 ? 9 : 8;
    __shared__ V P[1 << KT], L[1 << KT];
    __shared__ double red[ADJ_SLOTS][NW];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (uint32_t i = tid; i < ADJ_SLOTS * NW; i += NT) (&red[0][0])[i] = 0;
    const int k = a.k;
    const uint32_t tsize = 1u << k;
    V* phi = reinterpret_cast<V*>(a.phi);
    V* lam = reinterpret_cast<V*>(a.lam);
    const KSV_CONST uint64_t* ops = to_const(reinterpret_cast<const uint64_t*>(a.ops));
    // copies: workgroup w serves copy w / gc and walks that copy's tiles
    const uint64_t gc = (uint64_t)gridDim.x >> a.cb;
    const uint64_t copy = blockIdx.x / gc;
    const uint64_t tc = a.ntiles >> a.cb;
    const KSV_CONST double* mats = to_const(a.mats + copy * a.mstride);

    uint64_t g_t = 0;   // physical offset of the thread bits
#pragma unroll
    for (int b = 0; b < LNT; b++)
        if (b < k) g_t |= (uint64_t)((tid >> b) & 1) << byte_at(b < 8 ? a.tq_lo : a.tq_hi, b & 7);

    for (uint64_t t = blockIdx.x % gc; t < tc; t += gc) {
        uint64_t base = copy * tc + t;
#pragma unroll
        for (int i = 0; i < 16; i++)
            if (i < k) base = ins0_64(base, byte_at(i < 8 ? a.tq_lo : a.tq_hi, i & 7));
        for (uint32_t xh = 0; xh < tsize; xh += NT) {
            const uint32_t l = tid + xh;
            if (l < tsize) {
                uint64_t g_h = 0;
#pragma unroll
                for (int b = LNT; b < KT; b++)
                    if (b < k && ((xh >> b) & 1)) g_h |= 1ull << byte_at(a.tq_hi, b - 8);
                P[l] = phi[base | g_t | g_h];
                if (!a.single) L[l] = lam[base | g_t | g_h];
            }
        }
        __syncthreads();
        for (int o = 0; o < a.nops; o++) {
            AOp op;
            {
                uint64_t w[sizeof(AOp) / 8];
#pragma unroll
                for (int i = 0; i < (int)(sizeof(AOp) / 8); i++) w[i] = ops[(sizeof(AOp) / 8) * o + i];
                __builtin_memcpy(&op, w, sizeof op);
            }
            if (op.kind == AK_GRP) {
                const KSV_CONST uint64_t* memw = ops + (sizeof(AOp) / 8) * (o + 1);
                const uint32_t ne = tsize >> op.nins;
                if (op.k == 2) {
                    if (ne >= 2 * NT) adj_group<V, 2, 2, NT>(P, L, op, memw, op.run, mats, ne, tid, &red[0][0], lane, wave);
                    else adj_group<V, 2, 1, NT>(P, L, op, memw, op.run, mats, ne, tid, &red[0][0], lane, wave);
                } else {
                    adj_group<V, 3, 1, NT>(P, L, op, memw, op.run, mats, ne, tid, &red[0][0], lane, wave);
                }
                o += op.run;
                __syncthreads();
                continue;
            }
            if ((base & op.gctl) != op.gctl) {   // a chain shares its head's controls
                if (op.run > 1) o += op.run - 1;
                continue;
            }
            uint32_t ne = tsize >> op.nins;
            if (op.kind == AK_X && op.run <= 1) {
                const uint32_t tb = 1u << op.t[0];
                for (uint32_t e = tid; e < ne; e += NT) {
                    const uint32_t l0 = aop_elem(e, op), l1 = l0 | tb;
                    const V x0 = P[l0], x1 = P[l1], y0 = L[l0], y1 = L[l1];
                    P[l0] = x1; P[l1] = x0;
                    L[l0] = y1; L[l1] = y0;
                }
            } else if (op.kind == AK_SWAP) {
                const uint32_t ba = 1u << op.t[0], bb = 1u << op.t[1];
                for (uint32_t e = tid; e < ne; e += NT) {
                    const uint32_t l = aop_elem(e, op), la = l | ba, lb = l | bb;
                    const V x0 = P[la], x1 = P[lb], y0 = L[la], y1 = L[lb];
                    P[la] = x1; P[lb] = x0;
                    L[la] = y1; L[lb] = y0;
                }
            } else {
                // a chain of dense ops from here, possibly headed by an X with local
                // controls (its members are then uncontrolled one-qubit ops on the X's
                // target, enumerated by the first member)
                const int xh = op.kind == AK_X ? 1 : 0;
                const uint32_t xctl = xh ? op.lctl : 0u;
                const int span = op.run;
                const int run = span - xh;
                const KSV_CONST uint64_t* opw = ops + (sizeof(AOp) / 8) * (o + xh);
                if (xh) {
                    uint64_t w[sizeof(AOp) / 8];
#pragma unroll
                    for (int i = 0; i < (int)(sizeof(AOp) / 8); i++) w[i] = opw[i];
                    __builtin_memcpy(&op, w, sizeof op);
                    ne = tsize >> op.nins;
                }
                double acc[ACHAIN_1Q][3];
#pragma unroll
                for (int c = 0; c < ACHAIN_1Q; c++)
#pragma unroll
                    for (int j = 0; j < 3; j++) acc[c][j] = 0;
                constexpr int G1 = (1 << (KT - 1)) / NT >= 4 ? 4 : (1 << (KT - 1)) / NT;
                constexpr int G2 = (1 << (KT - 2)) / NT >= 2 ? 2 : 1;
                if (op.k == 1) adj_chain_g<V, 1, G1, ACHAIN_1Q, NT>(P, L, op, opw, run, mats, ne, tid, acc, xctl);
                else if (op.k == 2) adj_chain_g<V, 2, G2, ACHAIN_2Q, NT>(P, L, op, opw, run, mats, ne, tid, acc, 0u);
                else adj_chain<V, 3, 1, 1, NT>(P, L, op, opw, run, mats, ne, tid, acc, 0u);
#pragma unroll
                for (int c = 0; c < ACHAIN_1Q; c++) {
                    if (c >= run) break;
                    AOp m;
                    {
                        uint64_t w[sizeof(AOp) / 8];
#pragma unroll
                        for (int i = 0; i < (int)(sizeof(AOp) / 8); i++) w[i] = opw[(sizeof(AOp) / 8) * c + i];
                        __builtin_memcpy(&m, w, sizeof m);
                    }
#pragma unroll
                    for (int j = 0; j < 3; j++) {
                        if (j >= m.nrec) break;
                        const double v = wave_sum_last(acc[c][j]);
                        if (lane == 63) red[m.slot[j]][wave] += v;
                    }
                }
                o += span - 1;
            }
            __syncthreads();
        }
        for (uint32_t xh = 0; xh < tsize; xh += NT) {
            const uint32_t l = tid + xh;
            if (l < tsize) {
                uint64_t g_h = 0;
#pragma unroll
                for (int b = LNT; b < KT; b++)
                    if (b < k && ((xh >> b) & 1)) g_h |= 1ull << byte_at(a.tq_hi, b - 8);
                phi[base | g_t | g_h] = P[l];
                if (!a.single) lam[base | g_t | g_h] = L[l];
            }
        }
    }
    __syncthreads();
    for (int s = (int)tid; s < a.nslot; s += NT) {
        double v = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) v += red[s][w];
        a.part[(size_t)a.slot_rec[s] * gridDim.x + blockIdx.x] = v;
    }
}

// out[r * C + c] = sum of part[r * G + g] over copy c's workgroups
// g in [c G / C, (c + 1) G / C), one workgroup per (record r, copy c).
__global__ __launch_bounds__(256) void k_recsum(const double* part, int G, double* out) {
    __shared__ double sh[4];
    const size_t r = blockIdx.x, C = gridDim.y, c = blockIdx.y;
    const int gc = G / (int)C;
    double v = 0;
    for (int g = threadIdx.x; g < gc; g += 256) v += part[r * G + c * gc + g];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[r * C + c] = sh[0] + sh[1] + sh[2] + sh[3];
}

// Re <a|b> per copy: block (x, c) sums its share of copy c's N amplitudes
// into part[c * gridDim.x + x]; k_rowsum then adds each copy's row in order.
template <typename V>
__global__ __launch_bounds__(256) void k_dotre_copies(const V* a, const V* b, uint64_t N, double* part) {
    __shared__ double sh[4];
    const uint64_t off = (uint64_t)blockIdx.y * N;
    double s = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x) {
        V x = a[off + i], y = b[off + i];
        s += (double)x.x * y.x + (double)x.y * y.y;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

__global__ void k_rowsum(const double* part, int cols, int rows, double* out) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < rows; r += gridDim.x * blockDim.x) {
        double v = 0;
        for (int c = 0; c < cols; c++) v += part[(size_t)r * cols + c];
        out[r] = v;
    }
}

template <typename V>
__global__ void k_fill_zero(V* p, uint64_t N) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x) {
        V z;
        z.x = 0;
        z.y = 0;
        p[i] = z;
    }
}

// lambda = sum_t c_t P_t psi for up to 8 terms per launch, accumulated into
// lam (the first launch overwrites it).
template <typename V>
__global__ __launch_bounds__(256) void k_hpsi(const V* psi, V* lam, uint64_t N, const uint64_t* xm,
                                              const uint64_t* zm, const double* cre, const double* cim, int nt,
                                              int first) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x) {
        double ar = 0, ai = 0;
        if (!first) { V o = lam[i]; ar = o.x; ai = o.y; }
        for (int t = 0; t < nt; t++) {
            uint64_t j = i ^ xm[t];
            V b = psi[j];
            double sg = (__popcll(j & zm[t]) & 1) ? -1.0 : 1.0;
            // (c * i^ny) folded into (cre, cim) by the host
            double pr = sg * cre[t], pi = sg * cim[t];
            ar += pr * b.x - pi * b.y;
            ai += pr * b.y + pi * b.x;
        }
        V o;
        o.x = ar;
        o.y = ai;
        lam[i] = o;
    }
}

template <typename V>
__global__ __launch_bounds__(256) void k_dotre(const V* a, const V* b, uint64_t N, double* part) {
    __shared__ double sh[4];
    double s = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x) {
        V x = a[i], y = b[i];
        s += (double)x.x * y.x + (double)x.y * y.y;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sh[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) r += sh[i];
        part[blockIdx.x] = r;
    }
}

}  // namespace ksv

using namespace ksv;

// LDS capacity (tile qubits) and workgroup size of the instantiated passes;
// a plan's tile k is at most the capacity.
static hipError_t launch_adjpass(int prec, int tile, int threads, int G, const AdjPassArgs& a, hipStream_t st) {
    const dim3 grid((unsigned)G);
    if (prec == KSV_F64) {
        if (tile == 11 && threads == 512) hipLaunchKernelGGL((k_adjpass<double2, 11, 512>), grid, dim3(512), 0, st, a);
        else if (tile == 11) hipLaunchKernelGGL((k_adjpass<double2, 11, 256>), grid, dim3(256), 0, st, a);
        else if (tile == 10) hipLaunchKernelGGL((k_adjpass<double2, 10, 256>), grid, dim3(256), 0, st, a);
        else return hipErrorInvalidValue;
    } else {
        if (tile == 12 && threads == 512) hipLaunchKernelGGL((k_adjpass<float2, 12, 512>), grid, dim3(512), 0, st, a);
        else if (tile == 12) hipLaunchKernelGGL((k_adjpass<float2, 12, 256>), grid, dim3(256), 0, st, a);
        else if (tile == 11) hipLaunchKernelGGL((k_adjpass<float2, 11, 256>), grid, dim3(256), 0, st, a);
        else if (tile == 10) hipLaunchKernelGGL((k_adjpass<float2, 10, 256>), grid, dim3(256), 0, st, a);
        else return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---- host side --------------------------------------------------------------

namespace ksv {
// Entry points of capi.hip used here.
int adj_context(ksv_state* s, int* n, int* prec, void** psi, void** psi2, hipStream_t* st, double** part,
                int* prune, int* tile, int* atile, int* athreads);
int adj_fail(int code, const std::string& msg);
int adj_run_tape(ksv_state* s, const ksv_op* ops, size_t nops, const double* mats, size_t nmats, int prune,
                 void* psi);
}  // namespace ksv

#define AHIP(expr)                                                                        \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return adj_fail(KSV_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {
struct DevBuf {
    void* p = nullptr;
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, std::max<size_t>(bytes, 256)); }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};
size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
}  // namespace

extern "C" int ksv_adjoint(ksv_state* s, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                           const ksv_param* params, size_t nrec, size_t nparam, const uint64_t* xmask,
                           const uint64_t* zmask, const double* coeff, size_t nterms, double* grad, double* energy) {
    int n, prec, prune, tile, atile, athreads;
    void *psi, *psi2;
    hipStream_t st;
    double* part;
    int rc0 = adj_context(s, &n, &prec, &psi, &psi2, &st, &part, &prune, &tile, &atile, &athreads);
    if (rc0) return rc0;
    const int adef = prec == KSV_F64 ? ADJ_TILE_F64 : ADJ_TILE_F32;
    if (athreads == 512 && atile != adef) return adj_fail(KSV_ERR_RANGE, "512-thread adjoint passes need the default tile");
    if (prec != KSV_F64 && prec != KSV_F32) return adj_fail(KSV_ERR_TYPE, "bad precision");
    if (nparam && !grad) return adj_fail(KSV_ERR_TYPE, "null gradient buffer");
    if (nterms && (!xmask || !zmask || !coeff)) return adj_fail(KSV_ERR_TYPE, "null term buffer");
    const uint64_t N = 1ull << n, all = N - 1;
    for (size_t t = 0; t < nterms; t++)
        if ((xmask[t] | zmask[t]) & ~all) return adj_fail(KSV_ERR_RANGE, "Pauli term acts outside the register");
    // validate records, group them per op
    std::map<int, std::vector<const ksv_param*>> recs;
    for (size_t r = 0; r < nrec; r++) {
        const ksv_param& p = params[r];
        if (p.op < 0 || (size_t)p.op >= nops) return adj_fail(KSV_ERR_TYPE, "parameter record names a missing op");
        if (p.param < 0 || (size_t)p.param >= nparam) return adj_fail(KSV_ERR_TYPE, "parameter index out of range");
        const ksv_op& o = ops[p.op];
        int dim = o.kind == KSV_SINGLE || o.kind == KSV_CONTROLLED ? 2 : (o.kind == KSV_TWO ? 4 : 0);
        if (!dim) return adj_fail(KSV_ERR_TYPE, "only single, controlled and two-qubit gates can carry parameters");
        if (p.dmat < 0 || (size_t)p.dmat + 2 * dim * dim > nmats) return adj_fail(KSV_ERR_TYPE, "bad derivative matrix offset");
        recs[p.op].push_back(&p);
        if (recs[p.op].size() > 3) return adj_fail(KSV_ERR_TYPE, "more than three parameter records on one op");
    }
    for (size_t i = 0; i < nparam; i++) grad[i] = 0;
    // forward
    int rc = adj_run_tape(s, ops, nops, mats, nmats, 0, psi);
    if (rc) return rc;
    // lambda = H psi
    {
        std::vector<double> cre(nterms), cim(nterms);
        for (size_t t = 0; t < nterms; t++) {
            int ny = __builtin_popcountll(xmask[t] & zmask[t]) & 3;
            const double c = coeff[t];
            // multiply c by i^ny
            const double re[4] = {c, 0, -c, 0}, im[4] = {0, c, 0, -c};
            cre[t] = re[ny];
            cim[t] = im[ny];
        }
        const size_t nt = std::max<size_t>(nterms, 1);
        DevBuf tb;      // the term arrays, freed on every return
        AHIP(tb.alloc(4 * 8 * nt));
        uint64_t* d_x = static_cast<uint64_t*>(tb.p);
        uint64_t* d_z = d_x + nt;
        double* d_cr = reinterpret_cast<double*>(d_z + nt);
        double* d_ci = d_cr + nt;
        if (nterms) {
            AHIP(hipMemcpyAsync(d_x, xmask, 8 * nterms, hipMemcpyHostToDevice, st));
            AHIP(hipMemcpyAsync(d_z, zmask, 8 * nterms, hipMemcpyHostToDevice, st));
            AHIP(hipMemcpyAsync(d_cr, cre.data(), 8 * nterms, hipMemcpyHostToDevice, st));
            AHIP(hipMemcpyAsync(d_ci, cim.data(), 8 * nterms, hipMemcpyHostToDevice, st));
        }
        int g = (int)std::min<uint64_t>((N + 255) / 256, 8192);
        if (nterms == 0) {
            AHIP(hipMemsetAsync(psi2, 0, (size_t)(prec == KSV_F64 ? 16 : 8) << n, st));
        }
        for (size_t t0 = 0; t0 < nterms; t0 += 8) {
            int m = (int)std::min<size_t>(8, nterms - t0);
            if (prec == KSV_F64)
                hipLaunchKernelGGL((k_hpsi<double2>), dim3(g), dim3(256), 0, st, (const double2*)psi, (double2*)psi2,
                                   N, d_x + t0, d_z + t0, d_cr + t0, d_ci + t0, m, t0 == 0 ? 1 : 0);
            else
                hipLaunchKernelGGL((k_hpsi<float2>), dim3(g), dim3(256), 0, st, (const float2*)psi, (float2*)psi2, N,
                                   d_x + t0, d_z + t0, d_cr + t0, d_ci + t0, m, t0 == 0 ? 1 : 0);
            AHIP(hipGetLastError());
        }
        int gb = (int)std::min<uint64_t>((N + 255) / 256, 1024);
        if (prec == KSV_F64)
            hipLaunchKernelGGL((k_dotre<double2>), dim3(gb), dim3(256), 0, st, (const double2*)psi, (const double2*)psi2, N, part);
        else
            hipLaunchKernelGGL((k_dotre<float2>), dim3(gb), dim3(256), 0, st, (const float2*)psi, (const float2*)psi2, N, part);
        AHIP(hipGetLastError());
        std::vector<double> hp(gb);
        AHIP(hipMemcpyAsync(hp.data(), part, 8 * gb, hipMemcpyDeviceToHost, st));
        AHIP(hipStreamSynchronize(st));
        double e = 0;
        for (int i = 0; i < gb; i++) e += hp[i];
        if (energy) *energy = e;
    }
    // backward sweep: the inverse tape in tile passes over phi (psi) and
    // lambda (psi2); every workgroup writes one partial per record, summed
    // per record at the end
    AdjPlan plan;
    {
        std::string err;
        rc = plan_adjoint(n, atile, ops, nops, mats, nmats, params, nrec, plan, err, true);
        if (rc) return adj_fail(rc, err);
    }
    const uint64_t ntiles = 1ull << (n - plan.k);
    const size_t nr = std::max<size_t>(nrec, 1);
    // workgroups per pass: enough to fill the chip, partials kept under 1 GiB
    const int G = (int)std::max<uint64_t>(1, std::min<uint64_t>({ntiles, 2048, (1ull << 27) / nr}));
    DevBuf blob;
    const size_t off_ops = 0;
    const size_t off_mats = align256(off_ops + sizeof(AOp) * plan.aops.size());
    const size_t off_slot = align256(off_mats + sizeof(double) * plan.mats.size());
    const size_t off_part = align256(off_slot + sizeof(int32_t) * plan.slot_rec.size());
    const size_t off_sum = align256(off_part + sizeof(double) * nr * G);
    const size_t bytes = off_sum + sizeof(double) * nr;
    AHIP(blob.alloc(bytes));
    {
        std::vector<char> h(off_part, 0);
        if (!plan.aops.empty()) std::memcpy(h.data() + off_ops, plan.aops.data(), sizeof(AOp) * plan.aops.size());
        if (!plan.mats.empty()) std::memcpy(h.data() + off_mats, plan.mats.data(), sizeof(double) * plan.mats.size());
        if (!plan.slot_rec.empty())
            std::memcpy(h.data() + off_slot, plan.slot_rec.data(), sizeof(int32_t) * plan.slot_rec.size());
        AHIP(hipMemcpyAsync(blob.p, h.data(), h.size(), hipMemcpyHostToDevice, st));
        AHIP(hipStreamSynchronize(st));   // the host copy goes out of scope
    }
    char* dev = static_cast<char*>(blob.p);
    double* d_part = reinterpret_cast<double*>(dev + off_part);
    double* d_sum = reinterpret_cast<double*>(dev + off_sum);
    for (const AdjStep& sp : plan.steps) {
        if (sp.big) {
            const ksv_op& w = plan.bops[sp.bop];
            rc = adj_run_tape(s, &w, 1, plan.bmats.data(), plan.bmats.size(), 0, psi);
            if (!rc) rc = adj_run_tape(s, &w, 1, plan.bmats.data(), plan.bmats.size(), 0, psi2);
            if (rc) return rc;
            continue;
        }
        AdjPassArgs a;
        std::memset(&a, 0, sizeof a);
        a.phi = psi;
        a.lam = psi2;
        a.ops = reinterpret_cast<const AOp*>(dev + off_ops) + sp.op_begin;
        a.mats = reinterpret_cast<const double*>(dev + off_mats);
        a.slot_rec = reinterpret_cast<const int32_t*>(dev + off_slot) + sp.slot_begin;
        a.part = d_part;
        a.nops = sp.op_count;
        a.nslot = sp.slot_count;
        a.k = sp.k;
        a.ntiles = ntiles;
        for (int i = 0; i < 8; i++) {
            a.tq_lo |= (uint64_t)(uint8_t)sp.tq[i] << (8 * i);
            a.tq_hi |= (uint64_t)(uint8_t)sp.tq[8 + i] << (8 * i);
        }
        AHIP(launch_adjpass(prec, atile, athreads, G, a, st));
    }
    if (nrec) {
        hipLaunchKernelGGL(k_recsum, dim3((unsigned)nrec, 1), dim3(256), 0, st, d_part, G, d_sum);
        AHIP(hipGetLastError());
    }
    std::vector<double> sums(nr, 0.0);
    AHIP(hipMemcpyAsync(sums.data(), d_sum, sizeof(double) * nr, hipMemcpyDeviceToHost, st));
    AHIP(hipStreamSynchronize(st));
    for (size_t r = 0; r < nrec; r++) grad[params[r].param] += 2.0 * params[r].coeff * sums[r];
    return KSV_OK;
}

// ---- batched adjoint: many parameter sets of one tape ------------------------
// The state holds B = 2^cb copies of n qubits (copy t = amplitudes
// [t 2^n, (t + 1) 2^n)); copy t runs the tape with its own matrix array.  Both
// sweeps are tile passes of k_adjpass over all copies at once: the forward
// sweep is the backward planner's output for the inverse tape (its passes
// apply G_1 .. G_N), on phi only; every workgroup serves one copy and reads
// that copy's matrices, so per-record partials and the per-copy sums stay
// apart.  The whole launch sequence is captured once in a HIP graph and
// replayed with each call's matrices.
struct ksv_adj_batch {
    ksv_state* s = nullptr;
    int n = 0, cb = 0, prec = KSV_F64, atile = 0, athreads = 256;
    uint64_t B = 1;
    void *psi = nullptr, *psi2 = nullptr;
    hipStream_t st = nullptr;
    std::vector<ksv_op> ops;
    std::vector<ksv_param> params;
    size_t nmats = 0, nrec = 0, nparam = 0, nterms = 0;
    AdjPlan fwd, bwd;
    size_t fm = 0, bm = 0;               // doubles of one copy's forward / backward matrices
    int G = 0, gd = 0;                   // adjoint-pass workgroups; dot blocks per copy
    void* blob = nullptr;
    size_t off_faops = 0, off_baops = 0, off_slot = 0, off_mats = 0, off_part = 0, off_sums = 0, off_epart = 0,
           off_energy = 0, off_terms = 0;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    std::vector<double> hmats;           // host staging of every copy's plan matrices

    ~ksv_adj_batch() {
        if (st) (void)hipStreamSynchronize(st);
        for (int g = 0; g < 2; g++) {
            if (exec[g]) (void)hipGraphExecDestroy(exec[g]);
            if (graph[g]) (void)hipGraphDestroy(graph[g]);
        }
        if (blob) (void)hipFree(blob);
    }
};

namespace {

// every copy's forward and backward plan matrices, copy t at hmats[t * (fm + bm)]
int batch_mats(ksv_adj_batch* b, const double* mats) {
    std::string err;
    std::vector<ksv_op> iops;
    std::vector<double> imats, f, g;
    b->hmats.resize(b->B * (b->fm + b->bm));
    for (uint64_t t = 0; t < b->B; t++) {
        const double* m = mats + t * b->nmats;
        int rc = invert_tape(b->ops.data(), b->ops.size(), m, b->nmats, iops, imats, err);
        if (!rc) rc = adjoint_plan_mats(b->fwd, iops.data(), iops.size(), imats.data(), imats.size(), nullptr, f, err);
        if (!rc) rc = adjoint_plan_mats(b->bwd, b->ops.data(), b->ops.size(), m, b->nmats, b->params.data(), g, err);
        if (rc) return adj_fail(rc, "copy " + std::to_string(t) + ": " + err);
        std::copy(f.begin(), f.end(), b->hmats.begin() + t * (b->fm + b->bm));
        std::copy(g.begin(), g.end(), b->hmats.begin() + t * (b->fm + b->bm) + b->fm);
    }
    return KSV_OK;
}

hipError_t batch_passes(ksv_adj_batch* b, const AdjPlan& plan, const AOp* d_aops, const double* d_mats, int single,
                        const int32_t* d_slot, double* d_part) {
    char* dev = static_cast<char*>(b->blob);
    (void)dev;
    for (const AdjStep& sp : plan.steps) {
        AdjPassArgs a;
        std::memset(&a, 0, sizeof a);
        a.phi = b->psi;
        a.lam = b->psi2;
        a.ops = d_aops + sp.op_begin;
        a.mats = d_mats;
        a.slot_rec = d_slot ? d_slot + sp.slot_begin : nullptr;
        a.part = d_part;
        a.nops = sp.op_count;
        a.nslot = single ? 0 : sp.slot_count;
        a.k = sp.k;
        a.single = single;
        a.ntiles = 1ull << (b->n + b->cb - sp.k);
        a.mstride = b->fm + b->bm;
        a.cb = b->cb;
        for (int i = 0; i < 8; i++) {
            a.tq_lo |= (uint64_t)(uint8_t)sp.tq[i] << (8 * i);
            a.tq_hi |= (uint64_t)(uint8_t)sp.tq[8 + i] << (8 * i);
        }
        hipError_t e = launch_adjpass(b->prec, b->atile, b->athreads, b->G, a, b->st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// the whole launch sequence of one evaluation (captured into a graph)
hipError_t batch_schedule(ksv_adj_batch* b, bool grad) {
    char* dev = static_cast<char*>(b->blob);
    const int ntot = b->n + b->cb;
    const uint64_t N = 1ull << ntot;
    const double* d_mats = reinterpret_cast<const double*>(dev + b->off_mats);
    hipError_t e = shots_basis(ntot, b->n, b->prec, b->psi, 0, b->st);
    if (e == hipSuccess)
        e = batch_passes(b, b->fwd, reinterpret_cast<const AOp*>(dev + b->off_faops), d_mats, 1, nullptr, nullptr);
    if (e != hipSuccess) return e;
    const int g = (int)std::min<uint64_t>((N + 255) / 256, 8192);
    uint64_t* d_x = reinterpret_cast<uint64_t*>(dev + b->off_terms);
    const size_t nt = std::max<size_t>(b->nterms, 1);
    uint64_t* d_z = d_x + nt;
    double* d_cr = reinterpret_cast<double*>(d_z + nt);
    double* d_ci = d_cr + nt;
    if (b->nterms == 0) {
        if (b->prec == KSV_F64) hipLaunchKernelGGL((k_fill_zero<double2>), dim3(g), dim3(256), 0, b->st, (double2*)b->psi2, N);
        else hipLaunchKernelGGL((k_fill_zero<float2>), dim3(g), dim3(256), 0, b->st, (float2*)b->psi2, N);
    }
    for (size_t t0 = 0; t0 < b->nterms; t0 += 8) {
        const int m = (int)std::min<size_t>(8, b->nterms - t0);
        if (b->prec == KSV_F64)
            hipLaunchKernelGGL((k_hpsi<double2>), dim3(g), dim3(256), 0, b->st, (const double2*)b->psi, (double2*)b->psi2,
                               N, d_x + t0, d_z + t0, d_cr + t0, d_ci + t0, m, t0 == 0 ? 1 : 0);
        else
            hipLaunchKernelGGL((k_hpsi<float2>), dim3(g), dim3(256), 0, b->st, (const float2*)b->psi, (float2*)b->psi2,
                               N, d_x + t0, d_z + t0, d_cr + t0, d_ci + t0, m, t0 == 0 ? 1 : 0);
    }
    double* d_epart = reinterpret_cast<double*>(dev + b->off_epart);
    double* d_energy = reinterpret_cast<double*>(dev + b->off_energy);
    const uint64_t Nc = 1ull << b->n;
    if (b->prec == KSV_F64)
        hipLaunchKernelGGL((k_dotre_copies<double2>), dim3(b->gd, (unsigned)b->B), dim3(256), 0, b->st,
                           (const double2*)b->psi, (const double2*)b->psi2, Nc, d_epart);
    else
        hipLaunchKernelGGL((k_dotre_copies<float2>), dim3(b->gd, (unsigned)b->B), dim3(256), 0, b->st,
                           (const float2*)b->psi, (const float2*)b->psi2, Nc, d_epart);
    hipLaunchKernelGGL(k_rowsum, dim3((unsigned)((b->B + 255) / 256)), dim3(256), 0, b->st, d_epart, b->gd, (int)b->B,
                       d_energy);
    e = hipGetLastError();
    if (e != hipSuccess || !grad) return e;
    double* d_part = reinterpret_cast<double*>(dev + b->off_part);
    e = batch_passes(b, b->bwd, reinterpret_cast<const AOp*>(dev + b->off_baops), d_mats + b->fm, 0,
                     reinterpret_cast<const int32_t*>(dev + b->off_slot), d_part);
    if (e != hipSuccess) return e;
    if (b->nrec)
        hipLaunchKernelGGL(k_recsum, dim3((unsigned)b->nrec, (unsigned)b->B), dim3(256), 0, b->st, d_part, b->G,
                           reinterpret_cast<double*>(dev + b->off_sums));
    return hipGetLastError();
}

}  // namespace

extern "C" int ksv_adjoint_batch_create(ksv_state* s, int n, const ksv_op* ops, size_t nops, const double* mats,
                                        size_t nmats, const ksv_param* params, size_t nrec, size_t nparam,
                                        const uint64_t* xmask, const uint64_t* zmask, const double* coeff,
                                        size_t nterms, ksv_adj_batch** out) {
    if (!out) return adj_fail(KSV_ERR_TYPE, "null output pointer");
    *out = nullptr;
    int ntot, prec, prune, tile, atile, athreads;
    void *psi, *psi2;
    hipStream_t st;
    double* part;
    int rc = adj_context(s, &ntot, &prec, &psi, &psi2, &st, &part, &prune, &tile, &atile, &athreads);
    if (rc) return rc;
    if (n < 1 || n > ntot) return adj_fail(KSV_ERR_RANGE, "copy width must be 1.." + std::to_string(ntot) + " qubits");
    if (ntot - n > 16) return adj_fail(KSV_ERR_RANGE, "at most 65536 copies");
    const int adef = prec == KSV_F64 ? ADJ_TILE_F64 : ADJ_TILE_F32;
    if (athreads == 512 && atile != adef) return adj_fail(KSV_ERR_RANGE, "512-thread adjoint passes need the default tile");
    if (!ops && nops) return adj_fail(KSV_ERR_TYPE, "null tape");
    if (!mats && nmats) return adj_fail(KSV_ERR_TYPE, "null matrix buffer");
    if (nrec && !params) return adj_fail(KSV_ERR_TYPE, "null parameter records");
    if (nterms && (!xmask || !zmask || !coeff)) return adj_fail(KSV_ERR_TYPE, "null term buffer");
    const uint64_t cmask = (1ull << n) - 1;
    for (size_t t = 0; t < nterms; t++)
        if ((xmask[t] | zmask[t]) & ~cmask) return adj_fail(KSV_ERR_RANGE, "Pauli term acts outside the copy's qubits");
    std::map<int, int> per_op;
    for (size_t r = 0; r < nrec; r++) {
        const ksv_param& p = params[r];
        if (p.op < 0 || (size_t)p.op >= nops) return adj_fail(KSV_ERR_TYPE, "parameter record names a missing op");
        if (p.param < 0 || (size_t)p.param >= nparam) return adj_fail(KSV_ERR_TYPE, "parameter index out of range");
        const ksv_op& o = ops[p.op];
        const int dim = o.kind == KSV_SINGLE || o.kind == KSV_CONTROLLED ? 2 : (o.kind == KSV_TWO ? 4 : 0);
        if (!dim) return adj_fail(KSV_ERR_TYPE, "only single, controlled and two-qubit gates can carry parameters");
        if (p.dmat < 0 || (size_t)p.dmat + 2 * dim * dim > nmats) return adj_fail(KSV_ERR_TYPE, "bad derivative matrix offset");
        if (++per_op[p.op] > 3) return adj_fail(KSV_ERR_TYPE, "more than three parameter records on one op");
    }
    std::unique_ptr<ksv_adj_batch> b(new ksv_adj_batch());
    b->s = s;
    b->n = n;
    b->cb = ntot - n;
    b->B = 1ull << b->cb;
    b->prec = prec;
    b->atile = atile;
    b->athreads = athreads;
    b->psi = psi;
    b->psi2 = psi2;
    b->st = st;
    b->ops.assign(ops, ops + nops);
    b->params.assign(params, params + nrec);
    b->nmats = nmats;
    b->nrec = nrec;
    b->nparam = nparam;
    b->nterms = nterms;
    {
        std::string err;
        std::vector<ksv_op> iops;
        std::vector<double> imats;
        rc = invert_tape(ops, nops, mats, nmats, iops, imats, err);
        if (!rc) rc = plan_adjoint(n, atile, iops.data(), iops.size(), imats.data(), imats.size(), nullptr, 0, b->fwd, err);
        if (!rc) rc = plan_adjoint(n, atile, ops, nops, mats, nmats, params, nrec, b->bwd, err);
        if (rc) return adj_fail(rc, err);
        for (const AdjStep& sp : b->fwd.steps)
            if (sp.big) return adj_fail(KSV_ERR_TYPE, "batched adjoint: unitaries of more than three qubits are not supported");
        for (const AdjStep& sp : b->bwd.steps)
            if (sp.big) return adj_fail(KSV_ERR_TYPE, "batched adjoint: unitaries of more than three qubits are not supported");
        // an op is treated as diagonal only if it is diagonal in every copy
        for (uint64_t t = 0; t < b->B; t++) {
            const double* m = mats + t * nmats;
            rc = invert_tape(ops, nops, m, nmats, iops, imats, err);
            if (rc) return adj_fail(rc, err);
            adjoint_plan_undiag(b->fwd, iops.data(), iops.size(), imats.data(), imats.size(), nullptr);
            adjoint_plan_undiag(b->bwd, ops, nops, m, nmats, params);
        }
    }
    b->fm = b->fwd.mats.size();
    b->bm = b->bwd.mats.size();
    const uint64_t tc = 1ull << (n - std::min(b->fwd.k, b->bwd.k));
    const uint64_t gc = std::max<uint64_t>(1, std::min<uint64_t>(tc, std::max<uint64_t>(1, 2048 / b->B)));
    const size_t nr = std::max<size_t>(nrec, 1);
    if ((uint64_t)nr * gc * b->B > (1ull << 27)) return adj_fail(KSV_ERR_RANGE, "too many parameter records for the batch");
    b->G = (int)(gc * b->B);
    b->gd = (int)std::min<uint64_t>(((1ull << n) + 255) / 256, 64);
    const size_t nt = std::max<size_t>(nterms, 1);
    size_t off = 0;
    b->off_faops = off; off = align256(off + sizeof(AOp) * std::max<size_t>(1, b->fwd.aops.size()));
    b->off_baops = off; off = align256(off + sizeof(AOp) * std::max<size_t>(1, b->bwd.aops.size()));
    b->off_slot = off; off = align256(off + 4 * std::max<size_t>(1, b->bwd.slot_rec.size()));
    b->off_terms = off; off = align256(off + 4 * 8 * nt);
    const size_t host_end = off;
    b->off_mats = off; off = align256(off + 8 * std::max<size_t>(1, b->B * (b->fm + b->bm)));
    b->off_part = off; off = align256(off + 8 * nr * (size_t)b->G);
    b->off_sums = off; off = align256(off + 8 * nr * b->B);
    b->off_epart = off; off = align256(off + 8 * (size_t)b->gd * b->B);
    b->off_energy = off; off = align256(off + 8 * b->B);
    if (hipMalloc(&b->blob, off) != hipSuccess) {
        b->blob = nullptr;
        return adj_fail(KSV_ERR_DEVICE, "out of device memory for the batched adjoint");
    }
    {
        std::vector<char> h(host_end, 0);
        if (!b->fwd.aops.empty()) std::memcpy(h.data() + b->off_faops, b->fwd.aops.data(), sizeof(AOp) * b->fwd.aops.size());
        if (!b->bwd.aops.empty()) std::memcpy(h.data() + b->off_baops, b->bwd.aops.data(), sizeof(AOp) * b->bwd.aops.size());
        if (!b->bwd.slot_rec.empty()) std::memcpy(h.data() + b->off_slot, b->bwd.slot_rec.data(), 4 * b->bwd.slot_rec.size());
        uint64_t* x = reinterpret_cast<uint64_t*>(h.data() + b->off_terms);
        double* cr = reinterpret_cast<double*>(x + 2 * nt);
        for (size_t t = 0; t < nterms; t++) {
            const int ny = __builtin_popcountll(xmask[t] & zmask[t]) & 3;
            const double c = coeff[t];
            const double re[4] = {c, 0, -c, 0}, im[4] = {0, c, 0, -c};   // c * i^ny
            x[t] = xmask[t];
            x[nt + t] = zmask[t];
            cr[t] = re[ny];
            cr[nt + t] = im[ny];
        }
        AHIP(hipMemcpyAsync(b->blob, h.data(), h.size(), hipMemcpyHostToDevice, st));
        AHIP(hipStreamSynchronize(st));
    }
    *out = b.release();
    return KSV_OK;
}

extern "C" int ksv_adjoint_batch_run(ksv_adj_batch* b, const double* mats, int want_grad, double* energy, double* grad) {
    if (!b) return adj_fail(KSV_ERR_HANDLE, "null batch");
    if (!mats && b->nmats) return adj_fail(KSV_ERR_TYPE, "null matrix buffer");
    if (want_grad && b->nparam && !grad) return adj_fail(KSV_ERR_TYPE, "null gradient buffer");
    int rc = batch_mats(b, mats);
    if (rc) return rc;
    char* dev = static_cast<char*>(b->blob);
    if (!b->hmats.empty())
        AHIP(hipMemcpyAsync(dev + b->off_mats, b->hmats.data(), 8 * b->hmats.size(), hipMemcpyHostToDevice, b->st));
    const int g = want_grad ? 1 : 0;
    if (!b->exec[g]) {
        AHIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
        hipError_t e = batch_schedule(b, want_grad != 0);
        hipGraph_t gr = nullptr;
        hipError_t e2 = hipStreamEndCapture(b->st, &gr);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (gr) (void)hipGraphDestroy(gr);
            return adj_fail(KSV_ERR_DEVICE, std::string("batched adjoint capture: ") +
                                                hipGetErrorString(e != hipSuccess ? e : e2));
        }
        b->graph[g] = gr;
        AHIP(hipGraphInstantiate(&b->exec[g], gr, nullptr, nullptr, 0));
    }
    AHIP(hipGraphLaunch(b->exec[g], b->st));
    std::vector<double> e(b->B), sums(want_grad ? b->nrec * b->B : 0);
    AHIP(hipMemcpyAsync(e.data(), dev + b->off_energy, 8 * b->B, hipMemcpyDeviceToHost, b->st));
    if (!sums.empty())
        AHIP(hipMemcpyAsync(sums.data(), dev + b->off_sums, 8 * sums.size(), hipMemcpyDeviceToHost, b->st));
    AHIP(hipStreamSynchronize(b->st));
    if (energy) std::copy(e.begin(), e.end(), energy);
    if (want_grad && b->nparam) {
        std::fill(grad, grad + b->B * b->nparam, 0.0);
        for (size_t r = 0; r < b->nrec; r++)
            for (uint64_t t = 0; t < b->B; t++)
                grad[t * b->nparam + b->params[r].param] += 2.0 * b->params[r].coeff * sums[r * b->B + t];
    }
    return KSV_OK;
}

extern "C" int ksv_adjoint_batch_destroy(ksv_adj_batch* b) {
    if (!b) return adj_fail(KSV_ERR_HANDLE, "null batch");
    delete b;
    return KSV_OK;
}
