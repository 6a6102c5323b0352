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

#include "ksv_internal.h"

namespace ksv {

template <typename V>
__device__ __forceinline__ void cfma(double& ar, double& ai, V m, V x) {
    ar += (double)m.x * x.x - (double)m.y * x.y;
    ai += (double)m.x * x.y + (double)m.y * x.x;
}

struct AdjArgs {
    void* phi;
    void* lam;
    int k;              // targets (1 or 2)
    int nrec;           // derivative records (<= 3)
    int8_t t[2];        // targets, t[0] = MSB of the local index
    int8_t ins[4];      // ascending bits to insert (targets + controls)
    int nins;
    uint64_t ctl;       // control bits (set after insertion)
    double ud[32];      // U^dag, row-major interleaved (dim x dim)
    double du[3][32];   // dU per record
};

template <typename V>
__global__ __launch_bounds__(256) void k_adj(AdjArgs a, uint64_t groups, double* part) {
    __shared__ double sh[4];
    V* phi = reinterpret_cast<V*>(a.phi);
    V* lam = reinterpret_cast<V*>(a.lam);
    const int dim = 1 << a.k;
    double acc[3] = {0, 0, 0};
    for (uint64_t g = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; g < groups; g += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t b = g;
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (i < a.nins) b = ((b >> a.ins[i]) << (a.ins[i] + 1)) | (b & ((1ull << a.ins[i]) - 1));
        b |= a.ctl;
        uint64_t id[4];
        V x[4], l[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            if (r < dim) {
                uint64_t gi = b;
                if (a.k == 1) { if (r & 1) gi |= 1ull << a.t[0]; }
                else {
                    if (r & 2) gi |= 1ull << a.t[0];
                    if (r & 1) gi |= 1ull << a.t[1];
                }
                id[r] = gi;
                x[r] = phi[gi];
                l[r] = lam[gi];
            }
        }
        V xn[4], ln[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            if (r < dim) {
                double xr = 0, xi = 0, lr = 0, li = 0;
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    if (c < dim) {
                        V m;
                        m.x = a.ud[2 * (r * dim + c)];
                        m.y = a.ud[2 * (r * dim + c) + 1];
                        cfma(xr, xi, m, x[c]);
                        cfma(lr, li, m, l[c]);
                    }
                }
                xn[r].x = xr; xn[r].y = xi;
                ln[r].x = lr; ln[r].y = li;
            }
        }
#pragma unroll
        for (int j = 0; j < 3; j++) {
            if (j < a.nrec) {
                double s = 0;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    if (r < dim) {
                        double tr = 0, ti = 0;
#pragma unroll
                        for (int c = 0; c < 4; c++) {
                            if (c < dim) {
                                V m;
                                m.x = a.du[j][2 * (r * dim + c)];
                                m.y = a.du[j][2 * (r * dim + c) + 1];
                                cfma(tr, ti, m, xn[c]);
                            }
                        }
                        s += (double)l[r].x * tr + (double)l[r].y * ti;   // Re(conj(l) t)
                    }
                }
                acc[j] += s;
            }
        }
#pragma unroll
        for (int r = 0; r < 4; r++)
            if (r < dim) { phi[id[r]] = xn[r]; lam[id[r]] = ln[r]; }
    }
    for (int j = 0; j < 3; j++) {
        double v = acc[j];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        if (lane == 0) sh[w] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = 0;
            for (int i = 0; i < (int)(blockDim.x >> 6); i++) r += sh[i];
            part[(size_t)blockIdx.x * 3 + j] = r;
        }
        __syncthreads();
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

// ---- host side --------------------------------------------------------------

namespace ksv {
// Entry points of capi.hip used here.
int adj_context(ksv_state* s, int* n, int* prec, void** psi, void** psi2, hipStream_t* st, double** part,
                int* prune, int* tile);
int adj_fail(int code, const std::string& msg);
int adj_run_tape(ksv_state* s, const ksv_op* ops, size_t nops, const double* mats, size_t nmats, int prune,
                 void* psi);
}  // namespace ksv

#define AHIP(expr)                                                                        \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) return adj_fail(KSV_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static void adjoint_of(const std::vector<cd>& m, int dim, std::vector<cd>& out) {
    out.assign((size_t)dim * dim, cd(0, 0));
    for (int r = 0; r < dim; r++)
        for (int c = 0; c < dim; c++) out[(size_t)r * dim + c] = std::conj(m[(size_t)c * dim + r]);
}

// Inverse of a tape segment: reversed ops with adjoint matrices.
static void invert_tape(const ksv_op* ops, size_t b, size_t e, const double* mats, std::vector<ksv_op>& out,
                        std::vector<double>& om) {
    out.clear();
    om.clear();
    static const int dims[] = {2, 0, 0, 4, 2, 0, 0, 0, 0, 0};
    for (size_t i = e; i-- > b;) {
        ksv_op o = ops[i];
        if (o.kind == KSV_CSRSWAP) {
            // inverse of the controlled sqrt(iSWAP): conj-transposed 4x4 with the control
            const double s = 1 / std::sqrt(2.0);
            double m[32] = {0};
            m[0] = 1;
            m[2 * 5] = s;
            m[2 * 6 + 1] = -s;
            m[2 * 9 + 1] = -s;
            m[2 * 10] = s;
            m[2 * 15] = 1;
            // as a KSV_UNITARY on (control, a, b) is wider; express as controlled via an 8x8
            double u[128] = {0};
            for (int r = 0; r < 4; r++) { u[2 * (r * 8 + r)] = 1; }
            for (int r = 0; r < 4; r++)
                for (int c = 0; c < 4; c++) {
                    u[2 * ((4 + r) * 8 + 4 + c)] = m[2 * (r * 4 + c)];
                    u[2 * ((4 + r) * 8 + 4 + c) + 1] = m[2 * (r * 4 + c) + 1];
                }
            ksv_op w;
            std::memset(&w, 0, sizeof w);
            w.kind = KSV_UNITARY;
            w.nq = 3;
            w.q[0] = o.q[0];
            w.q[1] = o.q[1];
            w.q[2] = o.q[2];
            w.mat = (int64_t)om.size();
            om.insert(om.end(), u, u + 128);
            out.push_back(w);
            continue;
        }
        int dim = 0;
        if (o.kind == KSV_UNITARY) dim = 1 << o.nq;
        else dim = dims[o.kind];
        if (dim) {
            std::vector<cd> m((size_t)dim * dim), a;
            for (size_t j = 0; j < m.size(); j++) m[j] = cd(mats[o.mat + 2 * j], mats[o.mat + 2 * j + 1]);
            adjoint_of(m, dim, a);
            o.mat = (int64_t)om.size();
            for (auto& z : a) { om.push_back(z.real()); om.push_back(z.imag()); }
        } else {
            o.mat = -1;   // permutations (cnot, swap, toffoli, cswap) are self-inverse
        }
        out.push_back(o);
    }
}

extern "C" int ksv_adjoint(ksv_state* s, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                           const ksv_param* params, size_t nrec, size_t nparam, const uint64_t* xmask,
                           const uint64_t* zmask, const double* coeff, size_t nterms, double* grad, double* energy) {
    int n, prec, prune, tile;
    void *psi, *psi2;
    hipStream_t st;
    double* part;
    int rc0 = adj_context(s, &n, &prec, &psi, &psi2, &st, &part, &prune, &tile);
    if (rc0) return rc0;
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
        uint64_t *d_x, *d_z;
        double *d_cr, *d_ci;
        size_t nt = std::max<size_t>(nterms, 1);
        AHIP(hipMalloc(&d_x, 8 * nt));
        AHIP(hipMalloc(&d_z, 8 * nt));
        AHIP(hipMalloc(&d_cr, 8 * nt));
        AHIP(hipMalloc(&d_ci, 8 * nt));
        if (nterms) {
            AHIP(hipMemcpyAsync(d_x, xmask, 8 * nterms, hipMemcpyHostToDevice, st));
            AHIP(hipMemcpyAsync(d_z, zmask, 8 * nterms, hipMemcpyHostToDevice, st));
            AHIP(hipMemcpyAsync(d_cr, cre.data(), 8 * nterms, hipMemcpyHostToDevice, st));
            AHIP(hipMemcpyAsync(d_ci, cim.data(), 8 * nterms, hipMemcpyHostToDevice, st));
        }
        int g = (int)std::min<uint64_t>((N + 255) / 256, 8192);
        if (nterms == 0) {
            AHIP(hipMemsetAsync(psi2, 0, (prec == KSV_F64 ? 16 : 8) << n, st));
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
        hipFree(d_x);
        hipFree(d_z);
        hipFree(d_cr);
        hipFree(d_ci);
    }
    // backward sweep: every parametrised op writes its per-block partial
    // sums into its own slice of `slots`; one read-back at the end
    std::vector<ksv_op> inv;
    std::vector<double> im;
    size_t end = nops;
    const int GMAX = 1024;
    double* slots = nullptr;
    const size_t nslots = std::max<size_t>(recs.size(), 1);
    AHIP(hipMalloc(&slots, sizeof(double) * 3 * GMAX * nslots));
    struct Pending { int op; int g; size_t slot; };
    std::vector<Pending> pend;
    while (end > 0) {
        // maximal run of parameter-free ops ending at `end`
        size_t b = end;
        while (b > 0 && !recs.count((int)(b - 1))) b--;
        if (b < end) {
            invert_tape(ops, b, end, mats, inv, im);
            rc = adj_run_tape(s, inv.data(), inv.size(), im.data(), im.size(), 0, psi);
            if (!rc) rc = adj_run_tape(s, inv.data(), inv.size(), im.data(), im.size(), 0, psi2);
            if (rc) {
                hipFree(slots);
                return rc;
            }
            end = b;
            if (end == 0) break;
        }
        // op end-1 carries parameters
        const int oi = (int)(end - 1);
        const ksv_op& o = ops[oi];
        const auto& rv = recs[oi];
        AdjArgs a;
        std::memset(&a, 0, sizeof a);
        a.phi = psi;
        a.lam = psi2;
        a.nrec = (int)rv.size();
        int dim;
        std::vector<int> insb;
        if (o.kind == KSV_SINGLE) { a.k = 1; a.t[0] = (int8_t)o.q[0]; dim = 2; insb = {o.q[0]}; }
        else if (o.kind == KSV_CONTROLLED) {
            a.k = 1; a.t[0] = (int8_t)o.q[1]; dim = 2; insb = {o.q[0], o.q[1]};
            a.ctl = 1ull << o.q[0];
        } else {
            if (o.q[0] == o.q[1]) {
                hipFree(slots);
                return adj_fail(KSV_ERR_TYPE, "two-qubit gate needs distinct qubits");
            }
            a.k = 2; a.t[0] = (int8_t)o.q[0]; a.t[1] = (int8_t)o.q[1]; dim = 4; insb = {o.q[0], o.q[1]};
        }
        std::sort(insb.begin(), insb.end());
        a.nins = (int)insb.size();
        for (size_t i = 0; i < insb.size(); i++) a.ins[i] = (int8_t)insb[i];
        std::vector<cd> m((size_t)dim * dim), ad;
        for (size_t j = 0; j < m.size(); j++) m[j] = cd(mats[o.mat + 2 * j], mats[o.mat + 2 * j + 1]);
        adjoint_of(m, dim, ad);
        for (size_t j = 0; j < ad.size(); j++) { a.ud[2 * j] = ad[j].real(); a.ud[2 * j + 1] = ad[j].imag(); }
        for (size_t r = 0; r < rv.size(); r++)
            for (int j = 0; j < 2 * dim * dim; j++) a.du[r][j] = mats[rv[r]->dmat + j];
        const uint64_t groups = N >> a.nins;
        int g = (int)std::min<uint64_t>((groups + 255) / 256, GMAX);
        if (g < 1) g = 1;
        double* out = slots + pend.size() * 3 * GMAX;
        if (prec == KSV_F64) hipLaunchKernelGGL((k_adj<double2>), dim3(g), dim3(256), 0, st, a, groups, out);
        else hipLaunchKernelGGL((k_adj<float2>), dim3(g), dim3(256), 0, st, a, groups, out);
        if (hipGetLastError() != hipSuccess) {
            hipFree(slots);
            return adj_fail(KSV_ERR_DEVICE, "adjoint kernel launch failed");
        }
        pend.push_back({oi, g, pend.size()});
        end = end - 1;
    }
    std::vector<double> hp(3 * GMAX * nslots);
    AHIP(hipMemcpyAsync(hp.data(), slots, 8 * hp.size(), hipMemcpyDeviceToHost, st));
    AHIP(hipStreamSynchronize(st));
    hipFree(slots);
    for (const Pending& pd : pend) {
        const auto& rv = recs[pd.op];
        const double* h = hp.data() + pd.slot * 3 * GMAX;
        for (size_t r = 0; r < rv.size(); r++) {
            double sum = 0;
            for (int b2 = 0; b2 < pd.g; b2++) sum += h[(size_t)b2 * 3 + r];
            grad[rv[r]->param] += 2.0 * rv[r]->coeff * sum;
        }
    }
    return KSV_OK;
}
