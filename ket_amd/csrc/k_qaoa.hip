// Batched QAOA for Max-Cut: many instances of uneven width in one set of
// launches (ket's qaoa() and maxCutHamiltonian(), src/algorithms.ts:231-276).
//
//   psi_0 = H^n |0>,   per layer l:  psi <- D_l psi,  psi <- M_l psi
//   D_l(x) = exp(i gamma_l S(x) / 2),  S(x) = |E| - 2 cut(x)      (cnot . rz(-gamma) . cnot per edge)
//   M_l    = prod_q Rx_q(2 beta_l) = exp(-i beta_l sum_q X_q)
//   E      = sum_x cut(x) |psi(x)|^2                             (H_C = sum_e (I - Z_u Z_v) / 2)
//
// Adjoint gradients (lambda = H_C psi, swept back through the layers with psi):
//   dE/dbeta_l  = 2 Im <lambda| sum_q X_q |psi>   (any point of layer l's mixer: X_q commutes with it)
//   dE/dgamma_l = -Im sum_x S(x) conj(lambda(x)) psi(x)          (between the mixer and the cost layer)
//
// Every instance's state is a slice of one buffer.  A launch covers all
// instances: work item w is a 12-qubit tile of instance winst[w].  Pass A
// tiles are qubits 0..11 (cost phase + mixer on them); pass B tiles are
// qubits 0..L-1 and the high qubits 12..n-1, L = 24 - n (mixer on the high
// ones; L >= 2 keeps HBM runs of at least 64 bytes).  Inside a tile 256
// threads hold 16 amplitudes each; the mixer runs in three register rounds (tile bits 8..11, 4..7, 0..3 in the register
// index) separated by XOR-swizzled LDS transposes, so every slot is a
// compile-time register index and the code is straight-line.
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "kcommon.h"

namespace ksv {

struct QInst {
    int32_t n;
    int32_t nedges;
    uint64_t off;          // first amplitude of the instance in the state buffers
    uint32_t adj[24];      // neighbour mask of every vertex
    int32_t lb;            // pass B: low qubits in the tile (L = 24 - n)
    uint32_t mixb;         // pass B: tile bits whose qubit gets the mixer (L..11)
    int32_t gbase;         // index of this instance's gamma/beta row
    int32_t pad;
};

struct QArgs {
    void* psi;
    void* lam;
    const QInst* inst;
    const uint16_t* winst;     // instance of every work item
    const uint64_t* wstart;    // first work item of every instance
    const double* gamma;
    const double* beta;
    double* part;              // partial sums of this launch (2 per work item)
    const uint8_t* ll;         // per instance: cut of the low 12 qubits for each of the 4096 tile indices
    const double* phs;         // per instance and layer: exp(i gamma S / 2) for each cut value (re, im)
    int layer;
    int p;
    int mode;                  // QM_*
};

enum QMode { QM_INIT = 0, QM_FWD_A = 1, QM_FWD_B = 2, QM_ENERGY = 3, QM_BWD_B = 4, QM_BWD_A = 5 };

constexpr int QNT = 256, QR = 16;
constexpr int QPMAX = 232;     // phase-table entries per (instance, layer): cut values 0..231 (n <= 22)

template <typename T> struct QV2;
template <> struct QV2<double> { using type = double2; };
template <> struct QV2<float> { using type = float2; };

__device__ __forceinline__ uint32_t qswz(uint32_t x) { return x ^ ((x >> 4) & 15); }

// tile index of (thread t, register j) in layout L (1: regs = tile bits 8..11,
// 2: regs = 4..7, 3: regs = 0..3)
template <int L>
__device__ __forceinline__ uint32_t qidx(uint32_t t, uint32_t j) {
    if (L == 1) return t | (j << 8);
    if (L == 2) return (t & 15) | (j << 4) | ((t >> 4) << 8);
    return (t << 4) | j;
}

template <typename T, int LW, int LR>
__device__ __forceinline__ void qexch(T (&re)[QR], T (&im)[QR], typename QV2<T>::type* lds, uint32_t tid_in) {
    using V = typename QV2<T>::type;
    // the LDS addresses are formed in every exchange: shared between the
    // exchanges of psi and lambda, they would hold 48 registers across the mixer
    uint32_t tid = tid_in;
    asm volatile("" : "+v"(tid));
    __syncthreads();
#pragma unroll
    for (int j = 0; j < QR; j++) {
        V v;
        v.x = re[j];
        v.y = im[j];
        lds[qswz(qidx<LW>(tid, j))] = v;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < QR; j++) {
        const V v = lds[qswz(qidx<LR>(tid, j))];
        re[j] = v.x;
        im[j] = v.y;
    }
}

// Rx(2 beta) = [[c, -i s], [-i s, c]] on slot S
template <typename T, int S>
__device__ __forceinline__ void qrx(T (&re)[QR], T (&im)[QR], T c, T s) {
#pragma unroll
    for (int j = 0; j < QR; j++) {
        if ((j >> S) & 1) continue;
        const int j1 = j | (1 << S);
        const T ar = re[j], ai = im[j], br = re[j1], bi = im[j1];
        re[j] = __builtin_fma(c, ar, s * bi);
        im[j] = __builtin_fma(c, ai, -(s * br));
        re[j1] = __builtin_fma(c, br, s * ai);
        im[j1] = __builtin_fma(c, bi, -(s * ar));
    }
}

// Im sum over pairs on slot S of conj(l_a) p_b + conj(l_b) p_a
template <typename T, int S>
__device__ __forceinline__ double qxpair(const T (&pr)[QR], const T (&pi)[QR], const T (&lr)[QR], const T (&li)[QR]) {
    double acc = 0;
#pragma unroll
    for (int j = 0; j < QR; j++) {
        if ((j >> S) & 1) continue;
        const int j1 = j | (1 << S);
        acc += (double)lr[j] * pi[j1] - (double)li[j] * pr[j1];
        acc += (double)lr[j1] * pi[j] - (double)li[j1] * pr[j];
    }
    return acc;
}

// mixer on the slots whose tile bit (lo + s) is in `mask`; with GRAD also
// the dE/dbeta pair sums (psi = p*, lambda = l*) before the rotation
template <typename T, bool BWD>
__device__ __forceinline__ void qmix_round(T (&pr)[QR], T (&pi)[QR], T (&lr)[QR], T (&li)[QR], uint32_t mask4, T c,
                                           T s, double& gb) {
#define QSLOT(S)                                                     \
    if ((mask4 >> S) & 1) {                                          \
        if (BWD) {                                                   \
            gb += qxpair<T, S>(pr, pi, lr, li);                      \
            qrx<T, S>(lr, li, c, s);                                 \
        }                                                            \
        qrx<T, S>(pr, pi, c, s);                                     \
    }
    QSLOT(0) QSLOT(1) QSLOT(2) QSLOT(3)
#undef QSLOT
}

// One kernel per mode (MODE = a QMode): register allocation is per kernel, and
// the backward modes hold phi and lambda, so a single kernel for every mode
// would give the forward ones the backward ones' registers.
template <typename T, int MODE>
__global__ __launch_bounds__(QNT) void k_qaoa(QArgs a) {
    using V = typename QV2<T>::type;
    __shared__ V lds[1 << 12];
    const uint32_t tid = threadIdx.x;
    const uint64_t w = blockIdx.x;
    const int ii = a.winst[w];
    const QInst& q = a.inst[ii];
    const uint64_t t = w - a.wstart[ii];
    uint32_t adj[24];
#pragma unroll
    for (int u = 0; u < 24; u++) adj[u] = q.adj[u];
    const int layer = a.layer;
    V* psi = reinterpret_cast<V*>(a.psi) + q.off;
    V* lam = reinterpret_cast<V*>(a.lam) + q.off;
    const bool passb = MODE == QM_FWD_B || MODE == QM_BWD_B;

    // physical index of tile index x: pass A tiles are qubits 0..11 (tile t
    // on qubits 12..n-1); pass B tiles are qubits 0..L-1 on tile bits 0..L-1
    // and qubits 12..n-1 on tile bits L..11 (tile t on qubits L..11)
    const int L = q.lb;
    auto phys = [&](uint32_t x) -> uint64_t {
        if (!passb) return (t << 12) | x;
        return ((uint64_t)(x >> L) << 12) | (t << L) | (x & ((1u << L) - 1));
    };

    // pass A tiling (FWD_A, ENERGY, BWD_A): cut(x) = ll[x_lo] + cut_hh + sum_u<12 (x_u ? a1_u : a0_u),
    // with the low-low cuts of all 4096 tile indices and the phases of every cut value in LDS
    __shared__ uint8_t ll[1 << 12];
    __shared__ V phs[256];
    int cut_hh = 0;
    uint32_t a01[12];                      // a0 in the low 16 bits, a1 in the high 16 bits
    const bool tab = MODE == QM_FWD_A || MODE == QM_ENERGY || MODE == QM_BWD_A;
    if (tab) {
        const uint32_t th = (uint32_t)(t << 12);      // the tile's high bits
        const uint32_t hm = (uint32_t)(((1ull << q.n) - 1) & ~0xFFFull);
        // the instance's low-low cut table (k_qll), 16 bytes per thread
        reinterpret_cast<uint4*>(ll)[tid] = reinterpret_cast<const uint4*>(a.ll + (size_t)ii * 4096)[tid];
        int hh = 0;
        for (int u = 12; u < q.n; u++) hh += __popc(adj[u] & hm & (((th >> u) & 1) ? ~th : th));
        cut_hh = hh >> 1;
#pragma unroll
        for (int u = 0; u < 12; u++) {
            const uint32_t a0 = (uint32_t)__popc(adj[u] & hm & th), a1 = (uint32_t)__popc(adj[u] & hm & ~th);
            a01[u] = a0 | (a1 << 16);
        }
        if (MODE != QM_ENERGY) {
            // this layer's phases (k_qphs); the backward pass takes their conjugates
            const double* ph = a.phs + ((size_t)q.gbase + a.layer) * QPMAX * 2;
            const double sg = MODE == QM_BWD_A ? -1.0 : 1.0;
            for (int c = (int)tid; c <= q.nedges; c += QNT) {
                V v;
                v.x = (T)ph[2 * c];
                v.y = (T)(sg * ph[2 * c + 1]);
                phs[c] = v;
            }
        }
        __syncthreads();
    }
    auto cut_of = [&](uint32_t x) -> int {
        int c = ll[x] + cut_hh;
#pragma unroll
        for (int u = 0; u < 12; u++) c += (int)(((x >> u) & 1) ? (a01[u] >> 16) : (a01[u] & 0xFFFF));
        return c;
    };

    T pr[QR], pi[QR], lr[QR], li[QR];
    if (MODE == QM_INIT) {
        const T amp = (T)(1.0 / sqrt((double)(1ull << q.n)));
#pragma unroll
        for (int j = 0; j < QR; j++) {
            V v;
            v.x = amp;
            v.y = 0;
            psi[phys(qidx<1>(tid, j))] = v;
        }
        return;
    }
    const bool bwd = MODE == QM_BWD_A || MODE == QM_BWD_B;
#pragma unroll
    for (int j = 0; j < QR; j++) {
        const uint64_t g = phys(qidx<1>(tid, j));
        const V v = psi[g];
        pr[j] = v.x;
        pi[j] = v.y;
        if (bwd) {
            const V u = lam[g];
            lr[j] = u.x;
            li[j] = u.y;
        }
    }
    if (MODE == QM_ENERGY) {
        double e = 0;
#pragma unroll
        for (int j = 0; j < QR; j++) {
            const uint32_t x = qidx<1>(tid, j);
            const uint64_t g = phys(x);
            const int c = cut_of(x);
            e += (double)c * ((double)pr[j] * pr[j] + (double)pi[j] * pi[j]);
            V u;
            u.x = (T)c * pr[j];
            u.y = (T)c * pi[j];
            lam[g] = u;
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) e += __shfl_down(e, o, 64);
        __shared__ double sh[4];
        if ((tid & 63) == 0) sh[tid >> 6] = e;
        __syncthreads();
        if (tid == 0) {
            a.part[2 * w] = sh[0] + sh[1] + sh[2] + sh[3];
            a.part[2 * w + 1] = 0;
        }
        return;
    }

    const double bet = a.beta[q.gbase + layer];
    const T c = (T)cos(bet), s = (T)(bwd ? -sin(bet) : sin(bet));
    const uint32_t mask = passb ? q.mixb : 0xFFFu;
    double gb = 0, gg = 0;

    // forward pass A: cost phase first
    if (MODE == QM_FWD_A) {
#pragma unroll
        for (int j = 0; j < QR; j++) {
            const V phv = phs[cut_of(qidx<1>(tid, j))];
            const T cr = phv.x, ci = phv.y;
            const T xr = pr[j], xi = pi[j];
            pr[j] = __builtin_fma(cr, xr, -(ci * xi));
            pi[j] = __builtin_fma(cr, xi, ci * xr);
        }
    }
    // round 1: tile bits 8..11 in registers
    if (bwd) qmix_round<T, true>(pr, pi, lr, li, (mask >> 8) & 15, c, s, gb);
    else qmix_round<T, false>(pr, pi, lr, li, (mask >> 8) & 15, c, s, gb);
    if (mask & 0xFF) {
        qexch<T, 1, 2>(pr, pi, lds, tid);
        if (bwd) qexch<T, 1, 2>(lr, li, lds, tid);
        if (bwd) qmix_round<T, true>(pr, pi, lr, li, (mask >> 4) & 15, c, s, gb);
        else qmix_round<T, false>(pr, pi, lr, li, (mask >> 4) & 15, c, s, gb);
        qexch<T, 2, 3>(pr, pi, lds, tid);
        if (bwd) qexch<T, 2, 3>(lr, li, lds, tid);
        if (bwd) qmix_round<T, true>(pr, pi, lr, li, mask & 15, c, s, gb);
        else qmix_round<T, false>(pr, pi, lr, li, mask & 15, c, s, gb);
        qexch<T, 3, 1>(pr, pi, lds, tid);
        if (bwd) qexch<T, 3, 1>(lr, li, lds, tid);
    }
    // backward pass A: the mixer is undone, now dE/dgamma and the inverse cost phase
    if (MODE == QM_BWD_A) {
#pragma unroll
        for (int j = 0; j < QR; j++) {
            const int cut = cut_of(qidx<1>(tid, j));
            const int S = q.nedges - 2 * cut;
            gg -= (double)S * ((double)lr[j] * pi[j] - (double)li[j] * pr[j]);
            const V phv = phs[cut];
            const T cr = phv.x, ci = phv.y;
            T xr = pr[j], xi = pi[j];
            pr[j] = __builtin_fma(cr, xr, -(ci * xi));
            pi[j] = __builtin_fma(cr, xi, ci * xr);
            xr = lr[j];
            xi = li[j];
            lr[j] = __builtin_fma(cr, xr, -(ci * xi));
            li[j] = __builtin_fma(cr, xi, ci * xr);
        }
    }
    // the store addresses are formed again here: kept from the loads, they would
    // hold 32 registers across the whole mixer
    uint32_t tid_st = tid;
    asm volatile("" : "+v"(tid_st));
#pragma unroll
    for (int j = 0; j < QR; j++) {
        const uint64_t g = phys(qidx<1>(tid_st, j));
        V v;
        v.x = pr[j];
        v.y = pi[j];
        psi[g] = v;
        if (bwd) {
            V u;
            u.x = lr[j];
            u.y = li[j];
            lam[g] = u;
        }
    }
    if (bwd) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            gb += __shfl_down(gb, o, 64);
            gg += __shfl_down(gg, o, 64);
        }
        __shared__ double sb[4], sg[4];
        if ((tid & 63) == 0) {
            sb[tid >> 6] = gb;
            sg[tid >> 6] = gg;
        }
        __syncthreads();
        if (tid == 0) {
            a.part[2 * w] = sb[0] + sb[1] + sb[2] + sb[3];
            a.part[2 * w + 1] = sg[0] + sg[1] + sg[2] + sg[3];
        }
    }
}

// Cut of the low 12 qubits (the edges among them) for each of the 4096
// tile indices of instance blockIdx.x.
__global__ __launch_bounds__(256) void k_qll(const QInst* inst, uint8_t* ll) {
    const QInst& q = inst[blockIdx.x];
    uint32_t adj[12];
#pragma unroll
    for (int u = 0; u < 12; u++) adj[u] = q.adj[u] & 0xFFFu;
    for (uint32_t e = threadIdx.x; e < 4096; e += 256) {
        int c = 0;
#pragma unroll
        for (int u = 0; u < 12; u++) c += __popc(adj[u] & (((e >> u) & 1) ? ~e : e));
        ll[(size_t)blockIdx.x * 4096 + e] = (uint8_t)(c >> 1);
    }
}

// exp(i gamma_l S / 2), S = |E| - 2 cut, for every cut value of instance
// blockIdx.x and layer blockIdx.y (S / 2 the phase of cnot . rz(-gamma) . cnot per edge).
__global__ __launch_bounds__(256) void k_qphs(const QInst* inst, const double* gamma, double* phs) {
    const QInst& q = inst[blockIdx.x];
    const double g = gamma[q.gbase + blockIdx.y] * 0.5;
    double* out = phs + ((size_t)q.gbase + blockIdx.y) * QPMAX * 2;
    for (int c = (int)threadIdx.x; c <= q.nedges; c += 256) {
        double sn, cs;
        sincos(g * (double)(q.nedges - 2 * c), &sn, &cs);
        out[2 * c] = cs;
        out[2 * c + 1] = sn;
    }
}

// out[i * stride + k], k < 2: per-instance sums of the two partials of a launch
__global__ __launch_bounds__(256) void k_qseg(const double* part, const uint64_t* wstart, int ninst, double* out,
                                              int stride) {
    const int i = blockIdx.x;
    if (i >= ninst) return;
    const uint64_t b = wstart[i], e = wstart[i + 1];
    double s0 = 0, s1 = 0;
    for (uint64_t w = b + threadIdx.x; w < e; w += blockDim.x) {
        s0 += part[2 * w];
        s1 += part[2 * w + 1];
    }
    __shared__ double r0[256], r1[256];
    r0[threadIdx.x] = s0;
    r1[threadIdx.x] = s1;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            r0[threadIdx.x] += r0[threadIdx.x + o];
            r1[threadIdx.x] += r1[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[(size_t)i * stride] = r0[0];
        out[(size_t)i * stride + 1] = r1[0];
    }
}

}  // namespace ksv

using namespace ksv;

struct ksv_batch {
    int count = 0, p = 0, prec = KSV_F64;
    uint64_t total_amps = 0, nwork = 0;
    hipStream_t st = nullptr;
    void* psi = nullptr;
    void* lam = nullptr;
    QInst* d_inst = nullptr;
    uint16_t* d_winst = nullptr;
    uint64_t* d_wstart = nullptr;
    double *d_gamma = nullptr, *d_beta = nullptr;
    double* d_part = nullptr;     // (2p + 1) launches x 2 x nwork
    uint8_t* d_ll = nullptr;      // count x 4096 low-qubit cut table
    double* d_phs = nullptr;      // count x p x QPMAX phases (re, im)
    double* d_out = nullptr;      // count x (2p + 1) x 2
    std::vector<QInst> inst;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};

    // releases whatever was allocated, so every error return of create / run frees it
    ~ksv_batch() {
        if (st) (void)hipStreamSynchronize(st);
        for (int g = 0; g < 2; g++) {
            if (exec[g]) (void)hipGraphExecDestroy(exec[g]);
            if (graph[g]) (void)hipGraphDestroy(graph[g]);
        }
        void* bufs[] = {psi, lam, d_inst, d_winst, d_wstart, d_gamma, d_beta, d_part, d_out, d_ll, d_phs};
        for (void* ptr : bufs)
            if (ptr) (void)hipFree(ptr);
        if (st) (void)hipStreamDestroy(st);
    }
};

static thread_local std::string q_err;
static int qfail(int code, const std::string& m) {
    q_err = m;
    return code;
}
#define QHIP(expr)                                                                              \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) return qfail(KSV_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace ksv {
int adj_fail(int code, const std::string& msg);
}

static int qlaunch(ksv_batch* b, int mode, int layer, double* part) {
    QArgs a;
    std::memset(&a, 0, sizeof a);
    a.psi = b->psi;
    a.lam = b->lam;
    a.inst = b->d_inst;
    a.winst = b->d_winst;
    a.wstart = b->d_wstart;
    a.gamma = b->d_gamma;
    a.beta = b->d_beta;
    a.part = part;
    a.ll = b->d_ll;
    a.phs = b->d_phs;
    a.layer = layer;
    a.p = b->p;
    a.mode = mode;
    const dim3 grid((unsigned)b->nwork), blk(QNT);
#define QL(M)                                                                         \
    if (b->prec == KSV_F64) hipLaunchKernelGGL((k_qaoa<double, M>), grid, blk, 0, b->st, a); \
    else hipLaunchKernelGGL((k_qaoa<float, M>), grid, blk, 0, b->st, a);
    switch (mode) {
    case QM_INIT: QL(QM_INIT) break;
    case QM_FWD_A: QL(QM_FWD_A) break;
    case QM_FWD_B: QL(QM_FWD_B) break;
    case QM_ENERGY: QL(QM_ENERGY) break;
    case QM_BWD_B: QL(QM_BWD_B) break;
    case QM_BWD_A: QL(QM_BWD_A) break;
    default: return qfail(KSV_ERR_RANGE, "bad QAOA kernel mode");
    }
#undef QL
    QHIP(hipGetLastError());
    return KSV_OK;
}

// the whole forward (+ backward) schedule; launch slot k writes d_part + 2 nwork k
static int qschedule(ksv_batch* b, bool grad) {
    const int P = b->p;
    hipLaunchKernelGGL(k_qphs, dim3(b->count, P), dim3(256), 0, b->st, b->d_inst, b->d_gamma, b->d_phs);
    QHIP(hipGetLastError());
    int rc = qlaunch(b, QM_INIT, 0, b->d_part);
    for (int l = 0; l < P && !rc; l++) {
        rc = qlaunch(b, QM_FWD_A, l, b->d_part);
        if (!rc) rc = qlaunch(b, QM_FWD_B, l, b->d_part);
    }
    if (rc) return rc;
    const uint64_t W = b->nwork;
    rc = qlaunch(b, QM_ENERGY, 0, b->d_part);
    if (rc) return rc;
    hipLaunchKernelGGL(k_qseg, dim3(b->count), dim3(256), 0, b->st, b->d_part, b->d_wstart, b->count, b->d_out,
                       2 * (2 * P + 1));
    QHIP(hipGetLastError());
    if (!grad) return KSV_OK;
    for (int l = P - 1; l >= 0; l--) {
        double* pb = b->d_part + 2 * W * (1 + 2 * (P - 1 - l));
        double* pa = pb + 2 * W;
        rc = qlaunch(b, QM_BWD_B, l, pb);
        if (!rc) rc = qlaunch(b, QM_BWD_A, l, pa);
        if (rc) return rc;
        const int slot = 1 + 2 * (P - 1 - l);
        hipLaunchKernelGGL(k_qseg, dim3(b->count), dim3(256), 0, b->st, pb, b->d_wstart, b->count,
                           b->d_out + 2 * slot, 2 * (2 * P + 1));
        hipLaunchKernelGGL(k_qseg, dim3(b->count), dim3(256), 0, b->st, pa, b->d_wstart, b->count,
                           b->d_out + 2 * (slot + 1), 2 * (2 * P + 1));
        QHIP(hipGetLastError());
    }
    return KSV_OK;
}

extern "C" {

int ksv_batch_last_error(char* buf, size_t len) {
    if (buf && len) {
        size_t m = std::min(len - 1, q_err.size());
        std::memcpy(buf, q_err.data(), m);
        buf[m] = 0;
    }
    return (int)q_err.size();
}

int ksv_qaoa_batch_create(int count, const int32_t* nq, const int64_t* edge_off, const int32_t* edges, int p,
                          int precision, ksv_batch** out) {
    if (!out) return qfail(KSV_ERR_TYPE, "null output pointer");
    *out = nullptr;
    if (count < 1 || count > 65535) return qfail(KSV_ERR_RANGE, "instance count must be 1..65535");
    if (p < 1 || p > 64) return qfail(KSV_ERR_RANGE, "layer count must be 1..64");
    if (!nq || !edge_off || (edge_off[count] > 0 && !edges)) return qfail(KSV_ERR_TYPE, "null instance buffer");
    if (precision != KSV_F64 && precision != KSV_F32) return qfail(KSV_ERR_TYPE, "unknown precision");
    std::unique_ptr<ksv_batch> b(new ksv_batch());
    b->count = count;
    b->p = p;
    b->prec = precision;
    std::vector<uint16_t> winst;
    std::vector<uint64_t> wstart(count + 1, 0);
    uint64_t off = 0;
    for (int i = 0; i < count; i++) {
        const int n = nq[i];
        if (n < 12 || n > 22) return qfail(KSV_ERR_RANGE, "batched QAOA instances need 12..22 qubits");
        QInst q;
        std::memset(&q, 0, sizeof q);
        q.n = n;
        q.off = off;
        q.gbase = i * p;
        for (int64_t e = edge_off[i]; e < edge_off[i + 1]; e++) {
            const int u = edges[2 * e], v = edges[2 * e + 1];
            if (u < 0 || v < 0 || u >= n || v >= n || u == v) return qfail(KSV_ERR_RANGE, "edge endpoint out of range");
            // the cut counts each vertex pair once (adjacency bits), while qaoa() and
            // maxCutHamiltonian() apply every listed edge: a repeat would be silently dropped
            if (q.adj[u] & (1u << v)) return qfail(KSV_ERR_RANGE, "repeated edge (u, v) or (v, u) in one instance");
            q.adj[u] |= 1u << v;
            q.adj[v] |= 1u << u;
            q.nedges++;
        }
        if (q.nedges >= QPMAX) return qfail(KSV_ERR_RANGE, "too many edges for the phase table");
        // pass B tile: qubits 0..L-1 and 12..n-1 (L = 24 - n); the mixer on tile bits L..11
        q.lb = 24 - n;
        q.mixb = 0xFFFu & ~((1u << q.lb) - 1);
        b->inst.push_back(q);
        const uint64_t tiles = 1ull << (n - 12);
        wstart[i] = winst.size();
        for (uint64_t k = 0; k < tiles; k++) winst.push_back((uint16_t)i);
        off += 1ull << n;
    }
    wstart[count] = winst.size();
    b->total_amps = off;
    b->nwork = winst.size();
    const size_t amp = precision == KSV_F64 ? 16 : 8;
    QHIP(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));
    if (hipMalloc(&b->psi, amp * off) != hipSuccess || hipMalloc(&b->lam, amp * off) != hipSuccess)
        return qfail(KSV_ERR_DEVICE, "out of device memory for the batch states");
    QHIP(hipMalloc(&b->d_inst, sizeof(QInst) * count));
    QHIP(hipMalloc(&b->d_winst, 2 * winst.size()));
    QHIP(hipMalloc(&b->d_wstart, 8 * (count + 1)));
    QHIP(hipMalloc(&b->d_gamma, 8 * (size_t)count * p));
    QHIP(hipMalloc(&b->d_beta, 8 * (size_t)count * p));
    QHIP(hipMalloc(&b->d_part, 8 * 2 * b->nwork * (2 * (size_t)p + 1)));
    QHIP(hipMalloc(&b->d_out, 8 * 2 * (size_t)count * (2 * p + 1)));
    QHIP(hipMemcpy(b->d_inst, b->inst.data(), sizeof(QInst) * count, hipMemcpyHostToDevice));
    QHIP(hipMemcpy(b->d_winst, winst.data(), 2 * winst.size(), hipMemcpyHostToDevice));
    QHIP(hipMemcpy(b->d_wstart, wstart.data(), 8 * (count + 1), hipMemcpyHostToDevice));
    QHIP(hipMalloc(&b->d_ll, 4096 * (size_t)count));
    QHIP(hipMalloc(&b->d_phs, 16 * (size_t)count * p * QPMAX));
    hipLaunchKernelGGL(k_qll, dim3(count), dim3(256), 0, b->st, b->d_inst, b->d_ll);
    QHIP(hipGetLastError());
    QHIP(hipStreamSynchronize(b->st));
    *out = b.release();
    return KSV_OK;
}

int ksv_qaoa_batch_run(ksv_batch* b, const double* gamma, const double* beta, int want_grad, double* energy,
                       double* dgamma, double* dbeta) {
    if (!b) return qfail(KSV_ERR_HANDLE, "null batch");
    if (!gamma || !beta) return qfail(KSV_ERR_TYPE, "null angle buffer");
    if (want_grad && (!dgamma || !dbeta)) return qfail(KSV_ERR_TYPE, "null gradient buffer");
    const size_t np = (size_t)b->count * b->p;
    QHIP(hipMemcpyAsync(b->d_gamma, gamma, 8 * np, hipMemcpyHostToDevice, b->st));
    QHIP(hipMemcpyAsync(b->d_beta, beta, 8 * np, hipMemcpyHostToDevice, b->st));
    const int g = want_grad ? 1 : 0;
    if (!b->exec[g]) {
        // capture the whole launch sequence once; replays read the angle buffers
        QHIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
        int rc = qschedule(b, want_grad != 0);
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(b->st, &gr);
        if (rc || e != hipSuccess) {
            if (gr) (void)hipGraphDestroy(gr);
            if (rc) return rc;
            return qfail(KSV_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
        }
        b->graph[g] = gr;
        QHIP(hipGraphInstantiate(&b->exec[g], gr, nullptr, nullptr, 0));
    }
    QHIP(hipGraphLaunch(b->exec[g], b->st));
    const int P = b->p, stride = 2 * (2 * P + 1);
    std::vector<double> out((size_t)b->count * stride);
    QHIP(hipMemcpyAsync(out.data(), b->d_out, 8 * out.size(), hipMemcpyDeviceToHost, b->st));
    QHIP(hipStreamSynchronize(b->st));
    for (int i = 0; i < b->count; i++) {
        const double* o = out.data() + (size_t)i * stride;
        if (energy) energy[i] = o[0];
        if (!want_grad) continue;
        for (int l = 0; l < P; l++) {
            const int slot = 1 + 2 * (P - 1 - l);
            const double gb = o[2 * slot] + o[2 * (slot + 1)];   // pass B and pass A mixer sums
            const double gg = o[2 * (slot + 1) + 1];
            dbeta[(size_t)i * P + l] = 2.0 * gb;
            dgamma[(size_t)i * P + l] = gg;
        }
    }
    return KSV_OK;
}

int ksv_qaoa_batch_amplitudes(ksv_batch* b, int inst, uint64_t offset, uint64_t count, double* out) {
    if (!b) return qfail(KSV_ERR_HANDLE, "null batch");
    if (inst < 0 || inst >= b->count) return qfail(KSV_ERR_RANGE, "instance index out of range");
    const uint64_t N = 1ull << b->inst[inst].n;
    if (offset > N || count > N - offset) return qfail(KSV_ERR_RANGE, "amplitude range out of bounds");
    if (!count) return KSV_OK;
    if (!out) return qfail(KSV_ERR_TYPE, "null output buffer");
    const uint64_t o = b->inst[inst].off + offset;
    if (b->prec == KSV_F64) {
        QHIP(hipMemcpy(out, (char*)b->psi + 16 * o, 16 * count, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> f(2 * count);
        QHIP(hipMemcpy(f.data(), (char*)b->psi + 8 * o, 8 * count, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < 2 * count; i++) out[i] = f[i];
    }
    return KSV_OK;
}

int ksv_qaoa_batch_info(const ksv_batch* b, uint64_t* total_amplitudes, uint64_t* work_items) {
    if (!b) return qfail(KSV_ERR_HANDLE, "null batch");
    if (total_amplitudes) *total_amplitudes = b->total_amps;
    if (work_items) *work_items = b->nwork;
    return KSV_OK;
}

int ksv_qaoa_batch_destroy(ksv_batch* b) {
    if (!b) return qfail(KSV_ERR_HANDLE, "null batch");
    delete b;
    return KSV_OK;
}

}  // extern "C"
