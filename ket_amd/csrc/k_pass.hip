// Fused LDS-tile pass kernel and the standalone wide-unitary kernel.
//
// One workgroup owns one tile: the 2^k amplitudes whose non-tile bits equal
// the workgroup's base index.  The tile always contains the lowest state
// bits, so each wavefront load/store is a run of contiguous amplitudes
// (coalesced 16-byte loads).  Every op of the pass runs on the LDS copy;
// HBM sees exactly one read and one write per amplitude per pass.
//
// All loops over per-op bit lists are unrolled to their static maximum with
// a guard, so kernel arguments and op fields stay in scalar registers
// (a dynamically indexed argument array would be spilled to scratch).
#include "kcommon.h"

namespace ksv {

// Element index of the e-th element of an op: insert zero bits at the op's
// ascending positions, then set the local control bits.
__device__ __forceinline__ uint32_t op_elem(uint32_t e, const DOp& op) {
    uint32_t l = e;
#pragma unroll
    for (int i = 0; i < 8; i++)
        if (i < op.nins) l = ins0(l, op.ins[i]);
    return l | op.lctl;
}

// Global offset of tile-local bits [lo, hi) of l.
__device__ __forceinline__ uint64_t scatter(uint32_t l, const PassArgs& a, int lo, int hi) {
    uint64_t off = 0;
#pragma unroll
    for (int i = 0; i < 16; i++)
        if (i >= lo && i < hi) off |= (uint64_t)((l >> i) & 1) << a.tq[i];
    return off;
}

template <typename V> struct PScal;
template <> struct PScal<double2> { using type = double; };
template <> struct PScal<float2> { using type = float; };

template <typename V, int KT>
__global__ __launch_bounds__(256) void k_pass(PassArgs a) {
    __shared__ V T[1 << KT];
    __shared__ uint64_t offj[64];
    const int nt = blockDim.x;              // power of two
    const int lnt = __ffs(nt) - 1;
    const int tid = threadIdx.x;
    const int k = a.k;
    const uint32_t tsize = 1u << k;
    V* psi = reinterpret_cast<V*>(a.psi);
    // op records, matrices and copy bitmaps through the constant address
    // space: uniform reads become scalar loads
    using S = typename PScal<V>::type;
    const KSV_CONST S* cm = to_const(reinterpret_cast<const S*>(a.mats));
    auto mat = [&](int i) {
        V r;
        r.x = cm[2 * i];
        r.y = cm[2 * i + 1];
        return r;
    };
    const KSV_CONST uint64_t* cops = to_const(reinterpret_cast<const uint64_t*>(a.ops));
    const KSV_CONST uint32_t* cbits = to_const(a.bits);

    uint64_t base = blockIdx.x;
#pragma unroll
    for (int i = 0; i < 16; i++)
        if (i < k) base = ins0_64(base, a.tq[i]);

    // element l = tid + j * nt: offset = scatter(tid) + scatter(j * nt)
    const int kj = k > lnt ? k - lnt : 0;
    if (tid < (1 << kj)) offj[tid] = scatter((uint32_t)tid << lnt, a, lnt, k);
    const uint64_t offt = base + scatter((uint32_t)tid, a, 0, lnt < k ? lnt : k);
    __syncthreads();
    for (uint32_t l = tid, j = 0; l < tsize; l += nt, j++) T[l] = psi[offt + offj[j]];
    __syncthreads();

    for (int o = 0; o < a.nops; o++) {
        DOp op;
        {
            uint64_t w[sizeof(DOp) / 8];
#pragma unroll
            for (int i = 0; i < (int)(sizeof(DOp) / 8); i++) w[i] = cops[(sizeof(DOp) / 8) * o + i];
            __builtin_memcpy(&op, w, sizeof op);
        }
        if ((base & op.gctl) != op.gctl) continue;
        if (op.cond) {      // only on the tiles of the copies in the op's bitmap
            const uint64_t copy = base >> a.nsys;
            if (!((cbits[op.cond - 1 + (copy >> 5)] >> (copy & 31)) & 1u)) continue;
        }
        const uint32_t ne = tsize >> op.nins;
        switch (op.kind) {
        case DK_M1: {
            const V m0 = mat(op.mat), m1 = mat(op.mat + 1);
            const V m2 = mat(op.mat + 2), m3 = mat(op.mat + 3);
            const uint32_t tb = 1u << op.t[0];
            uint32_t e = tid;
            // four disjoint pairs per iteration: all their LDS reads before any write
            for (; e + 3 * nt < ne; e += 4 * nt) {
                uint32_t l0[4], l1[4];
                V x0[4], x1[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    l0[u] = op_elem(e + u * nt, op);
                    l1[u] = l0[u] | tb;
                    x0[u] = T[l0[u]];
                    x1[u] = T[l1[u]];
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    T[l0[u]] = prune_amp(cmadd2(m0, x0[u], m1, x1[u]), a.prune);
                    T[l1[u]] = prune_amp(cmadd2(m2, x0[u], m3, x1[u]), a.prune);
                }
            }
            for (; e < ne; e += nt) {
                uint32_t l0 = op_elem(e, op), l1 = l0 | tb;
                V x0 = T[l0], x1 = T[l1];
                T[l0] = prune_amp(cmadd2(m0, x0, m1, x1), a.prune);
                T[l1] = prune_amp(cmadd2(m2, x0, m3, x1), a.prune);
            }
            break;
        }
        case DK_M2: {
            const uint32_t ba = 1u << op.t[0], bb = 1u << op.t[1];
            V m[16];
#pragma unroll
            for (int i = 0; i < 16; i++) m[i] = mat(op.mat + i);
            for (uint32_t e = tid; e < ne; e += nt) {
                uint32_t l = op_elem(e, op);
                const uint32_t id[4] = {l, l | bb, l | ba, l | ba | bb};
                V x[4];
#pragma unroll
                for (int c = 0; c < 4; c++) x[c] = T[id[c]];
#pragma unroll
                for (int r = 0; r < 4; r++) {       // ((m0 x0 + m1 x1) + m2 x2) + m3 x3, ket's order
                    V p = cmadd2(m[4 * r], x[0], m[4 * r + 1], x[1]);
                    kacc(p.x, p.y, m[4 * r + 2].x, m[4 * r + 2].y, x[2].x, x[2].y);
                    kacc(p.x, p.y, m[4 * r + 3].x, m[4 * r + 3].y, x[3].x, x[3].y);
                    T[id[r]] = prune_amp(p, a.prune);
                }
            }
            break;
        }
        case DK_MK: {
            const int kk = op.k, dim = 1 << kk;
            for (uint32_t e = tid; e < ne; e += nt) {
                uint32_t l = op_elem(e, op);
                uint32_t id[1 << MK_MAX];
                V x[1 << MK_MAX];
#pragma unroll
                for (int r = 0; r < (1 << MK_MAX); r++) {
                    if (r < dim) {
                        uint32_t g = l;
#pragma unroll
                        for (int b = 0; b < MK_MAX; b++)
                            if (b < kk && ((r >> (kk - 1 - b)) & 1)) g |= 1u << op.t[b];
                        id[r] = g;
                        x[r] = T[g];
                    }
                }
#pragma unroll
                for (int r = 0; r < (1 << MK_MAX); r++) {
                    if (r < dim) {
                        V acc;
                        acc.x = 0;
                        acc.y = 0;
#pragma unroll
                        for (int c = 0; c < (1 << MK_MAX); c++) {
                            if (c < dim) {
                                V mm = mat(op.mat + r * dim + c);
                                kacc(acc.x, acc.y, mm.x, mm.y, x[c].x, x[c].y);
                            }
                        }
                        T[id[r]] = prune_amp(acc, a.prune);
                    }
                }
            }
            break;
        }
        case DK_DIAG: {
            const int kk = op.k;
            uint32_t rg = 0;   // row bits contributed by global qubits
#pragma unroll
            for (int b = 0; b < DIAG_MAX; b++)
                if (b < kk && op.t[b] < 0 && ((base >> (-op.t[b] - 1)) & 1)) rg |= 1u << (kk - 1 - b);
            if (kk <= 2) {
                // at most four entries: uniform loads, selected per element
                const V d0 = mat(op.mat);
                const V d1 = kk >= 1 ? mat(op.mat + 1) : d0;
                const V d2 = kk == 2 ? mat(op.mat + 2) : d0;
                const V d3 = kk == 2 ? mat(op.mat + 3) : d0;
                for (uint32_t e = tid; e < ne; e += nt) {
                    uint32_t l = op_elem(e, op);
                    uint32_t r = rg;
#pragma unroll
                    for (int b = 0; b < 2; b++)
                        if (b < kk && op.t[b] >= 0 && ((l >> op.t[b]) & 1)) r |= 1u << (kk - 1 - b);
                    const V dv = r == 0 ? d0 : r == 1 ? d1 : r == 2 ? d2 : d3;
                    T[l] = prune_amp(cmul(dv, T[l]), a.prune);
                }
                break;
            }
            for (uint32_t e = tid; e < ne; e += nt) {
                uint32_t l = op_elem(e, op);
                uint32_t r = rg;
#pragma unroll
                for (int b = 0; b < DIAG_MAX; b++)
                    if (b < kk && op.t[b] >= 0 && ((l >> op.t[b]) & 1)) r |= 1u << (kk - 1 - b);
                T[l] = prune_amp(cmul(mat(op.mat + (int)r), T[l]), a.prune);
            }
            break;
        }
        case DK_PERMX: {
            const uint32_t tb = 1u << op.t[0];
            for (uint32_t e = tid; e < ne; e += nt) {
                uint32_t l0 = op_elem(e, op), l1 = l0 | tb;
                V x0 = T[l0];
                T[l0] = T[l1];
                T[l1] = x0;
            }
            break;
        }
        case DK_SWAP: {
            const uint32_t ba = 1u << op.t[0], bb = 1u << op.t[1];
            for (uint32_t e = tid; e < ne; e += nt) {
                uint32_t l = op_elem(e, op);
                V x0 = T[l | ba];
                T[l | ba] = T[l | bb];
                T[l | bb] = x0;
            }
            break;
        }
        }
        __syncthreads();
    }

    for (uint32_t l = tid, j = 0; l < tsize; l += nt, j++) psi[offt + offj[j]] = T[l];
}

// Standalone 2^k x 2^k unitary (5 <= k <= 12) for ket's wide `unitary`
// ops: one workgroup per group of 2^k amplitudes staged in LDS.
struct BigArgs {
    void* psi;
    const void* m;
    int k;
    int prune;
    int8_t tg[12];    // targets, tg[0] = MSB of the local index
    int8_t srt[12];   // the same, ascending
};

template <typename V>
__global__ __launch_bounds__(256) void k_big(BigArgs a) {
    extern __shared__ unsigned char smem[];
    V* x = reinterpret_cast<V*>(smem);
    V* psi = reinterpret_cast<V*>(a.psi);
    const V* m = reinterpret_cast<const V*>(a.m);
    const int k = a.k, dim = 1 << k;
    uint64_t base = blockIdx.x;
#pragma unroll
    for (int i = 0; i < 12; i++)
        if (i < k) base = ins0_64(base, a.srt[i]);
    auto gidx = [&](int r) {
        uint64_t g = base;
#pragma unroll
        for (int b = 0; b < 12; b++)
            if (b < k && ((r >> (k - 1 - b)) & 1)) g |= 1ull << a.tg[b];
        return g;
    };
    for (int r = threadIdx.x; r < dim; r += blockDim.x) x[r] = psi[gidx(r)];
    __syncthreads();
    for (int r = threadIdx.x; r < dim; r += blockDim.x) {
        V acc;
        acc.x = 0;
        acc.y = 0;
        for (int c = 0; c < dim; c++) {
            V mm = m[(size_t)r * dim + c];
            kacc(acc.x, acc.y, mm.x, mm.y, x[c].x, x[c].y);
        }
        psi[gidx(r)] = prune_amp(acc, a.prune);
    }
}

template <typename V>
static hipError_t launch_pass_t(const Program& p, const Step& s, void* psi, hipStream_t st) {
    PassArgs a;
    a.psi = psi;
    a.ops = p.d_ops + s.op_begin;
    a.mats = p.d_mats;
    a.nops = s.op_count;
    a.k = s.k;
    a.prune = p.prune;
    a.nsys = p.nsys;
    a.bits = p.d_bits;
    for (int i = 0; i < 16; i++) a.tq[i] = s.tq[i];
    const uint64_t blocks = 1ull << (p.n - s.k);
    int nt = 256;
    while (nt > 64 && nt > (1 << s.k) / 2) nt >>= 1;
    dim3 grid((unsigned)blocks), block(nt);
    const int kt = s.k;
    if (kt <= 8) hipLaunchKernelGGL((k_pass<V, 8>), grid, block, 0, st, a);
    else if (kt <= 10) hipLaunchKernelGGL((k_pass<V, 10>), grid, block, 0, st, a);
    else if (kt <= 11) hipLaunchKernelGGL((k_pass<V, 11>), grid, block, 0, st, a);
    else if (kt <= 12) hipLaunchKernelGGL((k_pass<V, 12>), grid, block, 0, st, a);
    else if constexpr (sizeof(V) == 16) hipLaunchKernelGGL((k_pass<V, 13>), grid, block, 0, st, a);
    else {
        if (kt <= 13) hipLaunchKernelGGL((k_pass<V, 13>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_pass<V, 14>), grid, block, 0, st, a);
    }
    return hipGetLastError();
}

hipError_t launch_pass(const Program& p, const Step& s, void* psi, hipStream_t st) {
    if (p.precision == KSV_F64) return launch_pass_t<double2>(p, s, psi, st);
    return launch_pass_t<float2>(p, s, psi, st);
}

template <typename V>
static hipError_t launch_big_t(const Program& p, const Step& s, void* psi, int prune, hipStream_t st) {
    const BigOp& b = p.big_ops[s.big_op];
    BigArgs a;
    a.psi = psi;
    a.m = reinterpret_cast<const V*>(p.d_mats) + b.mat;
    a.k = (int)b.tg.size();
    a.prune = prune;
    for (int i = 0; i < 12; i++) a.tg[i] = a.srt[i] = 0;
    for (int i = 0; i < a.k; i++) a.tg[i] = a.srt[i] = (int8_t)b.tg[i];
    for (int i = 1; i < a.k; i++)
        for (int j = i; j > 0 && a.srt[j - 1] > a.srt[j]; j--) std::swap(a.srt[j], a.srt[j - 1]);
    const uint64_t blocks = 1ull << (p.n - a.k);
    size_t shm = sizeof(V) << a.k;
    hipLaunchKernelGGL((k_big<V>), dim3((unsigned)blocks), dim3(256), shm, st, a);
    return hipGetLastError();
}

hipError_t launch_big(const Program& p, const Step& s, void* psi, int prune, hipStream_t st) {
    if (p.precision == KSV_F64) return launch_big_t<double2>(p, s, psi, prune, st);
    return launch_big_t<float2>(p, s, psi, prune, st);
}

}  // namespace ksv
