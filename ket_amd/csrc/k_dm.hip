// Density matrices the way ket computes them (src/density.ts).
//
// rho is dense in HBM: entry key = (row << n) | col, the key of ket's sparse
// map, as complex128 (an absent entry is +0).  Beside every entry sits its
// place in ket's Map order as a 64-bit order key (absent: ~0).  ket's results
// depend on that order in one place: depolarize2 accumulates its sixteen
// contributions per entry through dmAcc, which deletes a running sum the
// moment it is negligible, so an early tiny contribution is lost or kept
// depending on which source entry the Map reaches first.  Tracking the order
// reproduces ket bit for bit, and the order itself is what the read-outs
// (purity, Bloch angles) sum in.
//
// Every channel maps a group of entries to the same group, so each kernel
// owns its groups and works in place:
//   single gate / depolarize1   4 entries (qubit q in row and column), a thread
//   two-qubit gate, unitary k<=2 16 / 4 entries, a thread
//   unitary 3 <= k <= 6         4^k entries, a workgroup through LDS
//   depolarize2                 the 4 entries an X on (a, b) maps between, a thread
//   csrswap                     16 entries of one control block, a thread
//   cnot / swap / toffoli / cswap  pairs exchanged by the permutation
// The arithmetic is ket's step by step (rows by U first, then columns by U^+;
// every product and sum rounded on its own: contraction is off in this file).
//
// Order keys: a channel that walks groups emits a group's entries when ket's
// loop first meets the group, i.e. at the smallest order key among the group's
// present entries, in the loop's fixed sequence j; so the new key is
// min_key * G + j (G the group size).  depolarize2 emits an entry when its
// running sum is (re)created: key = source_key * 16 + contribution index.
// Keys grow by a few bits per channel; before they could pass 63 bits the
// present entries are sorted by key and renumbered 0, 1, 2, ... (ranks).
#pragma clang fp contract(off)

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "kcommon.h"

namespace ksv {
int adj_fail(int code, const std::string& msg);
}

using namespace ksv;

namespace {

constexpr uint64_t ABSENT = ~0ull;
constexpr int DM_MAXN = 15;         // 4^15 entries: 16 GiB of rho + 8 GiB of order keys
constexpr int DM_UMAX = 6;          // widest unitary: 4^6 entries per group through LDS

__device__ __forceinline__ bool negligible(double2 v) { return knorm2(v.x, v.y) < 1e-14; }

__device__ __forceinline__ double2 cz() { return make_double2(0.0, 0.0); }

__device__ __forceinline__ double2 conjd(double2 a) { return make_double2(a.x, -a.y); }

// v + m * x as add(v, mul(m, x))
__device__ __forceinline__ double2 macc(double2 v, double2 m, double2 x) {
    kacc(v.x, v.y, m.x, m.y, x.x, x.y);
    return v;
}

// dmSet: a negligible value leaves the entry absent
__device__ __forceinline__ void put(double2* v, uint64_t* key, uint64_t i, double2 x, uint64_t k) {
    if (negligible(x)) {
        v[i] = cz();
        key[i] = ABSENT;
    } else {
        v[i] = x;
        key[i] = k;
    }
}

__device__ __forceinline__ uint64_t ins2(uint64_t t, int lo, int hi) { return ins0_64(ins0_64(t, lo), hi); }

struct M2 {
    double2 a, b, c, d;
};

// rho -> U rho U^+ on qubit q (applySingle, src/density.ts:45-77)
__global__ __launch_bounds__(256) void k_dm_single(double2* v, uint64_t* key, int n, int q, M2 g, uint64_t groups) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= groups) return;
    const uint64_t cm = 1ull << q, rm = 1ull << (q + n);
    const uint64_t i00 = ins2(t, q, q + n), i01 = i00 | cm, i10 = i00 | rm, i11 = i00 | rm | cm;
    const double2 p00 = v[i00], p01 = v[i01], p10 = v[i10], p11 = v[i11];
    const uint64_t k0 = min(min(key[i00], key[i01]), min(key[i10], key[i11]));
    const double2 t00 = cmadd2(g.a, p00, g.b, p10), t01 = cmadd2(g.a, p01, g.b, p11);
    const double2 t10 = cmadd2(g.c, p00, g.d, p10), t11 = cmadd2(g.c, p01, g.d, p11);
    const double2 ca = conjd(g.a), cb = conjd(g.b), cc = conjd(g.c), cd = conjd(g.d);
    const uint64_t kb = k0 == ABSENT ? 0 : k0 * 4;
    put(v, key, i00, cmadd2(ca, t00, cb, t01), kb + 0);
    put(v, key, i01, cmadd2(cc, t00, cd, t01), kb + 1);
    put(v, key, i10, cmadd2(ca, t10, cb, t11), kb + 2);
    put(v, key, i11, cmadd2(cc, t10, cd, t11), kb + 3);
}

// depolarize1 (src/density.ts:193-224): same parity (1 - 2p/3, 2p/3), cross parity 1 - 4p/3
__global__ __launch_bounds__(256) void k_dm_dep1(double2* v, uint64_t* key, int n, int q, double same, double flip,
                                                 double cross, uint64_t groups) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= groups) return;
    const uint64_t cm = 1ull << q, rm = 1ull << (q + n);
    const uint64_t i00 = ins2(t, q, q + n), i01 = i00 | cm, i10 = i00 | rm, i11 = i00 | rm | cm;
    const double2 p00 = v[i00], p01 = v[i01], p10 = v[i10], p11 = v[i11];
    const uint64_t k0 = min(min(key[i00], key[i01]), min(key[i10], key[i11]));
    const uint64_t kb = k0 == ABSENT ? 0 : k0 * 4;
    double2 o;
    o.x = rn_add(rn_mul(same, p00.x), rn_mul(flip, p11.x));
    o.y = rn_add(rn_mul(same, p00.y), rn_mul(flip, p11.y));
    put(v, key, i00, o, kb + 0);
    o.x = rn_add(rn_mul(same, p11.x), rn_mul(flip, p00.x));
    o.y = rn_add(rn_mul(same, p11.y), rn_mul(flip, p00.y));
    put(v, key, i11, o, kb + 1);
    put(v, key, i01, make_double2(rn_mul(cross, p01.x), rn_mul(cross, p01.y)), kb + 2);
    put(v, key, i10, make_double2(rn_mul(cross, p10.x), rn_mul(cross, p10.y)), kb + 3);
}

// U rho U^+ for a 2^K x 2^K U on qubits tg (tg[0] the most significant local
// bit), K = 1, 2 per thread: t = U p then t U^+, every sum started from zero
// (applyTwo, src/density.ts:80-118; applyUnitaryN :132-178)
struct BlockArgs {
    double2* v;
    uint64_t* key;
    const double2* u;       // 2^K x 2^K, row-major
    uint64_t groups;
    int n;
    int k;
    int8_t tg[DM_UMAX];     // qubits, MSB first
    int8_t srt[2 * DM_UMAX];// the group's 2k key bits, ascending
};

__device__ __forceinline__ uint64_t local_bits(const BlockArgs& a, int i, int shift) {
    uint64_t b = 0;
    for (int j = 0; j < a.k; j++)
        if ((i >> (a.k - 1 - j)) & 1) b |= 1ull << (a.tg[j] + shift);
    return b;
}

template <int K>
__global__ __launch_bounds__(256) void k_dm_block_thread(BlockArgs a) {
    constexpr int D = 1 << K;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.groups) return;
    uint64_t base = t;
#pragma unroll
    for (int i = 0; i < 2 * K; i++) base = ins0_64(base, a.srt[i]);
    uint64_t rows[D], cols[D];
#pragma unroll
    for (int i = 0; i < D; i++) {
        rows[i] = local_bits(a, i, a.n);
        cols[i] = local_bits(a, i, 0);
    }
    double2 p[D][D];
    uint64_t k0 = ABSENT;
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) {
            const uint64_t g = base | rows[i] | cols[j];
            p[i][j] = a.v[g];
            k0 = min(k0, a.key[g]);
        }
    double2 u[D][D];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) u[i][j] = a.u[i * D + j];
    double2 tt[D][D];
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) {
            double2 s = cz();
#pragma unroll
            for (int l = 0; l < D; l++) s = macc(s, u[i][l], p[l][j]);
            tt[i][j] = s;
        }
    const uint64_t kb = k0 == ABSENT ? 0 : k0 * (D * D);
#pragma unroll
    for (int i = 0; i < D; i++)
#pragma unroll
        for (int j = 0; j < D; j++) {
            double2 s = cz();
#pragma unroll
            for (int l = 0; l < D; l++) s = macc(s, tt[i][l], conjd(u[j][l]));
            put(a.v, a.key, base | rows[i] | cols[j], s, kb + (uint64_t)(i * D + j));
        }
}

// the same for 3 <= K <= 6: one workgroup per group, p and t in LDS
template <int K>
__global__ __launch_bounds__(256) void k_dm_block_lds(BlockArgs a) {
    constexpr int D = 1 << K, E = D * D;
    __shared__ double2 p[E];
    __shared__ double2 tt[E];
    __shared__ uint64_t kmin[256 / 64];
    uint64_t base = blockIdx.x;
#pragma unroll
    for (int i = 0; i < 2 * K; i++) base = ins0_64(base, a.srt[i]);
    auto gidx = [&](int e) { return base | local_bits(a, e / D, a.n) | local_bits(a, e % D, 0); };
    uint64_t k0 = ABSENT;
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
        const uint64_t g = gidx(e);
        p[e] = a.v[g];
        k0 = min(k0, a.key[g]);
    }
    for (int o = 32; o > 0; o >>= 1) k0 = min(k0, (uint64_t)__shfl_xor((unsigned long long)k0, o));
    if ((threadIdx.x & 63) == 0) kmin[threadIdx.x >> 6] = k0;
    __syncthreads();
    k0 = ABSENT;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) k0 = min(k0, kmin[w]);
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
        const int i = e / D, j = e % D;
        double2 s = cz();
        for (int l = 0; l < D; l++) s = macc(s, a.u[i * D + l], p[l * D + j]);
        tt[e] = s;
    }
    __syncthreads();
    const uint64_t kb = k0 == ABSENT ? 0 : k0 * E;
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
        const int i = e / D, j = e % D;
        double2 s = cz();
        for (int l = 0; l < D; l++) s = macc(s, tt[i * D + l], conjd(a.u[j * D + l]));
        put(a.v, a.key, gidx(e), s, kb + (uint64_t)e);
    }
}

// depolarize2 (src/density.ts:239-272).  The group is the 4 entries the X on a,
// on b and on both (applied to row and column alike) map between; every source
// entry sends its 16 weighted copies inside the group.  ket walks the sources
// in Map order and adds each copy into its target with dmAcc; here the group's
// present sources are put in key order and the same 64 additions replayed.
__global__ __launch_bounds__(256) void k_dm_dep2(double2* v, uint64_t* key, int n, int a, int b, double keep,
                                                 double w, uint64_t groups) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= groups) return;
    const int lo = min(a, b) + n, hi = max(a, b) + n;
    const uint64_t base = ins2(t, lo, hi);    // row bits a and b clear
    const uint64_t xa = (1ull << (a + n)) | (1ull << a), xb = (1ull << (b + n)) | (1ull << b);
    uint64_t g[4];
    double2 x[4];
    uint64_t k[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        g[i] = base ^ ((i & 1) ? xb : 0) ^ ((i & 2) ? xa : 0);
        x[i] = v[g[i]];
        k[i] = key[g[i]];
    }
    // the row and column bits of a (of b) differ alike in every member
    const bool da = ((base >> (a + n)) ^ (base >> a)) & 1, db = ((base >> (b + n)) ^ (base >> b)) & 1;
    int ord[4] = {0, 1, 2, 3};
#pragma unroll
    for (int i = 1; i < 4; i++)
        for (int j = i; j > 0 && k[ord[j - 1]] > k[ord[j]]; j--) {
            const int s = ord[j];
            ord[j] = ord[j - 1];
            ord[j - 1] = s;
        }
    double2 acc[4];
    bool have[4] = {false, false, false, false};
    uint64_t made[4] = {ABSENT, ABSENT, ABSENT, ABSENT};
#pragma unroll
    for (int i = 0; i < 4; i++) acc[i] = cz();
    for (int si = 0; si < 4; si++) {
        const int s = ord[si];
        if (k[s] == ABSENT) break;
        const double2 xs = x[s];
        for (int j = 0; j < 16; j++) {
            // j = 0: (1 - p) rho; j = 1..15: P_a (x) P_b with b fastest over I, X, Y, Z
            const int pa = j >> 2, pb = j & 3;
            const bool fa = pa == 1 || pa == 2, fb = pb == 1 || pb == 2;
            const bool za = pa >= 2, zb = pb >= 2;
            const int tgt = s ^ (fb ? 1 : 0) ^ (fa ? 2 : 0);
            const double sc = j == 0 ? keep : (((za && da) != (zb && db)) ? -w : w);
            const double2 c = make_double2(rn_mul(sc, xs.x), rn_mul(sc, xs.y));
            double2 nx = c;
            if (have[tgt]) {
                nx.x = rn_add(acc[tgt].x, c.x);
                nx.y = rn_add(acc[tgt].y, c.y);
            }
            if (negligible(nx)) {
                have[tgt] = false;
            } else {
                if (!have[tgt]) made[tgt] = k[s] * 16 + (uint64_t)j;
                have[tgt] = true;
                acc[tgt] = nx;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        v[g[i]] = have[i] ? acc[i] : cz();
        key[g[i]] = have[i] ? made[i] : ABSENT;
    }
}

// controlled sqrt(iSWAP) (src/density.ts:521-570): rho splits into four blocks by
// the control bit of row and column; blocks (1,1) and (1,0) get V rho V^+ on
// (a, b), block (0,1) the same on its transpose, block (0,0) stays; the result
// lists block (0,0), then the new (0,1), (1,0), (1,1) entries.  A thread per
// (group, block) of 16 entries.  `shift` = bits of the incoming keys + 4.
__global__ __launch_bounds__(256) void k_dm_csrswap(double2* v, uint64_t* key, int n, int c, int a, int b,
                                                    int shift, uint64_t groups) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= groups * 4) return;
    const int blk = (int)(t & 3);               // (row control << 1) | column control
    uint64_t base = t >> 2;
    int srt[6] = {a, b, c, a + n, b + n, c + n};
    for (int i = 1; i < 6; i++)
        for (int j = i; j > 0 && srt[j - 1] > srt[j]; j--) {
            const int s = srt[j];
            srt[j] = srt[j - 1];
            srt[j - 1] = s;
        }
    for (int i = 0; i < 6; i++) base = ins0_64(base, srt[i]);
    if (blk & 2) base |= 1ull << (c + n);
    if (blk & 1) base |= 1ull << c;
    const uint64_t tag = (uint64_t)blk << shift;
    uint64_t rows[4], cols[4];
    for (int i = 0; i < 4; i++) {
        rows[i] = ((i & 2) ? 1ull << (a + n) : 0) | ((i & 1) ? 1ull << (b + n) : 0);
        cols[i] = ((i & 2) ? 1ull << a : 0) | ((i & 1) ? 1ull << b : 0);
    }
    if (blk == 0) {
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
                const uint64_t g = base | rows[i] | cols[j];
                if (key[g] != ABSENT) key[g] |= tag;
            }
        return;
    }
    const bool tr = blk == 1;                   // block (0,1) works on its transpose
    auto gid = [&](int i, int j) { return tr ? (base | rows[j] | cols[i]) : (base | rows[i] | cols[j]); };
    const double h = 1.0 / sqrt(2.0);
    double2 u[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) u[i][j] = cz();
    u[0][0] = make_double2(1.0, 0.0);
    u[3][3] = make_double2(1.0, 0.0);
    u[1][1] = u[2][2] = make_double2(h, 0.0);
    u[1][2] = u[2][1] = make_double2(0.0, h);
    double2 p[4][4];
    uint64_t k0 = ABSENT;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            p[i][j] = v[gid(i, j)];
            k0 = min(k0, key[gid(i, j)]);
        }
    double2 tt[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double2 s = cz();
            for (int l = 0; l < 4; l++) s = macc(s, u[i][l], p[l][j]);
            tt[i][j] = s;
        }
    const uint64_t kb = k0 == ABSENT ? 0 : k0 * 16;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double2 s = cz();
            for (int l = 0; l < 4; l++) s = macc(s, tt[i][l], conjd(u[j][l]));
            put(v, key, gid(i, j), s, tag | (kb + (uint64_t)(i * 4 + j)));
        }
}

// cnot / swap / toffoli / cswap on rows and columns alike (applyPerm,
// src/density.ts:121-129): the Map keeps its order, so values and order keys
// move together.  Every one of these maps is its own inverse: swap pairs.
__global__ __launch_bounds__(256) void k_dm_perm(double2* v, uint64_t* key, int n, int kind, int q0, int q1, int q2,
                                                 uint64_t total) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    auto f = [&](uint64_t x) -> uint64_t {
        const bool b0 = (x >> q0) & 1, b1 = (x >> q1) & 1;
        switch (kind) {
        case KSV_CNOT: return b0 ? x ^ (1ull << q1) : x;
        case KSV_SWAP: return b0 == b1 ? x : x ^ (1ull << q0) ^ (1ull << q1);
        case KSV_TOFFOLI: return (b0 && b1) ? x ^ (1ull << q2) : x;
        default: {      // cswap: control q0, swapped q1, q2
            const bool b2 = (x >> q2) & 1;
            return (!b0 || b1 == b2) ? x : x ^ (1ull << q1) ^ (1ull << q2);
        }
        }
    };
    const uint64_t mask = (1ull << n) - 1;
    const uint64_t j = (f(i >> n) << n) | f(i & mask);
    if (j <= i) return;
    const double2 t = v[i];
    v[i] = v[j];
    v[j] = t;
    const uint64_t kt = key[i];
    key[i] = key[j];
    key[j] = kt;
}

__global__ void k_dm_iota(uint32_t* idx, uint64_t total) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) idx[i] = (uint32_t)i;
}

// after the sort: rank i to the i-th present entry; the count of present entries
__global__ void k_dm_rank(const uint64_t* sorted, const uint32_t* idx, uint64_t* key, uint64_t total,
                          unsigned long long* count) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total || sorted[i] == ABSENT) return;
    key[idx[i]] = i;
    if (i + 1 == total || sorted[i + 1] == ABSENT) *count = i + 1;
}

__global__ void k_dm_init(double2* v, uint64_t* key, uint64_t total) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    v[i] = i == 0 ? make_double2(1.0, 0.0) : cz();
    key[i] = i == 0 ? 0 : ABSENT;
}

dim3 grid_for(uint64_t threads) { return dim3((unsigned)((threads + 255) / 256)); }

}  // namespace

struct ksv_dm {
    int n = 0;
    int bits = 0;                   // every present order key is < 2^bits
    hipStream_t st = nullptr;
    double2* v = nullptr;
    uint64_t* key = nullptr;
    // sort scratch (allocated on the first renumbering)
    uint64_t* sorted = nullptr;
    uint32_t* idx = nullptr;
    uint32_t* idx2 = nullptr;
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    unsigned long long* d_count = nullptr;
    double2* d_u = nullptr;         // matrix of the current block op
    size_t u_cap = 0;
    uint64_t total() const { return 1ull << (2 * n); }
    ~ksv_dm() {
        for (void* p : {(void*)v, (void*)key, (void*)sorted, (void*)idx, (void*)idx2, tmp, (void*)d_count,
                        (void*)d_u})
            if (p) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
    }
};

#define DMCHK(expr)                                                                               \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return adj_fail(KSV_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// sort the present entries by order key and renumber them 0, 1, ...
static int dm_rank(ksv_dm* d, unsigned long long* present) {
    const uint64_t N = d->total();
    if (!d->sorted) {
        DMCHK(hipMalloc(&d->sorted, 8 * N));
        DMCHK(hipMalloc(&d->idx, 4 * N));
        DMCHK(hipMalloc(&d->idx2, 4 * N));
        DMCHK(hipMalloc(&d->d_count, 8));
        size_t bytes = 0;
        DMCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, d->key, d->sorted, d->idx, d->idx2, (int)N, 0, 64,
                                                 d->st));
        DMCHK(hipMalloc(&d->tmp, bytes));
        d->tmp_bytes = bytes;
    }
    hipLaunchKernelGGL(k_dm_iota, grid_for(N), dim3(256), 0, d->st, d->idx, N);
    DMCHK(hipGetLastError());
    // present keys are < 2^bits, an absent key has bit `bits` set: bits + 1 sort bits suffice
    const int end_bit = std::min(64, d->bits + 1);
    size_t bytes = d->tmp_bytes;
    DMCHK(hipcub::DeviceRadixSort::SortPairs(d->tmp, bytes, d->key, d->sorted, d->idx, d->idx2, (int)N, 0, end_bit,
                                             d->st));
    DMCHK(hipMemsetAsync(d->d_count, 0, 8, d->st));
    hipLaunchKernelGGL(k_dm_rank, grid_for(N), dim3(256), 0, d->st, d->sorted, d->idx2, d->key, N, d->d_count);
    DMCHK(hipGetLastError());
    unsigned long long cnt = 0;
    DMCHK(hipMemcpyAsync(&cnt, d->d_count, 8, hipMemcpyDeviceToHost, d->st));
    DMCHK(hipStreamSynchronize(d->st));
    d->bits = 1;
    while ((1ull << d->bits) < std::max<unsigned long long>(cnt, 1)) d->bits++;
    if (present) *present = cnt;
    return KSV_OK;
}

// make room for `grow` more key bits
static int dm_room(ksv_dm* d, int grow) {
    if (d->bits + grow <= 62) return KSV_OK;
    return dm_rank(d, nullptr);
}

static int dm_upload_u(ksv_dm* d, const std::vector<double2>& u) {
    if (u.size() > d->u_cap) {
        if (d->d_u) DMCHK(hipFree(d->d_u));
        d->d_u = nullptr;
        DMCHK(hipMalloc(&d->d_u, sizeof(double2) * u.size()));
        d->u_cap = u.size();
    }
    DMCHK(hipMemcpyAsync(d->d_u, u.data(), sizeof(double2) * u.size(), hipMemcpyHostToDevice, d->st));
    // the launch that reads it follows on the same stream; the host vector must
    // outlive the copy, so wait for it here
    DMCHK(hipStreamSynchronize(d->st));
    return KSV_OK;
}

static int dm_block(ksv_dm* d, const int* qs, int k, const std::vector<double2>& u) {
    int rc = dm_room(d, 2 * k);
    if (rc) return rc;
    rc = dm_upload_u(d, u);
    if (rc) return rc;
    BlockArgs a;
    std::memset(&a, 0, sizeof a);
    a.v = d->v;
    a.key = d->key;
    a.u = d->d_u;
    a.n = d->n;
    a.k = k;
    a.groups = 1ull << (2 * (d->n - k));
    std::vector<int> srt;
    for (int i = 0; i < k; i++) {
        a.tg[i] = (int8_t)qs[i];
        srt.push_back(qs[i]);
        srt.push_back(qs[i] + d->n);
    }
    std::sort(srt.begin(), srt.end());
    for (int i = 0; i < 2 * k; i++) a.srt[i] = (int8_t)srt[i];
    switch (k) {
    case 1: hipLaunchKernelGGL(k_dm_block_thread<1>, grid_for(a.groups), dim3(256), 0, d->st, a); break;
    case 2: hipLaunchKernelGGL(k_dm_block_thread<2>, grid_for(a.groups), dim3(256), 0, d->st, a); break;
    case 3: hipLaunchKernelGGL(k_dm_block_lds<3>, dim3((unsigned)a.groups), dim3(64), 0, d->st, a); break;
    case 4: hipLaunchKernelGGL(k_dm_block_lds<4>, dim3((unsigned)a.groups), dim3(256), 0, d->st, a); break;
    case 5: hipLaunchKernelGGL(k_dm_block_lds<5>, dim3((unsigned)a.groups), dim3(256), 0, d->st, a); break;
    default: hipLaunchKernelGGL(k_dm_block_lds<6>, dim3((unsigned)a.groups), dim3(256), 0, d->st, a); break;
    }
    DMCHK(hipGetLastError());
    d->bits += 2 * k;
    return KSV_OK;
}

static int dm_dep2(ksv_dm* d, int a, int b, double p) {
    if (!(p > 0)) return KSV_OK;
    int rc = dm_room(d, 4);
    if (rc) return rc;
    const uint64_t groups = d->total() / 4;
    hipLaunchKernelGGL(k_dm_dep2, grid_for(groups), dim3(256), 0, d->st, d->v, d->key, d->n, a, b, 1 - p, p / 15,
                       groups);
    DMCHK(hipGetLastError());
    d->bits += 4;
    return KSV_OK;
}

static int dm_perm(ksv_dm* d, int kind, const int* q) {
    const uint64_t N = d->total();
    hipLaunchKernelGGL(k_dm_perm, grid_for(N), dim3(256), 0, d->st, d->v, d->key, d->n, kind, q[0], q[1],
                       kind == KSV_TOFFOLI || kind == KSV_CSWAP ? q[2] : 0, N);
    DMCHK(hipGetLastError());
    return KSV_OK;
}

static double2 cell(const double* mats, int64_t off, int i) { return make_double2(mats[off + 2 * i], mats[off + 2 * i + 1]); }

extern "C" {

int ksv_dm_create(int n, ksv_dm** out) {
    if (!out) return adj_fail(KSV_ERR_TYPE, "null output pointer");
    *out = nullptr;
    if (n < 1 || n > DM_MAXN)
        return adj_fail(KSV_ERR_RANGE, "dm(): " + std::to_string(n) + " qubits is outside 1.." +
                                           std::to_string(DM_MAXN));
    auto d = std::make_unique<ksv_dm>();
    d->n = n;
    DMCHK(hipStreamCreateWithFlags(&d->st, hipStreamNonBlocking));
    const uint64_t N = d->total();
    if (hipMalloc(&d->v, 16 * N) != hipSuccess || hipMalloc(&d->key, 8 * N) != hipSuccess)
        return adj_fail(KSV_ERR_DEVICE, "out of device memory for the density matrix");
    hipLaunchKernelGGL(k_dm_init, grid_for(N), dim3(256), 0, d->st, d->v, d->key, N);
    DMCHK(hipGetLastError());
    DMCHK(hipStreamSynchronize(d->st));
    d->bits = 1;
    *out = d.release();
    return KSV_OK;
}

int ksv_dm_apply(ksv_dm* d, const ksv_op* ops, size_t nops, const double* mats, size_t nmats, double p1, double p2) {
    if (!d) return adj_fail(KSV_ERR_HANDLE, "null density matrix");
    if (nops && !ops) return adj_fail(KSV_ERR_TYPE, "null op buffer");
    {
        // the statevector path's checks: kinds, operand counts and ranges, distinct
        // qubits, matrix offsets
        std::vector<LOp> scratch;
        std::string err;
        int rc = lower_tape(d->n, ops, nops, mats, nmats, scratch, err);
        if (rc) return adj_fail(rc, err);
        for (size_t i = 0; i < nops; i++)
            if (ops[i].kind == KSV_UNITARY && ops[i].nq > DM_UMAX)
                return adj_fail(KSV_ERR_RANGE, "dm(): a unitary on " + std::to_string(ops[i].nq) +
                                                   " qubits exceeds the " + std::to_string(DM_UMAX) + "-qubit limit");
    }
    const int n = d->n;
    const uint64_t g4 = d->total() / 4;
    const double same = 1 - 2 * p1 / 3, flip = 2 * p1 / 3, cross = 1 - 4 * p1 / 3;
    int rc = KSV_OK;
    for (size_t i = 0; i < nops && !rc; i++) {
        const ksv_op& op = ops[i];
        const int* q = op.q;
        switch (op.kind) {
        case KSV_SINGLE: {
            rc = dm_room(d, 2);
            if (rc) break;
            M2 g{cell(mats, op.mat, 0), cell(mats, op.mat, 1), cell(mats, op.mat, 2), cell(mats, op.mat, 3)};
            hipLaunchKernelGGL(k_dm_single, grid_for(g4), dim3(256), 0, d->st, d->v, d->key, n, q[0], g, g4);
            if (hipGetLastError() != hipSuccess) { rc = adj_fail(KSV_ERR_DEVICE, "dm single-qubit launch failed"); break; }
            d->bits += 2;
            if (p1 > 0) {
                rc = dm_room(d, 2);
                if (rc) break;
                hipLaunchKernelGGL(k_dm_dep1, grid_for(g4), dim3(256), 0, d->st, d->v, d->key, n, q[0], same, flip,
                                   cross, g4);
                if (hipGetLastError() != hipSuccess) { rc = adj_fail(KSV_ERR_DEVICE, "dm depolarize1 launch failed"); break; }
                d->bits += 2;
            }
            break;
        }
        case KSV_CNOT:
        case KSV_SWAP:
            rc = dm_perm(d, op.kind, q);
            if (!rc) rc = dm_dep2(d, q[0], q[1], p2);
            break;
        case KSV_TWO: {
            std::vector<double2> u(16);
            for (int e = 0; e < 16; e++) u[e] = cell(mats, op.mat, e);
            rc = dm_block(d, q, 2, u);
            if (!rc) rc = dm_dep2(d, q[0], q[1], p2);
            break;
        }
        case KSV_CONTROLLED: {
            // controlledGate (src/mps.ts:53-60): identity on control 0, the gate on control 1
            std::vector<double2> u(16, make_double2(0.0, 0.0));
            u[0] = u[5] = make_double2(1.0, 0.0);
            u[10] = cell(mats, op.mat, 0);
            u[11] = cell(mats, op.mat, 1);
            u[14] = cell(mats, op.mat, 2);
            u[15] = cell(mats, op.mat, 3);
            rc = dm_block(d, q, 2, u);
            if (!rc) rc = dm_dep2(d, q[0], q[1], p2);
            break;
        }
        case KSV_TOFFOLI:
        case KSV_CSWAP:
            rc = dm_perm(d, op.kind, q);
            break;
        case KSV_CSRSWAP: {
            rc = dm_room(d, 6);
            if (rc) break;
            const uint64_t groups = d->total() >> 6;
            hipLaunchKernelGGL(k_dm_csrswap, grid_for(groups * 4), dim3(256), 0, d->st, d->v, d->key, n, q[0], q[1],
                               q[2], d->bits + 4, groups);
            if (hipGetLastError() != hipSuccess) { rc = adj_fail(KSV_ERR_DEVICE, "dm csrswap launch failed"); break; }
            d->bits += 6;
            break;
        }
        case KSV_UNITARY: {
            const int k = op.nq, dim = 1 << k;
            std::vector<double2> u((size_t)dim * dim);
            for (int e = 0; e < dim * dim; e++) u[e] = cell(mats, op.mat, e);
            rc = dm_block(d, q, k, u);
            break;
        }
        default:
            break;      // barrier
        }
    }
    if (rc) return rc;
    DMCHK(hipStreamSynchronize(d->st));
    return KSV_OK;
}

int ksv_dm_read(ksv_dm* d, double* rho, uint64_t* order, uint64_t* count) {
    if (!d) return adj_fail(KSV_ERR_HANDLE, "null density matrix");
    const uint64_t N = d->total();
    if (rho) {
        DMCHK(hipMemcpyAsync(rho, d->v, 16 * N, hipMemcpyDeviceToHost, d->st));
        DMCHK(hipStreamSynchronize(d->st));
    }
    if (order || count) {
        unsigned long long present = 0;
        int rc = dm_rank(d, &present);
        if (rc) return rc;
        if (count) *count = present;
        if (order && present) {
            std::vector<uint32_t> idx(present);
            DMCHK(hipMemcpy(idx.data(), d->idx2, 4 * present, hipMemcpyDeviceToHost));
            for (uint64_t i = 0; i < present; i++) order[i] = idx[i];
        }
    }
    return KSV_OK;
}

int ksv_dm_destroy(ksv_dm* d) {
    delete d;
    return KSV_OK;
}

}  // extern "C"
