// Fused-block pass kernel: dense 2^BQ x 2^BQ blocks applied on MFMA.
//
// A workgroup of 256 threads stages one K-qubit tile (the amplitudes whose
// non-tile bits equal its base index) in LDS.  For every block of the pass
// the tile is a complex matrix X with 2^BQ rows (the block's local index,
// bit i = block qubit i) and 2^(K-BQ) columns (the other tile bits), and the
// block computes Y = U X with Gauss's three real products
//
//     T1 = Ur Xr,  T2 = Ui Xi,  T3 = (Ur + Ui)(Xr + Xi)
//     Yr = T1 - T2,  Yi = T3 - T1 - T2
//
// (three quarters of the MFMA work of the real embedding
// [[Ur, -Ui], [Ui, Ur]]) on v_mfma_{f64,f32}_16x16x4: each wave takes
// 16-column slices, reads its B values (lane l: column l%16, rows 4t + l/16)
// from LDS, accumulates 2^BQ/16 output tiles of each product over 2^BQ/4
// k-steps, and writes the result back in place (lane l, element v: row
// 4v + l/16 for f64, 4(l/16) + v for f32).  A fragments come pre-arranged
// per lane from the host.  A block whose ops are all diagonal is instead one
// phase per row: its fragments hold the diagonal and each lane multiplies
// the amplitudes it reads in place.  Columns differ between waves, so blocks need a
// barrier only between them; within a block a wave reads its next slice from LDS
// while the current one runs on MFMA.  The LDS address of tile index x is x with its
// low four bits XORed by a per-pass word of its higher bits, chosen so that
// each 16-lane group hits 16 distinct 16-byte bank groups.  HBM sees one
// coalesced read and one coalesced write of the tile per pass.
//
// FP64 passes of 3-4 blocks run as k_bpass_pipe instead: two 512-thread
// workgroups per CU each walk a strided list of tiles, keep the pass's
// fragments in LDS, and load the next tile into registers while the current
// one's blocks run (barriers order LDS only, so those loads stay in flight).
#include <algorithm>
#include <cstring>

#include "kcommon.h"

namespace ksv {

template <typename T> struct BVec2;
template <> struct BVec2<double> { using type = double2; };
template <> struct BVec2<float> { using type = float2; };

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef double d2_t __attribute__((ext_vector_type(2)));
typedef float f2_t __attribute__((ext_vector_type(2)));
// the amplitude as a plain vector: an array of these held across a loop stays in
// registers, where an array of double2 structs ends up in scratch
template <typename T> struct Ext2;
template <> struct Ext2<double> { using type = d2_t; };
template <> struct Ext2<float> { using type = f2_t; };

// Output layout of 16x16x4 (measured on gfx950): D[i][j] sits in lane
// j + 16 * (i % 4), element i / 4 for f64, and in lane j + 16 * (i / 4),
// element i % 4 for f32.  A[i][k] is lane i + 16k, B[k][j] lane j + 16k.
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    static constexpr bool lane_is_row_low = true;
    using acc = d4_t;
    static __device__ __forceinline__ acc step(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<float> {
    static constexpr bool lane_is_row_low = false;
    using acc = f4_t;
    static __device__ __forceinline__ acc step(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

__device__ __forceinline__ uint32_t swz_part(uint32_t x, uint64_t w, int lo, int hi) {
    // XOR of the swizzle words of the set bits p in [lo, hi) of x (p >= 4)
    uint32_t g = 0;
#pragma unroll
    for (int p = 4; p < 16; p++)
        if (p >= lo && p < hi && ((x >> p) & 1)) g ^= (uint32_t)((w >> (4 * (p - 4))) & 15);
    return g;
}

// Tile index x = tid + i * NT of a workgroup: the thread part and the uniform
// part of its physical offset and of its swizzled LDS address.
template <int K, int NT>
struct TileMap {
    static constexpr int LNT = NT == 512 ? 9 : 8;
    static constexpr int PER = (1 << K) / NT;   // amplitudes per thread at load / store
    uint64_t g_t;
    uint32_t l_t;
    __device__ __forceinline__ TileMap(const BPassArgs& a, uint32_t tid) {
        g_t = 0;
#pragma unroll
        for (int b = 0; b < LNT; b++)
            g_t |= (uint64_t)((tid >> b) & 1) << byte_at(b < 8 ? a.tq_lo : a.tq_hi, b & 7);
        l_t = tid ^ swz_part(tid, a.swz, 4, LNT);
    }
    static __device__ __forceinline__ uint64_t base(const BPassArgs& a, uint64_t tile) {
        uint64_t b = tile;
#pragma unroll
        for (int i = 0; i < 16; i++)
            if (i < K) b = ins0_64(b, byte_at(i < 8 ? a.tq_lo : a.tq_hi, i & 7));
        return b;
    }
    __device__ __forceinline__ uint64_t gaddr(const BPassArgs& a, uint64_t base, int i) const {
        const uint32_t xh = (uint32_t)i * NT;
        uint64_t g_h = 0;
#pragma unroll
        for (int b = LNT; b < K; b++)
            if ((xh >> b) & 1) g_h |= 1ull << byte_at(a.tq_hi, b - 8);
        return base | g_t | g_h;
    }
    __device__ __forceinline__ uint32_t laddr(const BPassArgs& a, int i) const {
        const uint32_t xh = (uint32_t)i * NT;
        return l_t ^ xh ^ swz_part(xh, a.swz, LNT, K);
    }
};

// One block of a pass on the tile in LDS (record w, fragments af).  MONO: the
// build has the monomial-block path (the pipelined walk leaves it out: its
// registers would cost that kernel its second workgroup per CU).
template <typename T, int BQ, int K, int NT, bool MONO = true>
__device__ __forceinline__ void bpass_block(typename BVec2<T>::type* lds, const KSV_CONST uint64_t* w,
                                            const T (&af)[3 * ((1 << BQ) / 16) * ((1 << BQ) / 4)], uint32_t lane,
                                            uint32_t wave) {
    using V = typename BVec2<T>::type;
    using M = Mfma<T>;
    using acc_t = typename M::acc;
    constexpr int NW = NT / 64;
    constexpr int R = 1 << BQ;
    constexpr int MB = R / 16;
    constexpr int KS = R / 4;
    constexpr int NC = 1 << (K - BQ);
    constexpr int NBW = NC / 16 / NW;
    static_assert(NBW >= 1, "tile too small for the block width");
    uint16_t ar[8], ac[16];
    {
        const uint64_t r0 = w[0], r1 = w[1], r2 = w[2], r3 = w[3], r4 = w[4], r5 = w[5];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            ar[i] = (uint16_t)(r0 >> (16 * i));
            ar[4 + i] = (uint16_t)(r1 >> (16 * i));
            ac[i] = (uint16_t)(r2 >> (16 * i));
            ac[4 + i] = (uint16_t)(r3 >> (16 * i));
            ac[8 + i] = (uint16_t)(r4 >> (16 * i));
            ac[12 + i] = (uint16_t)(r5 >> (16 * i));
        }
    }
    const bool diag = (w[7] & BREC_DIAG) != 0;     // a phase per row (fragments hold the diagonal)
    const uint32_t rl = lane >> 4, cl = lane & 15;
    // read rows r = 4t + rl: bits 0,1 from rl, the rest from t
    const uint32_t rd_l = ((rl & 1) ? ar[0] : 0u) ^ ((rl & 2) ? ar[1] : 0u);
    // write rows r = 16 h + (4 v + rl) (f64) or 16 h + (4 rl + v) (f32)
    constexpr bool LOW = M::lane_is_row_low;
    const uint32_t wr_l = LOW ? (((rl & 1) ? ar[0] : 0u) ^ ((rl & 2) ? ar[1] : 0u))
                              : (((rl & 1) ? ar[2] : 0u) ^ ((rl & 2) ? ar[3] : 0u));
    const uint32_t col_l = ((cl & 1) ? ac[0] : 0u) ^ ((cl & 2) ? ac[1] : 0u) ^ ((cl & 4) ? ac[2] : 0u) ^
                           ((cl & 8) ? ac[3] : 0u);
    auto col_of = [&](int m) {
        const uint32_t nb = wave + NW * m;
        uint32_t col = col_l;
#pragma unroll
        for (int b = 0; b < K - BQ - 4; b++)
            if ((nb >> b) & 1) col ^= ac[4 + b];
        return col;
    };
    uint32_t ra[R / 4];
#pragma unroll
    for (int t = 0; t < R / 4; t++) {
        ra[t] = rd_l;
#pragma unroll
        for (int b = 0; b < BQ - 2; b++)
            if ((t >> b) & 1) ra[t] ^= ar[2 + b];
    }
    if (diag) {
#pragma unroll
        for (int m = 0; m < NBW; m++) {
            const uint32_t col = col_of(m);
#pragma unroll
            for (int t = 0; t < R / 4; t++) {
                const V x = lds[ra[t] ^ col];
                const T pr = af[2 * t], pi = af[2 * t + 1];
                V y;
                y.x = x.x * pr - x.y * pi;
                y.y = x.x * pi + x.y * pr;
                lds[ra[t] ^ col] = y;
            }
        }
        return;
    }
    if constexpr (BQ == 4 && MONO) {
    if (w[7] & BREC_MONO) {
        // one entry per column: row r moves to row pi(r) (nibble r of w[1]) times
        // U[pi(r)][r].  The wave reads all rows of its columns before writing any,
        // and columns belong to one wave, so no write can overtake a read.
        const uint64_t perm = w[1];
        uint32_t wa[R / 4];
#pragma unroll
        for (int t = 0; t < R / 4; t++) {
            const uint32_t pr = (uint32_t)(perm >> (4 * (4 * t + rl))) & 15u;
            wa[t] = ((pr & 1) ? ar[0] : 0u) ^ ((pr & 2) ? ar[1] : 0u) ^ ((pr & 4) ? ar[2] : 0u) ^
                    ((pr & 8) ? ar[3] : 0u);
        }
#pragma unroll
        for (int m = 0; m < NBW; m++) {
            const uint32_t col = col_of(m);
            V x[R / 4];
#pragma unroll
            for (int t = 0; t < R / 4; t++) x[t] = lds[ra[t] ^ col];
#pragma unroll
            for (int t = 0; t < R / 4; t++) {
                const T pr = af[2 * t], pi = af[2 * t + 1];
                V y;
                y.x = x[t].x * pr - x[t].y * pi;
                y.y = x[t].x * pi + x[t].y * pr;
                lds[wa[t] ^ col] = y;
            }
        }
        return;
    }
    }
    // the slices own disjoint columns: slice m + 1 is read from LDS while
    // slice m runs on MFMA
    V xn[R / 4];
#pragma unroll
    for (int t = 0; t < R / 4; t++) xn[t] = lds[ra[t] ^ col_of(0)];
#pragma unroll
    for (int m = 0; m < NBW; m++) {
        const uint32_t col = col_of(m);
        V x[R / 4];
#pragma unroll
        for (int t = 0; t < R / 4; t++) x[t] = xn[t];
        if (m + 1 < NBW) {
            const uint32_t cn = col_of(m + 1);
#pragma unroll
            for (int t = 0; t < R / 4; t++) xn[t] = lds[ra[t] ^ cn];
        }
        acc_t t1[MB], t2[MB], t3[MB];
#pragma unroll
        for (int mb = 0; mb < MB; mb++) t1[mb] = t2[mb] = t3[mb] = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int kk = 0; kk < KS; kk++) {
            const T xr = x[kk].x, xi = x[kk].y, xs = xr + xi;
#pragma unroll
            for (int mb = 0; mb < MB; mb++) {
                t1[mb] = M::step(af[mb * KS + kk], xr, t1[mb]);
                t2[mb] = M::step(af[(MB + mb) * KS + kk], xi, t2[mb]);
                t3[mb] = M::step(af[(2 * MB + mb) * KS + kk], xs, t3[mb]);
            }
        }
#pragma unroll
        for (int h = 0; h < MB; h++) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                uint32_t wa = wr_l ^ (LOW ? (((v & 1) ? ar[2] : 0u) ^ ((v & 2) ? ar[3] : 0u))
                                          : (((v & 1) ? ar[0] : 0u) ^ ((v & 2) ? ar[1] : 0u)));
                if (h) wa ^= ar[4];
                V y;
                y.x = t1[h][v] - t2[h][v];
                y.y = t3[h][v] - t1[h][v] - t2[h][v];
                lds[wa ^ col] = y;
            }
        }
    }
}

// One tile per workgroup: load, every block with a barrier after it, store.
template <typename T, int BQ, int K, int NT>
__global__ __launch_bounds__(NT) void k_bpass(BPassArgs a) {
    using V = typename BVec2<T>::type;
    constexpr int NF = 3 * ((1 << BQ) / 16) * ((1 << BQ) / 4);   // A fragments: Ur, Ui, Ur + Ui
    using TM = TileMap<K, NT>;
    __shared__ V lds[1 << K];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    V* psi = reinterpret_cast<V*>(a.psi);
    const T* frags = reinterpret_cast<const T*>(a.frags);
    const KSV_CONST uint64_t* recs = to_const(reinterpret_cast<const uint64_t*>(a.recs));
    const TM tm(a, tid);
    const uint64_t base = TM::base(a, blockIdx.x);
#pragma unroll
    for (int i = 0; i < TM::PER; i++) lds[tm.laddr(a, i)] = psi[tm.gaddr(a, base, i)];
    // fragments of block b + 1 are loaded while block b runs
    T af[NF];
    if (a.nblk > 0) {
        const int64_t foff = (int64_t)recs[6];
#pragma unroll
        for (int f = 0; f < NF; f++) af[f] = frags[foff + f * 64 + lane];
    }
    __syncthreads();
    for (int blk = 0; blk < a.nblk; blk++) {
        const KSV_CONST uint64_t* w = recs + 8 * blk;
        T an[NF];
        const int64_t fn = (int64_t)recs[8 * (blk + 1 < a.nblk ? blk + 1 : blk) + 6];
#pragma unroll
        for (int f = 0; f < NF; f++) an[f] = frags[fn + f * 64 + lane];
        bpass_block<T, BQ, K, NT>(lds, w, af, lane, wave);
#pragma unroll
        for (int f = 0; f < NF; f++) af[f] = an[f];
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < TM::PER; i++) psi[tm.gaddr(a, base, i)] = lds[tm.laddr(a, i)];
}

// Workgroup barrier that orders LDS accesses only: outstanding global loads
// (the next tile) and stores (the previous one) stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Fragments the pipelined kernel keeps in LDS per block: FP64 keeps Ur and Ui
// and adds Ur + Ui in registers (the same double addition the host makes, so
// the fragments are identical); FP32 keeps all three (the host rounds the FP64
// sum to FP32, which a sum of the rounded parts would not always equal).
template <typename T, int BQ> struct PipeFrags {
    static constexpr int NL = (sizeof(T) == 8 ? 2 : 3) * ((1 << BQ) / 16) * ((1 << BQ) / 4);
};

// Pipelined form: each workgroup walks tiles blockIdx.x, blockIdx.x + gridDim.x, ...
// The pass's fragments sit in LDS beside the tile (read once per workgroup), so
// the only global traffic inside the loop is the tiles themselves: tile t + G is
// loaded into registers while the blocks of tile t run on MFMA, and the stores
// of tile t drain while tile t + G is computed.
template <typename T, int BQ, int K, int NT>
__global__ __launch_bounds__(NT) void k_bpass_pipe(BPassArgs a, uint64_t ntiles) {
    using V = typename BVec2<T>::type;
    constexpr int NF = 3 * ((1 << BQ) / 16) * ((1 << BQ) / 4);
    constexpr int NL = PipeFrags<T, BQ>::NL;   // fragments kept in LDS per block
    using TM = TileMap<K, NT>;
    extern __shared__ __align__(16) unsigned char smem[];
    V* lds = reinterpret_cast<V*>(smem);
    T* flds = reinterpret_cast<T*>(smem + sizeof(V) * (1 << K));
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const T* frags = reinterpret_cast<const T*>(a.frags);
    const KSV_CONST uint64_t* recs = to_const(reinterpret_cast<const uint64_t*>(a.recs));
    for (int b = 0; b < a.nblk; b++) {
        const int64_t foff = (int64_t)recs[8 * b + 6];
        for (int r = tid; r < NL * 64; r += NT) flds[b * NL * 64 + r] = frags[foff + r];
    }
    using E = typename Ext2<T>::type;
    E* psie = reinterpret_cast<E*>(a.psi);
    E* ldse = reinterpret_cast<E*>(smem);
    const TM tm(a, tid);
    E nxt[TM::PER];
    {
        const uint64_t base = TM::base(a, blockIdx.x);
#pragma unroll
        for (int i = 0; i < TM::PER; i++) nxt[i] = psie[tm.gaddr(a, base, i)];
#pragma unroll
        for (int i = 0; i < TM::PER; i++) ldse[tm.laddr(a, i)] = nxt[i];
    }
    lds_barrier();
    const uint64_t g = gridDim.x;
    const uint64_t cnt = (ntiles - blockIdx.x + g - 1) / g;   // tiles of this workgroup
    for (uint64_t j = 0; j < cnt; j++) {
        const uint64_t t = blockIdx.x + j * g;
        // the next tile (the last one re-reads itself, so the loads need no branch)
        const uint64_t bn = TM::base(a, j + 1 < cnt ? t + g : t);
#pragma unroll
        for (int i = 0; i < TM::PER; i++) nxt[i] = psie[tm.gaddr(a, bn, i)];
        for (int blk = 0; blk < a.nblk; blk++) {
            const KSV_CONST uint64_t* w = recs + 8 * blk;
            T af[NF];
#pragma unroll
            for (int f = 0; f < NL; f++) af[f] = flds[(blk * NL + f) * 64 + lane];
#pragma unroll
            for (int f = NL; f < NF; f++) af[f] = af[f - NL] + af[f - NL / 2];   // Ur + Ui, as the host forms it
            bpass_block<T, BQ, K, NT, false>(lds, w, af, lane, wave);
            lds_barrier();
        }
        const uint64_t base = TM::base(a, t);
#pragma unroll
        for (int i = 0; i < TM::PER; i++) psie[tm.gaddr(a, base, i)] = ldse[tm.laddr(a, i)];
        lds_barrier();
#pragma unroll
        for (int i = 0; i < TM::PER; i++) ldse[tm.laddr(a, i)] = nxt[i];
        lds_barrier();
    }
}

template <typename T, int BQ, int K, int NT>
static hipError_t launch_pipe(const BPassArgs& a, int n, hipStream_t st, bool* done) {
    constexpr int NL = PipeFrags<T, BQ>::NL;
    using V = typename BVec2<T>::type;
    const size_t bytes = sizeof(V) * ((size_t)1 << K) + sizeof(T) * (size_t)a.nblk * NL * 64;
    int dev = 0, cus = 0, maxlds = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    if (e != hipSuccess) return e;
    *done = false;
    if (bytes > (size_t)maxlds) return hipSuccess;   // too many blocks: the one-tile kernel runs the pass
    auto kern = k_bpass_pipe<T, BQ, K, NT>;
    e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return e;
    int occ = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, NT, bytes);
    if (e != hipSuccess) return e;
    if (occ < 2) return hipSuccess;   // one workgroup per CU: the one-tile kernel is faster
    const uint64_t ntiles = 1ull << (n - K);
    const uint64_t g = std::min<uint64_t>(ntiles, (uint64_t)occ * (uint64_t)cus);
    hipLaunchKernelGGL(kern, dim3((unsigned)g), dim3(NT), bytes, st, a, ntiles);
    *done = true;
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_bpass_t(const Program& p, const Step& s, void* psi, hipStream_t st) {
    BPassArgs a;
    std::memset(&a, 0, sizeof a);
    a.psi = psi;
    a.recs = p.d_brecs + s.op_begin;
    a.frags = p.d_frags;
    a.nblk = s.op_count;
    a.k = s.k;
    for (int i = 0; i < 8; i++) {
        a.tq_lo |= (uint64_t)(uint8_t)s.tq[i] << (8 * i);
        a.tq_hi |= (uint64_t)(uint8_t)s.tq[8 + i] << (8 * i);
    }
    a.swz = s.swz;
    const dim3 grid((unsigned)(1ull << (p.n - s.k)));
    const bool wide = p.bthreads == 512;
    // FP64 passes of 3 or 4 dense blocks (diagonal blocks not counted): their MFMA work is about the tile's HBM
    // time, which the pipelined walk overlaps (synthetic 30-qubit passes: 7.5 -> 7.1 ms
    // at 3 blocks, 9.8 -> 8.8 ms at 4).  With fewer blocks the one-tile kernel streams
    // faster; with more, the fragments leave room for one workgroup per CU only, and
    // FP32 passes gain nothing.
    int ndense = 0, nmono = 0;   // diagonal and monomial blocks are phases, cheap next to the tile's HBM time
    for (int i = 0; i < s.op_count; i++) {
        const int64_t f = p.brecs[s.op_begin + i].flags;
        ndense += (f & (BREC_DIAG | BREC_MONO)) ? 0 : 1;
        nmono += (f & BREC_MONO) ? 1 : 0;
    }
    if (p.bpipe && sizeof(T) == 8 && s.bq == 4 && s.k == 12 && ndense >= 3 && ndense <= 4 && nmono == 0) {
        bool done = false;
        hipError_t e = launch_pipe<T, 4, 12, 512>(a, p.n, st, &done);
        if (e != hipSuccess || done) return e;
    }
    if (s.bq == 4 && s.k == 11) hipLaunchKernelGGL((k_bpass<T, 4, 11, 256>), grid, dim3(256), 0, st, a);
    else if (s.bq == 4 && s.k == 12 && !wide) hipLaunchKernelGGL((k_bpass<T, 4, 12, 256>), grid, dim3(256), 0, st, a);
    else if (s.bq == 4 && s.k == 12) hipLaunchKernelGGL((k_bpass<T, 4, 12, 512>), grid, dim3(512), 0, st, a);
    else if (s.bq == 4 && s.k == 13) hipLaunchKernelGGL((k_bpass<T, 4, 13, 256>), grid, dim3(256), 0, st, a);
    else if (s.bq == 5 && s.k == 13) hipLaunchKernelGGL((k_bpass<T, 5, 13, 256>), grid, dim3(256), 0, st, a);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_bpass(const Program& p, const Step& s, void* psi, hipStream_t st) {
    if (p.precision == KSV_F64) return launch_bpass_t<double>(p, s, psi, st);
    return launch_bpass_t<float>(p, s, psi, st);
}

}  // namespace ksv
