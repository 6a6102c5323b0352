// Register-round pass kernel.
//
// A workgroup of 2^TB threads owns a tile of 2^(TB+4) amplitudes (the
// amplitudes whose non-tile bits equal the tile's base index).  Each thread
// holds 16 of them in registers: in every round four tile qubits are the bits
// of the register index j ("slots") and the other TB tile qubits are the bits
// of the thread index.  One-qubit gates, permutations and diagonals on slots
// are straight-line FP code on registers indexed by compile-time constants;
// the op stream is dispatched by a small switch over (kind, slot, control
// form).  Dense 2..4-qubit gates stage the thread's 16 amplitudes in LDS and
// read their inputs back with run-time indices, so they need no register
// temporaries.  An R_EXCH op writes the tile to LDS under the current
// assignment and reads it back under the next one (XOR-swizzled so that both
// sides are free of bank conflicts).
//
// One tile per workgroup; the workgroups resident on a CU overlap each
// other's HBM traffic and arithmetic.  (Persistent workgroups that prefetch
// their next tile into registers measured slower: the second register set
// pushes the kernel to 256 VGPRs and the compiler waits on the prefetch
// before the first gate.)  HBM sees one read and one write of every
// amplitude per pass, coalesced: thread bits 0..2 are physical qubits 0..2
// at load and store, so each 8 lanes move a full 128-byte run.
//
// Op records are 128 bytes, matrix inline.  An ordinary pass reads them
// through the constant address space (uniform scalar loads; its op stream fits
// the scalar cache).  A copy program (batched trajectories) stages them through
// LDS in chunks of 2^TB / 8 records (16 bytes per thread, coalesced),
// double-buffered so that the next chunk's global loads are in flight while
// the current one runs; each op then reads its record from LDS (uniform
// address) into scalar registers.  Read straight through the scalar cache
// instead, every op waited on its record from L2 (a copy program's op stream
// far exceeds that cache).  The
// copy-bitmap words of conditional ops come with the chunk, one per lane of
// every wave, and a ballot turns them into a 64-bit per-chunk mask of the ops
// that run on the tile (unconditional ops have all-ones words); the op loop
// walks its set bits, so an op skipped by a copy costs nothing.  (Looping over
// every op instead, each skipped op paid the loop's copy of the 32-register
// state from one register set to the other.)
#include <algorithm>
#include <cstring>
#include <type_traits>

#include "kcommon.h"

namespace ksv {

template <typename T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };

template <typename T>
struct CM {        // one complex matrix entry in scalar registers
    T x, y;
};

template <typename T>
__device__ __forceinline__ CM<T> ldm(const KSV_CONST T* m, int i) {
    CM<T> r;
    r.x = m[2 * i];
    r.y = m[2 * i + 1];
    return r;
}

// A value every lane holds alike (read from LDS at a uniform address) moved to
// scalar registers, so that branches on it are uniform and FMAs take it as
// their scalar operand.
__device__ __forceinline__ uint32_t sreg(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ uint64_t sreg(uint64_t v) {
    return (uint64_t)sreg((uint32_t)v) | ((uint64_t)sreg((uint32_t)(v >> 32)) << 32);
}
__device__ __forceinline__ double sreg(double v) {
    return __builtin_bit_cast(double, sreg(__builtin_bit_cast(uint64_t, v)));
}
__device__ __forceinline__ float sreg(float v) { return __builtin_bit_cast(float, sreg(__builtin_bit_cast(uint32_t, v))); }

// word i of an op record: staged in LDS (moved to scalar registers) or read
// through the constant address space (already scalar)
__device__ __forceinline__ uint64_t rec_word(const uint64_t* w, int i) { return sreg(w[i]); }
__device__ __forceinline__ uint64_t rec_word(const KSV_CONST uint64_t* w, int i) { return w[i]; }
__device__ __forceinline__ const double* rec_pay(const uint64_t* w, double*) { return reinterpret_cast<const double*>(w + 4); }
__device__ __forceinline__ const float* rec_pay(const uint64_t* w, float*) { return reinterpret_cast<const float*>(w + 4); }
__device__ __forceinline__ const KSV_CONST double* rec_pay(const KSV_CONST uint64_t* w, double*) {
    return reinterpret_cast<const KSV_CONST double*>(w + 4);
}
__device__ __forceinline__ const KSV_CONST float* rec_pay(const KSV_CONST uint64_t* w, float*) {
    return reinterpret_cast<const KSV_CONST float*>(w + 4);
}

// matrix entry i of an op record staged in LDS
template <typename T>
__device__ __forceinline__ CM<T> ldm(const T* m, int i) {
    CM<T> r;
    r.x = sreg(m[2 * i]);
    r.y = sreg(m[2 * i + 1]);
    return r;
}

template <typename T, bool PR>
__device__ __forceinline__ void prune2(T& r, T& i) {
    if (PR) {
        if (knorm2((double)r, (double)i) < 1e-14) { r = 0; i = 0; }
    }
}

// ket's accumulate test, deferred to the end of an op.  max(|re|, |im|) >= PRUNE_T
// implies norm2 >= 1e-14 as ket rounds it (fl(T * T) >= 1e-14 and the rounded sum
// of two non-negative squares is at least the larger one), so only amplitudes
// below PRUNE_T are candidates: an op tests each output with one max and one
// compare, and runs the exact test (prune_all) over the amplitudes it wrote only
// where some lane of the wave holds a candidate.
constexpr double PRUNE_T = 1.0000001e-7;

template <typename T>
__device__ __forceinline__ bool pcand(T r, T i) {
    return (double)__builtin_fmax(__builtin_fabs(r), __builtin_fabs(i)) < PRUNE_T;
}

// the op wrote register j when j has every bit of rctl (and, with a thread-bit
// control, when this thread is selected)
template <typename T, bool PR, bool SEL>
__device__ __forceinline__ void prune_all(T (&re)[RREGS], T (&im)[RREGS], bool cand, uint32_t rctl, bool lp) {
    if (PR && cand && (!SEL || lp)) {
#pragma unroll
        for (int j = 0; j < RREGS; j++) {
            if (((uint32_t)j & rctl) != rctl) continue;
            prune2<T, PR>(re[j], im[j]);
        }
    }
}

// (ore, oim) = a * x0 + b * x1.  PR (ket's per-gate pruning): ket's
// add(mul(a, x0), mul(b, x1)), every product and sum rounded on its own
// (src/statevector.ts:61-62), so results and prune decisions are ket's bit
// for bit; exact mode: four fused multiply-adds per component.
template <typename T, bool PR>
__device__ __forceinline__ void mac2(T& ore, T& oim, CM<T> a, T x0r, T x0i, CM<T> b, T x1r, T x1i) {
    if (PR) {
        T pr, pi, qr, qi;
        kmul(a.x, a.y, x0r, x0i, pr, pi);
        kmul(b.x, b.y, x1r, x1i, qr, qi);
        ore = rn_add(pr, qr);
        oim = rn_add(pi, qi);
    } else {
        ore = __builtin_fma(a.x, x0r, __builtin_fma(-a.y, x0i, __builtin_fma(b.x, x1r, -(b.y * x1i))));
        oim = __builtin_fma(a.x, x0i, __builtin_fma(a.y, x0r, __builtin_fma(b.x, x1i, b.y * x1r)));
    }
}

// (yr, yi) = d * x: ket's mul under PR (a diagonal gate's other products are
// zeros, which leave ket's sum unchanged), fused otherwise
template <typename T, bool PR>
__device__ __forceinline__ void dmul(T& yr, T& yi, T dx, T dy, T xr, T xi) {
    if (PR) {
        kmul(dx, dy, xr, xi, yr, yi);
    } else {
        yr = __builtin_fma(dx, xr, -(dy * xi));
        yi = __builtin_fma(dx, xi, dy * xr);
    }
}

#define RREGS_ARGS T (&re)[RREGS], T (&im)[RREGS]

// 2x2 on slot S.  SEL: a thread-bit control exists (lp: this thread is
// selected).  RC: slot controls exist (rctl).  Under PR, `form` (ROP_REAL,
// ROP_RXF) drops the products with a zero factor from ket's complex products:
// mul(a, x) with a.y == 0 is (a.x x.re, a.x x.im) and with a.x == 0 it is
// (-(a.y x.im), a.y x.re), each exact except for the sign of a zero, so the
// sums below equal ket's add(mul(a, x0), mul(b, x1)) bit for bit wherever the
// result is non-zero, and zeros are pruned either way.
template <typename T, int S, bool PR, bool SEL, bool RC, typename P>
__device__ __forceinline__ void r_m1(RREGS_ARGS, P m, uint32_t rctl, bool lp, uint32_t form) {
    const CM<T> m0 = ldm(m, 0), m1 = ldm(m, 1), m2 = ldm(m, 2), m3 = ldm(m, 3);
    bool cand = false;
#pragma unroll
    for (int j = 0; j < RREGS; j++) {
        if ((j >> S) & 1) continue;
        if (RC && ((uint32_t)j & rctl) != rctl) continue;
        const int j1 = j | (1 << S);
        const T x0r = re[j], x0i = im[j], x1r = re[j1], x1i = im[j1];
        T y0r, y0i, y1r, y1i;
        if (PR && form == ROP_REAL) {
            y0r = rn_add(rn_mul(m0.x, x0r), rn_mul(m1.x, x1r));
            y0i = rn_add(rn_mul(m0.x, x0i), rn_mul(m1.x, x1i));
            y1r = rn_add(rn_mul(m2.x, x0r), rn_mul(m3.x, x1r));
            y1i = rn_add(rn_mul(m2.x, x0i), rn_mul(m3.x, x1i));
        } else if (PR && form == ROP_RXF) {
            y0r = rn_sub(rn_mul(m0.x, x0r), rn_mul(m1.y, x1i));
            y0i = rn_add(rn_mul(m0.x, x0i), rn_mul(m1.y, x1r));
            y1r = rn_sub(rn_mul(m3.x, x1r), rn_mul(m2.y, x0i));
            y1i = rn_add(rn_mul(m2.y, x0r), rn_mul(m3.x, x1i));
        } else {
            mac2<T, PR>(y0r, y0i, m0, x0r, x0i, m1, x1r, x1i);
            mac2<T, PR>(y1r, y1i, m2, x0r, x0i, m3, x1r, x1i);
        }
        if (PR) cand = cand | pcand(y0r, y0i) | pcand(y1r, y1i);
        if (SEL) {
            y0r = lp ? y0r : x0r; y0i = lp ? y0i : x0i;
            y1r = lp ? y1r : x1r; y1i = lp ? y1i : x1i;
        }
        re[j] = y0r; im[j] = y0i;
        re[j1] = y1r; im[j1] = y1i;
    }
    prune_all<T, PR, SEL>(re, im, cand, RC ? rctl : 0u, lp);
}

template <typename T, int S, bool SEL>
__device__ __forceinline__ void r_permx(RREGS_ARGS, uint32_t rctl, bool lp) {
#pragma unroll
    for (int j = 0; j < RREGS; j++) {
        if ((j >> S) & 1) continue;
        if (((uint32_t)j & rctl) != rctl) continue;
        const int j1 = j | (1 << S);
        const T ar = re[j], ai = im[j], br = re[j1], bi = im[j1];
        re[j] = SEL ? (lp ? br : ar) : br;
        im[j] = SEL ? (lp ? bi : ai) : bi;
        re[j1] = SEL ? (lp ? ar : br) : ar;
        im[j1] = SEL ? (lp ? ai : bi) : ai;
    }
}

template <typename T, int A, int B, bool SEL>
__device__ __forceinline__ void r_swap(RREGS_ARGS, uint32_t rctl, bool lp) {
#pragma unroll
    for (int j = 0; j < RREGS; j++) {
        if (!((j >> A) & 1) || ((j >> B) & 1)) continue;
        if (((uint32_t)j & rctl) != rctl) continue;
        const int j1 = j ^ (1 << A) ^ (1 << B);
        const T ar = re[j], ai = im[j], br = re[j1], bi = im[j1];
        re[j] = SEL ? (lp ? br : ar) : br;
        im[j] = SEL ? (lp ? bi : ai) : bi;
        re[j1] = SEL ? (lp ? ar : br) : ar;
        im[j1] = SEL ? (lp ? ai : bi) : ai;
    }
}

// One-qubit diagonal diag(d0, d1).  S >= 0: the qubit is slot S, so each
// register's entry is fixed at compile time; S < 0: the qubit is a thread bit
// or a bit outside the tile, `hi` says which entry this thread takes.  The
// same arithmetic as r_diag's, with the row selection out of the loop.
template <typename T, int S, bool PR, bool SEL, typename P>
__device__ __forceinline__ void r_diag1(RREGS_ARGS, uint32_t rctl, P pay, bool hi, bool lp) {
    const CM<T> d0 = ldm(pay, 0), d1 = ldm(pay, 1);
    CM<T> dt;
    dt.x = hi ? d1.x : d0.x;
    dt.y = hi ? d1.y : d0.y;
    bool cand = false;
#pragma unroll
    for (int j = 0; j < RREGS; j++) {
        if (((uint32_t)j & rctl) != rctl) continue;
        const CM<T> d = S >= 0 ? (((j >> (S < 0 ? 0 : S)) & 1) ? d1 : d0) : dt;
        T yr, yi;
        dmul<T, PR>(yr, yi, d.x, d.y, re[j], im[j]);
        if (PR) cand = cand | pcand(yr, yi);
        if (SEL) { yr = lp ? yr : re[j]; yi = lp ? yi : im[j]; }
        re[j] = yr;
        im[j] = yi;
    }
    prune_all<T, PR, SEL>(re, im, cand, rctl, lp);
}

// Diagonal over k <= 6 qubits.  Row bit (k-1-i) of qubit i comes from a
// thread bit, a slot or a physical bit outside the tile (loc[i]).  For
// k <= 2 the entries are inline (scalar registers) and selected per
// amplitude; wider diagonals are looked up per amplitude.
template <typename T, bool PR, bool SEL, typename P>
__device__ __forceinline__ void r_diag(RREGS_ARGS, uint32_t k, uint64_t locw, uint32_t rctl, P pay,
                                       const T* wide, uint32_t tid, uint64_t base, bool lp) {
    uint32_t rl = 0, s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
    for (int i = 0; i < DIAG_MAX; i++) {
        if ((uint32_t)i >= k) continue;
        const uint32_t rb = 1u << (k - 1 - i);
        const uint32_t L = (uint32_t)(locw >> (8 * i)) & 0xFF;
        if (L < 16) {
            if ((tid >> L) & 1) rl |= rb;
        } else if (L < 32) {
            const uint32_t s = L - 16;
            if (s == 0) s0 |= rb;
            else if (s == 1) s1 |= rb;
            else if (s == 2) s2 |= rb;
            else s3 |= rb;
        } else if ((base >> (L - 32)) & 1) {
            rl |= rb;
        }
    }
    bool cand = false;
    if (k <= 2) {
        const CM<T> d0 = ldm(pay, 0), d1 = ldm(pay, 1), d2 = ldm(pay, 2), d3 = ldm(pay, 3);
#pragma unroll
        for (int j = 0; j < RREGS; j++) {
            if (((uint32_t)j & rctl) != rctl) continue;
            const uint32_t row = rl | ((j & 1) ? s0 : 0) | ((j & 2) ? s1 : 0) | ((j & 4) ? s2 : 0) | ((j & 8) ? s3 : 0);
            const T dx = (row & 2) ? ((row & 1) ? d3.x : d2.x) : ((row & 1) ? d1.x : d0.x);
            const T dy = (row & 2) ? ((row & 1) ? d3.y : d2.y) : ((row & 1) ? d1.y : d0.y);
            T yr, yi;
            dmul<T, PR>(yr, yi, dx, dy, re[j], im[j]);
            if (PR) cand = cand | pcand(yr, yi);
            if (SEL) { yr = lp ? yr : re[j]; yi = lp ? yi : im[j]; }
            re[j] = yr;
            im[j] = yi;
        }
    } else {
#pragma unroll
        for (int j = 0; j < RREGS; j++) {
            if (((uint32_t)j & rctl) != rctl) continue;
            const uint32_t row = rl | ((j & 1) ? s0 : 0) | ((j & 2) ? s1 : 0) | ((j & 4) ? s2 : 0) | ((j & 8) ? s3 : 0);
            const T dx = wide[2 * row], dy = wide[2 * row + 1];
            T yr, yi;
            dmul<T, PR>(yr, yi, dx, dy, re[j], im[j]);
            if (PR) cand = cand | pcand(yr, yi);
            if (SEL) { yr = lp ? yr : re[j]; yi = lp ? yi : im[j]; }
            re[j] = yr;
            im[j] = yi;
        }
    }
    prune_all<T, PR, SEL>(re, im, cand, rctl, lp);
}

// Dense 2^k x 2^k (k = 2..4) on slots ts[0] (MSB) .. ts[k-1].  The thread's
// amplitudes go to its own LDS column; output j reads its 2^k inputs back.
template <typename T, int TB, bool PR>
__device__ __forceinline__ void r_dense(RREGS_ARGS, typename Vec2<T>::type* lds, uint32_t k, uint64_t locw,
                                        uint32_t rctl, const KSV_CONST T* m, uint32_t tid, bool sel, bool lp) {
    using V = typename Vec2<T>::type;
    constexpr int NT = 1 << TB;
    const uint32_t dim = 1u << k;
    uint32_t tmask = 0;
    for (uint32_t i = 0; i < k; i++) tmask |= 1u << ((locw >> (8 * i)) & 0xFF);
    __syncthreads();        // other threads may still read the last exchange
#pragma unroll
    for (int j = 0; j < RREGS; j++) {
        V v;
        v.x = re[j];
        v.y = im[j];
        lds[j * NT + tid] = v;
    }
#pragma unroll
    for (int j = 0; j < RREGS; j++) {
        if (((uint32_t)j & rctl) != rctl) continue;
        uint32_t row = 0;
        for (uint32_t i = 0; i < k; i++) row |= (((uint32_t)j >> ((locw >> (8 * i)) & 0xFF)) & 1) << (k - 1 - i);
        const uint32_t jb = (uint32_t)j & ~tmask;
        T ar = 0, ai = 0;
        for (uint32_t c = 0; c < dim; c++) {
            uint32_t jc = jb;
            for (uint32_t i = 0; i < k; i++) jc |= ((c >> (k - 1 - i)) & 1) << ((locw >> (8 * i)) & 0xFF);
            const V x = lds[jc * NT + tid];
            const CM<T> e = ldm(m, (int)(row * dim + c));
            if (PR) {       // add(out, mul(m, x)) in column order (src/statevector.ts:125-126, :263-264)
                kacc(ar, ai, e.x, e.y, x.x, x.y);
            } else {
                ar = __builtin_fma(e.x, x.x, __builtin_fma(-e.y, x.y, ar));
                ai = __builtin_fma(e.x, x.y, __builtin_fma(e.y, x.x, ai));
            }
        }
        prune2<T, PR>(ar, ai);
        if (!sel || lp) {
            re[j] = ar;
            im[j] = ai;
        }
    }
}

// LDS transpose from the current slot/thread assignment to the next one.
// `xw`: the record's 32 16-bit entries (write address of thread bit 0..11,
// of slot 0..3, read address likewise) as eight words.
template <typename T, int TB, typename P>
__device__ __forceinline__ void r_exch(RREGS_ARGS, typename Vec2<T>::type* lds, P xw, uint32_t tid) {
    using V = typename Vec2<T>::type;
    uint64_t xs[8];
#pragma unroll
    for (int i = 0; i < 8; i++) xs[i] = rec_word(xw, i);
    auto x = [&](int i) -> uint32_t { return (uint32_t)(xs[i >> 2] >> (16 * (i & 3))) & 0xFFFF; };
    uint32_t wb = 0, rb = 0;
#pragma unroll
    for (int i = 0; i < TB; i++) {
        const uint32_t bit = (tid >> i) & 1;
        wb ^= bit ? x(i) : 0u;
        rb ^= bit ? x(16 + i) : 0u;
    }
    const uint32_t w0 = x(12), w1 = x(13), w2 = x(14), w3 = x(15);
    const uint32_t r0 = x(28), r1 = x(29), r2 = x(30), r3 = x(31);
    __syncthreads();        // the previous LDS use of every thread is finished
#pragma unroll
    for (int j = 0; j < RREGS; j++) {
        const uint32_t a = wb ^ ((j & 1) ? w0 : 0) ^ ((j & 2) ? w1 : 0) ^ ((j & 4) ? w2 : 0) ^ ((j & 8) ? w3 : 0);
        V v;
        v.x = re[j];
        v.y = im[j];
        lds[a] = v;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RREGS; j++) {
        const uint32_t a = rb ^ ((j & 1) ? r0 : 0) ^ ((j & 2) ? r1 : 0) ^ ((j & 4) ? r2 : 0) ^ ((j & 8) ? r3 : 0);
        const V v = lds[a];
        re[j] = v.x;
        im[j] = v.y;
    }
}

// The same exchange through a real array of half the size: real parts, then
// imaginary parts (two more barriers).  Builds without the dense path use it,
// so that their workgroups take 16 KiB of LDS (plus staged records) at
// 11-qubit tiles and three waves fit on a SIMD.
template <typename T, int TB, typename P>
__device__ __forceinline__ void r_exch_half(RREGS_ARGS, T* lds, P xw, uint32_t tid) {
    uint64_t xs[8];
#pragma unroll
    for (int i = 0; i < 8; i++) xs[i] = rec_word(xw, i);
    auto x = [&](int i) -> uint32_t { return (uint32_t)(xs[i >> 2] >> (16 * (i & 3))) & 0xFFFF; };
    uint32_t wb = 0, rb = 0;
#pragma unroll
    for (int i = 0; i < TB; i++) {
        const uint32_t bit = (tid >> i) & 1;
        wb ^= bit ? x(i) : 0u;
        rb ^= bit ? x(16 + i) : 0u;
    }
    const uint32_t w0 = x(12), w1 = x(13), w2 = x(14), w3 = x(15);
    const uint32_t r0 = x(28), r1 = x(29), r2 = x(30), r3 = x(31);
    auto wa = [&](int j) { return wb ^ ((j & 1) ? w0 : 0) ^ ((j & 2) ? w1 : 0) ^ ((j & 4) ? w2 : 0) ^ ((j & 8) ? w3 : 0); };
    auto ra = [&](int j) { return rb ^ ((j & 1) ? r0 : 0) ^ ((j & 2) ? r1 : 0) ^ ((j & 4) ? r2 : 0) ^ ((j & 8) ? r3 : 0); };
    __syncthreads();        // the previous LDS use of every thread is finished
#pragma unroll
    for (int j = 0; j < RREGS; j++) lds[wa(j)] = re[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RREGS; j++) re[j] = lds[ra(j)];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RREGS; j++) lds[wa(j)] = im[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RREGS; j++) im[j] = lds[ra(j)];
}

#define RSLOT_CASES(SLOT, FN) \
    switch (SLOT) {           \
    case 0: FN(0); break;     \
    case 1: FN(1); break;     \
    case 2: FN(2); break;     \
    default: FN(3); break;    \
    }

#define RPAIR_CASES(A, B, FN)          \
    switch ((A) * 4 + (B)) {           \
    case 4: FN(1, 0); break;           \
    case 8: FN(2, 0); break;           \
    case 9: FN(2, 1); break;           \
    case 12: FN(3, 0); break;          \
    case 13: FN(3, 1); break;          \
    default: FN(3, 2); break;          \
    }

__device__ __forceinline__ uint64_t phys_offset(uint32_t bits, int nbits, uint64_t lo, uint64_t hi) {
    uint64_t off = 0;
#pragma unroll
    for (int i = 0; i < 12; i++)
        if (i < nbits) off |= (uint64_t)((bits >> i) & 1) << byte_at(i < 8 ? lo : hi, i & 7);
    return off;
}

__device__ __forceinline__ uint64_t tile_base(uint64_t tile, int K, uint64_t tq_lo, uint64_t tq_hi) {
    uint64_t base = tile;
#pragma unroll
    for (int i = 0; i < 16; i++)
        if (i < K) base = ins0_64(base, byte_at(i < 8 ? tq_lo : tq_hi, i & 7));
    return base;
}

template <typename T>
__device__ __forceinline__ uint64_t slot_offset(int j, uint32_t rmap) {
    uint64_t ro = 0;
#pragma unroll
    for (int s = 0; s < RSLOTS; s++)
        if ((j >> s) & 1) ro |= 1ull << byte_at(rmap, s);
    return ro;
}

// DN: the kernel has the dense-op path.  Register allocation is global to the
// kernel, so leaving that path out changes the code of every other op; steps
// without dense ops run the build without it (42.9 against 44.5 ms for the
// 12-qubit noisy batch of 10 240 copies; 2880 against 2888 ms for the 30-qubit
// random circuit in 11-qubit register rounds).
// STG: op records staged through LDS with the per-chunk run mask (copy
// programs: long op streams, copy-conditional ops); without it each op reads
// its record through the scalar cache, which holds an ordinary pass's short
// op stream (2888 against 3367 ms for the 30-qubit random circuit in 11-qubit
// register rounds).
template <typename T, int TB, bool PR, bool DN, bool STG>
__device__ __forceinline__ void rpass_body(const RPassArgs& a) {
    using V = typename Vec2<T>::type;
    constexpr int K = TB + RSLOTS;
    constexpr int NT = 1 << TB;
    constexpr int C = NT / 8;                 // op records per staged chunk: 16 bytes per thread
    static_assert(C <= 64 && sizeof(ROp) == 128, "one wave's ballot covers a chunk");
    // the dense path stages complex amplitudes; without it the exchange moves
    // real and imaginary parts in turn through half the space
    __shared__ typename std::conditional<DN, V, T>::type lds[1 << K];
    __shared__ uint4 stage[STG ? 2 : 1][STG ? NT : 1];   // two chunks of op records
    const uint32_t tid = threadIdx.x;
    V* psi = reinterpret_cast<V*>(a.psi);
    const T* wide_mats = reinterpret_cast<const T*>(a.mats);
    const KSV_CONST T* cmats = to_const(reinterpret_cast<const T*>(a.mats));
    const uint4* gops = reinterpret_cast<const uint4*>(a.ops);

    uint64_t tile = blockIdx.x;
    if (tile >= a.ntiles) return;
    const uint64_t ld_off = phys_offset(tid, TB, a.ld_t_lo, a.ld_t_hi);
    const uint64_t st_off = phys_offset(tid, TB, a.st_t_lo, a.st_t_hi);
    uint64_t base = tile_base(tile, K, a.tq_lo, a.tq_hi);

    // Chunk c of the op stream: this thread's 16 bytes of its records, and (lanes
    // 0..C-1 of every wave) the copy-bitmap word of op c*C + lane for this
    // tile's copy.  Chunk c+1 is loaded while chunk c runs.
    const uint32_t nops = (uint32_t)a.nops, nch = (nops + C - 1) / C;
    const uint64_t copy = base >> a.nsys;
    const uint32_t lane = tid & 63;
    const uint32_t* crow = a.rbits ? a.rbits + (size_t)(copy >> 5) * a.rstride + a.op_base : nullptr;
    auto fetch_rec = [&](uint32_t c) -> uint4 {
        const size_t i = (size_t)c * NT + tid;
        return i < (size_t)nops * 8 ? gops[i] : make_uint4(0, 0, 0, 0);
    };
    auto fetch_cw = [&](uint32_t c) -> uint32_t {
        const uint32_t o = c * C + lane;
        return (crow && lane < (uint32_t)C && o < nops) ? crow[o] : 0u;
    };
    auto cmask_of = [&](uint32_t cw) -> uint64_t {
        return __builtin_amdgcn_ballot_w64(((cw >> (copy & 31)) & 1u) != 0);
    };

    T re[RREGS], im[RREGS];
#pragma unroll
    for (int j = 0; j < RREGS; j++) {
        const V t = psi[base | ld_off | slot_offset<T>(j, a.ld_r)];
        re[j] = t.x;
        im[j] = t.y;
    }

    // one op; `w`: its record (LDS or constant address space)
    auto op_body = [&](auto w) {
        const uint64_t h0 = rec_word(w, 0), h1 = rec_word(w, 1), h2 = rec_word(w, 2), h3 = rec_word(w, 3);
        const uint32_t kind = (uint32_t)h0 & 0xFF;
        const uint32_t sa = (uint32_t)(h0 >> 8) & 0xFF, sb = (uint32_t)(h0 >> 16) & 0xFF;
        const uint32_t kq = (uint32_t)(h0 >> 24) & 0xFF;
        const uint64_t locw = (h0 >> 32) | ((h1 & 0xFFFF) << 32);
        const uint32_t tctl = (uint32_t)(h1 >> 32);
        const uint32_t rctl = (uint32_t)h2;
        const int32_t mat = (int32_t)(h2 >> 32);
        const uint64_t gctl = h3;
        const auto pay = rec_pay(w, (T*)nullptr);
        if (kind == R_EXCH) {
            if constexpr (DN) r_exch<T, TB>(re, im, lds, w + 4, tid);
            else r_exch_half<T, TB>(re, im, lds, w + 4, tid);
            return;
        }
        if constexpr (DN) {
        if (kind == R_DENSE) {
            // every thread reaches the barrier inside, whatever gctl says
            const bool on = (base & gctl) == gctl;
            const bool lp = on && ((tid & tctl) == tctl);
            r_dense<T, TB, PR>(re, im, lds, kq, locw, rctl, cmats + 2 * mat, tid, true, lp);
            return;
        }
        }
        if ((base & gctl) != gctl) return;
        const bool lp = (tid & tctl) == tctl;
        const bool sel = tctl != 0;
        switch (kind) {
        case R_M1: {
            // copy programs (STG) keep the general form: the reduced forms' code
            // costs that build registers (noisy batch 142k -> 112k shots/s)
            const uint32_t form = STG ? 0u : (uint32_t)(h1 >> 16) & (ROP_REAL | ROP_RXF);
            if (rctl == 0) {
#define F(S) (sel ? r_m1<T, S, PR, true, false>(re, im, pay, 0, lp, form) : r_m1<T, S, PR, false, false>(re, im, pay, 0, lp, form))
                RSLOT_CASES(sa, F)
#undef F
            } else {
#define F(S) (sel ? r_m1<T, S, PR, true, true>(re, im, pay, rctl, lp, form) : r_m1<T, S, PR, false, true>(re, im, pay, rctl, lp, form))
                RSLOT_CASES(sa, F)
#undef F
            }
            break;
        }
        case R_DIAG:
            if (kq == 1) {
                const uint32_t L = (uint32_t)locw & 0xFF;
                if (L >= 16 && L < 32) {
#define F(S) (sel ? r_diag1<T, S, PR, true>(re, im, rctl, pay, false, lp) : r_diag1<T, S, PR, false>(re, im, rctl, pay, false, lp))
                    RSLOT_CASES(L - 16, F)
#undef F
                } else {
                    const bool hi = L < 16 ? ((tid >> L) & 1) != 0 : ((base >> (L - 32)) & 1) != 0;
                    if (sel) r_diag1<T, -1, PR, true>(re, im, rctl, pay, hi, lp);
                    else r_diag1<T, -1, PR, false>(re, im, rctl, pay, hi, lp);
                }
                break;
            }
            if (sel) r_diag<T, PR, true>(re, im, kq, locw, rctl, pay, wide_mats + 2 * mat, tid, base, lp);
            else r_diag<T, PR, false>(re, im, kq, locw, rctl, pay, wide_mats + 2 * mat, tid, base, lp);
            break;
        case R_PERMX:
#define F(S) (sel ? r_permx<T, S, true>(re, im, rctl, lp) : r_permx<T, S, false>(re, im, rctl, lp))
            RSLOT_CASES(sa, F)
#undef F
            break;
        case R_SWAP:
#define F(A, B) (sel ? r_swap<T, A, B, true>(re, im, rctl, lp) : r_swap<T, A, B, false>(re, im, rctl, lp))
            RPAIR_CASES(sa, sb, F)
#undef F
            break;
        default:
            break;
        }
    };

    if constexpr (!STG) {
        const KSV_CONST uint64_t* cops = to_const(reinterpret_cast<const uint64_t*>(a.ops));
        for (int o = 0; o < a.nops; o++) op_body(cops + 16 * o);
    } else {
        stage[0][tid] = fetch_rec(0);
        uint64_t cmask = cmask_of(fetch_cw(0));
        uint4 pre = fetch_rec(1);
        uint32_t pre_cw = fetch_cw(1);
        __syncthreads();
        for (uint32_t c = 0; c < nch; c++) {
            const uint64_t* recs = reinterpret_cast<const uint64_t*>(stage[c & 1]);
            const uint32_t cn = nops - c * C < (uint32_t)C ? nops - c * C : (uint32_t)C;
            // the chunk's ops that run on this tile: the ballot of the copy-bitmap
            // words (all ones for unconditional ops), else every op
            uint64_t run = crow ? cmask : (cn >= 64 ? ~0ull : ((1ull << cn) - 1));
            while (run) {
                const uint32_t oi = (uint32_t)__builtin_ctzll(run);
                run &= run - 1;
                op_body(recs + 16 * oi);
            }
            if (c + 1 < nch) {
                // chunk c-1's buffer: every thread finished it before the last barrier
                stage[(c + 1) & 1][tid] = pre;
                cmask = cmask_of(pre_cw);
                pre = fetch_rec(c + 2);
                pre_cw = fetch_cw(c + 2);
                __syncthreads();
            }
        }
    }

#pragma unroll
    for (int j = 0; j < RREGS; j++) {
        V t;
        t.x = re[j];
        t.y = im[j];
        psi[base | st_off | slot_offset<T>(j, a.st_r)] = t;
    }
}

template <typename T, int TB, bool PR, bool DN, bool STG>
__global__ __launch_bounds__(1 << TB) void k_rpass(RPassArgs a) {
    rpass_body<T, TB, PR, DN, STG>(a);
}

// FP64 rounds under ket's pruning without the dense path (the library's default
// mode on circuits of one- and two-qubit gates, and the copies of batched
// trajectories): held to 168 registers, three waves per SIMD instead of two (the
// half-size exchange leaves LDS for them).
template <int TB, bool STG>
__global__ __launch_bounds__(1 << TB) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_rpass_prune3(RPassArgs a) {
    rpass_body<double, TB, true, false, STG>(a);
}

template <typename T, bool PR, bool DN, bool STG>
static void launch_tb(const RPassArgs& a, int tb, hipStream_t st) {
    if constexpr (sizeof(T) == 8 && PR && !DN) {
        // 13-qubit rounds spill too much at 168 registers: they keep two waves
        const dim3 grid((unsigned)a.ntiles);
        if (tb == 7) {
            hipLaunchKernelGGL((k_rpass_prune3<7, STG>), grid, dim3(1 << 7), 0, st, a);
            return;
        }
        if (tb == 8) {
            hipLaunchKernelGGL((k_rpass_prune3<8, STG>), grid, dim3(1 << 8), 0, st, a);
            return;
        }
    }
    const dim3 grid((unsigned)a.ntiles);
    if (tb == 7) hipLaunchKernelGGL((k_rpass<T, 7, PR, DN, STG>), grid, dim3(1 << 7), 0, st, a);
    else if (tb == 8) hipLaunchKernelGGL((k_rpass<T, 8, PR, DN, STG>), grid, dim3(1 << 8), 0, st, a);
    else hipLaunchKernelGGL((k_rpass<T, 9, PR, DN, STG>), grid, dim3(1 << 9), 0, st, a);
}

// copy programs: staged records; every other program: records through the
// scalar cache; each with the dense path only when the step has dense ops
template <typename T, bool PR>
static void launch_dn(const RPassArgs& a, int tb, bool copies, bool dense, hipStream_t st) {
    if (!copies) {
        if (dense) launch_tb<T, PR, true, false>(a, tb, st);
        else launch_tb<T, PR, false, false>(a, tb, st);
    } else if (dense) launch_tb<T, PR, true, true>(a, tb, st);
    else launch_tb<T, PR, false, true>(a, tb, st);
}

hipError_t launch_rpass(const Program& p, const Step& s, void* psi, hipStream_t st) {
    RPassArgs a;
    std::memset(&a, 0, sizeof a);
    a.psi = psi;
    a.ops = p.d_rops + s.op_begin;
    a.mats = p.d_mats;
    a.nops = s.op_count;
    a.k = s.k;
    a.ntiles = 1ull << (p.n - s.k);
    auto pack = [](const int8_t* b, int n) {
        uint64_t w = 0;
        for (int i = 0; i < n && i < 8; i++) w |= (uint64_t)(uint8_t)b[i] << (8 * i);
        return w;
    };
    const int tb = s.k - RSLOTS;
    if (tb < 7 || tb > 9) return hipErrorInvalidValue;
    a.tq_lo = pack(s.tq, 8);
    a.tq_hi = pack(s.tq + 8, 8);
    a.ld_t_lo = pack(s.ld_t, std::min(tb, 8));
    a.ld_t_hi = tb > 8 ? pack(s.ld_t + 8, tb - 8) : 0;
    a.st_t_lo = pack(s.st_t, std::min(tb, 8));
    a.st_t_hi = tb > 8 ? pack(s.st_t + 8, tb - 8) : 0;
    a.ld_r = (uint32_t)pack(s.ld_r, RSLOTS);
    a.st_r = (uint32_t)pack(s.st_r, RSLOTS);
    a.rbits = p.d_rbits;
    a.rstride = (uint32_t)p.rops.size();
    a.op_base = (uint32_t)s.op_begin;
    a.nsys = p.nsys;
    const bool copies = p.nsys > 0;
    bool dense = false;
    for (int i = 0; i < s.op_count && !dense; i++) dense = p.rops[(size_t)s.op_begin + i].kind == R_DENSE;
    if (p.precision == KSV_F64) {
        if (p.prune) launch_dn<double, true>(a, tb, copies, dense, st);
        else launch_dn<double, false>(a, tb, copies, dense, st);
    } else {
        if (p.prune) launch_dn<float, true>(a, tb, copies, dense, st);
        else launch_dn<float, false>(a, tb, copies, dense, st);
    }
    return hipGetLastError();
}

}  // namespace ksv
