// Register-tile pass kernel.
//
// One wavefront owns a tile of 2^(6+RL) amplitudes: lane l, register j holds
// the amplitude whose lane-bit qubits read l and whose register-slot qubits
// read j.  Gates whose targets sit in register slots are plain FP math on
// registers (no LDS, no barriers, no index arithmetic); a gate on a
// lane-bit qubit is preceded by W_LSWAP, which exchanges that lane bit with a
// register slot (v_permlane32/16_swap for lane bits 5/4, DPP or ds_swizzle
// for lane bits 0..3).  Lane bits 0..2 hold state bits 0..2 at load and
// store, so every 8 consecutive lanes move one contiguous 128-byte run.  Each
// amplitude crosses HBM once in and once out per pass.
//
// The tile lives in two scalar arrays (real, imaginary) indexed only by
// compile-time constants, so it is promoted to VGPRs; every op is dispatched
// to a template specialised on its slot numbers.  Op records and matrices
// are read through the constant address space (scalar loads into SGPRs).
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
__device__ __forceinline__ CM<T> ldc(const KSV_CONST T* m, int i) {
    CM<T> r;
    r.x = m[2 * i];
    r.y = m[2 * i + 1];
    return r;
}

// (ore, oim) = a * (x0) + b * (x1)
template <typename T>
__device__ __forceinline__ void mac2(T& ore, T& oim, CM<T> a, T x0r, T x0i, CM<T> b, T x1r, T x1i) {
    ore = (a.x * x0r - a.y * x0i) + (b.x * x1r - b.y * x1i);
    oim = (a.x * x0i + a.y * x0r) + (b.x * x1i + b.y * x1r);
}

template <typename T, bool PR>
__device__ __forceinline__ void prune2(T& r, T& i) {
    if (PR) {
        if (((double)r * r + (double)i * i) < 1e-14) { r = 0; i = 0; }
    }
}

#define REGS T* re, T* im

template <typename T, int RL, int S, bool LP, bool PR>
__device__ __forceinline__ void w_m1(REGS, const WOp& op, const KSV_CONST T* mats, bool lp) {
    constexpr int R = 1 << RL;
    const KSV_CONST T* m = mats + 2 * op.mat;
    const CM<T> m0 = ldc(m, 0), m1 = ldc(m, 1), m2 = ldc(m, 2), m3 = ldc(m, 3);
#pragma unroll
    for (int j = 0; j < R; j++) {
        if ((j >> S) & 1) continue;
        const int j1 = j | (1 << S);
        if (((uint32_t)j & op.rctl) != op.rctl) continue;
        const T x0r = re[j], x0i = im[j], x1r = re[j1], x1i = im[j1];
        T y0r, y0i, y1r, y1i;
        mac2(y0r, y0i, m0, x0r, x0i, m1, x1r, x1i);
        mac2(y1r, y1i, m2, x0r, x0i, m3, x1r, x1i);
        prune2<T, PR>(y0r, y0i);
        prune2<T, PR>(y1r, y1i);
        if (LP) {
            y0r = lp ? y0r : x0r; y0i = lp ? y0i : x0i;
            y1r = lp ? y1r : x1r; y1i = lp ? y1i : x1i;
        }
        re[j] = y0r; im[j] = y0i;
        re[j1] = y1r; im[j1] = y1i;
    }
}

template <typename T, int RL, int SA, int SB, bool LP, bool PR>
__device__ __forceinline__ void w_m2(REGS, const WOp& op, const KSV_CONST T* mats, bool lp) {
    constexpr int R = 1 << RL;
    const KSV_CONST T* mm = mats + 2 * op.mat;
#pragma unroll
    for (int j = 0; j < R; j++) {
        if (((j >> SA) & 1) || ((j >> SB) & 1)) continue;
        if (((uint32_t)j & op.rctl) != op.rctl) continue;
        const int id[4] = {j, j | (1 << SB), j | (1 << SA), j | (1 << SA) | (1 << SB)};
        T xr[4], xi[4];
#pragma unroll
        for (int c = 0; c < 4; c++) { xr[c] = re[id[c]]; xi[c] = im[id[c]]; }
#pragma unroll
        for (int r = 0; r < 4; r++) {
            T pr, pi, qr, qi;
            mac2(pr, pi, ldc(mm, 4 * r), xr[0], xi[0], ldc(mm, 4 * r + 1), xr[1], xi[1]);
            mac2(qr, qi, ldc(mm, 4 * r + 2), xr[2], xi[2], ldc(mm, 4 * r + 3), xr[3], xi[3]);
            pr += qr;
            pi += qi;
            prune2<T, PR>(pr, pi);
            re[id[r]] = LP ? (lp ? pr : xr[r]) : pr;
            im[id[r]] = LP ? (lp ? pi : xi[r]) : pi;
        }
    }
}

template <typename T, int RL, int S, bool LP>
__device__ __forceinline__ void w_permx(REGS, const WOp& op, bool lp) {
    constexpr int R = 1 << RL;
#pragma unroll
    for (int j = 0; j < R; j++) {
        if ((j >> S) & 1) continue;
        const int j1 = j | (1 << S);
        if (((uint32_t)j & op.rctl) != op.rctl) continue;
        const T ar = re[j], ai = im[j], br = re[j1], bi = im[j1];
        re[j] = LP ? (lp ? br : ar) : br;
        im[j] = LP ? (lp ? bi : ai) : bi;
        re[j1] = LP ? (lp ? ar : br) : ar;
        im[j1] = LP ? (lp ? ai : bi) : ai;
    }
}

template <typename T, int RL, int SA, int SB, bool LP>
__device__ __forceinline__ void w_swapr(REGS, const WOp& op, bool lp) {
    constexpr int R = 1 << RL;
#pragma unroll
    for (int j = 0; j < R; j++) {
        if (!((j >> SA) & 1) || ((j >> SB) & 1)) continue;
        if (((uint32_t)j & op.rctl) != op.rctl) continue;
        const int j1 = j ^ (1 << SA) ^ (1 << SB);
        const T ar = re[j], ai = im[j], br = re[j1], bi = im[j1];
        re[j] = LP ? (lp ? br : ar) : br;
        im[j] = LP ? (lp ? bi : ai) : bi;
        re[j1] = LP ? (lp ? ar : br) : ar;
        im[j1] = LP ? (lp ? ai : bi) : ai;
    }
}

// ---- lane-bit exchange ---------------------------------------------------
template <int B>
__device__ __forceinline__ uint32_t lane_xor(uint32_t x) {
    if constexpr (B == 0) return __builtin_amdgcn_update_dpp(0u, x, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    else if constexpr (B == 1) return __builtin_amdgcn_update_dpp(0u, x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    else if constexpr (B == 2) return __builtin_amdgcn_ds_swizzle(x, 0x101F);                    // lanes ^ 4
    else return __builtin_amdgcn_update_dpp(0u, x, 0x128, 0xF, 0xF, false);                      // row_ror:8 = lanes ^ 8
}

// Swap 32-bit words so that lanes with bit B = 1 take in `lo` the word their
// partner held in `hi`, and lanes with bit B = 0 take in `hi` the word their
// partner held in `lo`.
template <int B>
__device__ __forceinline__ void swap_words(uint32_t& lo, uint32_t& hi) {
    if constexpr (B >= 4) {
        auto r = B == 5 ? __builtin_amdgcn_permlane32_swap(lo, hi, false, false)
                        : __builtin_amdgcn_permlane16_swap(lo, hi, false, false);
        lo = r[0];
        hi = r[1];
    } else {
        const bool lb = (__lane_id() >> B) & 1;
        const uint32_t send = lb ? lo : hi;
        const uint32_t recv = lane_xor<B>(send);
        lo = lb ? recv : lo;
        hi = lb ? hi : recv;
    }
}

template <int B>
__device__ __forceinline__ void swap_scalar(double& lo, double& hi) {
    uint32_t a[2], b[2];
    __builtin_memcpy(a, &lo, 8);
    __builtin_memcpy(b, &hi, 8);
    swap_words<B>(a[0], b[0]);
    swap_words<B>(a[1], b[1]);
    __builtin_memcpy(&lo, a, 8);
    __builtin_memcpy(&hi, b, 8);
}
template <int B>
__device__ __forceinline__ void swap_scalar(float& lo, float& hi) {
    uint32_t a, b;
    __builtin_memcpy(&a, &lo, 4);
    __builtin_memcpy(&b, &hi, 4);
    swap_words<B>(a, b);
    __builtin_memcpy(&lo, &a, 4);
    __builtin_memcpy(&hi, &b, 4);
}

// Exchange lane bit B with register slot S.
template <typename T, int RL, int S, int B>
__device__ __forceinline__ void w_lswap_b(REGS) {
    constexpr int R = 1 << RL;
#pragma unroll
    for (int j = 0; j < R; j++) {
        if ((j >> S) & 1) continue;
        swap_scalar<B>(re[j], re[j | (1 << S)]);
        swap_scalar<B>(im[j], im[j | (1 << S)]);
    }
}

template <typename T, int RL, int S>
__device__ __forceinline__ void w_lswap(REGS, int b) {
    switch (b) {
    case 0: w_lswap_b<T, RL, S, 0>(re, im); break;
    case 1: w_lswap_b<T, RL, S, 1>(re, im); break;
    case 2: w_lswap_b<T, RL, S, 2>(re, im); break;
    case 3: w_lswap_b<T, RL, S, 3>(re, im); break;
    case 4: w_lswap_b<T, RL, S, 4>(re, im); break;
    default: w_lswap_b<T, RL, S, 5>(re, im); break;
    }
}

template <typename T, int RL, bool LP, bool PR>
__device__ __forceinline__ void w_phase(REGS, const WOp& op, const KSV_CONST T* mats, bool lp) {
    constexpr int R = 1 << RL;
    const CM<T> ph = ldc(mats + 2 * op.mat, 0);
#pragma unroll
    for (int j = 0; j < R; j++) {
        if (((uint32_t)j & op.rctl) != op.rctl) continue;
        T yr = ph.x * re[j] - ph.y * im[j];
        T yi = ph.x * im[j] + ph.y * re[j];
        prune2<T, PR>(yr, yi);
        re[j] = LP ? (lp ? yr : re[j]) : yr;
        im[j] = LP ? (lp ? yi : im[j]) : yi;
    }
}

// Diagonal over op.k <= 2 qubits; row bit i (MSB first) comes from a lane
// bit, a slot or a global bit as op.loc[i] says.  The (at most four) entries
// are uniform and sit in scalar registers.
template <typename T, int RL, bool PR>
__device__ __forceinline__ void w_diag(REGS, const WOp& op, const KSV_CONST T* mats, bool lp, int lane,
                                       uint64_t base) {
    constexpr int R = 1 << RL;
    const int k = op.k;
    uint32_t rfix = 0;
    uint32_t sb0 = 0, sm0 = 0, sb1 = 0, sm1 = 0;
    int ns = 0;
#pragma unroll
    for (int i = 0; i < 2; i++) {
        if (i >= k) continue;
        const uint32_t rb = 1u << (k - 1 - i);
        const int L = op.loc[i];
        if (L < 8) { if ((lane >> L) & 1) rfix |= rb; }
        else if (L < 16) {
            if (ns == 0) { sb0 = 1u << (L - 8); sm0 = rb; ns = 1; }
            else { sb1 = 1u << (L - 8); sm1 = rb; }
        } else if ((base >> (L - 16)) & 1) rfix |= rb;
    }
    const KSV_CONST T* d = mats + 2 * op.mat;
    const CM<T> d0 = ldc(d, 0), d1 = k > 0 ? ldc(d, 1) : d0;
    const CM<T> d2 = k > 1 ? ldc(d, 2) : d0, d3 = k > 1 ? ldc(d, 3) : d0;
#pragma unroll
    for (int j = 0; j < R; j++) {
        if (((uint32_t)j & op.rctl) != op.rctl) continue;
        uint32_t row = rfix;
        if ((uint32_t)j & sb0) row |= sm0;
        if ((uint32_t)j & sb1) row |= sm1;
        const T dx = row == 0 ? d0.x : (row == 1 ? d1.x : (row == 2 ? d2.x : d3.x));
        const T dy = row == 0 ? d0.y : (row == 1 ? d1.y : (row == 2 ? d2.y : d3.y));
        T yr = dx * re[j] - dy * im[j];
        T yi = dx * im[j] + dy * re[j];
        prune2<T, PR>(yr, yi);
        re[j] = lp ? yr : re[j];
        im[j] = lp ? yi : im[j];
    }
}

#define SLOT_CASES(RL, SLOT, FN)                            \
    switch (SLOT) {                                         \
    case 0: FN(0); break;                                   \
    case 1: if constexpr (RL > 1) { FN(1); } break;         \
    case 2: if constexpr (RL > 2) { FN(2); } break;         \
    case 3: if constexpr (RL > 3) { FN(3); } break;         \
    case 4: if constexpr (RL > 4) { FN(4); } break;         \
    case 5: if constexpr (RL > 5) { FN(5); } break;         \
    default: break;                                         \
    }

#define PAIR_CASES(RL, A, B, FN)                                                    \
    do {                                                                            \
        const int ab_ = (A) * 8 + (B);                                              \
        switch (ab_) {                                                              \
        case 8: FN(1, 0); break;                                                    \
        case 16: FN(2, 0); break;                                                   \
        case 17: FN(2, 1); break;                                                   \
        case 24: if constexpr (RL > 3) { FN(3, 0); } break;                         \
        case 25: if constexpr (RL > 3) { FN(3, 1); } break;                         \
        case 26: if constexpr (RL > 3) { FN(3, 2); } break;                         \
        case 32: if constexpr (RL > 4) { FN(4, 0); } break;                         \
        case 33: if constexpr (RL > 4) { FN(4, 1); } break;                         \
        case 34: if constexpr (RL > 4) { FN(4, 2); } break;                         \
        case 35: if constexpr (RL > 4) { FN(4, 3); } break;                         \
        case 40: if constexpr (RL > 5) { FN(5, 0); } break;                         \
        case 41: if constexpr (RL > 5) { FN(5, 1); } break;                         \
        case 42: if constexpr (RL > 5) { FN(5, 2); } break;                         \
        case 43: if constexpr (RL > 5) { FN(5, 3); } break;                         \
        case 44: if constexpr (RL > 5) { FN(5, 4); } break;                         \
        default: break;                                                             \
        }                                                                           \
    } while (0)

template <typename T, int RL, bool PR>
__global__ __launch_bounds__(256) void k_wpass(WaveArgs a) {
    using V = typename Vec2<T>::type;
    constexpr int R = 1 << RL;
    const int lane = threadIdx.x & 63;
    V* psi = reinterpret_cast<V*>(a.psi);
    const KSV_CONST T* mats = to_const(reinterpret_cast<const T*>(a.mats));
    const KSV_CONST uint64_t* ops = to_const(reinterpret_cast<const uint64_t*>(a.ops));
    const uint64_t tile = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    uint64_t base = tile;
#pragma unroll
    for (int i = 0; i < 16; i++)
        if (i < a.k) base = ins0_64(base, byte_at(i < 8 ? a.tq_lo : a.tq_hi, i & 7));

    uint64_t lo = base;
#pragma unroll
    for (int b = 0; b < WLANE_BITS; b++) lo |= (uint64_t)((lane >> b) & 1) << byte_at(a.lane_in, b);
    T re[R], im[R];
#pragma unroll
    for (int j = 0; j < R; j++) {
        uint64_t ro = 0;
#pragma unroll
        for (int s = 0; s < RL; s++)
            if ((j >> s) & 1) ro |= 1ull << byte_at(a.reg_in, s);
        const V t = psi[lo | ro];
        re[j] = t.x;
        im[j] = t.y;
    }

    for (int o = 0; o < a.nops; o++) {
        WOp op;
        {
            const KSV_CONST uint64_t* w = ops + 4 * o;
            const uint64_t t[4] = {w[0], w[1], w[2], w[3]};
            __builtin_memcpy(&op, t, sizeof(WOp));
        }
        if ((base & op.gctl) != op.gctl) continue;
        const bool lp = ((uint32_t)lane & op.lctl) == op.lctl;
        const bool LPn = op.lctl != 0;
        switch (op.kind) {
        case W_M1:
#define F(S) (LPn ? w_m1<T, RL, S, true, PR>(re, im, op, mats, lp) : w_m1<T, RL, S, false, PR>(re, im, op, mats, lp))
            SLOT_CASES(RL, op.a, F)
#undef F
            break;
        case W_M2:
#define F(A, B) (LPn ? w_m2<T, RL, A, B, true, PR>(re, im, op, mats, lp) : w_m2<T, RL, A, B, false, PR>(re, im, op, mats, lp))
            PAIR_CASES(RL, op.a, op.b, F);
#undef F
            break;
        case W_PHASE:
            if (LPn) w_phase<T, RL, true, PR>(re, im, op, mats, lp);
            else w_phase<T, RL, false, PR>(re, im, op, mats, lp);
            break;
        case W_DIAG:
            w_diag<T, RL, PR>(re, im, op, mats, lp, lane, base);
            break;
        case W_PERMX:
#define F(S) (LPn ? w_permx<T, RL, S, true>(re, im, op, lp) : w_permx<T, RL, S, false>(re, im, op, lp))
            SLOT_CASES(RL, op.a, F)
#undef F
            break;
        case W_SWAPR:
#define F(A, B) (LPn ? w_swapr<T, RL, A, B, true>(re, im, op, lp) : w_swapr<T, RL, A, B, false>(re, im, op, lp))
            PAIR_CASES(RL, op.a, op.b, F);
#undef F
            break;
        case W_LSWAP:
#define F(S) w_lswap<T, RL, S>(re, im, op.a)
            SLOT_CASES(RL, op.b, F)
#undef F
            break;
        default: break;
        }
    }

    lo = base;
#pragma unroll
    for (int b = 0; b < WLANE_BITS; b++) lo |= (uint64_t)((lane >> b) & 1) << byte_at(a.lane_out, b);
#pragma unroll
    for (int j = 0; j < R; j++) {
        uint64_t ro = 0;
#pragma unroll
        for (int s = 0; s < RL; s++)
            if ((j >> s) & 1) ro |= 1ull << byte_at(a.reg_out, s);
        V t;
        t.x = re[j];
        t.y = im[j];
        psi[lo | ro] = t;
    }
}

template <typename T, int RL>
static void launch_rl(const WaveArgs& a, uint64_t tiles, int prune, hipStream_t st) {
    const unsigned blocks = (unsigned)(tiles / 4);
    if (prune) hipLaunchKernelGGL((k_wpass<T, RL, true>), dim3(blocks), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_wpass<T, RL, false>), dim3(blocks), dim3(256), 0, st, a);
}

hipError_t launch_wave(const Program& p, const Step& s, void* psi, hipStream_t st) {
    WaveArgs a;
    a.psi = psi;
    a.ops = p.d_wops + s.op_begin;
    a.mats = p.d_mats;
    a.nops = s.op_count;
    a.k = s.k;
    auto pack = [](const int8_t* b, int n) {
        uint64_t w = 0;
        for (int i = 0; i < n; i++) w |= (uint64_t)(uint8_t)b[i] << (8 * i);
        return w;
    };
    a.lane_in = pack(s.lane_in, 6);
    a.lane_out = pack(s.lane_out, 6);
    a.reg_in = pack(s.reg_in, 8);
    a.reg_out = pack(s.reg_out, 8);
    a.tq_lo = pack(s.tq, 8);
    a.tq_hi = pack(s.tq + 8, 8);
    const uint64_t tiles = 1ull << (p.n - s.k);
    if (p.precision == KSV_F64) {
        if (s.rlog == 4) launch_rl<double, 4>(a, tiles, p.prune, st);
        else launch_rl<double, 5>(a, tiles, p.prune, st);
    } else {
        if (s.rlog == 5) launch_rl<float, 5>(a, tiles, p.prune, st);
        else launch_rl<float, 6>(a, tiles, p.prune, st);
    }
    return hipGetLastError();
}

}  // namespace ksv
