// Device helpers shared by the pass kernels.
#pragma once

#include "ksv_internal.h"

namespace ksv {

// ket's arithmetic, rounded step by step.  JavaScript rounds every product
// and every sum on its own (add / mul / norm2, src/complex.ts:10-21); clang
// contracts a * b + c into one fused multiply-add by default, and
// __dmul_rn / __dadd_rn are plain * and + in this toolchain, contracted like
// any other expression.  An operation written under contract(off) carries no
// contraction flag, so it stays unfused after inlining into any kernel.
template <typename T>
__device__ __forceinline__ T rn_mul(T a, T b) {
#pragma clang fp contract(off)
    return a * b;
}
template <typename T>
__device__ __forceinline__ T rn_add(T a, T b) {
#pragma clang fp contract(off)
    return a + b;
}
template <typename T>
__device__ __forceinline__ T rn_sub(T a, T b) {
#pragma clang fp contract(off)
    return a - b;
}

// norm2 (src/complex.ts:21) in FP64: re * re + im * im, each step rounded
__device__ __forceinline__ double knorm2(double re, double im) { return rn_add(rn_mul(re, re), rn_mul(im, im)); }

// mul (src/complex.ts:13-14): (a.re b.re - a.im b.im, a.re b.im + a.im b.re)
template <typename T>
__device__ __forceinline__ void kmul(T ar, T ai, T br, T bi, T& re, T& im) {
    re = rn_sub(rn_mul(ar, br), rn_mul(ai, bi));
    im = rn_add(rn_mul(ar, bi), rn_mul(ai, br));
}

// out += m * x, as add(out, mul(m, x)) (applyTwo / applyUnitary's inner loop,
// src/statevector.ts:125-126, :263-264)
template <typename T>
__device__ __forceinline__ void kacc(T& ore, T& oim, T mr, T mi, T xr, T xi) {
    T pr, pi;
    kmul(mr, mi, xr, xi, pr, pi);
    ore = rn_add(ore, pr);
    oim = rn_add(oim, pi);
}

template <typename V>
__device__ __forceinline__ V cmul(V a, V b) {
    V r;
    kmul(a.x, a.y, b.x, b.y, r.x, r.y);
    return r;
}

// add(mul(m0, x0), mul(m1, x1)) (applySingle, src/statevector.ts:61-62)
template <typename V>
__device__ __forceinline__ V cmadd2(V m0, V x0, V m1, V x1) {
    V a = cmul(m0, x0), b = cmul(m1, x1), r;
    r.x = rn_add(a.x, b.x);
    r.y = rn_add(a.y, b.y);
    return r;
}

// ket's accumulate(): an amplitude with |a|^2 < 1e-14 becomes 0
// (isNegligible, src/complex.ts:24; src/statevector.ts:19-27)
template <typename V>
__device__ __forceinline__ V prune_amp(V v, int prune) {
    if (prune && knorm2((double)v.x, (double)v.y) < 1e-14) { v.x = 0; v.y = 0; }
    return v;
}

__device__ __forceinline__ uint32_t ins0(uint32_t p, int q) {
    return ((p >> q) << (q + 1)) | (p & ((1u << q) - 1));
}
__device__ __forceinline__ uint64_t ins0_64(uint64_t p, int q) {
    return ((p >> q) << (q + 1)) | (p & ((1ull << q) - 1));
}

template <typename V>
__device__ __forceinline__ V sel(bool c, V a, V b) {
    V r;
    r.x = c ? a.x : b.x;
    r.y = c ? a.y : b.y;
    return r;
}

__host__ __device__ __forceinline__ int byte_at(uint64_t w, int i) { return (int)((w >> (8 * i)) & 0xFF); }

// Pointers into the constant address space: uniform loads through them
// become scalar (s_load) loads into SGPRs.
#define KSV_CONST __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const KSV_CONST T* to_const(const T* p) {
    return (const KSV_CONST T*)(p);
}

// Sum of v over the 64 lanes of a wave, valid in lane 63, by DPP moves (VALU
// data paths between lanes) instead of LDS-crossbar permutes: within quads,
// half rows and rows, then row 0 into row 1, row 2 into row 3 (row_bcast15)
// and rows 0..1 into rows 2..3 (row_bcast31).  Lanes a DPP move does not
// write, or whose source lies outside the wave, take 0.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROWS, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}
__device__ __forceinline__ double wave_sum_last(double v) {
    v += dpp_f64<0xB1, 0xF>(v);    // quad_perm [1, 0, 3, 2]
    v += dpp_f64<0x4E, 0xF>(v);    // quad_perm [2, 3, 0, 1]
    v += dpp_f64<0x141, 0xF>(v);   // row_half_mirror
    v += dpp_f64<0x140, 0xF>(v);   // row_mirror
    v += dpp_f64<0x142, 0xA>(v);   // row_bcast15 into rows 1 and 3
    v += dpp_f64<0x143, 0xC>(v);   // row_bcast31 into rows 2 and 3
    return v;
}

}  // namespace ksv
