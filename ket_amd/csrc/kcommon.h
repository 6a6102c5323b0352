// Device helpers shared by the pass kernels.
#pragma once

#include "ksv_internal.h"

namespace ksv {

template <typename V>
__device__ __forceinline__ V cmul(V a, V b) {
    V r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

// m0 * x0 + m1 * x1 (complex)
template <typename V>
__device__ __forceinline__ V cmadd2(V m0, V x0, V m1, V x1) {
    V r;
    r.x = (m0.x * x0.x - m0.y * x0.y) + (m1.x * x1.x - m1.y * x1.y);
    r.y = (m0.x * x0.y + m0.y * x0.x) + (m1.x * x1.y + m1.y * x1.x);
    return r;
}

// ket's accumulate(): an amplitude with |a|^2 < 1e-14 becomes 0
template <typename V>
__device__ __forceinline__ V prune_amp(V v, int prune) {
    if (prune && ((double)v.x * v.x + (double)v.y * v.y) < 1e-14) { v.x = 0; v.y = 0; }
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

}  // namespace ksv
