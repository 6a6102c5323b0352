// Batched trajectories: B shots of an n-qubit circuit held as one state of
// n + b qubits, copy t being amplitudes [t 2^n, (t + 1) 2^n).  Gates reach
// every copy through the ordinary pass kernels (they act on qubits < n);
// the kernels here do what differs per copy: the initial basis state, the
// Pauli errors of ket's depolarising noise (src/circuit.ts:289-304),
// measurement and reset (collapseQubit, src/circuit.ts:55-70) and the final
// draw (sampleSV, src/circuit.ts:74-81).  Each per-copy reduction repeats,
// copy by copy, the exact addition order of its single-state counterpart in
// k_reduce.hip, and the final draw is sampleSV's sequential sum itself, so a
// copy's outcome and draw equal those of the same shot simulated alone.
#include "kcommon.h"

namespace ksv {

template <typename V>
__device__ __forceinline__ double sprob(V a) {
    return knorm2((double)a.x, (double)a.y);     // each product rounded, as ket's norm2
}
__device__ __forceinline__ uint64_t sins0(uint64_t p, int q) { return ((p >> q) << (q + 1)) | (p & ((1ull << q) - 1)); }

static int sgrid(uint64_t work, int maxb) {
    uint64_t b = (work + 255) / 256;
    if (b < 1) b = 1;
    if (b > (uint64_t)maxb) b = maxb;
    return (int)b;
}

#define SDISPATCH(prec, ...)                                                 \
    do {                                                                     \
        if ((prec) == KSV_F64) { using V = double2; __VA_ARGS__; }           \
        else { using V = float2; __VA_ARGS__; }                              \
    } while (0)

// every copy := |idx>
template <typename V>
__global__ void k_shots_basis(V* psi, uint64_t total, int n, uint64_t idx) {
    const uint64_t mask = (1ull << n) - 1;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        V v;
        v.x = (i & mask) == idx ? 1 : 0;
        v.y = 0;
        psi[i] = v;
    }
}

// copy t := |t> (2^n copies): applying a circuit then leaves column t of
// its matrix in copy t
template <typename V>
__global__ void k_shots_identity(V* psi, uint64_t total, int n) {
    const uint64_t mask = (1ull << n) - 1;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
        V v;
        v.x = (i & mask) == (i >> n) ? 1 : 0;
        v.y = 0;
        psi[i] = v;
    }
}

// Pauli `pauli[e]` (1 X, 2 Y, 3 Z) on qubit q of copy `copy[e]`; blockIdx.y = e.
// The entries are exact sign flips and swaps, as ket's applySingle gives
// them for these matrices.
template <typename V>
__global__ void k_shots_pauli(V* psi, int n, int q, const uint32_t* copy, const uint8_t* pauli) {
    const uint64_t half = 1ull << (n - 1), mk = 1ull << q;
    V* c = psi + ((uint64_t)copy[blockIdx.y] << n);
    const int P = pauli[blockIdx.y];
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < half; p += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i0 = sins0(p, q), i1 = i0 | mk;
        V a = c[i0], b = c[i1];
        if (P == 1) {
            c[i0] = b;
            c[i1] = a;
        } else if (P == 2) {          // Y = [[0, -i], [i, 0]]
            V o0, o1;
            o0.x = b.y;
            o0.y = -b.x;
            o1.x = -a.y;
            o1.y = a.x;
            c[i0] = o0;
            c[i1] = o1;
        } else if (P == 3) {
            b.x = -b.x;
            b.y = -b.y;
            c[i1] = b;
        }
    }
}

// Per-copy partial sums of P(qubit q = 1): block (x, t) walks copy t with
// the single-state grid and stride (k_bit1 in k_reduce.hip).
template <typename V>
__global__ __launch_bounds__(256) void k_shots_bit1(const V* psi, int n, int q, double* part) {
    __shared__ double sh[4];
    const uint64_t half = 1ull << (n - 1), mk = 1ull << q;
    const V* c = psi + ((uint64_t)blockIdx.y << n);
    double s = 0;
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < half; p += (uint64_t)gridDim.x * blockDim.x)
        s += sprob(c[sins0(p, q) | mk]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sh[w] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0;
        for (int i = 0; i < 4; i++) r += sh[i];
        part[(size_t)blockIdx.x * gridDim.y + blockIdx.y] = r;
    }
}

// Keep the amplitudes of copy t whose qubit q equals outcome[t], scaled by
// scale[t]; with reset[t] and outcome 1 the kept half moves to q = 0 (the
// X that ket's reset applies after collapsing).
template <typename V>
__global__ void k_shots_project(V* psi, int n, int q, const uint8_t* outcome, const double* scale,
                                const uint8_t* reset, uint64_t total) {
    const uint64_t half_all = total >> 1, mk = 1ull << q;
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < half_all; p += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i0 = sins0(p, q), i1 = i0 | mk;
        const uint64_t t = i0 >> n;
        const int o = outcome[t];
        const double sc = scale[t];
        V keep = o ? psi[i1] : psi[i0];
        keep.x = (decltype(keep.x))((double)keep.x * sc);
        keep.y = (decltype(keep.y))((double)keep.y * sc);
        V zero;
        zero.x = 0;
        zero.y = 0;
        if (o && reset[t]) {
            psi[i0] = keep;
            psi[i1] = zero;
        } else if (o) {
            psi[i0] = zero;
            psi[i1] = keep;
        } else {
            psi[i0] = keep;
            psi[i1] = zero;
        }
    }
}

template <typename V>
__device__ __forceinline__ double spfilt(const V* c, uint64_t N, uint64_t i, double thresh) {
    if (i >= N) return 0.0;
    double p = sprob(c[i]);
    return p > thresh ? p : 0.0;
}

// One wave per copy: ket's sampleSV (src/circuit.ts:74-81) on the copy, i.e.
// cum += p over its entries (p > thresh) in ascending index, one rounding per
// addition, and the first entry whose running sum reaches r[t]; the last entry
// when none does, index 0 for a copy without entries.  64 coalesced loads per
// step, the step's additions lane by lane on a value every lane holds alike;
// the walk stops at the step that holds the draw.
__device__ __forceinline__ double s_lane_value(double v, int k) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, k);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), k);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

template <typename V>
__global__ __launch_bounds__(64) void k_shots_seq_sample(const V* psi, int n, double thresh, const double* rr,
                                                         uint64_t* out) {
    const uint64_t N = 1ull << n;
    const V* c = psi + ((uint64_t)blockIdx.x << n);
    const double r = rr[blockIdx.x];
    const int lane = threadIdx.x;
    double run = 0.0;
    unsigned long long last = ~0ull, hit = ~0ull;
    for (uint64_t b = 0; b < N && hit == ~0ull; b += 64) {
        const double p = spfilt(c, N, b + lane, thresh);
        const unsigned long long m = __ballot(p > 0.0);
        if (!m) continue;
        last = b + 63 - (unsigned long long)__clzll((long long)m);
#pragma unroll
        for (int k = 0; k < 64; k++) {
            run = rn_add(run, s_lane_value(p, k));
            if (hit == ~0ull && ((m >> k) & 1) && r <= run) hit = b + k;
        }
    }
    if (lane == 0) out[blockIdx.x] = hit != ~0ull ? hit : last != ~0ull ? last : 0;
}

// Sum the per-block partials of each copy in block order (k_final's order).
__global__ void k_shots_final(const double* part, int nblk, int copies, double* out) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < copies; t += gridDim.x * blockDim.x) {
        double s = 0;
        for (int b = 0; b < nblk; b++) s += part[(size_t)b * copies + t];
        out[t] = s;
    }
}

// ---- launchers -------------------------------------------------------------
hipError_t shots_basis(int ntot, int n, int prec, void* psi, uint64_t idx, hipStream_t st) {
    const uint64_t total = 1ull << ntot;
    SDISPATCH(prec, hipLaunchKernelGGL((k_shots_basis<V>), dim3(sgrid(total, 8192)), dim3(256), 0, st, (V*)psi,
                                       total, n, idx));
    return hipGetLastError();
}

hipError_t shots_identity(int ntot, int n, int prec, void* psi, hipStream_t st) {
    const uint64_t total = 1ull << ntot;
    SDISPATCH(prec, hipLaunchKernelGGL((k_shots_identity<V>), dim3(sgrid(total, 8192)), dim3(256), 0, st, (V*)psi,
                                       total, n));
    return hipGetLastError();
}

hipError_t shots_pauli(int n, int prec, void* psi, int q, const uint32_t* d_copy, const uint8_t* d_pauli,
                       uint32_t count, hipStream_t st) {
    const uint64_t half = 1ull << (n - 1);
    const dim3 grid((unsigned)sgrid(half, std::max<uint64_t>(1, 4096 / count + 1)), count);
    SDISPATCH(prec, hipLaunchKernelGGL((k_shots_pauli<V>), grid, dim3(256), 0, st, (V*)psi, n, q, d_copy, d_pauli));
    return hipGetLastError();
}

hipError_t shots_bit1(int n, int prec, const void* psi, int q, uint32_t copies, int nblk, double* d_part,
                      int* grid_x, hipStream_t st) {
    const int g = sgrid(1ull << (n - 1), nblk);
    *grid_x = g;
    SDISPATCH(prec, hipLaunchKernelGGL((k_shots_bit1<V>), dim3(g, copies), dim3(256), 0, st, (const V*)psi, n, q,
                                       d_part));
    return hipGetLastError();
}

hipError_t shots_final(const double* d_part, int nblk, uint32_t copies, double* d_out, hipStream_t st) {
    hipLaunchKernelGGL(k_shots_final, dim3((copies + 255) / 256), dim3(256), 0, st, d_part, nblk, (int)copies, d_out);
    return hipGetLastError();
}

hipError_t shots_project(int ntot, int n, int prec, void* psi, int q, const uint8_t* d_out, const double* d_scale,
                         const uint8_t* d_reset, hipStream_t st) {
    const uint64_t total = 1ull << ntot;
    SDISPATCH(prec, hipLaunchKernelGGL((k_shots_project<V>), dim3(sgrid(total >> 1, 8192)), dim3(256), 0, st,
                                       (V*)psi, n, q, d_out, d_scale, d_reset, total));
    return hipGetLastError();
}

hipError_t shots_seq_sample(int n, int prec, const void* psi, uint32_t copies, double thresh, const double* d_r,
                            uint64_t* d_out, hipStream_t st) {
    SDISPATCH(prec, hipLaunchKernelGGL((k_shots_seq_sample<V>), dim3(copies), dim3(64), 0, st, (const V*)psi, n,
                                       thresh, d_r, d_out));
    return hipGetLastError();
}

}  // namespace ksv
