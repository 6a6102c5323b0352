// Read-out kernels: FP64 reductions (norm, marginals, one-qubit rho, Pauli
// expectations, parity sums), basis init / projection, ordered compaction of
// probabilities, the chunked CDF used by shot sampling, and radix-select
// histograms for top-k.  Every sum accumulates in FP64 whatever the state
// precision.  Reductions use a fixed grid and a fixed combine order, so a
// result is bit-reproducible run to run.
#include <algorithm>

#include "kcommon.h"

namespace ksv {

template <typename V>
__device__ __forceinline__ double prob(V a) {
    return knorm2((double)a.x, (double)a.y);     // each product rounded, as ket's norm2
}
__device__ __forceinline__ uint64_t ins0_64r(uint64_t p, int q) {
    return ((p >> q) << (q + 1)) | (p & ((1ull << q) - 1));
}

// Sum over a 256-thread block; result valid in thread 0.
__device__ double block_sum(double v, double* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; i++) r += sh[i];
    }
    __syncthreads();
    return r;
}

// Sum `m` partial vectors (nblk x m, row per block) in block order.
__global__ void k_final(const double* part, int nblk, int m, double* out) {
    for (int j = threadIdx.x; j < m; j += blockDim.x) {
        double s = 0;
        for (int b = 0; b < nblk; b++) s += part[(size_t)b * m + j];
        out[j] = s;
    }
}

template <typename V>
__global__ __launch_bounds__(256) void k_norm2(const V* psi, uint64_t N, double* part) {
    __shared__ double sh[4];
    double s = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x)
        s += prob(psi[i]);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

template <typename V>
__global__ __launch_bounds__(256) void k_marg(const V* psi, int n, uint64_t N, double* part) {
    __shared__ double sh[4];
    double acc[40];
#pragma unroll
    for (int q = 0; q < 40; q++) acc[q] = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x) {
        double p = prob(psi[i]);
#pragma unroll
        for (int q = 0; q < 40; q++)
            if (q < n && ((i >> q) & 1)) acc[q] += p;
    }
#pragma unroll
    for (int q = 0; q < 40; q++) {
        if (q < n) {
            double r = block_sum(acc[q], sh);
            if (threadIdx.x == 0) part[(size_t)blockIdx.x * n + q] = r;
        }
    }
}

template <typename V>
__global__ __launch_bounds__(256) void k_1q(const V* psi, int q, uint64_t half, double* part) {
    __shared__ double sh[4];
    double r00 = 0, r11 = 0, re = 0, im = 0;
    const uint64_t mk = 1ull << q;
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < half; p += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t i0 = ins0_64r(p, q);
        V a = psi[i0], b = psi[i0 | mk];
        double ar = a.x, ai = a.y, br = b.x, bi = b.y;
        r00 += ar * ar + ai * ai;
        r11 += br * br + bi * bi;
        re += ar * br + ai * bi;
        im += ai * br - ar * bi;
    }
    double v[4] = {r00, r11, re, im};
#pragma unroll
    for (int j = 0; j < 4; j++) {
        double r = block_sum(v[j], sh);
        if (threadIdx.x == 0) part[(size_t)blockIdx.x * 4 + j] = r;
    }
}

// Up to 8 Pauli terms per launch: sum_j conj(psi[j ^ x]) (-1)^{|j & z|} psi[j]
// (re, im) per term.
template <typename V>
__global__ __launch_bounds__(256) void k_pauli(const V* psi, uint64_t N, const uint64_t* xm, const uint64_t* zm,
                                               int nt, double* part) {
    __shared__ double sh[4];
    double are[8], aim[8];
    uint64_t x[8], z[8];
#pragma unroll
    for (int t = 0; t < 8; t++) {
        are[t] = aim[t] = 0;
        x[t] = t < nt ? xm[t] : 0;
        z[t] = t < nt ? zm[t] : 0;
    }
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < N; j += (uint64_t)gridDim.x * blockDim.x) {
        V a = psi[j];
        double ar = a.x, ai = a.y;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            if (t < nt) {
                V b = psi[j ^ x[t]];
                double br = b.x, bi = b.y;
                double sg = (__popcll(j & z[t]) & 1) ? -1.0 : 1.0;
                are[t] += sg * (br * ar + bi * ai);
                aim[t] += sg * (br * ai - bi * ar);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 8; t++) {
        if (t < nt) {
            double r1 = block_sum(are[t], sh);
            double r2 = block_sum(aim[t], sh);
            if (threadIdx.x == 0) {
                part[(size_t)blockIdx.x * 16 + 2 * t] = r1;
                part[(size_t)blockIdx.x * 16 + 2 * t + 1] = r2;
            }
        }
    }
}

// sum over p > thresh of (-1)^{|i & mask|} p   (ket's parity read-out)
template <typename V>
__global__ __launch_bounds__(256) void k_parity(const V* psi, uint64_t N, uint64_t mask, double thresh, double* part) {
    __shared__ double sh[4];
    double s = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x) {
        double p = prob(psi[i]);
        if (p > thresh) s += (__popcll(i & mask) & 1) ? -p : p;
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

template <typename V>
__global__ __launch_bounds__(256) void k_bit1(const V* psi, int q, uint64_t half, double* part) {
    __shared__ double sh[4];
    double s = 0;
    const uint64_t mk = 1ull << q;
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < half; p += (uint64_t)gridDim.x * blockDim.x)
        s += prob(psi[ins0_64r(p, q) | mk]);
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

template <typename V>
__global__ void k_basis(V* psi, uint64_t N, uint64_t idx) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x) {
        V v;
        v.x = i == idx ? 1 : 0;
        v.y = 0;
        psi[i] = v;
    }
}

template <typename V>
__global__ void k_proj(V* psi, uint64_t N, int q, int outcome, double scale) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x) {
        V v = psi[i];
        if ((int)((i >> q) & 1) == outcome) {
            v.x = (decltype(v.x))((double)v.x * scale);
            v.y = (decltype(v.y))((double)v.y * scale);
        } else {
            v.x = 0;
            v.y = 0;
        }
        psi[i] = v;
    }
}

static int grid_for(uint64_t work, int maxb) {
    uint64_t b = (work + 255) / 256;
    if (b < 1) b = 1;
    if (b > (uint64_t)maxb) b = maxb;
    return (int)b;
}

#define DISPATCH(prec, ...)                                                  \
    do {                                                                     \
        if ((prec) == KSV_F64) { using V = double2; __VA_ARGS__; }           \
        else { using V = float2; __VA_ARGS__; }                              \
    } while (0)

hipError_t red_norm2(int n, int prec, const void* psi, double* d_out, double* d_part, int nblk, hipStream_t st) {
    uint64_t N = 1ull << n;
    int g = grid_for(N, nblk);
    DISPATCH(prec, hipLaunchKernelGGL((k_norm2<V>), dim3(g), dim3(256), 0, st, (const V*)psi, N, d_part));
    hipLaunchKernelGGL(k_final, dim3(1), dim3(64), 0, st, d_part, g, 1, d_out);
    return hipGetLastError();
}

hipError_t red_marginals(int n, int prec, const void* psi, double* d_out, double* d_part, int nblk, hipStream_t st) {
    uint64_t N = 1ull << n;
    int g = grid_for(N, nblk);
    DISPATCH(prec, hipLaunchKernelGGL((k_marg<V>), dim3(g), dim3(256), 0, st, (const V*)psi, n, N, d_part));
    hipLaunchKernelGGL(k_final, dim3(1), dim3(64), 0, st, d_part, g, n, d_out);
    return hipGetLastError();
}

hipError_t red_1q(int n, int prec, const void* psi, int q, double* d_out, double* d_part, int nblk, hipStream_t st) {
    uint64_t half = 1ull << (n - 1);
    int g = grid_for(half, nblk);
    DISPATCH(prec, hipLaunchKernelGGL((k_1q<V>), dim3(g), dim3(256), 0, st, (const V*)psi, q, half, d_part));
    hipLaunchKernelGGL(k_final, dim3(1), dim3(64), 0, st, d_part, g, 4, d_out);
    return hipGetLastError();
}

hipError_t red_pauli(int n, int prec, const void* psi, const uint64_t* d_x, const uint64_t* d_z, int nterms,
                     double* d_out, double* d_part, int nblk, hipStream_t st) {
    uint64_t N = 1ull << n;
    int g = grid_for(N, nblk);
    DISPATCH(prec, hipLaunchKernelGGL((k_pauli<V>), dim3(g), dim3(256), 0, st, (const V*)psi, N, d_x, d_z, nterms, d_part));
    hipLaunchKernelGGL(k_final, dim3(1), dim3(64), 0, st, d_part, g, 16, d_out);
    return hipGetLastError();
}

hipError_t red_parity(int n, int prec, const void* psi, uint64_t mask, double thresh, double* d_out, double* d_part,
                      int nblk, hipStream_t st) {
    uint64_t N = 1ull << n;
    int g = grid_for(N, nblk);
    DISPATCH(prec, hipLaunchKernelGGL((k_parity<V>), dim3(g), dim3(256), 0, st, (const V*)psi, N, mask, thresh, d_part));
    hipLaunchKernelGGL(k_final, dim3(1), dim3(64), 0, st, d_part, g, 1, d_out);
    return hipGetLastError();
}

hipError_t red_bit1(int n, int prec, const void* psi, int q, double* d_out, double* d_part, int nblk, hipStream_t st) {
    uint64_t half = 1ull << (n - 1);
    int g = grid_for(half, nblk);
    DISPATCH(prec, hipLaunchKernelGGL((k_bit1<V>), dim3(g), dim3(256), 0, st, (const V*)psi, q, half, d_part));
    hipLaunchKernelGGL(k_final, dim3(1), dim3(64), 0, st, d_part, g, 1, d_out);
    return hipGetLastError();
}

hipError_t k_init_basis(int n, int prec, void* psi, uint64_t idx, hipStream_t st) {
    uint64_t N = 1ull << n;
    int g = grid_for(N, 8192);
    DISPATCH(prec, hipLaunchKernelGGL((k_basis<V>), dim3(g), dim3(256), 0, st, (V*)psi, N, idx));
    return hipGetLastError();
}

hipError_t k_project(int n, int prec, void* psi, int q, int outcome, double scale, hipStream_t st) {
    uint64_t N = 1ull << n;
    int g = grid_for(N, 8192);
    DISPATCH(prec, hipLaunchKernelGGL((k_proj<V>), dim3(g), dim3(256), 0, st, (V*)psi, N, q, outcome, scale));
    return hipGetLastError();
}

// Device-to-device state copy: each thread moves four 16-byte words per step
// (all four loads issued before the stores), consecutive threads take
// consecutive words, each workgroup walks its own contiguous span.  Its
// read + write rate is the streaming ceiling the pass kernels are measured
// against (a pass also reads and writes every amplitude once).
__global__ __launch_bounds__(256) void k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                    uint64_t words, uint64_t span) {
    const uint64_t b0 = (uint64_t)blockIdx.x * span;
    const uint64_t b1 = b0 + span < words ? b0 + span : words;
    uint64_t i = b0 + threadIdx.x;
    for (; i + 768 < b1; i += 1024) {
        const uint4 a = src[i], b = src[i + 256], c = src[i + 512], d = src[i + 768];
        dst[i] = a;
        dst[i + 256] = b;
        dst[i + 512] = c;
        dst[i + 768] = d;
    }
    for (; i < b1; i += 256) dst[i] = src[i];
}

hipError_t k_copy(int n, int prec, void* dst, const void* src, hipStream_t st) {
    const uint64_t words = ((uint64_t)(prec == KSV_F64 ? 16 : 8) << n) / 16;   // n >= 1: at least one word
    const uint64_t blocks = std::min<uint64_t>(8192, (words + 1023) / 1024);
    const uint64_t span = ((words + blocks - 1) / blocks + 1023) / 1024 * 1024;
    hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)((words + span - 1) / span)), dim3(256), 0, st,
                       (const uint4*)src, (uint4*)dst, words, span);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Chunked scans.  A chunk is 2^CHUNK_LOG consecutive amplitudes handled by
// one 256-thread block: wave w owns the contiguous quarter [w*Q, (w+1)*Q),
// walked as 64-lane segments in index order.  The chunk-local inclusive
// prefix of p is defined by one fixed sequence of FP64 additions
// (chunk_prefix below); the sums kernel and the search kernel both evaluate
// it, so a chunk total equals the search's last prefix bit for bit and the
// global CDF (host-scanned chunk totals + local prefix) is monotone.
// ---------------------------------------------------------------------------
constexpr int CH = 1 << CHUNK_LOG;
constexpr int QW = CH / 4;          // amplitudes per wave
constexpr int SEGS = QW / 64;       // segments per wave

template <typename V>
__device__ __forceinline__ double pfilt(const V* psi, uint64_t N, uint64_t i, double thresh) {
    if (i >= N) return 0.0;
    double p = prob(psi[i]);
    return p > thresh ? p : 0.0;
}

__device__ __forceinline__ double wave_incl_scan(double v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        double t = __shfl_up(v, o, 64);
        if (lane >= o) v = __dadd_rn(v, t);
    }
    return v;
}

// Per-wave running offsets `run[s]` before segment s and the wave totals.
// sh_wt[4] receives the wave totals, sh_woff[4] the wave offsets.
template <typename V>
__device__ void chunk_wave_totals(const V* psi, uint64_t N, uint64_t c0, double thresh, double* sh_wt,
                                  double* sh_woff) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double run = 0;
    for (int s = 0; s < SEGS; s++) {
        double v = wave_incl_scan(pfilt(psi, N, c0 + (uint64_t)w * QW + s * 64 + lane, thresh));
        run = __dadd_rn(run, __shfl(v, 63, 64));
    }
    if (lane == 0) sh_wt[w] = run;
    __syncthreads();
    if (threadIdx.x == 0) {
        double o = 0;
        for (int i = 0; i < 4; i++) { sh_woff[i] = o; o = __dadd_rn(o, sh_wt[i]); }
    }
    __syncthreads();
}

template <typename V>
__global__ __launch_bounds__(256) void k_chunk_sums(const V* psi, uint64_t N, double thresh, double* sums) {
    __shared__ double wt[4], woff[4];
    const uint64_t c0 = (uint64_t)blockIdx.x * CH;
    chunk_wave_totals(psi, N, c0, thresh, wt, woff);
    if (threadIdx.x == 0) sums[blockIdx.x] = __dadd_rn(woff[3], wt[3]);
}

// One block per shot: pick the chunk by binary search over the inclusive
// chunk CDF, then the first amplitude whose global CDF value reaches r.
template <typename V>
__global__ __launch_bounds__(256) void k_sample(const V* psi, uint64_t N, double thresh, const double* ccum,
                                                int64_t nch, const double* rr, uint64_t last_idx, uint64_t* out) {
    __shared__ double wt[4], woff[4];
    __shared__ unsigned long long best;
    const double r = rr[blockIdx.x];
    // first chunk with ccum[c] >= r  (ccum = inclusive chunk CDF)
    int64_t lo = 0, hi = nch;   // search in [0, nch); nch means "beyond"
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (ccum[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    if (lo >= nch) {
        if (threadIdx.x == 0) out[blockIdx.x] = last_idx;
        return;
    }
    const int64_t c = lo;
    const double cbase = c > 0 ? ccum[c - 1] : 0.0;
    const uint64_t c0 = (uint64_t)c * CH;
    if (threadIdx.x == 0) best = ~0ull;
    chunk_wave_totals(psi, N, c0, thresh, wt, woff);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double run = 0;
    for (int s = 0; s < SEGS; s++) {
        uint64_t i = c0 + (uint64_t)w * QW + s * 64 + lane;
        double v = wave_incl_scan(pfilt(psi, N, i, thresh));
        double cum = __dadd_rn(cbase, __dadd_rn(woff[w], __dadd_rn(run, v)));
        if (cum >= r && i < N) atomicMin(&best, (unsigned long long)i);
        run = __dadd_rn(run, __shfl(v, 63, 64));
    }
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = best == ~0ull ? last_idx : best;
}

// ---------------------------------------------------------------------------
// ket's CDF in ket's order (run()'s fast path, src/circuit.ts:3427-3436):
// cum += p over the entries with p > thresh in ascending index, each addition
// rounded on its own.  The additions are one dependent chain, so one wave
// walks the state: 64 coalesced loads per step, then that step's 64
// additions lane by lane on a value every lane holds alike.  cum[i] is the
// JavaScript loop's running sum after index i (unchanged over zero entries),
// so a draw's index is ket's whatever its distance to a CDF boundary.
// ends[0], ends[1]: first and last index with p > thresh (~0 when none).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double lane_value(double v, int k) {
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, k);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), k);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

template <typename V>
__global__ __launch_bounds__(64) void k_seq_cdf(const V* psi, uint64_t N, double thresh, double* cum,
                                                unsigned long long* ends) {
    const int lane = threadIdx.x;
    double run = 0.0;
    unsigned long long first = ~0ull, last = ~0ull;
    for (uint64_t b = 0; b < N; b += 64) {
        const uint64_t i = b + lane;
        const double p = pfilt(psi, N, i, thresh);
        double mine = 0.0;
#pragma unroll
        for (int k = 0; k < 64; k++) {
            run = rn_add(run, lane_value(p, k));
            mine = lane == k ? run : mine;
        }
        const unsigned long long m = __ballot(p > 0.0);
        if (m) {
            if (first == ~0ull) first = b + (unsigned long long)(__ffsll((long long)m) - 1);
            last = b + 63 - (unsigned long long)__clzll((long long)m);
        }
        if (i < N) cum[i] = mine;
    }
    if (lane == 0) {
        ends[0] = first;
        ends[1] = last;
    }
}

// One thread per draw: ket's binary search over its CDF (first entry with
// cumP >= r, the last entry's cumP set to 1.0), on the dense running sums of
// k_seq_cdf.  The sum only grows at entries, so the first dense index whose
// sum reaches r > 0 is an entry; the search stops below the last entry, which
// takes every draw the sums leave (ket's 1.0).  Draws r <= 0 take the first
// entry, a state without entries gives index 0 (ket's `?? 0n`).
__global__ void k_cdf_search(const double* cum, const unsigned long long* ends, const double* rr, uint64_t shots,
                             uint64_t* out) {
    const uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (s >= shots) return;
    const unsigned long long first = ends[0], last = ends[1];
    if (first == ~0ull) {
        out[s] = 0;
        return;
    }
    const double r = rr[s];
    if (!(r > 0.0)) {
        out[s] = first;
        return;
    }
    uint64_t lo = first, hi = last;
    while (lo < hi) {
        const uint64_t mid = lo + ((hi - lo) >> 1);
        if (cum[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    out[s] = lo;
}

// Predicate of the ordered compaction: mode 0 p > thresh, mode 1 key == kv,
// mode 2 key > kv (key = bit pattern of p, monotone for p >= 0).
struct Pred {
    int mode;
    double thresh;
    uint64_t kv;
};
template <typename V>
__device__ __forceinline__ bool pred_ok(const V* psi, uint64_t N, uint64_t i, const Pred& pr, double& p) {
    if (i >= N) return false;
    p = prob(psi[i]);
    if (pr.mode == 0) return p > pr.thresh;
    uint64_t key = (uint64_t)__double_as_longlong(p);
    return pr.mode == 1 ? key == pr.kv : key > pr.kv;
}

template <typename V>
__global__ __launch_bounds__(256) void k_count(const V* psi, uint64_t N, Pred pr, uint32_t* cnt) {
    __shared__ uint32_t sh[4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t c0 = (uint64_t)blockIdx.x * CH;
    uint32_t c = 0;
    for (int s = 0; s < SEGS; s++) {
        double p;
        bool ok = pred_ok(psi, N, c0 + (uint64_t)w * QW + s * 64 + lane, pr, p);
        c += __popcll(__ballot(ok));
    }
    if (lane == 0) sh[w] = c;
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

template <typename V>
__global__ __launch_bounds__(256) void k_compact(const V* psi, uint64_t N, Pred pr, const uint64_t* off,
                                                 uint64_t* oidx, double* op, uint64_t cap) {
    __shared__ uint32_t sh[4];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t c0 = (uint64_t)blockIdx.x * CH;
    uint32_t c = 0;
    for (int s = 0; s < SEGS; s++) {
        double p;
        bool ok = pred_ok(psi, N, c0 + (uint64_t)w * QW + s * 64 + lane, pr, p);
        c += __popcll(__ballot(ok));
    }
    if (lane == 0) sh[w] = c;
    __syncthreads();
    uint64_t pos = off[blockIdx.x];
    for (int i = 0; i < w; i++) pos += sh[i];
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int s = 0; s < SEGS; s++) {
        uint64_t i = c0 + (uint64_t)w * QW + s * 64 + lane;
        double p = 0;
        bool ok = pred_ok(psi, N, i, pr, p);
        uint64_t m = __ballot(ok);
        if (ok) {
            uint64_t at = pos + __popcll(m & lt);
            if (at < cap) {
                if (oidx) oidx[at] = i;
                if (op) op[at] = p;
            }
        }
        pos += __popcll(m);
    }
}

// 12-bit radix histogram of the keys whose bits above `shift_hi` equal `prefix`.
template <typename V>
__global__ __launch_bounds__(256) void k_hist(const V* psi, uint64_t N, uint64_t prefix, int shift_hi, int shift,
                                              uint32_t* hist) {
    __shared__ uint32_t h[4096];
    for (int i = threadIdx.x; i < 4096; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const uint64_t dmask = shift_hi - shift >= 12 ? 0xFFFull : ((1ull << (shift_hi - shift)) - 1);
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < N; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t key = (uint64_t)__double_as_longlong(prob(psi[i]));
        if (shift_hi >= 64 || (key >> shift_hi) == prefix) atomicAdd(&h[(key >> shift) & dmask], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 4096; i += blockDim.x)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

template <typename V>
__global__ __launch_bounds__(256) void k_lastnz(const V* psi, uint64_t N, double thresh, uint64_t c0,
                                                unsigned long long* out) {
    for (uint64_t i = c0 + threadIdx.x; i < c0 + CH && i < N; i += blockDim.x) {
        double p = prob(psi[i]);
        if (p > thresh) atomicMax(out, (unsigned long long)i);
    }
}

static uint64_t nchunks(int n) { return ((1ull << n) + CH - 1) / CH; }

hipError_t chunk_counts(int n, int prec, const void* psi, double thresh, uint32_t* d_cnt, hipStream_t st) {
    Pred pr{0, thresh, 0};
    DISPATCH(prec, hipLaunchKernelGGL((k_count<V>), dim3((unsigned)nchunks(n)), dim3(256), 0, st, (const V*)psi,
                                      1ull << n, pr, d_cnt));
    return hipGetLastError();
}

hipError_t chunk_compact(int n, int prec, const void* psi, double thresh, const uint64_t* d_off, uint64_t* d_idx,
                         double* d_p, uint64_t cap, hipStream_t st) {
    Pred pr{0, thresh, 0};
    DISPATCH(prec, hipLaunchKernelGGL((k_compact<V>), dim3((unsigned)nchunks(n)), dim3(256), 0, st, (const V*)psi,
                                      1ull << n, pr, d_off, d_idx, d_p, cap));
    return hipGetLastError();
}

hipError_t chunk_counts_pred(int n, int prec, const void* psi, int mode, uint64_t kv, uint32_t* d_cnt,
                             hipStream_t st) {
    Pred pr{mode, 0, kv};
    DISPATCH(prec, hipLaunchKernelGGL((k_count<V>), dim3((unsigned)nchunks(n)), dim3(256), 0, st, (const V*)psi,
                                      1ull << n, pr, d_cnt));
    return hipGetLastError();
}

hipError_t chunk_compact_pred(int n, int prec, const void* psi, int mode, uint64_t kv, const uint64_t* d_off,
                              uint64_t* d_idx, double* d_p, uint64_t cap, hipStream_t st) {
    Pred pr{mode, 0, kv};
    DISPATCH(prec, hipLaunchKernelGGL((k_compact<V>), dim3((unsigned)nchunks(n)), dim3(256), 0, st, (const V*)psi,
                                      1ull << n, pr, d_off, d_idx, d_p, cap));
    return hipGetLastError();
}

hipError_t chunk_sums(int n, int prec, const void* psi, double thresh, double* d_sum, hipStream_t st) {
    DISPATCH(prec, hipLaunchKernelGGL((k_chunk_sums<V>), dim3((unsigned)nchunks(n)), dim3(256), 0, st,
                                      (const V*)psi, 1ull << n, thresh, d_sum));
    return hipGetLastError();
}

hipError_t seq_cdf(int n, int prec, const void* psi, double thresh, double* d_cum, unsigned long long* d_ends,
                   hipStream_t st) {
    DISPATCH(prec, hipLaunchKernelGGL((k_seq_cdf<V>), dim3(1), dim3(64), 0, st, (const V*)psi, 1ull << n, thresh,
                                      d_cum, d_ends));
    return hipGetLastError();
}

hipError_t cdf_search(const double* d_cum, const unsigned long long* d_ends, const double* d_r, uint64_t shots,
                      uint64_t* d_out, hipStream_t st) {
    hipLaunchKernelGGL(k_cdf_search, dim3((unsigned)((shots + 255) / 256)), dim3(256), 0, st, d_cum, d_ends, d_r,
                       shots, d_out);
    return hipGetLastError();
}

hipError_t sample_search(int n, int prec, const void* psi, double thresh, const double* d_chunk_cum, int64_t nch,
                         const double* d_r, uint64_t shots, uint64_t last_idx, uint64_t* d_out, hipStream_t st) {
    DISPATCH(prec, hipLaunchKernelGGL((k_sample<V>), dim3((unsigned)shots), dim3(256), 0, st, (const V*)psi,
                                      1ull << n, thresh, d_chunk_cum, nch, d_r, last_idx, d_out));
    return hipGetLastError();
}

hipError_t key_histogram(int n, int prec, const void* psi, uint64_t prefix, int shift_hi, int shift, uint32_t* d_hist,
                         hipStream_t st) {
    uint64_t N = 1ull << n;
    int g = grid_for(N, 2048);
    DISPATCH(prec, hipLaunchKernelGGL((k_hist<V>), dim3(g), dim3(256), 0, st, (const V*)psi, N, prefix, shift_hi,
                                      shift, d_hist));
    return hipGetLastError();
}

hipError_t last_nonzero_chunk(int n, int prec, const void* psi, double thresh, int64_t chunk, uint64_t* d_out,
                              hipStream_t st) {
    DISPATCH(prec, hipLaunchKernelGGL((k_lastnz<V>), dim3(1), dim3(256), 0, st, (const V*)psi, 1ull << n, thresh,
                                      (uint64_t)chunk * CH, (unsigned long long*)d_out));
    return hipGetLastError();
}

}  // namespace ksv
