// C ABI of libketsv: state handles, programs, read-out and measurement.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "ksv_internal.h"

namespace ksv {
hipError_t chunk_counts_pred(int n, int prec, const void* psi, int mode, uint64_t kv, uint32_t* d_cnt,
                             hipStream_t st);
hipError_t chunk_compact_pred(int n, int prec, const void* psi, int mode, uint64_t kv, const uint64_t* d_off,
                              uint64_t* d_idx, double* d_p, uint64_t cap, hipStream_t st);
}  // namespace ksv

using namespace ksv;

static thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(KSV_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int NBLK = 1024;          // reduction grid
constexpr int MAXN = 34;            // 2^34 complex64 = 128 GiB

struct ksv_state {
    int n = 0;
    int prec = KSV_F64;
    int dev = 0;
    hipStream_t st = nullptr;
    void* psi = nullptr;
    void* psi2 = nullptr;           // second state buffer, allocated on first use
    void* ws = nullptr;             // read-out scratch: grows on demand, freed with the state
    size_t ws_bytes = 0;
    double* d_part = nullptr;       // NBLK x 64 partials
    double* d_small = nullptr;      // 64 results
    int prune = 1;
    int tile = 0;                   // 0 = default
    int timing = 0;
    int fuse = 0;
    int kernel = 0;                 // KSV_OPT_PASS_KERNEL
    int rtile = 0;                  // KSV_OPT_ROUND_TILE
    int bq = 0;                     // KSV_OPT_BLOCK_QUBITS
    int btile = 0;                  // KSV_OPT_BLOCK_TILE
    int bthreads = 256;             // KSV_OPT_BLOCK_THREADS
    int bpipe = 1;                  // KSV_OPT_BLOCK_PIPE
    int atile = 0;                  // KSV_OPT_ADJ_TILE
    int athreads = 256;             // KSV_OPT_ADJ_THREADS
    ksv_stats stats;
    std::vector<hipEvent_t> ev;
};

struct ksv_program : public Program {
    void* d_blob = nullptr;     // one device allocation holding every array below (owned)
    bool borrowed = false;      // arrays live in a caller's blob: nothing to free
};

static size_t amp_bytes(const ksv_state* s) { return s->prec == KSV_F64 ? 16 : 8; }

// Read-out scratch.  A call adds up what it needs (wsz), reserves it in the
// state's workspace, and carves 256-byte-aligned pieces from it (Carve).  The
// workspace only grows and stays with the state, so a read-out allocates
// nothing once it is large enough, and an error return leaves nothing behind.
static size_t wsz(size_t bytes) { return (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255; }

static int ws_reserve(ksv_state* s, size_t bytes) {
    if (bytes <= s->ws_bytes) return KSV_OK;
    hipError_t e = hipStreamSynchronize(s->st);       // the old workspace may still be in use
    if (e != hipSuccess) return fail(KSV_ERR_DEVICE, std::string("read-out scratch: ") + hipGetErrorString(e));
    if (s->ws) (void)hipFree(s->ws);
    s->ws = nullptr;
    s->ws_bytes = 0;
    const size_t want = std::max<size_t>(bytes, (size_t)1 << 20);
    if (hipMalloc(&s->ws, want) != hipSuccess) {
        s->ws = nullptr;
        return fail(KSV_ERR_DEVICE, "out of device memory for read-out scratch (" + std::to_string(want) + " bytes)");
    }
    s->ws_bytes = want;
    return KSV_OK;
}

struct Carve {
    char* base;
    size_t off = 0;
    template <typename T> T* take(size_t count) {
        T* p = reinterpret_cast<T*>(base + off);
        off += wsz(count * sizeof(T));
        return p;
    }
};
static size_t state_bytes(const ksv_state* s) { return amp_bytes(s) << s->n; }

static int default_tile(const ksv_state* s) {
    if (s->tile > 0) return s->tile;
    return s->prec == KSV_F64 ? 12 : 13;
}

extern "C" {

int ksv_last_error(char* buf, size_t len) {
    if (buf && len) {
        size_t m = std::min(len - 1, g_err.size());
        std::memcpy(buf, g_err.data(), m);
        buf[m] = 0;
    }
    return (int)g_err.size();
}

int ksv_device_count(int* count) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    if (count) *count = c;
    return KSV_OK;
}

int ksv_create(int n, int precision, ksv_state** out) {
    if (!out) return fail(KSV_ERR_TYPE, "ksv_create: null output pointer");
    *out = nullptr;
    if (n < 1 || n > MAXN) return fail(KSV_ERR_RANGE, "qubit count " + std::to_string(n) + " is out of range (1.." + std::to_string(MAXN) + ")");
    if (precision != KSV_F64 && precision != KSV_F32) return fail(KSV_ERR_TYPE, "unknown precision");
    int cnt = 0;
    if (hipGetDeviceCount(&cnt) != hipSuccess || cnt < 1) return fail(KSV_ERR_DEVICE, "no HIP device available");
    std::unique_ptr<ksv_state> s(new ksv_state());
    s->n = n;
    s->prec = precision;
    std::memset(&s->stats, 0, sizeof s->stats);
    HIPCHK(hipGetDevice(&s->dev));
    HIPCHK(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
    hipError_t e = hipMalloc(&s->psi, state_bytes(s.get()));
    const bool oom = e != hipSuccess;
    if (e == hipSuccess) e = hipMalloc(&s->d_part, sizeof(double) * NBLK * 64);
    if (e == hipSuccess) e = hipMalloc(&s->d_small, sizeof(double) * 64);
    if (e == hipSuccess) e = k_init_basis(n, precision, s->psi, 0, s->st);
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    if (e != hipSuccess) {      // release what was allocated
        (void)hipFree(s->psi);
        (void)hipFree(s->d_part);
        (void)hipFree(s->d_small);
        (void)hipStreamDestroy(s->st);
        if (oom) return fail(KSV_ERR_DEVICE, "out of device memory allocating a " + std::to_string(n) + "-qubit state");
        return fail(KSV_ERR_DEVICE, std::string("state set-up: ") + hipGetErrorString(e));
    }
    *out = s.release();
    return KSV_OK;
}

int ksv_destroy(ksv_state* s) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    (void)hipStreamSynchronize(s->st);
    (void)hipFree(s->psi);
    if (s->psi2) (void)hipFree(s->psi2);
    if (s->ws) (void)hipFree(s->ws);
    (void)hipFree(s->d_part);
    (void)hipFree(s->d_small);
    for (auto e : s->ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(s->st);
    delete s;
    return KSV_OK;
}

int ksv_set_option(ksv_state* s, int option, double value) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    switch (option) {
    case KSV_OPT_PRUNE: s->prune = value != 0; return KSV_OK;
    case KSV_OPT_TILE_QUBITS: {
        int t = (int)value;
        int tmax = s->prec == KSV_F64 ? KMAX_TILE : KMAX_TILE + 1;
        if (t != 0 && (t < 8 || t > tmax)) return fail(KSV_ERR_RANGE, "tile qubits must be 0 or 8.." + std::to_string(tmax));
        s->tile = t;
        return KSV_OK;
    }
    case KSV_OPT_TIMING: s->timing = value != 0; return KSV_OK;
    case KSV_OPT_FUSE: s->fuse = value != 0; return KSV_OK;
    case KSV_OPT_PASS_KERNEL: {
        int k = (int)value;
        if (k < 0 || k > 3) return fail(KSV_ERR_RANGE, "pass kernel must be 0, 1, 2 or 3");
        s->kernel = k;
        return KSV_OK;
    }
    case KSV_OPT_BLOCK_QUBITS: {
        int b = (int)value;
        if (b != 0 && b != 4 && b != 5) return fail(KSV_ERR_RANGE, "fused-block width must be 0, 4 or 5");
        s->bq = b;
        return KSV_OK;
    }
    case KSV_OPT_BLOCK_THREADS: {
        int t = (int)value;
        if (t != 0 && t != 256 && t != 512) return fail(KSV_ERR_RANGE, "fused-block workgroup size must be 0, 256 or 512");
        s->bthreads = t ? t : 256;
        return KSV_OK;
    }
    case KSV_OPT_BLOCK_PIPE:
        if (value != 0 && value != 1) return fail(KSV_ERR_RANGE, "fused-block pipelining must be 0 or 1");
        s->bpipe = (int)value;
        return KSV_OK;
    case KSV_OPT_ADJ_TILE: {
        int t = (int)value;
        const int hi = s->prec == KSV_F64 ? ADJ_TILE_F64 : ADJ_TILE_F32;
        if (t != 0 && (t < 10 || t > hi)) return fail(KSV_ERR_RANGE, "adjoint tile qubits must be 0 or 10..default");
        s->atile = t;
        return KSV_OK;
    }
    case KSV_OPT_ADJ_THREADS: {
        int t = (int)value;
        if (t != 0 && t != 256 && t != 512) return fail(KSV_ERR_RANGE, "adjoint workgroup size must be 0, 256 or 512");
        s->athreads = t ? t : 256;
        return KSV_OK;
    }
    case KSV_OPT_BLOCK_TILE: {
        int t = (int)value;
        if (t != 0 && (t < 11 || t > 13)) return fail(KSV_ERR_RANGE, "fused-block tile qubits must be 0 or 11..13");
        s->btile = t;
        return KSV_OK;
    }
    case KSV_OPT_ROUND_TILE: {
        int r = (int)value;
        if (r != 0 && (r < 11 || r > 13)) return fail(KSV_ERR_RANGE, "register-round tile qubits must be 0 or 11..13");
        s->rtile = r;
        return KSV_OK;
    }
    default: return fail(KSV_ERR_TYPE, "unknown option " + std::to_string(option));
    }
}

int ksv_init_basis(ksv_state* s, uint64_t index) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    if (index >> s->n) return fail(KSV_ERR_RANGE, "basis index out of range");
    HIPCHK(k_init_basis(s->n, s->prec, s->psi, index, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return KSV_OK;
}

int ksv_copy_state(ksv_state* dst, const ksv_state* src) {
    if (!dst || !src) return fail(KSV_ERR_HANDLE, "null state");
    if (dst->n != src->n || dst->prec != src->prec) return fail(KSV_ERR_TYPE, "state shapes differ");
    HIPCHK(hipStreamSynchronize(src->st));
    HIPCHK(k_copy(dst->n, dst->prec, dst->psi, src->psi, dst->st));
    HIPCHK(hipStreamSynchronize(dst->st));
    return KSV_OK;
}

// --------------------------------------------------------------------------
// programs
// --------------------------------------------------------------------------
// Lower and schedule a tape into `p` (host side only).
static int plan_into(Program& p, int n, int prec, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                     int kernel, int tile, int rtile, bool fuse, int prune, int btile = 0, int bq = 0) {
    std::string err;
    int bqv = bq > 0 ? bq : 4, btv = btile > 0 ? btile : 12;
    if (kernel == 3 && bqv == 5 && btv != 13) return fail(KSV_ERR_RANGE, "5-qubit blocks need 13-qubit tiles");
    if (kernel == 3 && btv == 11 && bqv != 4) return fail(KSV_ERR_RANGE, "11-qubit tiles take 4-qubit blocks");
    int rc = plan_program(p, n, prec, ops, nops, mats, nmats, kernel, tile, rtile, fuse, prune, err, btile, bq);
    if (rc) return fail(rc, err);
    return KSV_OK;
}

// Serialise a planned program's device arrays (ROps, DOps, BRecs, block
// fragments, matrices in the state's precision, copy bitmaps) into `blob` at
// 256-byte aligned offsets; off[] receives the six offsets (SIZE_MAX = absent).
static void program_blob(const ksv_state* s, const Program& p, std::vector<char>& blob, std::vector<size_t>& off) {
    auto put = [&](const void* src, size_t bytes) -> size_t {
        if (!bytes) return SIZE_MAX;
        size_t at = (blob.size() + 255) & ~(size_t)255;
        blob.resize(at + bytes);
        std::memcpy(blob.data() + at, src, bytes);
        return at;
    };
    off.assign(7, SIZE_MAX);
    off[0] = put(p.rops.data(), sizeof(ROp) * p.rops.size());
    off[1] = put(p.dops.data(), sizeof(DOp) * p.dops.size());
    off[2] = put(p.brecs.data(), sizeof(BRec) * p.brecs.size());
    if (!p.brecs.empty()) {
        size_t total = 0;
        std::vector<int> bqs(p.brecs.size(), 4);
        for (const Step& st : p.steps)
            if (st.blk)
                for (int i = 0; i < st.op_count; i++) bqs[st.op_begin + i] = st.bq;
        for (size_t i = 0; i < p.brecs.size(); i++) {
            const int R = 1 << bqs[i];
            total = std::max(total, (size_t)p.brecs[i].frag + (size_t)(2 * R / 16) * (2 * R / 4) * 64);
        }
        std::vector<double> fr(total, 0.0);
        for (size_t i = 0; i < p.brecs.size(); i++)
            block_fragments(p.bmats.data() + p.bmat_off[i], bqs[i], p.brecs[i].flags,
                            fr.data() + p.brecs[i].frag);
        if (s->prec == KSV_F64) {
            off[3] = put(fr.data(), 8 * total);
        } else {
            std::vector<float> ff(total);
            for (size_t i = 0; i < total; i++) ff[i] = (float)fr[i];
            off[3] = put(ff.data(), 4 * total);
        }
    }
    if (!p.mats.empty()) {
        if (s->prec == KSV_F64) {
            off[4] = put(p.mats.data(), 16 * p.mats.size());
        } else {
            std::vector<float> f(2 * p.mats.size());
            for (size_t i = 0; i < p.mats.size(); i++) {
                f[2 * i] = (float)p.mats[i].real();
                f[2 * i + 1] = (float)p.mats[i].imag();
            }
            off[4] = put(f.data(), 8 * p.mats.size());
        }
    }
    off[5] = put(p.bits.data(), sizeof(uint32_t) * p.bits.size());
    off[6] = put(p.rbits.data(), sizeof(uint32_t) * p.rbits.size());
}

static void bind_blob(Program& p, char* base, const std::vector<size_t>& off) {
    auto at = [&](size_t o) -> void* { return o == SIZE_MAX ? nullptr : base + o; };
    p.d_rops = (ROp*)at(off[0]);
    p.d_ops = (DOp*)at(off[1]);
    p.d_brecs = (BRec*)at(off[2]);
    p.d_frags = at(off[3]);
    p.d_mats = at(off[4]);
    p.d_bits = (uint32_t*)at(off[5]);
    p.d_rbits = (uint32_t*)at(off[6]);
}

static int build_program(ksv_state* s, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                         int prune, bool fuse, ksv_program** out) {
    std::unique_ptr<ksv_program> p(new ksv_program());
    int rc = plan_into(*p, s->n, s->prec, ops, nops, mats, nmats, s->kernel, default_tile(s), s->rtile, fuse,
                       prune, s->btile, s->bq);
    if (rc) return rc;
    p->dev = s->dev;
    p->bthreads = s->bthreads;
    p->bpipe = s->bpipe;
    std::vector<char> blob;
    std::vector<size_t> off;
    program_blob(s, *p, blob, off);
    if (!blob.empty()) {
        HIPCHK(hipMalloc(&p->d_blob, blob.size()));
        HIPCHK(hipMemcpy(p->d_blob, blob.data(), blob.size(), hipMemcpyHostToDevice));
        bind_blob(*p, (char*)p->d_blob, off);
    }
    *out = p.release();
    return KSV_OK;
}

static int run_program_on(ksv_state* s, ksv_program* p, void* psi, bool sync = true) {
    const bool timing = s->timing != 0;
    const size_t nst = p->steps.size();
    if (timing) {
        while (s->ev.size() < nst + 1) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            s->ev.push_back(e);
        }
        HIPCHK(hipEventRecord(s->ev[0], s->st));
    }
    for (size_t i = 0; i < nst; i++) {
        const Step& st = p->steps[i];
        if (st.big) HIPCHK(launch_big(*p, st, psi, p->prune, s->st));
        else if (st.reg) HIPCHK(launch_rpass(*p, st, psi, s->st));
        else if (st.blk) HIPCHK(launch_bpass(*p, st, psi, s->st));
        else HIPCHK(launch_pass(*p, st, psi, s->st));
        if (timing) HIPCHK(hipEventRecord(s->ev[i + 1], s->st));
    }
    if (!sync && !timing) {
        s->stats.pass_launches += nst;
        s->stats.pass_bytes += 2.0 * (double)state_bytes(s) * nst;
        s->stats.gates_applied += p->tape_ops;
        return KSV_OK;
    }
    HIPCHK(hipStreamSynchronize(s->st));
    const double bytes = 2.0 * (double)state_bytes(s);
    for (size_t i = 0; i < nst; i++) {
        const Step& st = p->steps[i];
        float ms = 0;
        if (timing) HIPCHK(hipEventElapsedTime(&ms, s->ev[i], s->ev[i + 1]));
        if (st.big) {
            s->stats.other_launches++;
            s->stats.other_ms += ms;
            s->stats.other_bytes += bytes;
        } else {
            s->stats.pass_launches++;
            s->stats.pass_ms += ms;
            s->stats.pass_bytes += bytes;
        }
    }
    s->stats.gates_applied += p->tape_ops;
    return KSV_OK;
}

int ksv_program_create(ksv_state* s, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                       ksv_program** out) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    if (!out) return fail(KSV_ERR_TYPE, "null output pointer");
    return build_program(s, ops, nops, mats, nmats, s->prune, s->fuse != 0, out);
}

int ksv_program_run(ksv_state* s, ksv_program* p) {
    if (!s || !p) return fail(KSV_ERR_HANDLE, "null state or program");
    if (p->n != s->n || p->precision != s->prec) return fail(KSV_ERR_TYPE, "program was built for another state shape");
    return run_program_on(s, p, s->psi);
}

static int summarize(const Program& p, ksv_program_summary* out) {
    out->ops = p.tape_ops;
    out->passes = p.steps.size();
    out->device_ops = p.dops.size() + p.big_ops.size();
    int regs = 0, exch = 0, blks = 0;
    for (const Step& st : p.steps) {
        if (st.reg) { regs++; exch += st.exchanges; }
        if (st.blk) blks++;
    }
    out->blk_passes = blks;
    out->blocks = p.brecs.size();
    out->device_ops += p.brecs.size();
    for (const ROp& r : p.rops)
        if (r.kind != R_EXCH) out->device_ops++;
    out->tile_qubits = blks ? p.steps[0].k : (regs ? p.rtile : p.tile);
    for (const Step& st : p.steps)
        if (st.blk) { out->tile_qubits = st.k; break; }
    out->reg_passes = regs;
    out->exchanges = exch;
    return KSV_OK;
}

int ksv_program_set_matrix(ksv_program* p, size_t op, const double* mat, size_t nmat) {
    if (!p) return fail(KSV_ERR_HANDLE, "null program");
    if (!mat) return fail(KSV_ERR_TYPE, "null matrix");
    std::vector<int> rops, blocks;
    std::vector<std::pair<size_t, size_t>> ranges;
    std::string err;
    int rc = update_program(*p, op, mat, nmat, rops, blocks, ranges, err);
    if (rc) return fail(rc, err);
    for (int r : rops) HIPCHK(hipMemcpy(p->d_rops + r, &p->rops[r], sizeof(ROp), hipMemcpyHostToDevice));
    const bool f64 = p->precision == KSV_F64;
    for (const auto& rg : ranges) {
        if (f64) {
            HIPCHK(hipMemcpy((char*)p->d_mats + 16 * rg.first, p->mats.data() + rg.first, 16 * rg.second,
                             hipMemcpyHostToDevice));
        } else {
            std::vector<float> f(2 * rg.second);
            for (size_t i = 0; i < rg.second; i++) {
                f[2 * i] = (float)p->mats[rg.first + i].real();
                f[2 * i + 1] = (float)p->mats[rg.first + i].imag();
            }
            HIPCHK(hipMemcpy((char*)p->d_mats + 8 * rg.first, f.data(), 8 * rg.second, hipMemcpyHostToDevice));
        }
    }
    for (int b : blocks) {
        const int R = 1 << p->blk_bq[b];
        const size_t nf = (size_t)(2 * R / 16) * (2 * R / 4) * 64;
        std::vector<double> fr(nf);
        block_fragments(p->bmats.data() + p->bmat_off[b], p->blk_bq[b], p->brecs[b].flags, fr.data());
        if (f64) {
            HIPCHK(hipMemcpy((double*)p->d_frags + p->brecs[b].frag, fr.data(), 8 * nf, hipMemcpyHostToDevice));
        } else {
            std::vector<float> ff(nf);
            for (size_t i = 0; i < nf; i++) ff[i] = (float)fr[i];
            HIPCHK(hipMemcpy((float*)p->d_frags + p->brecs[b].frag, ff.data(), 4 * nf, hipMemcpyHostToDevice));
        }
    }
    return KSV_OK;
}

int ksv_program_info(const ksv_program* p, ksv_program_summary* out) {
    if (!p || !out) return fail(KSV_ERR_HANDLE, "null program");
    return summarize(*p, out);
}

int ksv_program_destroy(ksv_program* p) {
    if (!p) return fail(KSV_ERR_HANDLE, "null program");
    if (p->d_blob && !p->borrowed) (void)hipFree(p->d_blob);
    delete p;
    return KSV_OK;
}

int ksv_plan(int n, int precision, const ksv_op* ops, size_t nops, const double* mats, size_t nmats, int kernel,
             int tile, int rtile, int fuse, int prune, ksv_program_summary* out) {
    return ksv_plan_blocks(n, precision, ops, nops, mats, nmats, kernel, tile, rtile, 0, 0, fuse, prune, out);
}

int ksv_plan_blocks(int n, int precision, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                    int kernel, int tile, int rtile, int btile, int bq, int fuse, int prune, ksv_program_summary* out) {
    if (!out) return fail(KSV_ERR_TYPE, "null output pointer");
    if (n < 1 || n > MAXN) return fail(KSV_ERR_RANGE, "qubit count out of range");
    if (precision != KSV_F64 && precision != KSV_F32) return fail(KSV_ERR_TYPE, "unknown precision");
    if (kernel < 0 || kernel > 3) return fail(KSV_ERR_RANGE, "pass kernel must be 0, 1, 2 or 3");
    if (bq != 0 && bq != 4 && bq != 5) return fail(KSV_ERR_RANGE, "fused-block width must be 0, 4 or 5");
    if (btile != 0 && (btile < 11 || btile > 13)) return fail(KSV_ERR_RANGE, "fused-block tile qubits must be 0 or 11..13");
    Program p;
    if (rtile != 0 && (rtile < 11 || rtile > 13)) return fail(KSV_ERR_RANGE, "register-round tile qubits must be 0 or 11..13");
    int rc = plan_into(p, n, precision, ops, nops, mats, nmats, kernel, tile, rtile, fuse != 0, prune, btile, bq);
    if (rc) return rc;
    return summarize(p, out);
}

int ksv_apply(ksv_state* s, const ksv_op* ops, size_t nops, const double* mats, size_t nmats) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    ksv_program* p = nullptr;
    int rc = build_program(s, ops, nops, mats, nmats, s->prune, s->fuse != 0, &p);
    if (rc) return rc;
    rc = run_program_on(s, p, s->psi);
    ksv_program_destroy(p);
    return rc;
}

// --------------------------------------------------------------------------
// read-out
// --------------------------------------------------------------------------
int ksv_get_amplitudes(ksv_state* s, uint64_t offset, uint64_t count, double* out) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    if (offset > (1ull << s->n) || count > (1ull << s->n) - offset) return fail(KSV_ERR_RANGE, "amplitude range out of bounds");
    if (!count) return KSV_OK;
    if (!out) return fail(KSV_ERR_TYPE, "null output buffer");
    if (s->prec == KSV_F64) {
        HIPCHK(hipMemcpyAsync(out, (char*)s->psi + 16 * offset, 16 * count, hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
    } else {
        std::vector<float> f(2 * count);
        HIPCHK(hipMemcpyAsync(f.data(), (char*)s->psi + 8 * offset, 8 * count, hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
        for (uint64_t i = 0; i < 2 * count; i++) out[i] = f[i];
    }
    return KSV_OK;
}

int ksv_set_amplitudes(ksv_state* s, uint64_t offset, uint64_t count, const double* in) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    if (offset > (1ull << s->n) || count > (1ull << s->n) - offset) return fail(KSV_ERR_RANGE, "amplitude range out of bounds");
    if (!count) return KSV_OK;
    if (!in) return fail(KSV_ERR_TYPE, "null input buffer");
    if (s->prec == KSV_F64) {
        HIPCHK(hipMemcpyAsync((char*)s->psi + 16 * offset, in, 16 * count, hipMemcpyHostToDevice, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
    } else {
        std::vector<float> f(2 * count);
        for (uint64_t i = 0; i < 2 * count; i++) f[i] = (float)in[i];
        HIPCHK(hipMemcpyAsync((char*)s->psi + 8 * offset, f.data(), 8 * count, hipMemcpyHostToDevice, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
    }
    return KSV_OK;
}

int ksv_amplitude(ksv_state* s, uint64_t index, double out[2]) {
    return ksv_get_amplitudes(s, index, 1, out);
}

static int fetch_small(ksv_state* s, double* out, int m) {
    HIPCHK(hipMemcpyAsync(out, s->d_small, sizeof(double) * m, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return KSV_OK;
}

int ksv_norm2(ksv_state* s, double* out) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    HIPCHK(red_norm2(s->n, s->prec, s->psi, s->d_small, s->d_part, NBLK, s->st));
    return fetch_small(s, out, 1);
}

int ksv_marginals(ksv_state* s, double* out) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    HIPCHK(red_marginals(s->n, s->prec, s->psi, s->d_small, s->d_part, NBLK, s->st));
    return fetch_small(s, out, s->n);
}

int ksv_reduce_1q(ksv_state* s, int q, double out[4]) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    if (q < 0 || q >= s->n) return fail(KSV_ERR_RANGE, "qubit index " + std::to_string(q) + " is out of range for a " + std::to_string(s->n) + "-qubit circuit");
    HIPCHK(red_1q(s->n, s->prec, s->psi, q, s->d_small, s->d_part, NBLK, s->st));
    return fetch_small(s, out, 4);
}

// Ordered compaction of amplitudes satisfying a predicate into caller buffers.
static int compact(ksv_state* s, int mode, double thresh, uint64_t kv, uint64_t* idx, double* p, uint64_t cap,
                   uint64_t* total) {
    const uint64_t nch = ((1ull << s->n) + (1ull << CHUNK_LOG) - 1) >> CHUNK_LOG;
    int rc = ws_reserve(s, wsz(4 * nch));
    if (rc) return rc;
    uint32_t* d_cnt = Carve{(char*)s->ws}.take<uint32_t>(nch);
    if (mode == 0) HIPCHK(chunk_counts(s->n, s->prec, s->psi, thresh, d_cnt, s->st));
    else HIPCHK(chunk_counts_pred(s->n, s->prec, s->psi, mode, kv, d_cnt, s->st));
    std::vector<uint32_t> cnt(nch);
    std::vector<uint64_t> off(nch);
    HIPCHK(hipMemcpyAsync(cnt.data(), d_cnt, 4 * nch, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    uint64_t acc = 0;
    for (uint64_t c = 0; c < nch; c++) { off[c] = acc; acc += cnt[c]; }
    *total = acc;
    const uint64_t m = std::min(acc, cap);
    if (m) {
        rc = ws_reserve(s, wsz(8 * nch) + 2 * wsz(8 * m));
        if (rc) return rc;
        Carve c{(char*)s->ws};
        uint64_t* d_off = c.take<uint64_t>(nch);
        uint64_t* d_idx = c.take<uint64_t>(m);
        double* d_p = c.take<double>(m);
        HIPCHK(hipMemcpyAsync(d_off, off.data(), 8 * nch, hipMemcpyHostToDevice, s->st));
        if (mode == 0) HIPCHK(chunk_compact(s->n, s->prec, s->psi, thresh, d_off, d_idx, d_p, m, s->st));
        else HIPCHK(chunk_compact_pred(s->n, s->prec, s->psi, mode, kv, d_off, d_idx, d_p, m, s->st));
        if (idx) HIPCHK(hipMemcpyAsync(idx, d_idx, 8 * m, hipMemcpyDeviceToHost, s->st));
        if (p) HIPCHK(hipMemcpyAsync(p, d_p, 8 * m, hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
    }
    return KSV_OK;
}

int ksv_probs_filtered(ksv_state* s, double thresh, uint64_t* idx, double* p, size_t cap, size_t* count) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    uint64_t total = 0;
    int rc = compact(s, 0, thresh, 0, idx, p, cap, &total);
    if (rc) return rc;
    if (count) *count = (size_t)total;
    return KSV_OK;
}

int ksv_topk(ksv_state* s, size_t k, uint64_t* idx, double* p) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    const uint64_t N = 1ull << s->n;
    if (k == 0) return KSV_OK;
    if (k > N) return fail(KSV_ERR_RANGE, "k exceeds the number of amplitudes");
    if (!idx || !p) return fail(KSV_ERR_TYPE, "null output buffer");
    // radix-select the k-th largest key (bit pattern of p) with 12-bit digits
    int rc = ws_reserve(s, wsz(4 * 4096));
    if (rc) return rc;
    uint32_t* d_hist = Carve{(char*)s->ws}.take<uint32_t>(4096);
    std::vector<uint32_t> h(4096);
    uint64_t prefix = 0, above = 0;   // `above` = keys strictly greater than the current prefix range
    uint64_t want = k;                // rank (1-based) of the target within the prefix range
    const int rounds[6][2] = {{64, 52}, {52, 40}, {40, 28}, {28, 16}, {16, 4}, {4, 0}};
    for (auto& rd : rounds) {
        const int hi = rd[0], lo = rd[1];
        HIPCHK(hipMemsetAsync(d_hist, 0, 4 * 4096, s->st));
        HIPCHK(key_histogram(s->n, s->prec, s->psi, prefix, hi, lo, d_hist, s->st));
        HIPCHK(hipMemcpyAsync(h.data(), d_hist, 4 * 4096, hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
        const int nb = 1 << std::min(12, hi - lo);
        uint64_t cum = 0;
        int d = nb - 1;
        for (; d >= 0; d--) {
            if (cum + h[d] >= want) break;
            cum += h[d];
        }
        if (d < 0) d = 0;
        above += cum;
        want -= cum;
        prefix = (prefix << (hi - lo)) | (uint64_t)d;
    }
    const uint64_t kth = prefix;      // full 64-bit key of the k-th largest
    // keys > kth: `above` of them, all wanted; keys == kth: the first (k - above) by index
    std::vector<uint64_t> ig(above + 1), ie(k - above + 1);
    std::vector<double> pg(above + 1), pe(k - above + 1);
    uint64_t tg = 0, te = 0;
    rc = compact(s, 2, 0, kth, ig.data(), pg.data(), above, &tg);
    if (rc) return rc;
    rc = compact(s, 1, 0, kth, ie.data(), pe.data(), k - above, &te);
    if (rc) return rc;
    std::vector<std::pair<double, uint64_t>> all;
    all.reserve(k);
    for (uint64_t i = 0; i < std::min(tg, above); i++) all.push_back({pg[i], ig[i]});
    for (uint64_t i = 0; i < std::min(te, (uint64_t)(k - above)); i++) all.push_back({pe[i], ie[i]});
    std::sort(all.begin(), all.end(), [](const std::pair<double, uint64_t>& a, const std::pair<double, uint64_t>& b) {
        if (a.first != b.first) return a.first > b.first;
        return a.second < b.second;
    });
    for (size_t i = 0; i < k; i++) {
        idx[i] = i < all.size() ? all[i].second : 0;
        p[i] = i < all.size() ? all[i].first : 0;
    }
    return KSV_OK;
}

// ket's run() draw (src/circuit.ts:3425-3450) on ket's sequential CDF: one
// wave forms the running sums in ket's addition order (d_cum: 2^n doubles),
// then one thread per draw searches them.  d_r / d_out hold `shots` doubles /
// indices; the draws are uploaded and the indices read back on the state stream.
static int sample_seq(ksv_state* s, const double* r, size_t shots, uint64_t* out, double* d_cum,
                      unsigned long long* d_ends, double* d_r, uint64_t* d_out) {
    HIPCHK(seq_cdf(s->n, s->prec, s->psi, 1e-14, d_cum, d_ends, s->st));
    HIPCHK(hipMemcpyAsync(d_r, r, 8 * shots, hipMemcpyHostToDevice, s->st));
    HIPCHK(cdf_search(d_cum, d_ends, d_r, shots, d_out, s->st));
    HIPCHK(hipMemcpyAsync(out, d_out, 8 * shots, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return KSV_OK;
}

// ---- allocation-free sampling ---------------------------------------------
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc):
// ksv_sample's exact path with its scratch in one fixed device array per GPU,
// held until the process ends: 2^20 running sums (8 MiB), the two end
// indices, and 65536 draws and indices (1 MiB).  Host code reaches the array
// only through its device address, looked up once per device.
constexpr int NOALLOC_MAX_QUBITS = 20;
constexpr size_t NOALLOC_MAX_SHOTS = 65536;
constexpr size_t NOALLOC_CUM = (size_t)8 << NOALLOC_MAX_QUBITS;
constexpr size_t NOALLOC_ENDS = 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_CUM + NOALLOC_ENDS + 2 * 8 * NOALLOC_MAX_SHOTS;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Device address of `scratch` on device `dev` (looked up with that device
// current, once per device; each GPU holds its own copy at its own address).
static int scratch_on(int dev, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(KSV_ERR_DEVICE, "device index beyond the scratch table");
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        HIPCHK(hipGetDevice(&cur));
        if (cur != dev) HIPCHK(hipSetDevice(dev));
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != dev) (void)hipSetDevice(cur);
        if (e != hipSuccess || !ptr)
            return fail(KSV_ERR_DEVICE, std::string("scratch address: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return KSV_OK;
}

int ksv_sample(ksv_state* s, const double* r, size_t shots, uint64_t* out) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    if (!shots) return KSV_OK;
    if (!r || !out) return fail(KSV_ERR_TYPE, "null buffer");
    const double thresh = 1e-14;
    if (s->n <= SEQ_CDF_MAX_QUBITS) {
        const uint64_t N = 1ull << s->n;
        int rc = ws_reserve(s, wsz(8 * N) + wsz(16) + 2 * wsz(8 * shots));
        if (rc) return rc;
        Carve c{(char*)s->ws};
        double* d_cum = c.take<double>(N);
        unsigned long long* d_ends = c.take<unsigned long long>(2);
        double* d_r = c.take<double>(shots);
        uint64_t* d_out = c.take<uint64_t>(shots);
        return sample_seq(s, r, shots, out, d_cum, d_ends, d_r, d_out);
    }
    // beyond ket's 2^24-entry range: the CDF from a chunked scan (its sums are
    // associated differently from a sequential loop)
    const int64_t nch = (int64_t)(((1ull << s->n) + (1ull << CHUNK_LOG) - 1) >> CHUNK_LOG);
    int rc = ws_reserve(s, wsz(8 * nch) + wsz(8) + 2 * wsz(8 * shots));
    if (rc) return rc;
    Carve c{(char*)s->ws};
    double* d_sum = c.take<double>(nch);
    uint64_t* d_last = c.take<uint64_t>(1);
    double* d_r = c.take<double>(shots);
    uint64_t* d_out = c.take<uint64_t>(shots);
    HIPCHK(chunk_sums(s->n, s->prec, s->psi, thresh, d_sum, s->st));
    std::vector<double> cum(nch);
    HIPCHK(hipMemcpyAsync(cum.data(), d_sum, 8 * nch, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    int64_t lastc = -1;
    double acc = 0;
    for (int64_t k = 0; k < nch; k++) {
        if (cum[k] > 0) lastc = k;
        acc += cum[k];
        cum[k] = acc;
    }
    if (lastc < 0) {   // no probability above the filter: ket returns index 0
        for (size_t i = 0; i < shots; i++) out[i] = 0;
        return KSV_OK;
    }
    HIPCHK(hipMemsetAsync(d_last, 0, 8, s->st));
    HIPCHK(last_nonzero_chunk(s->n, s->prec, s->psi, thresh, lastc, d_last, s->st));
    uint64_t last_idx = 0;
    HIPCHK(hipMemcpyAsync(&last_idx, d_last, 8, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipMemcpyAsync(d_sum, cum.data(), 8 * nch, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemcpyAsync(d_r, r, 8 * shots, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    HIPCHK(sample_search(s->n, s->prec, s->psi, thresh, d_sum, nch, d_r, shots, last_idx, d_out, s->st));
    HIPCHK(hipMemcpyAsync(out, d_out, 8 * shots, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return KSV_OK;
}

int ksv_sample_noalloc(ksv_state* s, const double* r, size_t shots, uint64_t* out) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    if (!shots) return KSV_OK;
    if (!r || !out) return fail(KSV_ERR_TYPE, "null buffer");
    if (s->n > NOALLOC_MAX_QUBITS)
        return fail(KSV_ERR_RANGE, "ksv_sample_noalloc: " + std::to_string(s->n) + " qubits exceeds its scratch (at most " +
                                       std::to_string(NOALLOC_MAX_QUBITS) + ")");
    if (shots > NOALLOC_MAX_SHOTS)
        return fail(KSV_ERR_RANGE, "ksv_sample_noalloc: " + std::to_string(shots) + " draws exceeds its scratch (at most " +
                                       std::to_string(NOALLOC_MAX_SHOTS) + " per call)");
    unsigned char* base = nullptr;
    int rc = scratch_on(s->dev, &base);
    if (rc) return rc;
    double* d_cum = reinterpret_cast<double*>(base);
    unsigned long long* d_ends = reinterpret_cast<unsigned long long*>(base + NOALLOC_CUM);
    double* d_r = reinterpret_cast<double*>(base + NOALLOC_CUM + NOALLOC_ENDS);
    uint64_t* d_out = reinterpret_cast<uint64_t*>(base + NOALLOC_CUM + NOALLOC_ENDS + 8 * NOALLOC_MAX_SHOTS);
    // every running sum and both end indices are rewritten by this call: nothing to clear
    return sample_seq(s, r, shots, out, d_cum, d_ends, d_r, d_out);
}

static int ensure_psi2(ksv_state* s) {
    if (s->psi2) return KSV_OK;
    hipError_t e = hipMalloc(&s->psi2, state_bytes(s));
    if (e != hipSuccess) { s->psi2 = nullptr; return fail(KSV_ERR_DEVICE, "out of device memory for a second state"); }
    return KSV_OK;
}

// Rotation into the Z basis of one Pauli term (ket: X -> h, Y -> rx(pi/2)).
static void rotation_tape(int n, uint64_t x, uint64_t z, std::vector<ksv_op>& ops, std::vector<double>& mats) {
    const double sq2 = 1 / std::sqrt(2.0);
    const double h[8] = {sq2, 0, sq2, 0, sq2, 0, -sq2, 0};
    const double co = std::cos(M_PI / 2 / 2), si = std::sin(M_PI / 2 / 2);
    const double rx[8] = {co, 0, 0, -si, 0, -si, co, 0};
    mats.assign(h, h + 8);
    mats.insert(mats.end(), rx, rx + 8);
    ops.clear();
    for (int q = 0; q < n; q++) {
        bool bx = (x >> q) & 1, bz = (z >> q) & 1;
        if (!bx) continue;
        ksv_op o;
        std::memset(&o, 0, sizeof o);
        o.kind = KSV_SINGLE;
        o.nq = 1;
        o.q[0] = q;
        o.mat = bz ? 8 : 0;
        ops.push_back(o);
    }
}

int ksv_expect_pauli(ksv_state* s, const uint64_t* xmask, const uint64_t* zmask, const double* coeff, size_t nterms,
                     int mode, double* out, double* total) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    if (nterms && (!xmask || !zmask || !coeff || !out)) return fail(KSV_ERR_TYPE, "null term buffer");
    if (mode != 0 && mode != 1) return fail(KSV_ERR_TYPE, "unknown expectation mode");
    const uint64_t all = (1ull << s->n) - 1;
    for (size_t t = 0; t < nterms; t++)
        if ((xmask[t] | zmask[t]) & ~all) return fail(KSV_ERR_RANGE, "Pauli term acts outside the register");
    double tot = 0;
    if (mode == 0) {
        uint64_t *d_x = nullptr, *d_z = nullptr;
        if (nterms) {
            int rc = ws_reserve(s, 2 * wsz(8 * nterms));
            if (rc) return rc;
            Carve c{(char*)s->ws};
            d_x = c.take<uint64_t>(nterms);
            d_z = c.take<uint64_t>(nterms);
            HIPCHK(hipMemcpyAsync(d_x, xmask, 8 * nterms, hipMemcpyHostToDevice, s->st));
            HIPCHK(hipMemcpyAsync(d_z, zmask, 8 * nterms, hipMemcpyHostToDevice, s->st));
        }
        for (size_t t0 = 0; t0 < nterms; t0 += 8) {
            int m = (int)std::min<size_t>(8, nterms - t0);
            HIPCHK(red_pauli(s->n, s->prec, s->psi, d_x + t0, d_z + t0, m, s->d_small, s->d_part, NBLK, s->st));
            double v[16];
            int rc = fetch_small(s, v, 16);
            if (rc) return rc;
            for (int j = 0; j < m; j++) {
                size_t t = t0 + j;
                int ny = __builtin_popcountll(xmask[t] & zmask[t]) & 3;
                double re = v[2 * j], im = v[2 * j + 1], val;
                switch (ny) {
                case 0: val = re; break;
                case 1: val = -im; break;
                case 2: val = -re; break;
                default: val = im; break;
                }
                out[t] = coeff[t] * val;
                tot += out[t];
            }
        }
    } else {
        std::vector<ksv_op> ops;
        std::vector<double> mats;
        for (size_t t = 0; t < nterms; t++) {
            if (std::fabs(coeff[t]) < 1e-15) { out[t] = 0; continue; }
            if (!(xmask[t] | zmask[t])) { out[t] = coeff[t]; tot += out[t]; continue; }
            int rc = ensure_psi2(s);
            if (rc) return rc;
            HIPCHK(k_copy(s->n, s->prec, s->psi2, s->psi, s->st));
            rotation_tape(s->n, xmask[t], zmask[t], ops, mats);
            if (!ops.empty()) {
                ksv_program* p = nullptr;
                rc = build_program(s, ops.data(), ops.size(), mats.data(), mats.size(), 1, false, &p);
                if (rc) return rc;
                rc = run_program_on(s, p, s->psi2);
                ksv_program_destroy(p);
                if (rc) return rc;
            }
            HIPCHK(red_parity(s->n, s->prec, s->psi2, xmask[t] | zmask[t], 1e-14, s->d_small, s->d_part, NBLK, s->st));
            double v;
            rc = fetch_small(s, &v, 1);
            if (rc) return rc;
            out[t] = coeff[t] * v;
            tot += out[t];
        }
    }
    if (total) *total = tot;
    return KSV_OK;
}

int ksv_collapse(ksv_state* s, int q, double r, int* outcome) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    if (q < 0 || q >= s->n) return fail(KSV_ERR_RANGE, "qubit index " + std::to_string(q) + " is out of range for a " + std::to_string(s->n) + "-qubit circuit");
    HIPCHK(red_bit1(s->n, s->prec, s->psi, q, s->d_small, s->d_part, NBLK, s->st));
    double p1 = 0;
    int rc = fetch_small(s, &p1, 1);
    if (rc) return rc;
    int o = r < p1 ? 1 : 0;
    double scale = 1 / std::sqrt(o ? p1 : 1 - p1);
    HIPCHK(k_project(s->n, s->prec, s->psi, q, o, scale, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    if (outcome) *outcome = o;
    return KSV_OK;
}


}  // extern "C"

// ---- batched trajectories -------------------------------------------------
// Device scratch freed on every exit path.
struct DevBufs {
    std::vector<void*> p;
    ~DevBufs() {
        for (void* x : p) hipFree(x);
    }
    template <typename T> hipError_t get(T** out, size_t count) {
        void* x = nullptr;
        hipError_t e = hipMalloc(&x, std::max<size_t>(1, count) * sizeof(T));
        if (e == hipSuccess) p.push_back(x);
        *out = (T*)x;
        return e;
    }
};

static int shots_check(ksv_state* s, int n, uint32_t* copies) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    if (n < 1 || n > s->n) return fail(KSV_ERR_RANGE, "copy width must be 1.." + std::to_string(s->n) + " qubits");
    if (s->n - n > 16) return fail(KSV_ERR_RANGE, "at most 65536 copies");
    *copies = 1u << (s->n - n);
    return KSV_OK;
}

// xorshift32 (src/circuit.ts:813-819) is linear over GF(2)^32: one step is a
// 32x32 bit matrix M, stored as the images of the 32 unit vectors, and `steps`
// steps are M^steps, built by repeated squaring.
namespace {

uint32_t xs_step(uint32_t v) {
    v ^= v << 13;
    v ^= v >> 17;
    v ^= v << 5;
    return v;
}

struct Gf2Map {
    uint32_t col[32];
};

uint32_t gf2_apply(const Gf2Map& m, uint32_t v) {
    uint32_t r = 0;
    for (int i = 0; v; i++, v >>= 1)
        if (v & 1) r ^= m.col[i];
    return r;
}

uint32_t xs_jump(uint32_t v, uint64_t steps) {
    Gf2Map p;
    for (int i = 0; i < 32; i++) p.col[i] = xs_step(1u << i);
    while (steps) {
        if (steps & 1) v = gf2_apply(p, v);
        Gf2Map sq;
        for (int i = 0; i < 32; i++) sq.col[i] = gf2_apply(p, p.col[i]);
        p = sq;
        steps >>= 1;
    }
    return v;
}

// out[i] = the (i+1)-th state after v, divided by 2^32; returns the last state
uint32_t xs_fill(uint32_t v, double* out, size_t count) {
    for (size_t i = 0; i < count; i++) {
        v = xs_step(v);
        out[i] = v / 4294967296.0;
    }
    return v;
}

}  // namespace

extern "C" {

int ksv_xorshift32(uint32_t* state, size_t count, double* out) {
    if (!state || (count && !out)) return fail(KSV_ERR_TYPE, "null buffer");
    const uint32_t v0 = *state ? *state : 1u;
    // Long streams are cut into chunks, one per thread; each chunk starts from
    // the state xs_jump gives for its first index, so the output is the serial
    // stream's.
    const size_t chunk_min = size_t(1) << 20;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>({(size_t)std::min(hw, 16u), (count + chunk_min - 1) / chunk_min});
    if (nt <= 1) {
        *state = xs_fill(v0, out, count);
        return KSV_OK;
    }
    const size_t per = (count + nt - 1) / nt;
    std::vector<std::thread> pool;
    size_t covered = per;           // [0, covered): chunk 0 here, the others on started threads
    try {
        for (size_t t = 1; t < nt && covered < count; t++) {
            const size_t b = t * per, e = std::min(count, b + per);
            pool.emplace_back([=] { xs_fill(xs_jump(v0, b), out + b, e - b); });
            covered = e;
        }
    } catch (...) {                 // a thread that could not start: its chunks run below
    }
    xs_fill(v0, out, per);
    if (covered < count) xs_fill(xs_jump(v0, covered), out + covered, count - covered);
    for (auto& th : pool) th.join();
    *state = xs_jump(v0, count);
    return KSV_OK;
}

int ksv_shots_init(ksv_state* s, int n, uint64_t index) {
    uint32_t copies = 0;
    int rc = shots_check(s, n, &copies);
    if (rc) return rc;
    if (index >> n) return fail(KSV_ERR_RANGE, "basis index out of range for a " + std::to_string(n) + "-qubit copy");
    HIPCHK(shots_basis(s->n, n, s->prec, s->psi, index, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return KSV_OK;
}

int ksv_shots_init_identity(ksv_state* s, int n) {
    uint32_t copies = 0;
    int rc = shots_check(s, n, &copies);
    if (rc) return rc;
    if (s->n != 2 * n) return fail(KSV_ERR_RANGE, "an identity batch of " + std::to_string(n) + "-qubit copies needs a " + std::to_string(2 * n) + "-qubit state");
    HIPCHK(shots_identity(s->n, n, s->prec, s->psi, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return KSV_OK;
}

int ksv_shots_pauli(ksv_state* s, int n, int q, const uint32_t* copy, const uint8_t* pauli, size_t count) {
    uint32_t copies = 0;
    int rc = shots_check(s, n, &copies);
    if (rc) return rc;
    if (q < 0 || q >= n) return fail(KSV_ERR_RANGE, "qubit index " + std::to_string(q) + " is out of range for a " + std::to_string(n) + "-qubit circuit");
    if (!count) return KSV_OK;
    if (!copy || !pauli) return fail(KSV_ERR_TYPE, "null buffer");
    if (count > 65535) return fail(KSV_ERR_RANGE, "at most 65535 Pauli entries per call");
    for (size_t i = 0; i < count; i++) {
        if (copy[i] >= copies) return fail(KSV_ERR_RANGE, "copy index out of range");
        if (pauli[i] < 1 || pauli[i] > 3) return fail(KSV_ERR_TYPE, "Pauli code must be 1 (X), 2 (Y) or 3 (Z)");
    }
    DevBufs b;
    uint32_t* d_copy = nullptr;
    uint8_t* d_pauli = nullptr;
    HIPCHK(b.get(&d_copy, count));
    HIPCHK(b.get(&d_pauli, count));
    HIPCHK(hipMemcpyAsync(d_copy, copy, 4 * count, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemcpyAsync(d_pauli, pauli, count, hipMemcpyHostToDevice, s->st));
    HIPCHK(shots_pauli(n, s->prec, s->psi, q, d_copy, d_pauli, (uint32_t)count, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return KSV_OK;
}

}  // extern "C"

// Per-copy P(q = 1), collapseQubit's outcome rule and the projection (with
// the reset flip); reset == nullptr: no copy resets, reset_all: every copy.
static int shots_measure_impl(ksv_state* s, int n, uint32_t copies, int q, const double* r, const uint8_t* reset,
                              bool reset_all, uint8_t* outcome) {
    DevBufs b;
    double *d_part = nullptr, *d_p1 = nullptr, *d_scale = nullptr;
    uint8_t *d_o = nullptr, *d_reset = nullptr;
    const int gmax = (int)std::min<uint64_t>(NBLK, ((1ull << (n - 1)) + 255) / 256);
    HIPCHK(b.get(&d_part, (size_t)std::max(1, gmax) * copies));
    HIPCHK(b.get(&d_p1, copies));
    HIPCHK(b.get(&d_scale, copies));
    HIPCHK(b.get(&d_o, copies));
    HIPCHK(b.get(&d_reset, copies));
    int g = 0;
    HIPCHK(shots_bit1(n, s->prec, s->psi, q, copies, NBLK, d_part, &g, s->st));
    HIPCHK(shots_final(d_part, g, copies, d_p1, s->st));
    std::vector<double> p1(copies), scale(copies);
    HIPCHK(hipMemcpyAsync(p1.data(), d_p1, 8 * copies, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    std::vector<uint8_t> rs(copies, 0);
    for (uint32_t t = 0; t < copies; t++) {       // collapseQubit's rule, as ksv_collapse
        const int o = r[t] < p1[t] ? 1 : 0;
        outcome[t] = (uint8_t)o;
        scale[t] = 1 / std::sqrt(o ? p1[t] : 1 - p1[t]);
        rs[t] = reset_all ? 1 : (reset && reset[t]) ? 1 : 0;
    }
    HIPCHK(hipMemcpyAsync(d_o, outcome, copies, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemcpyAsync(d_scale, scale.data(), 8 * copies, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemcpyAsync(d_reset, rs.data(), copies, hipMemcpyHostToDevice, s->st));
    HIPCHK(shots_project(s->n, n, s->prec, s->psi, q, d_o, d_scale, d_reset, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return KSV_OK;
}

extern "C" {

int ksv_shots_measure(ksv_state* s, int n, int q, const double* r, const uint8_t* reset, uint8_t* outcome) {
    uint32_t copies = 0;
    int rc = shots_check(s, n, &copies);
    if (rc) return rc;
    if (q < 0 || q >= n) return fail(KSV_ERR_RANGE, "qubit index " + std::to_string(q) + " is out of range for a " + std::to_string(n) + "-qubit circuit");
    if (!r || !outcome) return fail(KSV_ERR_TYPE, "null buffer");
    return shots_measure_impl(s, n, copies, q, r, reset, false, outcome);
}

int ksv_shots_run(ksv_state* s, int n, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                  const ksv_shot_event* ev, size_t nev, const uint32_t* copy, const uint8_t* pauli, size_t npauli,
                  const double* r, size_t nr, uint8_t* outcome) {
    uint32_t copies = 0;
    int rc = shots_check(s, n, &copies);
    if (rc) return rc;
    if ((nops && !ops) || (nev && !ev)) return fail(KSV_ERR_TYPE, "null buffer");
    // validate everything before touching the state
    uint64_t at = 0;
    for (size_t i = 0; i < nev; i++) {
        const ksv_shot_event& e = ev[i];
        if (e.at < at || e.at > nops) return fail(KSV_ERR_RANGE, "event positions must be ascending and within the tape");
        at = e.at;
        if (e.q < 0 || e.q >= n) return fail(KSV_ERR_RANGE, "qubit index " + std::to_string(e.q) + " is out of range for a " + std::to_string(n) + "-qubit circuit");
        if (e.kind == KSV_SHOT_PAULI) {
            if (!copy || !pauli || e.off > npauli || e.count > npauli - e.off || e.count > 65535)
                return fail(KSV_ERR_RANGE, "Pauli event range out of bounds");
            for (uint64_t j = e.off; j < e.off + e.count; j++) {
                if (copy[j] >= copies) return fail(KSV_ERR_RANGE, "copy index out of range");
                if (pauli[j] < 1 || pauli[j] > 3) return fail(KSV_ERR_TYPE, "Pauli code must be 1 (X), 2 (Y) or 3 (Z)");
            }
        } else if (e.kind == KSV_SHOT_MEASURE || e.kind == KSV_SHOT_RESET) {
            if (!r || !outcome || e.off > nr || copies > nr - e.off) return fail(KSV_ERR_RANGE, "draw range out of bounds");
        } else {
            return fail(KSV_ERR_TYPE, "unknown shot event kind");
        }
    }
    // Per-copy mode: a copy of at most 13 qubits fits one tile, so each gate
    // segment is one pass (register rounds for 11..13 qubits, LDS below) in which
    // every workgroup runs one copy, the Pauli events inside it as ops on their
    // copies' tiles; only measurements and resets split segments.
    // Otherwise every event ends a segment and Paulis are launches of their own.
    const bool copy_mode = s->kernel <= 2 && n <= 13;
    const size_t nwords = (copies + 31) / 32;
    // plan every segment on the host, then one upload
    struct Action { int seg; int ev; };             // run program `seg`, or event `ev`
    std::vector<Action> acts;
    std::vector<std::unique_ptr<ksv_program>> seg;
    std::vector<std::vector<size_t>> segoff;
    std::vector<char> blob;
    std::vector<CopyPauli> pend;
    size_t cur = 0;
    auto close = [&](size_t upto) -> int {
        if (upto <= cur && pend.empty()) return KSV_OK;
        std::unique_ptr<ksv_program> p(new ksv_program());
        int prc;
        if (copy_mode) {
            std::string err;
            prc = plan_copies(*p, s->n, n, s->prec, ops + cur, upto - cur, mats, nmats, pend, s->prune, err);
            if (prc) return fail(prc, err);
        } else {
            prc = plan_into(*p, s->n, s->prec, ops + cur, upto - cur, mats, nmats, s->kernel, default_tile(s),
                            s->rtile, s->fuse != 0, s->prune, s->btile, s->bq);
            if (prc) return prc;
        }
        p->dev = s->dev;
        p->bthreads = s->bthreads;
        p->bpipe = s->bpipe;
        p->borrowed = true;
        std::vector<size_t> off;
        std::vector<char> one;
        program_blob(s, *p, one, off);
        const size_t base = (blob.size() + 255) & ~(size_t)255;
        blob.resize(base + one.size());
        std::memcpy(blob.data() + base, one.data(), one.size());
        for (size_t& o : off)
            if (o != SIZE_MAX) o += base;
        acts.push_back({(int)seg.size(), -1});
        seg.push_back(std::move(p));
        segoff.push_back(off);
        cur = upto;
        pend.clear();
        return KSV_OK;
    };
    for (size_t i = 0; i < nev; i++) {
        const ksv_shot_event& e = ev[i];
        if (copy_mode && e.kind == KSV_SHOT_PAULI) {
            // one conditional op per Pauli kind present, over a bitmap of its copies
            for (int pk = 1; pk <= 3; pk++) {
                CopyPauli c;
                c.at = (size_t)e.at - cur;
                c.q = e.q;
                c.pauli = pk;
                c.words.assign(nwords, 0u);
                bool any = false;
                for (uint64_t j = e.off; j < e.off + e.count; j++)
                    if (pauli[j] == pk) {
                        c.words[copy[j] >> 5] |= 1u << (copy[j] & 31);
                        any = true;
                    }
                if (any) pend.push_back(std::move(c));
            }
            continue;
        }
        if ((rc = close((size_t)e.at))) return rc;
        acts.push_back({-1, (int)i});
    }
    if ((rc = close(nops))) return rc;
    DevBufs b;
    char* d_blob = nullptr;
    uint32_t* d_copy = nullptr;
    uint8_t* d_pauli = nullptr;
    if (!blob.empty()) {
        HIPCHK(b.get(&d_blob, blob.size()));
        HIPCHK(hipMemcpy(d_blob, blob.data(), blob.size(), hipMemcpyHostToDevice));
    }
    for (size_t i = 0; i < seg.size(); i++) bind_blob(*seg[i], d_blob, segoff[i]);
    if (npauli && !copy_mode) {
        HIPCHK(b.get(&d_copy, npauli));
        HIPCHK(b.get(&d_pauli, npauli));
        HIPCHK(hipMemcpy(d_copy, copy, 4 * npauli, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(d_pauli, pauli, npauli, hipMemcpyHostToDevice));
    }
    for (const Action& act : acts) {
        if (act.seg >= 0) {
            rc = run_program_on(s, seg[act.seg].get(), s->psi, false);
            if (rc) return rc;
            continue;
        }
        const ksv_shot_event& e = ev[act.ev];
        if (e.kind == KSV_SHOT_PAULI) {
            if (e.count)
                HIPCHK(shots_pauli(n, s->prec, s->psi, e.q, d_copy + e.off, d_pauli + e.off, (uint32_t)e.count, s->st));
        } else {
            rc = shots_measure_impl(s, n, copies, e.q, r + e.off, nullptr, e.kind == KSV_SHOT_RESET, outcome + e.off);
            if (rc) return rc;
        }
    }
    HIPCHK(hipStreamSynchronize(s->st));
    return KSV_OK;
}

int ksv_shots_sample(ksv_state* s, int n, const double* r, uint64_t* out) {
    uint32_t copies = 0;
    int rc = shots_check(s, n, &copies);
    if (rc) return rc;
    if (!r || !out) return fail(KSV_ERR_TYPE, "null buffer");
    rc = ws_reserve(s, 2 * wsz(8 * (size_t)copies));
    if (rc) return rc;
    Carve c{(char*)s->ws};
    double* d_r = c.take<double>(copies);
    uint64_t* d_out = c.take<uint64_t>(copies);
    HIPCHK(hipMemcpyAsync(d_r, r, 8 * copies, hipMemcpyHostToDevice, s->st));
    // sampleSV takes every entry of the shot's map: the non-zero amplitudes
    HIPCHK(shots_seq_sample(n, s->prec, s->psi, copies, 0.0, d_r, d_out, s->st));
    HIPCHK(hipMemcpyAsync(out, d_out, 8 * copies, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return KSV_OK;
}

int ksv_get_stats(ksv_state* s, ksv_stats* out) {
    if (!s || !out) return fail(KSV_ERR_HANDLE, "null state");
    *out = s->stats;
    return KSV_OK;
}

int ksv_reset_stats(ksv_state* s) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    std::memset(&s->stats, 0, sizeof s->stats);
    return KSV_OK;
}

}  // extern "C"

// Helpers for the adjoint translation unit.
namespace ksv {
int adj_fail(int code, const std::string& msg) { return fail(code, msg); }

int adj_context(ksv_state* s, int* n, int* prec, void** psi, void** psi2, hipStream_t* st, double** part,
                int* prune, int* tile, int* atile, int* athreads) {
    if (!s) return fail(KSV_ERR_HANDLE, "null state");
    int rc = ensure_psi2(s);
    if (rc) return rc;
    *n = s->n;
    *prec = s->prec;
    *psi = s->psi;
    *psi2 = s->psi2;
    *st = s->st;
    *part = s->d_part;
    *prune = s->prune;
    *tile = default_tile(s);
    *atile = s->atile ? s->atile : (s->prec == KSV_F64 ? ADJ_TILE_F64 : ADJ_TILE_F32);
    *athreads = s->athreads;
    return KSV_OK;
}

int adj_run_tape(ksv_state* s, const ksv_op* ops, size_t nops, const double* mats, size_t nmats, int prune,
                 void* psi) {
    ksv_program* p = nullptr;
    int rc = build_program(s, ops, nops, mats, nmats, prune, false, &p);
    if (rc) return rc;
    rc = run_program_on(s, p, psi);
    ksv_program_destroy(p);
    return rc;
}
}  // namespace ksv
