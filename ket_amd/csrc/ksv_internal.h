// Internal declarations shared by the translation units of libketsv.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <complex>
#include <string>
#include <vector>

#include "../../include/ketsv.h"

namespace ksv {

using cd = std::complex<double>;

// ---------------------------------------------------------------------------
// Device op: one step of a fused pass, expressed in tile-local bit positions.
// ---------------------------------------------------------------------------
enum DKind : int32_t {
    DK_M1 = 0,   // 2x2 on local bit t[0]
    DK_M2 = 1,   // 4x4 on local bits t[0] (MSB), t[1]
    DK_MK = 2,   // 2^k x 2^k, k <= 4, on local bits t[0..k) (t[0] = MSB)
    DK_DIAG = 3, // diagonal over k <= 6 qubits; t[i] >= 0 local bit, t[i] < 0 global bit -(t+1)
    DK_PERMX = 4,// X on local bit t[0] (CNOT / Toffoli / X): pure permutation
    DK_SWAP = 5, // exchange local bits t[0], t[1] (SWAP / CSWAP): pure permutation
};

struct DOp {
    int32_t kind;
    int32_t k;        // targets (DIAG: qubits of the diagonal)
    int32_t nins;     // zero bits inserted to enumerate the op's elements
    int32_t mat;      // offset (complex entries) into the program matrix array
    int8_t  t[8];     // targets
    int8_t  ins[8];   // ascending local positions of the inserted bits
    uint32_t lctl;    // local control bits (set after insertion)
    uint32_t pad;
    uint64_t gctl;    // global control mask: the op applies to a tile only if all set
};
static_assert(sizeof(DOp) == 48, "DOp layout");

constexpr int KMAX_TILE = 13;      // widest tile (2^13 complex128 = 128 KiB LDS)
constexpr int MK_MAX = 4;          // widest dense in-pass matrix
constexpr int DIAG_MAX = 6;        // widest in-pass diagonal
constexpr int BIG_MAX = 12;        // widest standalone unitary

// ---------------------------------------------------------------------------
// Register-tile ("wave") pass: one wavefront owns 2^(6+r) amplitudes, lane l
// register j holding the amplitude whose tile-local lane bits are l and
// register-slot bits are j.  Ops act on register slots; a qubit on a lane
// bit is brought into a slot by W_LSWAP (cross-lane exchange).
// ---------------------------------------------------------------------------
enum WKind : uint8_t {
    W_M1 = 0,     // 2x2 on slot a
    W_M2 = 1,     // 4x4 on slots a (MSB), b
    W_PHASE = 2,  // multiply by mats[mat] where every control bit is 1
    W_DIAG = 3,   // diagonal over k qubits at loc[]
    W_PERMX = 4,  // X on slot a (controlled)
    W_SWAPR = 5,  // exchange slots a, b (controlled)
    W_LSWAP = 6,  // exchange lane bit a with slot b
};

struct WOp {
    uint8_t kind;
    uint8_t a, b;
    uint8_t k;
    uint8_t loc[6];    // W_DIAG: 0..5 lane bit, 8+s slot, 16+g global bit g; loc[0] = MSB of the row
    uint16_t pad;
    uint32_t lctl;     // lane-bit controls (mask of lane index bits)
    uint32_t rctl;     // slot controls (mask of register index bits)
    int32_t mat;       // complex offset into the program matrix array
    uint64_t gctl;     // global controls
};
static_assert(sizeof(WOp) == 32, "WOp layout");

constexpr int WLANE_BITS = 6;

// Byte i of a packed word holds the state bit of lane bit / register slot /
// tile bit i (arrays inside a by-value kernel argument would be copied to
// scratch memory, packed words stay in scalar registers).
struct WaveArgs {
    void* psi;
    const WOp* ops;
    const void* mats;
    int nops;
    int k;               // tile qubits = 6 + r
    uint64_t lane_in, lane_out;   // state bit of each lane bit at load / store
    uint64_t reg_in, reg_out;     // state bit of each register slot at load / store
    uint64_t tq_lo, tq_hi;        // tile state bits ascending (to build the base)
};

struct PassArgs {
    void* psi;
    const DOp* ops;
    const void* mats;
    int nops;
    int k;          // tile qubits of this pass
    int prune;
    int pad;
    int8_t tq[16];  // physical bit of tile-local bit i (ascending)
};

// ---------------------------------------------------------------------------
// Host-side logical op (a ket FlatOp lowered to one of four classes).
// ---------------------------------------------------------------------------
enum LClass { L_MAT = 0, L_DIAG = 1, L_PERMX = 2, L_SWAP = 3 };

struct LOp {
    int cls;
    std::vector<int> tg;     // targets, MSB first
    std::vector<int> ctl;    // controls (fire on 1)
    std::vector<cd> m;       // dense dim x dim (L_MAT) or dim diagonal (L_DIAG)
    uint64_t touch = 0;      // every qubit the op reads
    uint64_t need = 0;       // qubits that must be tile-local
    int src = -1;            // tape index
    bool big = false;        // applied by the standalone unitary kernel
};

struct Step {
    bool big = false;        // standalone unitary step
    bool wave = false;       // register-tile pass (WOps) instead of an LDS pass (DOps)
    int k = 0;               // tile qubits
    int8_t tq[16] = {0};
    int op_begin = 0, op_count = 0;   // DOps (LDS pass) or WOps (wave pass) of this step
    int big_op = -1;         // index into Program::big_ops
    int rlog = 0;            // wave pass: log2 registers per lane
    int8_t lane_in[6] = {0}, lane_out[6] = {0};
    int8_t reg_in[8] = {0}, reg_out[8] = {0};
};

struct BigOp {
    std::vector<int> tg;
    int mat = 0;             // complex offset into program mats
};

struct Program {
    int n = 0;
    int precision = KSV_F64;
    int prune = 1;
    int tile = 12;
    uint64_t tape_ops = 0;
    std::vector<Step> steps;
    std::vector<DOp> dops;
    std::vector<WOp> wops;
    std::vector<BigOp> big_ops;
    int rlog = 0;            // > 0: wave passes with 2^rlog registers per lane
    WOp* d_wops = nullptr;
    std::vector<cd> mats;
    // device copies
    DOp* d_ops = nullptr;
    void* d_mats = nullptr;
    int dev = 0;
};

// schedule.cpp
int lower_tape(int n, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
               std::vector<LOp>& out, std::string& err);
void schedule(int n, int tile, bool fuse, std::vector<LOp>& lops, Program& prog);
void schedule_wave(int n, int rlog, bool fuse, std::vector<LOp>& lops, Program& prog);

// kernels (k_pass.hip / k_reduce.hip)
hipError_t launch_pass(const Program& p, const Step& s, void* psi, hipStream_t st);
hipError_t launch_big(const Program& p, const Step& s, void* psi, int prune, hipStream_t st);
hipError_t launch_wave(const Program& p, const Step& s, void* psi, hipStream_t st);

struct Scratch {
    double* partial = nullptr;     // reduction partials
    size_t partial_n = 0;
};

hipError_t red_norm2(int n, int prec, const void* psi, double* d_out, double* d_part, int nblk,
                     hipStream_t st);
hipError_t red_marginals(int n, int prec, const void* psi, double* d_out, double* d_part, int nblk,
                         hipStream_t st);
hipError_t red_1q(int n, int prec, const void* psi, int q, double* d_out, double* d_part, int nblk,
                  hipStream_t st);
hipError_t red_pauli(int n, int prec, const void* psi, const uint64_t* d_x, const uint64_t* d_z,
                     int nterms, double* d_out, double* d_part, int nblk, hipStream_t st);
hipError_t red_parity(int n, int prec, const void* psi, uint64_t mask, double thresh, double* d_out,
                      double* d_part, int nblk, hipStream_t st);
hipError_t red_bit1(int n, int prec, const void* psi, int q, double* d_out, double* d_part, int nblk,
                    hipStream_t st);
hipError_t k_init_basis(int n, int prec, void* psi, uint64_t idx, hipStream_t st);
hipError_t k_project(int n, int prec, void* psi, int q, int outcome, double scale, hipStream_t st);
hipError_t k_copy(int n, int prec, void* dst, const void* src, hipStream_t st);

// probability compaction / sampling / top-k
constexpr int CHUNK_LOG = 12;     // amplitudes per block chunk in scans
hipError_t chunk_counts(int n, int prec, const void* psi, double thresh, uint32_t* d_cnt,
                        hipStream_t st);
hipError_t chunk_compact(int n, int prec, const void* psi, double thresh, const uint64_t* d_off,
                         uint64_t* d_idx, double* d_p, uint64_t cap, hipStream_t st);
hipError_t chunk_sums(int n, int prec, const void* psi, double thresh, double* d_sum,
                      hipStream_t st);
hipError_t sample_search(int n, int prec, const void* psi, double thresh, const double* d_chunk_cum,
                         int64_t nchunks, const double* d_r, uint64_t shots, uint64_t last_idx,
                         uint64_t* d_out, hipStream_t st);
hipError_t key_histogram(int n, int prec, const void* psi, uint64_t prefix, int shift_hi,
                         int shift, uint32_t* d_hist, hipStream_t st);
hipError_t last_nonzero_chunk(int n, int prec, const void* psi, double thresh, int64_t chunk,
                              uint64_t* d_out, hipStream_t st);
hipError_t select_ge(int n, int prec, const void* psi, uint64_t key_thr, uint64_t* d_idx,
                     double* d_p, unsigned int* d_count, uint64_t cap, hipStream_t st);

}  // namespace ksv
