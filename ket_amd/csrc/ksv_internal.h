// Internal declarations shared by the translation units of libketsv.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <array>
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
    uint32_t cond;    // 0, or 1 + word offset of a bitmap over copies: the op applies to a
                      // tile only if the bit of its copy (base >> nsys) is set
    uint64_t gctl;    // global control mask: the op applies to a tile only if all set
};
static_assert(sizeof(DOp) == 48, "DOp layout");

constexpr int KMAX_TILE = 13;      // widest tile (2^13 complex128 = 128 KiB LDS)
constexpr int MK_MAX = 4;          // widest dense in-pass matrix
constexpr int DIAG_MAX = 6;        // widest in-pass diagonal
constexpr int BIG_MAX = 12;        // widest standalone unitary

// ---------------------------------------------------------------------------
// Register-round pass: a workgroup of 2^TB threads owns a tile of 2^(TB+4)
// amplitudes; every thread holds 16 of them in registers.  In each round
// four tile qubits are "slots" (bits of the register index j) and the other
// TB tile qubits are thread bits.  Gates act on slots; an R_EXCH op
// transposes the tile through LDS into the next round's assignment.
// ---------------------------------------------------------------------------
enum RKind : uint8_t {
    R_M1 = 0,     // 2x2 on slot a; matrix inline
    R_DENSE = 1,  // 2^k x 2^k (k = 2..4) on slots loc[0] (MSB) .. loc[k-1], through LDS; matrix at `mat`
    R_DIAG = 2,   // diagonal over k <= 6 qubits at loc[] (k = 0: a phase under the controls);
                  // k <= 2: entries inline, else at `mat`
    R_PERMX = 3,  // X on slot a
    R_SWAP = 4,   // exchange slots a, b; a > b
    R_EXCH = 5,   // LDS transpose; its XRec inline
};

constexpr int RSLOTS = 4;          // qubits held in registers per round
constexpr int RREGS = 1 << RSLOTS; // amplitudes per thread

// LDS address (in amplitudes) of tile index x under an exchange's swizzle is
// linear in x, so it is the XOR of one word per set bit: wt/rt per thread bit
// of the writing/reading layout, wr/rr per slot.
struct XRec {
    uint16_t wt[12];
    uint16_t wr[4];
    uint16_t rt[12];
    uint16_t rr[4];
};
static_assert(sizeof(XRec) == 64, "XRec layout");

// 128-byte op record: a 32-byte header and an inline payload, so one
// scalar fetch brings both the op and its (small) matrix.
struct ROp {
    uint8_t kind;
    uint8_t a, b;
    uint8_t k;
    uint8_t loc[6];    // R_DIAG, MSB first: 0..15 thread bit, 16+s slot s, 32+g physical bit g outside the tile;
                       // R_DENSE: the target slots, MSB first
    uint16_t flags;    // ROP_COND: applied only to tiles whose copy (base >> nsys) has its bit
                       // set in the bitmap at word `mat` (R_M1, R_PERMX, R_DIAG with k <= 2)
    uint32_t tctl;     // thread-bit controls (mask over the thread index)
    uint32_t rctl;     // slot controls (mask over the register index)
    int32_t mat;       // complex offset into the program matrices (R_DENSE, R_DIAG with k > 2)
    uint64_t gctl;     // controls on physical bits outside the tile
    uint64_t pay[12];  // R_M1: 4 complex entries, R_DIAG k <= 2: 2^k entries, in the state's precision;
                       // R_EXCH: XRec
};
static_assert(sizeof(ROp) == 128, "ROp layout");
constexpr uint16_t ROP_COND = 1;
// R_M1 matrices whose entries are each purely real or purely imaginary: ket's
// complex products with a zero part reduce to single real products (exactly,
// up to the sign of zero results, which ket's 1e-14 prune then clears), so the
// kernel may skip the vanishing terms.  ROP_REAL: all four entries real (ry,
// h, z, ...); ROP_RXF: real diagonal, imaginary off-diagonal (rx, y).
constexpr uint16_t ROP_REAL = 2;
constexpr uint16_t ROP_RXF = 4;

// ---------------------------------------------------------------------------
// Fused-block pass: the circuit is fused on the host into dense blocks of
// BQ (4 or 5) qubits; a pass stages a K-qubit tile in LDS and applies its
// blocks as complex GEMMs on MFMA with Gauss's three real products:
// Yr = Ur Xr - Ui Xi, Yi = (Ur + Ui)(Xr + Xi) - Ur Xr - Ui Xi,
// rows = the block's local index (bit i = qubit q[i]), columns = the other
// K - BQ tile bits.
// ---------------------------------------------------------------------------
struct BRec {
    uint16_t ar[8];    // LDS address (swizzled) contribution of row bit i, i < BQ
    uint16_t ac[16];   // ... of column bit i, i < K - BQ
    int64_t frag;      // offset (elements) of the block's A fragments: [mb][kk][lane]
    int64_t flags;     // BREC_DIAG: every absorbed op is diagonal; the fragments hold the
                       // diagonal and the block is a phase per row instead of a GEMM.
                       // BREC_MONO (4-qubit blocks): every absorbed op is diagonal or a
                       // permutation, so U has one entry per column: row r goes to row
                       // pi(r) times U[pi(r)][r]; pi(r) is nibble r of ar[4..7] (unused
                       // row words of a 4-qubit block) and the fragments hold the phases
};
constexpr int64_t BREC_DIAG = 1;
constexpr int64_t BREC_MONO = 2;
static_assert(sizeof(BRec) == 64, "BRec layout");

struct BPassArgs {
    void* psi;
    const BRec* recs;
    const void* frags;
    int nblk;
    int k;
    uint64_t tq_lo, tq_hi;      // physical bit of tile bit i, ascending
    uint64_t swz;               // 4-bit XOR word of tile bit p (p >= 4) at bits 4(p-4)
};

// Byte i of a packed word holds a physical state bit (arrays inside a
// by-value kernel argument would be copied to scratch; packed words stay in
// scalar registers).
struct RPassArgs {
    void* psi;
    const ROp* ops;
    const void* mats;
    int nops;
    int k;                        // tile qubits
    uint64_t ntiles;              // 2^(n - k)
    uint64_t tq_lo, tq_hi;        // physical bit of tile bit i, ascending
    uint64_t ld_t_lo, ld_t_hi;    // physical bit of thread bit i at load
    uint64_t st_t_lo, st_t_hi;    // ... at store
    uint32_t ld_r, st_r;          // physical bit of slot s at load / store
    const uint32_t* rbits;        // word w of op g's copy bitmap at rbits[w * rstride + g] (null: no conditional ops)
    uint32_t rstride;             // ops of the whole program
    uint32_t op_base;             // program index of ops[0]
    int nsys;                     // copy of a tile = base >> nsys
};

struct PassArgs {
    void* psi;
    const DOp* ops;
    const void* mats;
    int nops;
    int k;          // tile qubits of this pass
    int prune;
    int nsys;       // copy of a tile = base >> nsys (conditional ops)
    int8_t tq[16];  // physical bit of tile-local bit i (ascending)
    const uint32_t* bits;   // copy bitmaps of conditional ops
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
    int cond = 0;            // 0, or 1 + word offset of the copy bitmap it is applied under (DOp::cond)
};

struct Step {
    bool big = false;        // standalone unitary step
    bool reg = false;        // register-round pass (ROps) instead of an LDS pass (DOps)
    bool blk = false;        // fused-block MFMA pass (BRecs)
    int bq = 0;              // fused-block pass: block width
    uint64_t swz = 0;        // fused-block pass: LDS swizzle word
    int k = 0;               // tile qubits
    int8_t tq[16] = {0};
    int op_begin = 0, op_count = 0;   // DOps (LDS pass) or ROps (register-round pass)
    int big_op = -1;         // index into Program::big_ops
    int exchanges = 0;       // register-round pass: LDS transposes
    int8_t ld_t[12] = {0}, st_t[12] = {0};   // physical bit of thread bit i at load / store
    int8_t ld_r[4] = {0}, st_r[4] = {0};     // physical bit of slot s at load / store
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
    std::vector<ROp> rops;
    std::vector<BigOp> big_ops;
    int rtile = 0;           // > 0: tile qubits of the register-round passes
    ROp* d_rops = nullptr;
    std::vector<BRec> brecs;         // fused-block passes
    std::vector<cd> bmats;           // their dense block matrices (2^bq x 2^bq each, row-major)
    std::vector<int64_t> bmat_off;   // per BRec: offset of its matrix in bmats
    BRec* d_brecs = nullptr;
    void* d_frags = nullptr;
    int bthreads = 256;              // fused-block passes: workgroup size (256 or 512)
    int bpipe = 1;                   // fused-block passes: pipelined tile walk (k_bpass_pipe) where it pays
    int nsys = 0;                    // batched copies: qubits of one copy (conditional ops)
    std::vector<uint32_t> bits;      // copy bitmaps of conditional ops
    uint32_t* d_bits = nullptr;
    // register-round copy programs: the same bitmaps by op, word w of rops[g]'s
    // at rbits[w * rops.size() + g] (all ones for unconditional ops), so that a
    // workgroup reads the words of consecutive ops from one row
    std::vector<uint32_t> rbits;
    uint32_t* d_rbits = nullptr;
    std::vector<cd> mats;
    // re-parameterisation (ksv_program_set_matrix): the tape, its lowered
    // ops and where each lowered op landed
    std::vector<ksv_op> tape;
    std::vector<LOp> lops;
    std::vector<std::vector<int>> lop_of_src;   // tape op -> lowered ops
    std::vector<int> rop_of_lop;                // register-round record of a lowered op, or -1
    std::vector<int> blk_of_lop;                // fused block (BRec index) of a lowered op, or -1
    std::vector<int64_t> mat_of_lop;            // LDS-pass / wide-unitary steps: offset of the op's matrix in mats
    std::vector<std::vector<int>> blk_lops;     // per BRec: its lowered ops in application order
    std::vector<std::vector<int>> blk_q;        // per BRec: its qubits, row bit i = qubit q[i]
    std::vector<int> blk_bq;                    // per BRec: block width
    // device copies
    DOp* d_ops = nullptr;
    void* d_mats = nullptr;
    int dev = 0;
};

// Host side of ksv_program_set_matrix: re-lower tape op `op` with its new
// matrix and recompute what depends on it.  Fills the ROps, BRecs and
// complex ranges of `mats` that must be uploaded again.
int update_program(Program& p, size_t op, const double* mat, size_t nmat, std::vector<int>& rops,
                   std::vector<int>& blocks, std::vector<std::pair<size_t, size_t>>& mat_ranges, std::string& err);

// schedule.cpp
int lower_tape(int n, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
               std::vector<LOp>& out, std::string& err);
void schedule(int n, int tile, bool fuse, std::vector<LOp>& lops, Program& prog);
void schedule_rounds(int n, int rtile, bool fuse, std::vector<LOp>& lops, Program& prog);
void schedule_blocks(int n, int btile, int bq, std::vector<LOp>& lops, Program& prog);
// A fragments of one block for MFMA 16x16x4 (2^(2 bq + 3) / 16 values, the
// size of the real embedding's fragments): Ur, Ui and Ur + Ui of the
// 2^bq x 2^bq matrix u (row-major), each [mb][kk][lane]; for a diagonal
// block, the diagonal (fragments 2t / 2t + 1).
void block_fragments(const cd* u, int bq, int64_t flags, double* out);
// Lower and schedule a tape into `p` (host side only).  kernel: KSV_OPT_PASS_KERNEL
// value; tile / rtile / btile: LDS, register-round and fused-block tile widths;
// bq: fused-block width (0 = defaults).
int plan_program(Program& p, int n, int prec, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                 int kernel, int tile, int rtile, bool fuse, int prune, std::string& err, int btile = 0,
                 int bq = 0);

// ---------------------------------------------------------------------------
// Adjoint backward pass: a workgroup stages the same K-qubit tile of phi and
// of lambda in LDS and walks a run of the inverse tape.  Every op is applied
// to both tiles; an op with parameter records also adds, per record j,
// Re <lambda| dG_j |G^dag phi> over the tile into a per-workgroup slot.
// Ops on disjoint qubits commute with each other's derivatives, so a pass
// may take an op ahead of skipped ops it does not touch.
// ---------------------------------------------------------------------------
enum AKind : int32_t {
    AK_DENSE = 0,   // 2^k x 2^k (k <= 3) G^dag on local bits t[0] (MSB) .. t[k-1]
    AK_X = 1,       // X on local bit t[0] (CNOT, Toffoli; self-inverse)
    AK_SWAP = 2,    // exchange local bits t[0], t[1] (SWAP, CSWAP; self-inverse)
    AK_GRP = 3,     // the next `run` records (one-qubit chains on distinct targets
                    // t[0] .. t[k-1], k = 2 or 3: X heads with local controls and
                    // uncontrolled one-qubit ops) run in one LDS round trip, each
                    // thread holding the 2^k elements of its groups over those targets
};

struct AOp {
    uint64_t gctl;     // controls on physical bits outside the tile
    int32_t kind;
    int32_t k;         // targets
    int32_t nins;      // local bits inserted to enumerate the op's elements
    int32_t nrec;      // parameter records (<= 3)
    uint32_t lctl;     // local control bits (set after insertion)
    int32_t mat;       // AK_DENSE: complex offset of G^dag in the pass matrices
    int8_t t[4];
    int8_t ins[8];     // ascending local positions of targets and local controls
    int32_t dmat[3];   // complex offset of each record's dG (2^k x 2^k), or -1 - P for a
                       // rotation about Pauli P (0 X, 1 Y, 2 Z; one qubit): dG G^dag = (-i/2) P
    int16_t slot[3];   // pass-local slot of each record
    uint8_t flags;     // AOP_DIAG: G^dag and every dG are diagonal
    uint8_t run;       // AK_DENSE: ops of the chain this op heads (0: inside a chain)
};
static_assert(sizeof(AOp) == 64, "AOp layout");
constexpr uint8_t AOP_DIAG = 1;
// A chain is a run of AK_DENSE ops on the same targets (same order) and the
// same controls, adjacent in the pass: the kernel applies the whole chain to
// each element group while it sits in registers.
constexpr int ACHAIN_1Q = 4;        // longest chain of 1-qubit ops
constexpr int ACHAIN_2Q = 2;        // ... of 2-qubit ops
constexpr int AGRP_MAX = 3;         // one-qubit chains sharing one LDS round trip

constexpr int ADJ_SLOTS = 128;      // parameter records per pass
constexpr int ADJ_DENSE_MAX = 3;    // widest dense op inside a pass
constexpr int ADJ_TILE_F64 = 11;    // tile qubits: phi and lambda tiles take 2 x 32 KiB of LDS
constexpr int ADJ_TILE_F32 = 12;

struct AdjStep {
    bool big = false;        // one wide inverse op through the general engine (bops[bop])
    int k = 0;
    int8_t tq[16] = {0};     // physical bit of tile bit i, ascending
    int op_begin = 0, op_count = 0;       // AOps
    int slot_begin = 0, slot_count = 0;   // slot_rec
    int bop = -1;
};

struct AdjPlan {
    int k = 0;
    std::vector<AdjStep> steps;
    std::vector<AOp> aops;
    std::vector<double> mats;        // interleaved complex: G^dag and dG matrices
    std::vector<int32_t> slot_rec;   // global parameter record of each pass slot
    std::vector<ksv_op> bops;        // wide inverse ops (KSV_UNITARY with conj-transposed matrix)
    std::vector<double> bmats;
    std::vector<int32_t> aop_src;    // tape op of every AOp
    std::vector<std::array<int32_t, 3>> aop_rec;   // its parameter records (-1: none)
};

// plan.mats rebuilt for another matrix array of the same tape (same layout, so
// the plan's AOps stay valid); KSV_ERR_TYPE if an op the plan treats as
// diagonal is not diagonal under the new matrices.
int adjoint_plan_mats(const AdjPlan& plan, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                      const ksv_param* params, std::vector<double>& out, std::string& err);
// Clear the diagonal flag of every AOp whose G^dag or dG is not diagonal under
// `mats` (a plan made for one matrix array, used for several).
void adjoint_plan_undiag(AdjPlan& plan, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                         const ksv_param* params);
// The inverse tape G_N^dag, ..., G_1^dag (csrswap as a 3-qubit unitary): planned
// with plan_adjoint, its passes apply G_1, ..., G_N, the forward evolution.
int invert_tape(const ksv_op* ops, size_t nops, const double* mats, size_t nmats, std::vector<ksv_op>& iops,
                std::vector<double>& imats, std::string& err);

// adjplan.cpp: the inverse of `ops` split into tile passes of K qubits.
// params[r].op names the tape op of record r (already validated).  With
// `pauli`, records of one-qubit Pauli rotations carry no derivative matrix
// (AOp::dmat); a plan whose matrices are refilled for other parameter sets
// (adjoint_plan_mats) is made without.
int plan_adjoint(int n, int K, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                 const ksv_param* params, size_t nrec, AdjPlan& plan, std::string& err, bool pauli = false);

// Batched copies (ksv_shots_run): a copy of nsys qubits fits one LDS tile, so one
// LDS pass over the ntot-qubit state runs the whole gate segment on every copy,
// each workgroup one copy.  A Pauli event (copies given by a bitmap of
// ceil(2^(ntot - nsys) / 32) words) is an op applied only to the tiles of its
// copies; it goes before segment op `at`.
struct CopyPauli {
    size_t at;
    int q;
    int pauli;                      // 1 X, 2 Y, 3 Z
    std::vector<uint32_t> words;
};
int plan_copies(Program& p, int ntot, int nsys, int prec, const ksv_op* ops, size_t nops, const double* mats,
                size_t nmats, const std::vector<CopyPauli>& paulis, int prune, std::string& err);

// kernels (k_pass.hip / k_reduce.hip)
hipError_t launch_pass(const Program& p, const Step& s, void* psi, hipStream_t st);
hipError_t launch_big(const Program& p, const Step& s, void* psi, int prune, hipStream_t st);
hipError_t launch_rpass(const Program& p, const Step& s, void* psi, hipStream_t st);
hipError_t launch_bpass(const Program& p, const Step& s, void* psi, hipStream_t st);

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

// batched trajectories (k_shots.hip): copies of an n-qubit state inside an
// ntot-qubit buffer
hipError_t shots_basis(int ntot, int n, int prec, void* psi, uint64_t idx, hipStream_t st);
hipError_t shots_identity(int ntot, int n, int prec, void* psi, hipStream_t st);
hipError_t shots_pauli(int n, int prec, void* psi, int q, const uint32_t* d_copy, const uint8_t* d_pauli,
                       uint32_t count, hipStream_t st);
hipError_t shots_bit1(int n, int prec, const void* psi, int q, uint32_t copies, int nblk, double* d_part,
                      int* grid_x, hipStream_t st);
hipError_t shots_final(const double* d_part, int nblk, uint32_t copies, double* d_out, hipStream_t st);
hipError_t shots_project(int ntot, int n, int prec, void* psi, int q, const uint8_t* d_out, const double* d_scale,
                         const uint8_t* d_reset, hipStream_t st);

hipError_t shots_seq_sample(int n, int prec, const void* psi, uint32_t copies, double thresh, const double* d_r,
                            uint64_t* d_out, hipStream_t st);

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
// ket's sequential CDF (exact running sums, first/last entry) and the draw search on it
constexpr int SEQ_CDF_MAX_QUBITS = 24;   // one wave walks 2^n amplitudes; ket's own path stops at 2^24 entries
hipError_t seq_cdf(int n, int prec, const void* psi, double thresh, double* d_cum, unsigned long long* d_ends,
                   hipStream_t st);
hipError_t cdf_search(const double* d_cum, const unsigned long long* d_ends, const double* d_r, uint64_t shots,
                      uint64_t* d_out, hipStream_t st);
hipError_t key_histogram(int n, int prec, const void* psi, uint64_t prefix, int shift_hi,
                         int shift, uint32_t* d_hist, hipStream_t st);
hipError_t last_nonzero_chunk(int n, int prec, const void* psi, double thresh, int64_t chunk,
                              uint64_t* d_out, hipStream_t st);
hipError_t select_ge(int n, int prec, const void* psi, uint64_t key_thr, uint64_t* d_idx,
                     double* d_p, unsigned int* d_count, uint64_t cap, hipStream_t st);

}  // namespace ksv
