/*
 * ketsv.h — C ABI of the MI355X dense statevector backend for ket.
 *
 * One shared library (libketsv.so, gfx950) replaces ket's sparse
 * Map<bigint, Complex> statevector path.  Every entry point is plain C:
 * pointers, sizes and status codes, no C++ or torch types.
 *
 * Index convention (unchanged from ket, src/statevector.ts:4-5): bit q of a
 * state index is qubit q (q0 = LSB).  Public q0-leftmost bitstrings are
 * reversed by the host before they reach this ABI.
 * Multi-qubit matrices are row-major 2^k x 2^k complex, interleaved
 * (re, im) doubles; the first qubit operand is the MSB of the local index
 * (src/statevector.ts:32, :104-109, :236-238).
 *
 * Status codes map onto the JavaScript error classes ket throws:
 *   KSV_ERR_RANGE -> RangeError (qubit index / size / capacity,
 *                    src/circuit.ts:1114-1117, :1469-1471)
 *   KSV_ERR_TYPE  -> TypeError  (malformed op, tape, bitstring, control == target,
 *                    src/circuit.ts:1118-1120, src/statevector.ts:213)
 *   KSV_ERR_DEVICE-> Error      (HIP failure, out of device memory)
 * ksv_last_error() returns the message of the last failure on the calling thread.
 *
 * Threading: one HIP stream per state handle; every call is synchronous
 * with respect to the host (results are in caller buffers on return).
 * Calls on one handle must not overlap.
 */
#ifndef KETSV_H
#define KETSV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KSV_OK           0
#define KSV_ERR_RANGE    1
#define KSV_ERR_TYPE     2
#define KSV_ERR_DEVICE   3
#define KSV_ERR_HANDLE   4

/* amplitude precision */
#define KSV_F64 0   /* complex128 state, FP64 arithmetic                       */
#define KSV_F32 1   /* complex64 state, FP32 arithmetic, FP64 accumulations    */

/* op kinds: one per ket FlatOp kind (src/circuit.ts:22-47) */
#define KSV_SINGLE      0   /* q = {q};               mat 2x2  (applySingle,   src/statevector.ts:47)  */
#define KSV_CNOT        1   /* q = {control, target}                (applyCNOT,     src/statevector.ts:70)  */
#define KSV_SWAP        2   /* q = {a, b}                           (applySWAP,     src/statevector.ts:85)  */
#define KSV_TWO         3   /* q = {a, b};            mat 4x4  (applyTwo,      src/statevector.ts:110) */
#define KSV_CONTROLLED  4   /* q = {control, target}; mat 2x2  (applyControlled, src/statevector.ts:212) */
#define KSV_TOFFOLI     5   /* q = {c1, c2, target}                 (applyToffoli,  src/statevector.ts:138) */
#define KSV_CSWAP       6   /* q = {control, a, b}                  (applyCSwap,    src/statevector.ts:153) */
#define KSV_CSRSWAP     7   /* q = {control, a, b}                  (applyCsrSwap,  src/statevector.ts:177) */
#define KSV_UNITARY     8   /* q = {qs[0] (MSB) ...}; mat 2^k x 2^k (applyUnitary, src/statevector.ts:240) */
#define KSV_BARRIER     9   /* no effect on the state (src/circuit.ts:155)                                 */

#define KSV_MAX_OPERANDS 12  /* widest KSV_UNITARY accepted */

/* One op of the tape.  `mat` is an offset, in doubles, into the tape's
 * matrix array (interleaved re/im), or -1 for kinds without a matrix. */
typedef struct ksv_op {
    int32_t kind;
    int32_t nq;
    int32_t q[KSV_MAX_OPERANDS];
    int64_t mat;
} ksv_op;

/* Gradient bookkeeping for ksv_adjoint: op `op` of the tape depends on
 * parameter `param` with d(op matrix)/d(param) = coeff * dmat, where `dmat`
 * is an offset (doubles) into the matrix array holding the derivative of the
 * op's matrix with respect to its own angle.  Several records may name the
 * same op (u3) or the same parameter (qaoa's gamma on every edge). */
typedef struct ksv_param {
    int32_t op;
    int32_t param;
    double  coeff;
    int64_t dmat;
} ksv_param;

typedef struct ksv_state   ksv_state;
typedef struct ksv_program ksv_program;

/* Options for ksv_set_option. */
#define KSV_OPT_PRUNE        0  /* 1: drop |a|^2 < 1e-14 after every gate like ket's accumulate
                                   (src/statevector.ts:19-27, src/complex.ts:24); 0: exact dense  */
#define KSV_OPT_TILE_QUBITS  1  /* qubits held per LDS tile in fused passes (0 = autotuned)      */
#define KSV_OPT_TIMING       2  /* 1: time every pass launch with HIP events on the state stream  */
#define KSV_OPT_FUSE         3  /* 1: multiply runs of gates on the same qubits into one matrix
                                   (exact mode only; ket's per-gate pruning forbids it)            */
#define KSV_OPT_PASS_KERNEL  4  /* 0 auto (3 in exact mode, 2 under ket pruning), 1 LDS-tile
                                   passes, 2 register-round passes, 3 fused-block passes on
                                   MFMA (exact mode; ket pruning falls back to 2)              */
#define KSV_OPT_ROUND_TILE   5  /* tile qubits of register-round passes: 0 = default (11 FP64,
                                   12 FP32) or 11..13; 2^(tile-4) threads x 16 amplitudes     */
#define KSV_OPT_BLOCK_QUBITS 6  /* fused-block width: 0 = default (4) or 4, 5                  */
#define KSV_OPT_BLOCK_TILE   7  /* fused-block tile qubits: 0 = default (12) or
                                   11..13 (5-qubit blocks need 13, 11 takes 4-qubit blocks)   */
#define KSV_OPT_BLOCK_THREADS 8 /* fused-block workgroup size: 0 = default (256) or 256, 512
                                   (512 only with 4-qubit blocks on 12-qubit tiles)            */
#define KSV_OPT_ADJ_TILE     9  /* ksv_adjoint backward-pass tile qubits (phi and lambda tiles in
                                   LDS): 0 = default (11 FP64, 12 FP32), or 10, 11 (FP64),
                                   10..12 (FP32)                                               */
#define KSV_OPT_ADJ_THREADS 10  /* ksv_adjoint backward-pass workgroup size: 0 = default (256),
                                   256 or 512 (512 on the default tile only)                   */
#define KSV_OPT_BLOCK_PIPE  11  /* FP64 fused-block passes of 3-4 blocks on 12-qubit tiles:
                                   1 (default) = workgroups walk tiles with the next tile loaded
                                   during the current one's blocks, 0 = one tile per workgroup  */

/* Per-state launch statistics (KSV_OPT_TIMING). */
typedef struct ksv_stats {
    uint64_t pass_launches;     /* fused LDS pass kernels launched          */
    double   pass_ms;           /* summed HIP-event time of those launches  */
    double   pass_bytes;        /* algorithmic HBM bytes of those launches  */
    uint64_t other_launches;    /* every other kernel                       */
    double   other_ms;
    double   other_bytes;
    uint64_t gates_applied;     /* tape ops applied                         */
} ksv_stats;

/* Program summary (ksv_program_info). */
typedef struct ksv_program_summary {
    uint64_t ops;               /* tape ops                                 */
    uint64_t passes;            /* HBM passes the schedule makes            */
    uint64_t device_ops;        /* ops after fusion                         */
    int32_t  tile_qubits;       /* qubits per tile of the main pass kind    */
    int32_t  reg_passes;        /* how many passes are register-round passes */
    uint64_t exchanges;         /* LDS transposes over all register-round passes */
    int32_t  blk_passes;        /* how many passes are fused-block passes    */
    int32_t  pad;
    uint64_t blocks;            /* fused dense blocks over all fused-block passes */
} ksv_program_summary;

/* ---- library ------------------------------------------------------------ */
int ksv_last_error(char* buf, size_t len);
int ksv_device_count(int* count);

/* ---- state lifetime ------------------------------------------------------ *
 * Replaces `zero(n)` / `svFromBitstring` (src/statevector.ts:16,
 * src/circuit.ts:136-140): ksv_create allocates 2^n amplitudes in HBM and
 * sets |0...0>; ksv_init_basis sets the basis state |index>.            */
int ksv_create(int n, int precision, ksv_state** out);
int ksv_destroy(ksv_state* s);
int ksv_set_option(ksv_state* s, int option, double value);
int ksv_init_basis(ksv_state* s, uint64_t index);
int ksv_copy_state(ksv_state* dst, const ksv_state* src);

/* ---- gate application ------------------------------------------------------ *
 * Replaces simulatePure's dispatch loop (src/circuit.ts:143-161) and the
 * apply* functions (src/statevector.ts:47-270).                        */
int ksv_apply(ksv_state* s, const ksv_op* ops, size_t nops, const double* mats, size_t nmats);
int ksv_program_create(ksv_state* s, const ksv_op* ops, size_t nops, const double* mats,
                       size_t nmats, ksv_program** out);
int ksv_program_run(ksv_state* s, ksv_program* p);
int ksv_program_info(const ksv_program* p, ksv_program_summary* out);
/* Give tape op `op` a new matrix (same layout as in the tape) without
 * rescheduling: the op's records (or the fused block holding it) are
 * recomputed and uploaded.  The matrix must keep the op's structure
 * (dense / diagonal / permutation); a program built with one-qubit fusion
 * (KSV_OPT_FUSE) cannot be updated.  Used for re-runs with new angles
 * (parameter shifts, variational loops); no counterpart in ket, which
 * rebuilds the circuit (src/circuit.ts:1108-1111). */
int ksv_program_set_matrix(ksv_program* p, size_t op, const double* mat, size_t nmat);
int ksv_program_destroy(ksv_program* p);
/* Schedule a tape without a device (dry run): the passes and ops
 * ksv_program_create would produce for a state of n qubits with the given
 * pass kernel (KSV_OPT_PASS_KERNEL values), LDS tile width and
 * register-round tile width (0 = defaults), fusion and pruning.  Used by the
 * autotuner. */
int ksv_plan(int n, int precision, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
             int kernel, int tile, int rtile, int fuse, int prune, ksv_program_summary* out);
/* ksv_plan with the fused-block tile width and block width (0 = defaults). */
int ksv_plan_blocks(int n, int precision, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                    int kernel, int tile, int rtile, int btile, int bq, int fuse, int prune,
                    ksv_program_summary* out);

/* ---- read-out ---------------------------------------------------------------- */
/* Dense amplitudes [offset, offset+count) as interleaved doubles (statevector(),
 * src/circuit.ts:1431-1441). */
int ksv_get_amplitudes(ksv_state* s, uint64_t offset, uint64_t count, double* out);
int ksv_set_amplitudes(ksv_state* s, uint64_t offset, uint64_t count, const double* in);
/* One amplitude (amplitude(), src/circuit.ts:1491-1495). */
int ksv_amplitude(ksv_state* s, uint64_t index, double out[2]);
/* sum |a|^2 in FP64 */
int ksv_norm2(ksv_state* s, double* out);
/* All (index, p) with p > thresh in ascending index (probabilities(),
 * src/statevector.ts:273-280; exactProbs, src/circuit.ts:3700-3711).
 * Writes min(count, cap) entries; *count receives the full count. */
int ksv_probs_filtered(ksv_state* s, double thresh, uint64_t* idx, double* p, size_t cap,
                       size_t* count);
/* The k most probable indices, p descending, ties by ascending index. */
int ksv_topk(ksv_state* s, size_t k, uint64_t* idx, double* p);
/* P(qubit q = 1) for every q (marginals(), src/circuit.ts:1508-1518). */
int ksv_marginals(ksv_state* s, double* out);
/* rho00, rho11, Re rho01, Im rho01 of qubit q (blochAngles(), src/circuit.ts:4102-4129). */
int ksv_reduce_1q(ksv_state* s, int q, double out[4]);
/* Shot sampling of run()'s fast path (src/circuit.ts:3425-3450): probabilities
 * p > 1e-14 in ascending index order, inclusive CDF with its last entry
 * forced to 1.0, and for each r the first index with cdf >= r.  The host
 * supplies the xorshift32 draws r (src/circuit.ts:813-819) in ket's order.
 * Up to 24 qubits (ket's own range) the CDF is ket's sequential sum, every
 * addition rounded in ket's order, so each draw in [0, 1) lands on ket's
 * index however close it is to a CDF boundary; above 24 qubits the CDF comes
 * from a chunked scan whose sums are associated differently. */
int ksv_sample(ksv_state* s, const double* r, size_t shots, uint64_t* out);
/* ksv_sample with no device allocation: no counterpart in ket; for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Its scratch is one fixed device array per GPU (9 MiB,
 * held until the process ends), so it takes states of at most 20 qubits and at
 * most 65536 draws per call (KSV_ERR_RANGE beyond).  Same results as ksv_sample.
 * Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (any states, any threads) race. */
int ksv_sample_noalloc(ksv_state* s, const double* r, size_t shots, uint64_t* out);

/* Pauli expectations.  Term t is coeff[t] * prod_q P_q with
 * P_q = X if xmask bit q only, Z if zmask bit q only, Y if both.
 * mode 0: exact <psi|P|psi>.
 * mode 1: ket's method (Circuit.expectation src/circuit.ts:3728-3753,
 *         vqe src/algorithms.ts:301-336): rotate a copy (X->H, Y->Rx(pi/2)),
 *         keep p > 1e-14, sum +-p by parity; identity terms give coeff.
 * out[t] = coeff[t] * <P_t>;  *total (may be NULL) = sum_t out[t]. */
int ksv_expect_pauli(ksv_state* s, const uint64_t* xmask, const uint64_t* zmask,
                     const double* coeff, size_t nterms, int mode, double* out, double* total);

/* Adjoint differentiation of E = <psi|H|psi>, psi = tape|init>, H = sum of
 * Pauli terms, w.r.t. the parameters named in `params` (replaces the
 * parameter-shift loop of gradient(), src/algorithms.ts:536-547).
 * `s` holds the initial state on entry; the forward pass runs on it and the
 * backward sweep un-computes it, so it holds the initial state again on
 * return (up to rounding).  Runs in exact mode (no pruning).
 * grad[0..nparam) receives dE/dparam. */
int ksv_adjoint(ksv_state* s, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                const ksv_param* params, size_t nrec, size_t nparam,
                const uint64_t* xmask, const uint64_t* zmask, const double* coeff, size_t nterms,
                double* grad, double* energy);

/* Batched adjoint: one tape with B = 2^(s->n - n) parameter sets at once (the
 * inner loop of vqe() / gradient() / minimize() over many parameter vectors,
 * src/algorithms.ts:301-336, :536-547; ket runs each set separately).  Copy t of
 * the state (amplitudes [t 2^n, (t + 1) 2^n)) runs the tape with its own
 * matrix array, mats + t * nmats, in the layout of the tape's offsets (gate and
 * derivative matrices alike); the records, the Pauli terms (on qubits < n) and
 * the tape's structure are shared.  Every run starts each copy from |0...0>,
 * evaluates E_t = <psi_t|H|psi_t> and, with want_grad, dE_t/dparam by the
 * adjoint method; the launch sequence is captured once in a HIP graph and
 * replayed with each run's matrices.  energy[t], grad[t * nparam + j].
 * Unitaries wider than three qubits are not supported.  An op whose matrices
 * were diagonal in every copy at creation must stay diagonal (KSV_ERR_TYPE
 * otherwise: create the batch again).  The batch uses the state's buffers:
 * destroy it before the state. */
typedef struct ksv_adj_batch ksv_adj_batch;
int ksv_adjoint_batch_create(ksv_state* s, int n, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                             const ksv_param* params, size_t nrec, size_t nparam, const uint64_t* xmask,
                             const uint64_t* zmask, const double* coeff, size_t nterms, ksv_adj_batch** out);
int ksv_adjoint_batch_run(ksv_adj_batch* b, const double* mats, int want_grad, double* energy, double* grad);
int ksv_adjoint_batch_destroy(ksv_adj_batch* b);

/* ---- measurement ---------------------------------------------------------------- */
/* collapseQubit (src/circuit.ts:55-70): outcome = r < p1; project and scale
 * by 1/sqrt(p_outcome) in FP64. */
int ksv_collapse(ksv_state* s, int q, double r, int* outcome);

/* ---- batched QAOA (Max-Cut) ----------------------------------------------------------- *
 * Many instances of ket's qaoa(n, edges, gamma, beta) with
 * maxCutHamiltonian(n, edges) (src/algorithms.ts:231-276) run together:
 * every launch covers all instances (uneven widths, 12..22 qubits), the
 * cost layer (all edges' cnot . rz(-gamma) . cnot) is one diagonal phase and
 * the mixer (rx(2 beta) on every qubit) runs in registers.  The launch
 * sequence is captured once in a HIP graph and replayed with new angles.
 * Instance i has nq[i] qubits and the edges edges[2e], edges[2e+1] for
 * e in [edge_off[i], edge_off[i+1]); gamma and beta are count x p, row-major.
 * ksv_qaoa_batch_run writes E_i = <H_C> (expected cut) and, with want_grad,
 * dE_i/dgamma_il and dE_i/dbeta_il by the adjoint method (no counterpart in
 * ket, which runs instances one by one and differentiates by parameter
 * shifts, src/algorithms.ts:536-547). */
typedef struct ksv_batch ksv_batch;
int ksv_qaoa_batch_create(int count, const int32_t* nq, const int64_t* edge_off, const int32_t* edges, int p,
                          int precision, ksv_batch** out);
int ksv_qaoa_batch_run(ksv_batch* b, const double* gamma, const double* beta, int want_grad, double* energy,
                       double* dgamma, double* dbeta);
/* The final state of one instance (after the last ksv_qaoa_batch_run's forward sweep;
 * with want_grad the backward sweep returns it to H^n|0>). */
int ksv_qaoa_batch_amplitudes(ksv_batch* b, int inst, uint64_t offset, uint64_t count, double* out);
int ksv_qaoa_batch_info(const ksv_batch* b, uint64_t* total_amplitudes, uint64_t* work_items);
int ksv_qaoa_batch_destroy(ksv_batch* b);
int ksv_batch_last_error(char* buf, size_t len);

/* ---- density matrices (circuit.dm()) ------------------------------------------------ *
 * Replaces runDM (src/density.ts:475-588) over ket's sparse map
 * (src/density.ts:5-35).  rho lives dense on the device, entry key
 * (row << n) | col, 1 <= n <= 15, starting as |0..0><0..0|.  Every channel
 * follows ket's arithmetic (U rho U^+ as rows by U then columns by U^+, every
 * product and sum rounded on its own, entries with |v|^2 < 1e-14 removed) and
 * every entry carries its place in ket's Map order, so depolarize2's
 * order-dependent partial-sum pruning (dmAcc, src/density.ts:29-35) comes out
 * as ket's.  ksv_dm_apply runs a pure tape (unitaries up to 6 qubits) with
 * depolarize1(p1) after every single-qubit gate and depolarize2(p2) after every
 * cnot / swap / two-qubit / controlled gate, as runDM does.  ksv_dm_read
 * writes rho (2 * 4^n doubles, interleaved re/im; NULL to skip), the keys of
 * the present entries in Map order (4^n capacity; NULL to skip) and their
 * count.                                                                     */
typedef struct ksv_dm ksv_dm;
int ksv_dm_create(int n, ksv_dm** out);
int ksv_dm_apply(ksv_dm* d, const ksv_op* ops, size_t nops, const double* mats, size_t nmats, double p1, double p2);
int ksv_dm_read(ksv_dm* d, double* rho, uint64_t* order, uint64_t* count);
int ksv_dm_destroy(ksv_dm* d);

/* ---- instrumentation -------------------------------------------------------------- */
/* ---- batched trajectories -------------------------------------------- *
 * Replaces run()'s per-shot loop (src/circuit.ts:3455-3481) over applyOps
 * (:164-207), dep1 / dep2 (:289-304), collapseQubit (:55-70) and the
 * final draw (:3425-3450), for circuits whose shots consume a fixed number
 * of PRNG draws (no `if`).  The B = 2^(s->n - n) shots live in one state,
 * copy t holding amplitudes [t 2^n, (t + 1) 2^n); gates reach every copy
 * through ksv_apply / programs (their qubits are < n).  r[] and outcome[]
 * have one entry per copy.  Outcomes, projections and draws repeat the
 * single-state arithmetic of ksv_collapse / ksv_sample copy by copy.      */
int ksv_shots_init(ksv_state* s, int n, uint64_t index);                 /* every copy := |index>       */
/* copy t := |t> with 2^n copies (s->n = 2n): one run of a circuit leaves
 * column t of its matrix in copy t.  Replaces circuitMatrix()'s loop of 2^n
 * simulations (src/circuit.ts:1463-1485).                                 */
int ksv_shots_init_identity(ksv_state* s, int n);
int ksv_shots_pauli(ksv_state* s, int n, int q, const uint32_t* copy,   /* pauli[i] (1 X, 2 Y, 3 Z) on */
                    const uint8_t* pauli, size_t count);                 /* qubit q of copy[i]          */
int ksv_shots_measure(ksv_state* s, int n, int q, const double* r,      /* outcome = r < P(q = 1);     */
                      const uint8_t* reset, uint8_t* outcome);           /* project; reset[t]: to |0>   */
/* one draw per copy by sampleSV (src/circuit.ts:74-81): cum += p over the copy's
 * non-zero amplitudes in ascending index, each addition rounded in ket's order;
 * the first index with r <= cum, else the last, else 0 (with n == s->n: the
 * single state, the per-shot driver's final draw) */
int ksv_shots_sample(ksv_state* s, int n, const double* r, uint64_t* out);

/* makePrng's stream (src/circuit.ts:813-819: xorshift32 <<13 >>>17 <<5,
 * r = s / 2^32) as `count` draws; *state advances.  Host-only: the batched
 * path draws every shot's numbers up front.  Fills of 2^20 draws or more are
 * cut into chunks on up to 16 host threads, each started by a GF(2)
 * jump-ahead; the output is the serial stream.                            */
int ksv_xorshift32(uint32_t* state, size_t count, double* out);

/* A whole batch in one call: the gate tape with events between its ops.
 * Event i happens after ops [0, at): a Pauli list (copy[off..off+count),
 * pauli[...], each copy at most once) on qubit q, or a measurement / reset
 * of qubit q with draws r[off..off+B) and outcomes written to
 * outcome[off..off+B).  Positions ascend; everything is validated before
 * the state changes.  When a copy fits one LDS tile (n <= 12 FP64, 13 FP32)
 * and the pass kernel is automatic or KSV_OPT_PASS_KERNEL 1, the gates
 * between measurements / resets run as one LDS pass per segment (each
 * workgroup one copy) with the Pauli lists as ops on their copies' tiles;
 * a single state of such a size takes the same LDS kernel with
 * KSV_OPT_PASS_KERNEL 1, which gives it bit-identical amplitudes.          */
#define KSV_SHOT_PAULI   0
#define KSV_SHOT_MEASURE 1
#define KSV_SHOT_RESET   2
typedef struct ksv_shot_event {
    uint64_t at;        /* tape position                                  */
    int32_t  kind;      /* KSV_SHOT_*                                     */
    int32_t  q;         /* qubit (< n)                                    */
    uint64_t off;       /* Pauli: first list entry; measure/reset: first draw */
    uint64_t count;     /* Pauli: list entries                             */
} ksv_shot_event;
int ksv_shots_run(ksv_state* s, int n, const ksv_op* ops, size_t nops, const double* mats, size_t nmats,
                  const ksv_shot_event* ev, size_t nev, const uint32_t* copy, const uint8_t* pauli,
                  size_t npauli, const double* r, size_t nr, uint8_t* outcome);

int ksv_get_stats(ksv_state* s, ksv_stats* out);
int ksv_reset_stats(ksv_state* s);

#ifdef __cplusplus
}
#endif
#endif /* KETSV_H */
