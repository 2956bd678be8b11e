/*
 * qarp_hip.h — C ABI of the MI355X statevector backend (libqarp_hip.so).
 *
 * Plain C: fixed-width integers, doubles and pointers only.  Host arrays are
 * passed by pointer + count; device state lives behind an opaque handle.
 * Every entry point returns a qhip_status; the message of the last failure
 * on the calling thread is read back with qhip_last_error().
 *
 * The entry points replace the native methods that OpenQARP's nanobind
 * module binds for `qx.QarpSimulator` (src/qarpx/python/bindings.cpp:1337-1517)
 * and the C++ kernels behind them; each declaration names the reference
 * function it stands in for.
 *
 * Conventions (src/qarpx/../docs/contracts/qarp_conventions.md §1-§3):
 *   - amplitude index i = sum_q b_q 2^q (qubit 0 is the least significant bit);
 *   - complex128 amplitudes are interleaved (re, im) doubles, complex64 (re, im)
 *     floats;
 *   - dense matrices are row-major, U[y*D + x], local bit b <-> qubits[b].
 */
#ifndef QARP_HIP_H
#define QARP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ── status codes ─────────────────────────────────────────────────────────
 * Mapped by the Python host onto the exception types the reference raises
 * (bindings.cpp:272-290): INVALID -> ValueError, RUNTIME -> RuntimeError,
 * CAPABILITY -> CapabilityError(ValueError), DEVICE/NOMEM -> RuntimeError. */
typedef enum qhip_status {
    QHIP_OK         = 0,
    QHIP_INVALID    = 1,  /* std::invalid_argument in the reference            */
    QHIP_RUNTIME    = 2,  /* std::runtime_error in the reference               */
    QHIP_CAPABILITY = 3,  /* qarpx::capability_error                           */
    QHIP_DEVICE     = 4,  /* a HIP runtime call failed                         */
    QHIP_NOMEM      = 5   /* device or host allocation failed                  */
} qhip_status;

/* ── gate identifiers: same order as qarpx::GateType (core/gates.h:8-57) ── */
typedef enum qhip_gate {
    QHIP_X = 0, QHIP_Y, QHIP_Z, QHIP_H, QHIP_S, QHIP_SDG, QHIP_T, QHIP_TDG,
    QHIP_SX, QHIP_SXDG, QHIP_ID,
    QHIP_RX, QHIP_RY, QHIP_RZ, QHIP_P,
    QHIP_U,
    QHIP_CX, QHIP_CY, QHIP_CZ, QHIP_SWAP, QHIP_ECR, QHIP_ISWAP, QHIP_ISWAPDG,
    QHIP_CH, QHIP_CS, QHIP_CSDG, QHIP_CSX, QHIP_CSXDG,
    QHIP_CRX, QHIP_CRY, QHIP_CRZ, QHIP_CP, QHIP_RZZ, QHIP_RXX, QHIP_RYY,
    QHIP_CU,
    QHIP_CCX, QHIP_CSWAP,
    QHIP_MCZ,
    QHIP_GPHASE,
    QHIP_BARRIER, QHIP_MEASURE, QHIP_RESET,
    QHIP_BRANCH_BEGIN, QHIP_BRANCH_ELSE, QHIP_BRANCH_END,
    QHIP_CUSTOM,
    QHIP_NUM_GATES
} qhip_gate;

enum { QHIP_MAX_PARAMS = 4, QHIP_MAX_FUSION_QUBITS = 8 };

/* Precision of a device state. */
enum { QHIP_C128 = 0, QHIP_C64 = 1 };

/* Op flags. */
enum {
    QHIP_OP_CONDITIONAL = 1  /* carries a classical condition (trajectory only) */
};

/* One command of a flat stream: the C image of qarpx::Command
 * (core/command.h:29-103).  Parameters are concrete values; the affine
 * dependence of a symbolic angle on the tracked parameters (what
 * gate_param_dependencies, qarp_simulator.cpp:1487-1507, derives from
 * affine_coefficients, param.cpp:462-497) is carried as (index, alpha)
 * pairs in the stream's dep pools, used by the gradient entry points only. */
typedef struct qhip_op {
    int32_t gate;            /* qhip_gate                                      */
    int32_t n_qubits;        /* operand count                                  */
    int64_t qubit_offset;    /* first operand in qhip_stream.qubits            */
    int32_t n_params;        /* 0..QHIP_MAX_PARAMS                             */
    int32_t flags;           /* QHIP_OP_*                                      */
    double  params[QHIP_MAX_PARAMS];
    int64_t unitary_offset;  /* Custom: first entry in qhip_stream.unitaries   */
                             /* (complex entries, row-major), else -1          */
    int64_t dep_offset;      /* gradient: first pair in dep_index / dep_alpha  */
    int32_t n_deps;          /* gradient: number of (index, alpha) pairs       */
    int32_t cbit;            /* Measure: recorded cbit, else -1                */
} qhip_op;

typedef struct qhip_stream {
    const qhip_op*  ops;
    int64_t         n_ops;
    const uint32_t* qubits;     /* operand pool                                 */
    const double*   unitaries;  /* complex128 pool, (re, im) interleaved        */
    const int32_t*  dep_index;  /* may be NULL when no op has n_deps > 0        */
    const double*   dep_alpha;
} qhip_stream;

/* One Pauli string of the observable ABI (simulator/pauli_expectation.h:15-16):
 * letters 'X' 'Y' 'Z' 'I' on the listed qubits; an empty string is the
 * identity term. */
typedef struct qhip_pauli_term {
    int64_t offset;     /* first (qubit, letter) in the observable pools */
    int32_t n_letters;
    int32_t reserved;
    double  coeff_re;
    double  coeff_im;
} qhip_pauli_term;

typedef struct qhip_observable {
    const qhip_pauli_term* terms;
    int64_t                n_terms;
    const uint32_t*        qubits;
    const char*            letters;
} qhip_observable;

/* Simulator knobs.  fusion_max_qubits / fusion_min_qubits have the meaning
 * of QarpSimulator::fusion_max_qubits / fusion_min_qubits
 * (simulator/qarp_simulator.h:41-61).  The rest steer the device schedule:
 * 0 (or -1 for use_mfma, relabel_qubits and stream_tiles) selects the default. */
typedef struct qhip_config {
    int32_t fusion_max_qubits;   /* 0..8                                       */
    int32_t fusion_min_qubits;   /* registers below this get 1q fusion only    */
    int32_t tile_qubits;         /* qubits held on chip per tile pass, 0=auto  */
    int32_t low_qubits;          /* contiguous low qubits per tile, 0=auto     */
    int32_t max_ops_per_pass;    /* 0 = auto                                   */
    int32_t use_mfma;            /* -1 auto, 0 never: dense k=3..5 on the matrix cores (FP64 or FP32 MFMA) */
    int32_t relabel_qubits;      /* -1 auto, 0 never: a pass's store may move qubits into the contiguous low
                                    tile positions for the next pass (the layout is restored before a
                                    schedule ends) */
    int32_t stream_tiles;        /* persistent workgroups that load their next tile while the current
                                    tile's ops run: 1 for single-state passes with matrix-core ops and
                                    for the widest tile, -1 for the widest tile only (which no other
                                    kernel runs), 0 never (the widest tile is then refused) */
} qhip_config;

/* What one apply did (filled when the caller passes a non-NULL pointer). */
typedef struct qhip_apply_stats {
    int64_t fused_ops;       /* ops after the fusion pass                      */
    int64_t passes;          /* full-state HBM passes launched                 */
    int64_t diagonal_ops;    /* ops applied as index-phase (no tile qubits)    */
    int64_t mfma_ops;        /* dense ops applied on the matrix cores          */
    double  flops;           /* real FP64/FP32 flops issued for matvecs        */
    int64_t stream_passes;   /* passes run on the streaming (prefetching) tile kernel */
} qhip_apply_stats;

typedef struct qhip_state qhip_state;   /* opaque device statevector */
typedef struct qhip_fused qhip_fused;   /* opaque host result of the fusion pass */

/* ── errors / device ─────────────────────────────────────────────────────── */

/* Message of the last failed call on this thread (empty when none).  Returns
 * the status code of that failure. */
int qhip_last_error(char* buf, size_t len);

/* Number of visible HIP devices (0 when none). */
int qhip_device_count(int32_t* out);

/* Synchronise the library's stream on the current device. */
int qhip_synchronize(void);

/* ── device states ────────────────────────────────────────────────────────
 * Stand in for csim allocate_quantum_state / initialize_quantum_state /
 * release_quantum_state (third_party/csim/memory_ops.cpp, init_ops_fill.cpp)
 * and the initial_state copy of QarpSimulator::statevector
 * (qarp_simulator.cpp:924-932). */
int qhip_state_create(int32_t n_qubits, int32_t precision, qhip_state** out);
int qhip_state_destroy(qhip_state* s);
int qhip_state_info(const qhip_state* s, int32_t* n_qubits, int32_t* precision);
/* |0...0>. */
int qhip_state_reset(qhip_state* s);
/* Copy 2^n interleaved amplitudes (doubles for C128, floats for C64). */
int qhip_state_upload(qhip_state* s, const void* host, uint64_t n_amps);
int qhip_state_download(const qhip_state* s, void* host, uint64_t n_amps);
/* Copy from a complex128 host array, converting for C64 states. */
int qhip_state_upload_c128(qhip_state* s, const double* host, uint64_t n_amps);
int qhip_state_download_c128(const qhip_state* s, double* host, uint64_t n_amps);
int qhip_state_copy(qhip_state* dst, const qhip_state* src);
/* FP64 sum |amp|^2 (validate_initial_state, qarp_simulator.cpp:731-746). */
int qhip_state_norm2(const qhip_state* s, double* out);
/* Device address of the amplitudes (for zero-copy interop). */
int qhip_state_device_ptr(const qhip_state* s, void** out);

/* ── fusion pass ──────────────────────────────────────────────────────────
 * QarpSimulator::fuse_for_dispatch (qarp_simulator.cpp:120-129) ->
 * fuse_for_simulation (simulator/fusion.cpp:120-268) or
 * fuse_single_qubit_gates (transpiler/fusion.cpp:196-301).  Exposed so the
 * fused stream itself can be checked; qhip_apply runs it internally. */
int qhip_fuse(const qhip_config* cfg, const qhip_stream* in, int32_t n_qubits,
              qhip_fused** out);
int qhip_fused_count(const qhip_fused* f, int64_t* n_ops);
/* Op i: gate id, operand count (<= 64 written to qubits), params (4), and
 * for Custom the 2^k x 2^k matrix into `matrix` (capacity in complex
 * entries; pass NULL to skip). */
int qhip_fused_get(const qhip_fused* f, int64_t i, int32_t* gate, int32_t* n_qubits,
                   uint32_t* qubits, double* params, double* matrix, int64_t capacity);
int qhip_fused_destroy(qhip_fused* f);

/* Host only (no device call): the device schedule of qhip_apply for a
 * register of n_qubits, as a sequence of global operations in execution
 * order — dense and diagonal blocks as Custom with their full matrix, mask
 * phases as MCZ, scalar phases as GPhase — with the pass each belongs to
 * (qhip_fused_pass).  For checking the scheduler; no reference counterpart. */
int qhip_schedule_inspect(const qhip_config* cfg, const qhip_stream* in, int32_t n_qubits,
                          int32_t precision, qhip_fused** out);
int qhip_fused_pass(const qhip_fused* f, int64_t i, int32_t* pass);

/* ── hot path: apply a command stream ─────────────────────────────────────
 * The gate loop of QarpSimulator::statevector (qarp_simulator.cpp:901-943):
 * fuse_for_dispatch, then every fused command on the device state in
 * order (apply_command, :356-596; apply_dense_block, dense_kernel.cpp:23-121).
 * Rejects trajectory commands (mid-circuit Measure with cbit, Reset,
 * conditions, branches) with QHIP_RUNTIME, as :906-916 does. */
int qhip_apply(const qhip_config* cfg, const qhip_stream* stream, qhip_state* s,
               qhip_apply_stats* stats);

/* ── structured-program kernels ──────────────────────────────────────────────
 * The non-gate kernels of qx.Program run by QarpSimulator::execute_program
 * (src/qarpx/src/simulator/qarp_simulator.cpp:789-808); gates and dense
 * kernels go through qhip_apply.
 *
 * Stands in for apply_permutation (src/qarpx/src/simulator/program.cpp:610-646):
 * |x> -> |table[x]> on `qubits` (local bit b <-> qubits[b]), k = 1..26 distinct
 * qubits, table a bijection of [0, 2^k).  Out of place through a library-owned
 * scratch state: a state created by qhip_state_create takes the permuted
 * buffer (its device pointer changes; re-read qhip_state_device_ptr), a state
 * wrapping caller memory gets the result copied back. */
int qhip_state_permute(qhip_state* s, const uint32_t* qubits, int32_t k, const uint64_t* table);

/* Stands in for apply_controlled_powers (program.cpp:648-712): matrix^exponents[j]
 * on `targets` wherever control j is |1> (the powers of one matrix commute).
 * matrix: 2^m x 2^m, row-major, complex128 (re, im) interleaved; the powers
 * are formed along one squaring chain as the reference does.  For m <= 7 the
 * chain runs on the host and each controlled power is a Custom op on
 * (targets..., control) of the fused apply; for 8 <= m <= 14 the chain runs
 * on the device and each controlled power is a matrix-core GEMM over the
 * groups whose control bit is set; m > 14 returns QHIP_CAPABILITY (the
 * reference accepts up to 26 targets, a 2^26 x 2^26 host matrix). */
int qhip_state_controlled_powers(const qhip_config* cfg, qhip_state* s, const uint32_t* controls,
                                 const uint64_t* exponents, int32_t n_controls, const uint32_t* targets,
                                 int32_t m, const double* matrix);

/* Stands in for permutation_table (program.cpp:714-724; Python binding
 * qx._permutation_table, bindings.cpp:1302-1315): the basis permutation the
 * stream applies to its k local qubits, exact including phase.  `local` must
 * already be remapped onto qubits [0, k) and hold only concrete, unconditioned
 * unitary commands (the caller rejects the rest, as to_local does,
 * program.cpp:404-436).  *found = 1 and table[0 .. 2^k) written when the
 * stream is a permutation the reference derives (classically gate by gate,
 * or with at most max_rest simulated qubits within 2^max_work_log2 x
 * |stream| work); *found = 0 otherwise.  Host computation, no device call. */
int qhip_permutation_table(const qhip_stream* local, int32_t k, uint32_t max_rest, uint32_t max_work_log2,
                           int32_t* found, uint64_t* table);

/* FNV-1a over the 8 little-endian bytes of each word, from `seed`: the mixing
 * step of commands_digest / local_commands_digest / parts_digest
 * (program.cpp:739-829); the caller lays a stream out as words in the
 * reference's order.  Host computation, no device call. */
int qhip_fnv1a_words(const uint64_t* words, int64_t n, uint64_t seed, uint64_t* out);

/* Batched basis states for transition columns: the state is read as blocks
 * of `stride` amplitudes (a power of two, at most 2^n); block b is set
 * to |labels[b]> for b < n_labels and to zero beyond.  Stands in for
 * set_basis_states (src/qarpx/src/simulator/transition_columns.cu:88-93)
 * after its memset of the batch. */
int qhip_state_set_basis(qhip_state* s, const uint64_t* labels, int64_t n_labels, uint64_t stride);

/* Compaction of transition columns (the loop of compact_columns,
 * transition_columns.cu:262-291, on the device): among the first n_amps
 * amplitudes, the indices whose magnitude hypot(re, im) exceeds `tolerance`
 * and the amplitudes (complex128, (re, im)), in ascending index order.  With
 * keys or values NULL only *n_out = the count is returned; otherwise up to
 * `capacity` entries are written (QHIP_INVALID if more are retained).
 * *max_pruned (when not NULL) receives the largest magnitude not retained,
 * 0 if none: the loop's call_max_pruned_amplitude. */
int qhip_compact(const qhip_state* s, uint64_t n_amps, double tolerance, int64_t* keys, double* values,
                 int64_t capacity, int64_t* n_out, double* max_pruned);

/* ── shot-batched trajectories ───────────────────────────────────────────────
 * The per-shot loop of QarpSimulator::run (src/qarpx/src/simulator/
 * qarp_simulator.cpp:1038-1211) with mid-circuit Measure, Reset, classical
 * conditions and branches (apply_command_trajectory, :600-689).  A batch of
 * shots lives in one state of n + b qubits, shot j in the block of `stride`
 * = 2^n amplitudes starting at j * stride; unitary runs go through qhip_apply
 * on that state (their qubits are < n).
 *
 * Every block := the n-qubit `src` (the shared unitary prefix). */
int qhip_shots_broadcast(qhip_state* dst, const qhip_state* src);
/* Blocks j with active[j] == 0 := the same blocks of `saved` (shots a
 * conditional run skipped; active has dim / stride entries). */
int qhip_shots_restore(qhip_state* dst, const qhip_state* saved, const uint8_t* active, uint64_t stride);
/* p0[j] = sum of |a|^2 over block j's amplitudes with `qubit` clear (M0_prob,
 * csim stat_ops_probability.cpp:11-30), FP64. */
int qhip_shots_prob0(const qhip_state* s, int32_t qubit, uint64_t stride, double* p0);
/* Per block: action -1 untouched; 0 keep |0> of `qubit` and scale; 1 keep |1>
 * and scale; 2 move |1> to |0> and scale (Reset after outcome 1).  P0_gate /
 * P1_gate, normalize(p) = scale by sqrt(1/p), X (qarp_simulator.cpp:636-685). */
int qhip_shots_project(qhip_state* s, int32_t qubit, uint64_t stride, const int8_t* action, const double* scale);
/* Noise channels per shot (apply_channel, src/qarpx/src/device/
 * noise_model.cpp:86-186, after each noisy gate, qarp_simulator.cpp:697-722):
 * blocks j with active[j] != 0 get the 2x2 complex matrix mats[8j .. 8j + 8)
 * (row-major (re, im)) on `qubit` — the sampled Pauli (X_gate / Y_gate /
 * Z_gate) or the sampled Kraus operator K_i / sqrt(p_i)
 * (single_qubit_dense_matrix_gate). */
int qhip_shots_apply1q(qhip_state* s, int32_t qubit, uint64_t stride, const double* mats, const uint8_t* active);
/* The reduced density matrix of `qubit` per block (reduce_density_matrix_1q,
 * noise_model.cpp:39-58): out[4j .. 4j + 4) = (rho00, rho11, Re rho01,
 * Im rho01), rho01 = sum a conj(b) over the pairs (a with the bit clear). */
int qhip_shots_rho1(const qhip_state* s, int32_t qubit, uint64_t stride, double* out);
/* Each shot's final draw (run()'s inline CDF walk, qarp_simulator.cpp:1151-1163):
 * for j < n_shots, out[j] = the first index of block j whose running sum of
 * |a|^2 (re*re + im*im, index order) reaches u[j] times the block's total,
 * else stride - 1.  One device thread per shot, sums rounded term by term. */
int qhip_shots_sample(const qhip_state* s, uint64_t stride, const double* u, int64_t n_shots, uint64_t* out);
/* n draws of std::uniform_real_distribution<double>(0, 1) from
 * std::mt19937(seed) — the per-shot stream mt19937(base_s + shot) of run(). */
int qhip_mt_uniforms(uint32_t seed, int64_t n, double* out);

/* Compiled stream: the fusion pass and the device schedule of qhip_apply done
 * once, op tables resident on the device, applied to any state of the given
 * width and precision.  The counterpart of compiling a circuit once and
 * dispatching it per call (QarpEngine._compile_one, src/qarp/engines/
 * _qarp_engine.py:283-300, with the program cache of qx.Program,
 * src/qarpx/include/qarpx/simulator/program.h:51-83). */
typedef struct qhip_plan qhip_plan;
int qhip_plan_create(const qhip_config* cfg, const qhip_stream* stream, int32_t n_qubits,
                     int32_t precision, qhip_plan** out);
int qhip_plan_apply(const qhip_plan* plan, qhip_state* s, qhip_apply_stats* stats);
/* As qhip_plan_apply, recording HIP events around every pass launch on the
 * library's stream: pass_ms[i] = duration of pass i (capacity entries). */
int qhip_plan_apply_timed(const qhip_plan* plan, qhip_state* s, double* pass_ms, int64_t capacity);
/* As qhip_plan_apply, replaying the pass launches as one hipGraph captured
 * on the first call for this state (a different state re-captures).  Same
 * kernels, same results as qhip_plan_apply; the launch cost of a plan with
 * many short passes drops to one graph launch.  No counterpart in the
 * reference (a CPU simulator has no launches to batch). */
int qhip_plan_apply_graph(qhip_plan* plan, qhip_state* s, qhip_apply_stats* stats);
/* Rebinds a plan to new parameter values of the same circuit (the stream is
 * re-fused and re-scheduled on the host): when the schedule keeps its
 * structure only the device coefficient table is rewritten and a captured
 * graph stays valid (*rebuilt = 0); otherwise the tables are rebuilt
 * (*rebuilt = 1).  The sweep counterpart of batch_run / batch_expectation
 * (qarp_simulator.cpp:1248-1267, 971-984). */
int qhip_plan_update(qhip_plan* plan, const qhip_config* cfg, const qhip_stream* stream, int32_t* rebuilt);
int qhip_plan_destroy(qhip_plan* plan);

/* As qhip_apply but recorded Measure / Barrier commands are skipped — the
 * terminal-measurement fast path of QarpSimulator::run (qarp_simulator.cpp:1017-1036).
 * Non-terminal measurements are rejected with QHIP_INVALID. */
int qhip_apply_terminal(const qhip_config* cfg, const qhip_stream* stream, qhip_state* s,
                        qhip_apply_stats* stats);

/* ── Pauli-sum expectation ────────────────────────────────────────────────
 * pauli_transition (simulator/pauli_expectation.cpp:159-167): <bra|H|ket>
 * as (re, im).  bra and ket may be the same handle.  n in [0, 62]. */
int qhip_transition(const qhip_state* bra, const qhip_state* ket,
                    const qhip_observable* obs, double out[2]);

/* qhip_transition with no device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * group table and reduction partials live in a fixed 1,081,856-byte device
 * array per GPU (held until the process ends), so the observable may have at
 * most 4096 distinct X/Y flip masks and 16384 terms, and the state at most
 * 2^38 amplitudes (QHIP_INVALID beyond).
 * Same arguments, errors and result (bit for bit) as qhip_transition.  Calls
 * share the array: the library serialises its own calls; concurrent use on one
 * GPU from outside that lock races. */
int qhip_transition_noalloc(const qhip_state* bra, const qhip_state* ket,
                            const qhip_observable* obs, double out[2]);

/* |out> = H |in> (apply_observable, qarp_simulator.cpp:1459-1479). */
int qhip_apply_observable(const qhip_state* in, const qhip_observable* obs, qhip_state* out);

/* ── terminal sampling ────────────────────────────────────────────────────
 * sample_terminal (qarp_simulator.cpp:751-783): n_shots draws of outcome
 * indices from |amp|^2, identical to std::discrete_distribution<uint64_t>
 * over std::mt19937(seed) as libstdc++ computes it (sequential normalising
 * sum, sequential partial sums, last entry 1.0, lower_bound per draw).
 * Outcomes are written in draw order. */
int qhip_sample(const qhip_state* s, int64_t n_shots, uint32_t seed, uint64_t* outcomes);
/* Same search for caller-supplied uniforms in [0, 1). */
int qhip_sample_uniforms(const qhip_state* s, const double* uniforms, int64_t n,
                         uint64_t* outcomes);
/* |amp|^2 of every index into a host array of 2^n doubles (the EXACT readout
 * of _exact_result, src/qarp/engines/_engine.py:79-116, before pruning). */
int qhip_probabilities(const qhip_state* s, double* out, uint64_t n);
/* EXACT readout of _exact_result (src/qarp/engines/_engine.py:79-116,
 * _EXACT_PRUNE_TOL at :76): every index with |amp|^2 > threshold, ascending,
 * with its probability, compacted on the device.  *count receives the number
 * retained; keys / probs (both NULL to query the count) must hold capacity >=
 * *count entries. */
int qhip_exact(const qhip_state* s, double threshold, int64_t* keys, double* probs, int64_t capacity,
               int64_t* count);

/* ── adjoint gradient ─────────────────────────────────────────────────────
 * run_gradient_core (qarp_simulator.cpp:1519-1605) with the observable
 * builder of run_gradient (:1607-1624): forward sweep from `psi`, phi = H psi, then
 * the reverse sweep un-applying every gate on psi and phi while all
 * generator inner products accumulate in FP64.  `psi` holds the root state on
 * entry and, the reverse sweep having un-applied every gate, again on exit.
 * grad has n_params entries (unused indices stay 0); energy receives
 * <psi|H|psi> of the evolved state (may be NULL).  `stream` is the concrete
 * stream with dep pools filled. */
int qhip_run_gradient(const qhip_config* cfg, const qhip_stream* stream, int32_t n_params,
                      qhip_state* psi, const qhip_observable* obs,
                      double* grad, double* energy);

/* run_gradient_phi (qarp_simulator.cpp:1626-1639): caller-supplied phi
 * (copied; `phi` is not modified). */
int qhip_run_gradient_phi(const qhip_config* cfg, const qhip_stream* stream, int32_t n_params,
                          qhip_state* psi, const qhip_state* phi, double* grad);

/* ── batched instances ────────────────────────────────────────────────────
 * Many independent circuits, each with its own register width, stream and
 * observable, evaluated together (replaces the sequential loops of
 * batch_expectation, qarp_simulator.cpp:971-984, and the per-parameter-set
 * sweep of QarpEngine._sweep, src/qarp/engines/_qarp_engine.py:450-503).
 * energies[i] = <psi_i|H_i|psi_i>; when grads is non-NULL the adjoint
 * gradient of instance i is written at grads + grad_offsets[i].
 * precision selects complex128 or complex64 states (FP64 reductions). */
int qhip_batch_energy_grad(const qhip_config* cfg, int32_t n_instances,
                           const int32_t* n_qubits, const qhip_stream* streams,
                           const qhip_observable* observables, const int32_t* n_params,
                           int32_t precision, double* energies,
                           double* grads, const int64_t* grad_offsets);
/* The same instances kept as a session for parameter sweeps (an optimizer's
 * iterations over batch_expectation / run_gradient, qarp_simulator.cpp:
 * 971-984, 1519-1624): plans, states and launch descriptors are built once;
 * qhip_batch_run executes a sweep (use_graph: replays the sweep's launches as
 * one hipGraph, captured on the first run); qhip_batch_update rebinds new
 * parameter values of the same circuits — only the coefficient table is
 * rewritten while every schedule keeps its structure (*rebuilt = 0),
 * otherwise the session is rebuilt (*rebuilt = 1).  streams and observables
 * must stay valid until the next update / destroy. */
typedef struct qhip_batch qhip_batch;
int qhip_batch_create(const qhip_config* cfg, int32_t n_instances, const int32_t* n_qubits,
                      const qhip_stream* streams, const qhip_observable* observables,
                      const int32_t* n_params, int32_t precision, int32_t gradients, qhip_batch** out);
int qhip_batch_run(qhip_batch* b, double* energies, double* grads, const int64_t* grad_offsets,
                   int32_t use_graph);
int qhip_batch_update(qhip_batch* b, const qhip_stream* streams, int32_t* rebuilt);
int qhip_batch_destroy(qhip_batch* b);
/* Algorithmic HBM bytes of one sweep of a session, from the schedules it
 * launches (no reference counterpart: measurement only).  B = bytes per
 * amplitude; per batched instance of n qubits: fwd_bytes += forward passes x
 * 2·B·2^n; obs_bytes += (x-groups + 1)·B·2^n for H|psi> (gathered reads and
 * the write of |phi>) + 2·B·2^n for the energy <psi|phi>; rev_bytes +=
 * reverse two-state passes x 4·B·2^n.  Launch counts are the tile-pass
 * launches of the sweep. */
typedef struct qhip_batch_info {
    double  fwd_bytes, obs_bytes, rev_bytes;
    int64_t fwd_launches, rev_launches, instances_batched, instances_single;
} qhip_batch_info;
int qhip_batch_stats(const qhip_batch* b, qhip_batch_info* out);

#ifdef __cplusplus
}
#endif

#endif /* QARP_HIP_H */
