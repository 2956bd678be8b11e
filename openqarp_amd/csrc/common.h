// common.h — internal declarations of libqarp_hip.so.
//
// Host-side runtime (fusion, scheduling, sampling resolve) is C++17; device
// code is HIP for gfx950.  Nothing here is part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <memory>
#include <complex>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qarp_hip.h"

namespace qhip {

using cplx = std::complex<double>;

// ── errors ──────────────────────────────────────────────────────────────────
// Library functions throw qhip::Error; the C entry points catch it, latch the
// message in a thread-local string and return the code.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void fail(int code, const std::string& msg);
void hip_check(hipError_t e, const char* what);
#define QHIP_CHECK(x) ::qhip::hip_check((x), #x)

// ── device context ──────────────────────────────────────────────────────────
// One context per process for the current device: its own non-blocking
// stream (default priority) and grow-only scratch buffers.
struct DeviceBuffer {
    void*  ptr = nullptr;
    size_t bytes = 0;
    void*  ensure(size_t n);
    ~DeviceBuffer();
};

struct Context {
    int          device = -1;
    hipStream_t  stream = nullptr;
    int          n_cu = 0;
    DeviceBuffer ops;       // device op tables
    DeviceBuffer coefs;     // gate coefficients
    DeviceBuffer red;       // reduction partials
    DeviceBuffer misc;      // sampling / adjoint scratch
    DeviceBuffer misc2;
    DeviceBuffer obs;       // prepared observable tables (adjoint sweep)
    DeviceBuffer phi;       // the adjoint's second state, up to 1 GiB
    DeviceBuffer perm;      // out-of-place permutation target (one state)
    DeviceBuffer pinned_dummy;
};
Context& ctx();   // initialises on first use (throws QHIP_DEVICE without a GPU)

// ── device states ───────────────────────────────────────────────────────────
}  // namespace qhip

struct qhip_state {
    int32_t n = 0;
    int32_t precision = QHIP_C128;
    void*   d = nullptr;      // 2^n amplitudes
    bool    owned = true;
    uint64_t dim() const { return uint64_t{1} << n; }
    size_t amp_bytes() const { return precision == QHIP_C128 ? 16 : 8; }
};

namespace qhip {

// ── host op representation after fusion ─────────────────────────────────────
enum OpKind : int32_t {
    OP_DENSE = 0,     // 2^k x 2^k matrix on k operands
    OP_DIAG = 1,      // 2^k diagonal on k operands (no locality requirement)
    OP_MASKPHASE = 2, // amplitude *= phase where (i & mask) == mask
    OP_SCALAR = 3,    // amplitude *= phase everywhere
    OP_IP_DENSE = 4,  // adjoint: accumulate <phi| G |psi> for a dense G
    OP_IP_DIAG = 5,   // adjoint: accumulate <phi| G |psi> for a diagonal G
    // Device-only: a register block of three tile positions (pos[0..2],
    // ascending) whose 8 amplitudes a thread holds while the lmask following
    // OP_SUB entries act on them; an OP_SUB's k is its pattern (0-2: one
    // qubit on block bit k; 3, 4, 5: two qubits on block bits 01, 02, 12, the
    // lower bit being matrix operand 0; 6-11: the same for a diagonal, the
    // lower block bit indexing entry bit 0; 12: a diagonal with no operand
    // on the block) and coef its matrix (or its diagonal entries).  A
    // diagonal member's operands outside the tile (spos bytes, count in
    // spos[7]) select a sub-table of its entries once per tile.
    OP_BLOCK = 6,
    OP_SUB = 7,
    // Device-only: dense op (complex128, k = 3..5) on the matrix cores; coef
    // points at the per-lane MFMA fragments of its real 2^(k+1)-square form.
    OP_MFMA = 8,
    // Device-only: diagonal op whose local operands (k <= 2, pos[], group
    // positions as for dense ops) index the entries within a group, while its
    // operands outside the tile (spos bytes 0..5: global qubits, byte 7:
    // their count) select a sub-table once per tile.  coef holds the 2^(k+kg)
    // entries at index (global bits << k) | local bits.  OP_IP_DIAGG is the
    // inner-product form (lmask: accumulator slot).
    OP_DIAGG = 9,
    OP_IP_DIAGG = 10,
    // Device-only: register-block member that accumulates <phi| G |psi> for
    // a generator on block bits (k: pattern as for OP_SUB, coef: its matrix,
    // lmask: accumulator slot); two-state passes only.
    OP_IPSUB = 11,
    // A run of diagonal phase ops merged into one (host and device): the
    // amplitude at index j is multiplied by exp(i theta(j)) with
    // theta(j) = c + sum_q h_q z_q + sum_{a<b} J_ab z_a z_b, z_q = 1 - 2 j_q
    // (HostOp::angle0 = c, HostOp::terms = the h and J terms).  On the
    // device coef points at exp(i c), then two records per term (exp(i w),
    // {operand code, 0}): the terms that touch no tile position >= 8, then
    // per row bit m = position 8 + m the terms with one operand there, then
    // (when lmask bits 16..31 are 1) the product of the terms with both
    // operands on row bits, one entry per row code; lmask / gmask hold the
    // list lengths (16 bits each: none-row, pair table, rows 0..5).
    OP_PHASE = 12,
    // Adjoint: the inner product <phi| D |psi> for a diagonal D(j) = k0 +
    // sum_q k_q z_q + sum_{a<b} k_ab z_a z_b with complex k (a run's
    // diagonal generators of one parameter, weighted by their prefactors and
    // dependency coefficients).  HostOp::phase = k0, terms hold k (w, wi).
    // On the device: records as OP_PHASE's (k0, then {k, {code, 0}} per term,
    // then the row-pair table, additive), lmask the accumulator slot, gmask
    // the lengths none-row | pair table << 16 | row 0 << 32 | row 1 << 48,
    // gpos bytes 0..7 rows 2..5 (16 bits each).
    OP_IP_PHASE = 13,
};

// Device op kinds that accumulate an inner product.
inline bool is_ip_kind(int32_t kind) {
    return kind == OP_IP_DENSE || kind == OP_IP_DIAG || kind == OP_IP_DIAGG || kind == OP_IPSUB || kind == OP_IP_PHASE;
}

// Inner-product members per register block (per-thread accumulators).
constexpr int kMaxBlockIp = 8;
// Register-block member patterns (DevOp::k of OP_SUB / OP_IPSUB), beyond
// 0-2 (2x2 on block bit B), 3-5 (4x4 on a bit pair), 6-8 / 9-11 (diagonal
// on one bit / a bit pair, sub-tables for operands outside the tile), 12 (no
// block operand): kReal2 + B a real 2x2 matrix, kXGen / kYGen / kZGen + B the
// Pauli generator on bit B (inner products only), kNegPair + p the sign flip
// where both bits of pair p (01, 02, 12) are set, kImag2 + B a 2x2 matrix
// with a real diagonal and an imaginary off-diagonal (Rx, Y).
constexpr int kReal2 = 13, kXGen = 16, kYGen = 19, kZGen = 22, kNegPair = 25, kImag2 = 28;

// One term of a merged phase op: weight w times z_a (b < 0) or z_a z_b
// (an inner-product op's weight is complex: w + i wi).
struct PhaseTerm {
    int32_t a = -1, b = -1;
    double w = 0.0, wi = 0.0;
};

struct HostOp {
    int32_t kind = OP_DENSE;
    int32_t k = 0;
    std::array<uint32_t, 8> q{};     // local bit b <-> q[b]
    std::vector<cplx> m;             // dense: D*D row-major; diag: D entries
    uint64_t mask = 0;               // maskphase operands
    cplx phase{1.0, 0.0};            // maskphase / scalar
    int32_t slot = -1;               // IP ops: accumulator slot
    double angle0 = 0.0;             // OP_PHASE: constant angle c
    std::vector<PhaseTerm> terms;    // OP_PHASE: h (b < 0) and J terms
    uint64_t qmask() const {
        if (kind == OP_MASKPHASE || kind == OP_PHASE || kind == OP_IP_PHASE) return mask;
        uint64_t r = 0;
        for (int b = 0; b < k; ++b) r |= uint64_t{1} << q[b];
        return r;
    }
    bool needs_local() const { return kind == OP_DENSE || kind == OP_IP_DENSE; }
};

// Matrix of one concrete named command (gate id, operands, params) over its
// own operands, local bit b <-> operand b.  Returns false for gates that are
// not a fixed-width unitary (markers, MCZ wider than 8, GPhase).
bool named_matrix(int gate, int k, const double* p, std::vector<cplx>& m);

// Stream view.
struct Stream {
    const qhip_stream* s;
    const qhip_op& op(int64_t i) const { return s->ops[i]; }
    const uint32_t* qubits(int64_t i) const { return s->qubits + s->ops[i].qubit_offset; }
    int64_t size() const { return s->n_ops; }
};

// Fusion pass of fuse_for_dispatch: emits HostOps (markers dropped when
// `drop_markers`).  Also used by qhip_fuse with the raw op/Custom split.
struct FusedItem {
    int64_t src = -1;                // verbatim input op, or -1 for a fused block
    int32_t k = 0;
    std::array<uint32_t, 8> q{};
    std::vector<cplx> m;             // fused block matrix
    // Schedule inspection only: gate id of the emitted operation, its pass,
    // mask-phase operands beyond 8 and the phase of mask / scalar ops.
    int32_t gate = QHIP_CUSTOM;
    int32_t pass = -1;
    std::vector<uint32_t> wide_q;
    cplx phase{1.0, 0.0};
};
std::vector<FusedItem> fuse_for_dispatch(const Stream& st, int n, int fusion_max, int fusion_min);

// Lower one input op (verbatim) or fused block into HostOps.
void lower_item(const Stream& st, const FusedItem& it, int n, std::vector<HostOp>& out,
                bool skip_terminal_markers);
void lower_op(const Stream& st, int64_t i, std::vector<HostOp>& out);

// Adjoint helpers: the dagger of op i as HostOps, and its generator.
void lower_dagger(const Stream& st, int64_t i, std::vector<HostOp>& out);
// Generator of a parametric op: fills G (dense or diagonal HostOp without
// slot) and prefactor c; returns false when the gate has no rule.
bool generator(const Stream& st, int64_t i, HostOp& G, cplx& c);

// Classify a dense HostOp as diagonal when every off-diagonal entry is 0.
void classify_diagonal(HostOp& op);

// ── device schedule ─────────────────────────────────────────────────────────
struct DevOp {          // 64 bytes, one per op per pass
    int32_t kind;
    int32_t k;
    int8_t  pos[8];     // dense/IP: local positions; diag: local pos or 32+global qubit
    int8_t  spos[8];    // dense/IP: local positions sorted ascending
    int64_t coef;       // first coefficient in the pass coefficient table
    uint64_t lmask;     // maskphase: local-position mask; IP ops: accumulator slot
    uint64_t gmask;     // maskphase: global-bit mask outside the tile;
                        // dense/IP dense: local positions of group bits >= 8
    int8_t  gpos[16];   // dense/IP dense: local position of group-number bit b
};
static_assert(sizeof(DevOp) == 64, "DevOp layout");

// Inner-product ops per pass (per-wave LDS slots of the adjoint tile kernel).
constexpr int kMaxIpPerPass = 64;

// Widest tile: 2^T amplitudes of each state must fit the LDS of one workgroup
// (160 KiB per CU; two states share it in adjoint passes).  The widest
// single-state tile (128 KiB: one workgroup per CU) runs only on the
// streaming kernel, whose prefetch overlaps its HBM traffic with its ops.
inline int max_tile_qubits(int precision, bool two_states = false) {
    return (precision == QHIP_C128 ? 12 : 13) + (two_states ? 0 : 1);
}

// Elements of one matrix-core op's fragments as the streaming tile kernel
// stages them in LDS: k = 3 as stored; k = 4, 5 in complex128 without the
// (Ur + Ui) pairs, which the kernel forms with the same double additions the
// host makes, and with them in complex64 (the host rounds the double sum to
// float once, which a float addition on the device would not reproduce).
__host__ __device__ constexpr int staged_fragment_elems(int k, bool f32) {
    return k == 3 ? 128 : k == 4 ? (f32 ? 384 : 256) : k == 5 ? (f32 ? 1536 : 1024) : 0;
}
// LDS the streaming tile kernel reserves for the fragments of one pass (ops
// beyond it read their fragments from global memory).
constexpr int kStreamFragBytes = 16384;
// Smallest tile array (2^TB0 amplitudes per state) of each amplitude type;
// the tile kernels are instantiated for TB0, TB0 + 1 and TB0 + 2.
constexpr int kTileArrayQubitsC128 = 10;
constexpr int kTileArrayQubitsC64 = 11;

// PassDesc::kmax flag: a pass with OP_PHASE / OP_IP_PHASE ops runs the
// kernel variant that has them (the other variants, the streaming kernel
// among them, leave them out and keep their registers and code).
constexpr int kPhaseKernel = 8;

struct PassDesc {
    int32_t n;           // register width
    int32_t T;           // tile qubits; 0 marks a wide pass (one dense op, k > 5)
    int32_t op_begin;
    int32_t op_end;
    int32_t kmax;        // widest VALU dense / IP op in the pass (matrix-core ops excluded), plus
                         // kPhaseKernel in passes that carry merged phase ops
    int32_t frag;        // elements of the pass's matrix-core fragments as the streaming kernel
                         // stages them in LDS; -1: the pass runs one tile per workgroup
    uint64_t tmask;      // global bits that are tile-local
    uint64_t hmask;      // global bits of local positions >= 8
    int32_t qpos[40];    // qpos[l] = global qubit of local position l
    // Store permutation: the amplitude written at output local index m is
    // the one at input local index sum_i bit_i(m) << sperm[i] (identity:
    // sperm[i] = i).  It relabels which qubit lives at which tile position
    // for the passes that follow.
    int8_t sperm[16];
};

// One instance's pass in a batched launch (tile_batch_kernel).
struct BatchItem {
    PassDesc P;
    void* psi;
    void* phi;             // two-state passes only
    const void* ops;       // DevOp table of the pass (indices in P are relative to it)
    const void* coefs;
    double2* ip_out;       // per-tile inner-product partials (two-state passes)
    int64_t n_tiles;
};

// Whether a single-state pass runs on the streaming tile kernel: its VALU
// dense ops are at most two-qubit, and its tile is the widest one (which
// only that kernel runs) or fills one of the other arrays and the pass has
// matrix-core ops (passes of VALU ops alone stream no faster than one tile
// per workgroup).
inline bool streaming_pass(const PassDesc& P, int precision) {
    const int tb0 = precision == QHIP_C128 ? kTileArrayQubitsC128 : kTileArrayQubitsC64;
    if (P.frag < 0 || (P.kmax & ~kPhaseKernel) > 2) return false;
    return P.T == tb0 + 3 || (P.T >= tb0 && P.T <= tb0 + 2 && P.frag > 0);
}

struct Plan {
    std::vector<PassDesc> passes;
    std::vector<DevOp> ops;
    std::vector<cplx> coefs;
    int64_t dense_ops = 0, diag_ops = 0, big_ops = 0, mfma_ops = 0, stream_passes = 0;
    double flops = 0.0;
};

struct ScheduleOptions {
    int tile_qubits = 0;     // 0 = default for the precision / two-state mode
    int low_qubits = 0;      // 0 = default
    int max_ops = 0;         // 0 = no cap
    double max_flops_per_amp = 0.0;  // 0 = no cap
    bool two_states = false; // adjoint passes hold psi and phi
    bool register_blocks = true;  // group k <= 2 dense ops into OP_BLOCKs
    int mfma = -1;           // dense k >= 3 on the matrix cores: -1 auto, 0 never, 1 where possible
    bool relabel = true;     // passes may move the next pass's qubits into the contiguous low positions
    int stream = -1;         // streaming (prefetching) tile kernel: -1 only for the widest tile,
                             // 0 never, 1 also for passes with matrix-core ops
    int precision = QHIP_C128;
};
Plan schedule(const std::vector<HostOp>& ops, int n, int precision, const ScheduleOptions& opt);
// Per-lane fragments of a dense op for the matrix-core path, and back.
// f32: fragments for v_mfma_f32_16x16x4 (complex64 states), whose result
// rows are laid out differently from the FP64 instruction's (k = 3 only).
void mfma_fragments(const std::vector<cplx>& U, int k, std::vector<cplx>& out, bool f32);
std::vector<cplx> mfma_matrix(const cplx* fragments, int k, bool f32);
int default_tile_qubits(int precision, bool two_states);

// A schedule with its op tables resident on the device.
struct CompiledPlan {
    Plan plan;
    int n = 0;
    int precision = QHIP_C128;
    void* d_tables = nullptr;
    void* d_ops = nullptr;
    void* d_coefs = nullptr;
    void* d_slots = nullptr;
    int max_ip = 0;
    int64_t fused_ops = 0;
    hipGraphExec_t graph = nullptr;   // captured pass launches (qhip_plan_apply_graph)
    void* graph_state = nullptr;      // the state the graph was captured for
};
CompiledPlan* compile_plan(Plan&& plan, int n, int precision);
void run_compiled_graph(CompiledPlan& cp, qhip_state* s);
// Rebinds a plan to a re-planned stream: when the schedule has the same
// structure only the coefficient table is rewritten (the captured graph stays
// valid); otherwise the device tables are rebuilt.  Returns true on rebuild.
bool update_compiled(CompiledPlan& cp, Plan&& plan, int64_t fused_ops);
void run_compiled(const CompiledPlan& cp, qhip_state* s, std::vector<double>* pass_ms);
void free_compiled(CompiledPlan* cp);

// ── device entry points (kernels_*.hip) ────────────────────────────────────
// One tile pass (T >= 1) of psi, or of psi and phi together when phi is set,
// on `stream`: tile_c128.hip / tile_c64.hip.
void launch_tile_pass(const PassDesc& P, double2* psi, double2* phi, double2* ip_part, const DevOp* d_ops,
                      const double2* d_coefs, int64_t n_tiles, hipStream_t stream);
void launch_tile_pass(const PassDesc& P, float2* psi, float2* phi, double2* ip_part, const DevOp* d_ops,
                      const float2* d_coefs, int64_t n_tiles, hipStream_t stream);
// One reduction job: *out = sum of n_tiles partials at in (tile order).
struct IpReduce {
    const double2* in;
    int64_t n_tiles;
    double2* out;
};
void launch_ip_reduce(const IpReduce* d_jobs, int n_jobs, hipStream_t stream);
// All items' tiles in one launch (items / first on the device; tmax and kmax
// bound the items' tile widths and VALU dense widths).
void launch_tile_batch(int precision, bool two, const BatchItem* d_items, const int64_t* d_first, int n_items,
                       int64_t total_tiles, int tmax, int kmax, hipStream_t stream);
void launch_plan(const Plan& plan, qhip_state* s, qhip_state* phi, void* d_ip_partials,
                 int n_slots);
void state_reset(qhip_state* s);
double state_norm2(const qhip_state* s);
void state_convert_upload(qhip_state* s, const double* host, uint64_t n);
void state_convert_download(const qhip_state* s, double* host, uint64_t n);

struct PauliTables;   // grouped observable on the device
cplx transition(const qhip_state* bra, const qhip_state* ket, const qhip_observable* obs);
// transition() with its group table and partials in a fixed module-scope
// device array (no device allocation); bit-equal to transition().
cplx transition_noalloc(const qhip_state* bra, const qhip_state* ket, const qhip_observable* obs);
void apply_observable(const qhip_state* in, const qhip_observable* obs, qhip_state* out);
// apply_observable in two steps: group and upload (into ctx().obs), then
// launch without host work (the adjoint overlaps its reverse schedule with
// the device's forward sweep and H|psi>).
struct PreparedObservable;
std::shared_ptr<PreparedObservable> prepare_observable(const qhip_observable* obs, int n);
void launch_prepared_observable(const PreparedObservable& po, const qhip_state* in, qhip_state* out);
cplx inner(const qhip_state* a, const qhip_state* b);   // <a|b> in FP64
// Grouped observables of many instances in one device allocation, and
// stream-ordered H|in> and <a|b> (results stay on the device).
struct ObsArena;
ObsArena* obs_arena_create(int n_obs, const int32_t* n_qubits, const qhip_observable* obs);
void obs_arena_destroy(ObsArena* a);
void obs_apply_async(const ObsArena* a, int i, const qhip_state* in, qhip_state* out);
int obs_arena_groups(const ObsArena* a, int i);   // flip-mask groups of observable i
int inner_blocks(const qhip_state* a);
void inner_async(const qhip_state* a, const qhip_state* b, double2* scratch, double2* d_out);
int inner_blocks(const qhip_state* a);   // partials inner_async needs in scratch

void sample(const qhip_state* s, const double* uniforms, int64_t n, uint64_t* out);
void probabilities(const qhip_state* s, double* out);
// Indices (ascending) and |amp|^2 above threshold; returns their number and
// fills keys / probs when both are set (capacity must cover the number).
int64_t exact(const qhip_state* s, double threshold, int64_t* keys, double* probs, int64_t capacity);

// IP partial reduction: sum tile partials per slot (deterministic order).
void reduce_ip(const void* d_partials, int64_t n_tiles, int n_slots, cplx* host_out);

void apply_stream(const qhip_config* cfg, const qhip_stream* stream, qhip_state* s,
                  qhip_apply_stats* stats, bool terminal);
// Structured-program kernels (kernels_program.hip).
void permute_state(qhip_state* s, const uint32_t* qubits, int k, const uint64_t* table);
void set_basis_states(qhip_state* s, const uint64_t* labels, int64_t n_labels, uint64_t stride);
// Shot-batched trajectories (kernels_traj.hip).
void shots_broadcast(qhip_state* dst, const qhip_state* src);
void shots_restore(qhip_state* dst, const qhip_state* saved, const uint8_t* active, uint64_t stride);
void shots_prob0(const qhip_state* s, int q, uint64_t stride, double* p0);
void shots_apply1q(qhip_state* s, int q, uint64_t stride, const double* mats, const uint8_t* active);
void shots_rho1(const qhip_state* s, int q, uint64_t stride, double* out);
void shots_project(qhip_state* s, int q, uint64_t stride, const int8_t* action, const double* scale);
void shots_sample(const qhip_state* s, uint64_t stride, const double* u, int64_t n_shots, uint64_t* out);
// run()'s inline CDF walk over a whole state with uniform u, exact in
// parallel (kernels_sample.hip).
uint64_t walk_draw(const qhip_state* s, double u);
// Entries of the first n_amps amplitudes with hypot(re, im) > tol, ascending
// (kernels_sample.hip); returns the count, fills keys / values when given.
int64_t compact_values(const qhip_state* s, uint64_t n_amps, double tol, int64_t* keys, double* values,
                       int64_t capacity, double* max_pruned);
// generate_canonical<double, 53> over std::mt19937(seed), which is also
// uniform_real_distribution<double>(0, 1) (kernels_sample.hip).
void mt_uniforms(uint32_t seed, int64_t n, double* out);
void controlled_powers(const qhip_config* cfg, qhip_state* s, const uint32_t* controls, const uint64_t* exponents,
                       int n_controls, const uint32_t* targets, int m, const double* matrix);
Plan plan_stream(const qhip_config* cfg, const qhip_stream* stream, int n, int precision,
                 bool terminal, int64_t* fused_ops);
// Adjoint pieces shared by run_gradient and the batched driver: the reverse
// op list (generator inner products and daggers) with its slots, and the
// gradient from the slots' inner products.
struct GradSlot {
    cplx c;        // generator prefactor
    int64_t op;    // stream op index
    // -1: the slot's inner product goes to the op's parameters through its
    // dependency coefficients; p >= 0: it adds 2 Re(c ip) to parameter p
    // directly (a merged run); kMergedSlot: folded into such a slot.
    int32_t param = -1;
};
constexpr int32_t kMergedSlot = -2;
void build_backward(const Stream& st, int n, std::vector<GradSlot>& slots, std::vector<HostOp>& back);
// Runs of diagonal gates in the reverse sweep whose generators share few
// parameters: their inner products are taken together before the run's
// un-applies (all diagonal, so they commute), one OP_IP_PHASE per
// parameter, and the un-applies merge into one OP_PHASE.
void merge_adjoint_runs(const Stream& st, int n, std::vector<GradSlot>& slots, std::vector<HostOp>& back);
void accumulate_gradient(const qhip_stream* stream, const std::vector<GradSlot>& slots, const cplx* ip,
                         int n_params, double* grad);
ScheduleOptions schedule_options(const qhip_config* cfg, bool two_states);
void run_gradient(const qhip_config* cfg, const qhip_stream* stream, int n_params,
                  qhip_state* psi, const qhip_observable* obs, const qhip_state* phi_in,
                  double* grad, double* energy);
struct BatchSession;
BatchSession* batch_create(const qhip_config* cfg, int n_instances, const int32_t* n_qubits,
                           const qhip_stream* streams, const qhip_observable* observables, const int32_t* n_params,
                           int precision, bool gradients);
void batch_run(BatchSession* s, double* energies, double* grads, const int64_t* grad_offsets, bool use_graph);
bool batch_update(BatchSession* s, const qhip_stream* streams);
void batch_destroy(BatchSession* s);
void batch_stats(const BatchSession* s, qhip_batch_info* out);

// Planner derivations (derive.cpp).
bool permutation_table(const qhip_stream* local, int k, uint32_t max_rest, uint32_t max_work_log2, uint64_t* out);
uint64_t fnv1a_words(const uint64_t* words, int64_t n, uint64_t h);
void batch_energy_grad(const qhip_config* cfg, int n_instances, const int32_t* n_qubits,
                       const qhip_stream* streams, const qhip_observable* observables,
                       const int32_t* n_params, int precision, double* energies, double* grads,
                       const int64_t* grad_offsets);

}  // namespace qhip
