// runtime.cpp — device context, orchestration and the C entry points of
// include/qarp_hip.h.
#include "common.h"

#include <algorithm>
#include <bit>
#include <cmath>
#include <memory>
#include <mutex>

namespace qhip {

void mt_uniforms(uint32_t seed, int64_t n, double* u);   // kernels_sample.hip

namespace {
thread_local std::string t_err;
thread_local int t_code = QHIP_OK;
std::mutex g_lock;   // one call on the device at a time (shared stream and scratch)
}  // namespace

[[noreturn]] void fail(int code, const std::string& msg) { throw Error(code, msg); }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) fail(QHIP_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

void* DeviceBuffer::ensure(size_t n) {
    if (n <= bytes && ptr) return ptr;
    if (ptr) {
        QHIP_CHECK(hipStreamSynchronize(ctx().stream));
        QHIP_CHECK(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
    }
    const size_t want = std::max<size_t>(n, 1 << 20);
    if (hipMalloc(&ptr, want) != hipSuccess) {
        (void)hipGetLastError();
        ptr = nullptr;
        fail(QHIP_NOMEM, "device scratch allocation of " + std::to_string(want) + " bytes failed");
    }
    bytes = want;
    return ptr;
}

DeviceBuffer::~DeviceBuffer() {
    // Process teardown: the runtime may already be gone; leak rather than fault.
}

Context& ctx() {
    static Context* c = nullptr;
    if (!c) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            (void)hipGetLastError();
            fail(QHIP_DEVICE, "no HIP device visible");
        }
        auto* nc = new Context();
        QHIP_CHECK(hipGetDevice(&nc->device));
        QHIP_CHECK(hipStreamCreateWithFlags(&nc->stream, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        QHIP_CHECK(hipGetDeviceProperties(&prop, nc->device));
        nc->n_cu = prop.multiProcessorCount;
        c = nc;
    }
    return *c;
}

// ── validation ──────────────────────────────────────────────────────────────

static bool needs_trajectory(const qhip_op& op) {
    return op.gate == QHIP_RESET || (op.gate == QHIP_MEASURE && op.cbit >= 0) ||
           (op.flags & QHIP_OP_CONDITIONAL) || op.gate == QHIP_BRANCH_BEGIN;
}

static void validate(const Stream& st, int n) {
    for (int64_t i = 0; i < st.size(); ++i) {
        const qhip_op& op = st.op(i);
        if (op.gate < 0 || op.gate >= QHIP_NUM_GATES)
            fail(QHIP_INVALID, "unknown gate id " + std::to_string(op.gate));
        if (op.n_params < 0 || op.n_params > QHIP_MAX_PARAMS)
            fail(QHIP_INVALID, "bad parameter count " + std::to_string(op.n_params));
        uint64_t seen = 0;
        for (int b = 0; b < op.n_qubits; ++b) {
            const uint32_t q = st.qubits(i)[b];
            if (q >= static_cast<uint32_t>(n))
                fail(QHIP_INVALID, "QarpSimulator: qubit " + std::to_string(q) + " out of range for a " +
                                       std::to_string(n) + "-qubit register");
            if ((seen >> q) & 1) fail(QHIP_INVALID, "QarpSimulator: repeated qubit " + std::to_string(q));
            seen |= uint64_t{1} << q;
        }
    }
}

// measurements_are_terminal (qarp_simulator.cpp:226-252)
static bool terminal_only(const Stream& st) {
    uint64_t measured = 0;
    for (int64_t i = 0; i < st.size(); ++i) {
        const qhip_op& op = st.op(i);
        if (op.gate == QHIP_RESET || (op.flags & QHIP_OP_CONDITIONAL)) return false;
        if (op.gate == QHIP_BRANCH_BEGIN || op.gate == QHIP_BRANCH_ELSE || op.gate == QHIP_BRANCH_END) return false;
        if (op.gate == QHIP_BARRIER) continue;
        if (op.gate == QHIP_MEASURE) {
            if (op.cbit >= 0) {
                if (op.n_qubits == 0 || st.qubits(i)[0] >= 64) return false;
                measured |= uint64_t{1} << st.qubits(i)[0];
            }
            continue;
        }
        for (int b = 0; b < op.n_qubits; ++b) {
            const uint32_t q = st.qubits(i)[b];
            if (q >= 64 || ((measured >> q) & 1)) return false;
        }
    }
    return true;
}

static ScheduleOptions options(const qhip_config* cfg, bool two) {
    ScheduleOptions o;
    o.two_states = two;
    if (cfg) {
        o.tile_qubits = cfg->tile_qubits;
        o.low_qubits = cfg->low_qubits;
        o.max_ops = cfg->max_ops_per_pass;
        o.mfma = cfg->use_mfma;
        o.relabel = cfg->relabel_qubits != 0;
        o.stream = cfg->stream_tiles < 0 ? -1 : cfg->stream_tiles > 0 ? 1 : 0;
    }
    return o;
}

Plan plan_stream(const qhip_config* cfg, const qhip_stream* stream, int n, int precision, bool terminal,
                 int64_t* fused_ops) {
    const Stream st{stream};
    validate(st, n);
    if (!terminal) {
        for (int64_t i = 0; i < st.size(); ++i)
            if (needs_trajectory(st.op(i)))
                fail(QHIP_RUNTIME,
                     "QarpSimulator::statevector: circuit contains mid-circuit measurement, Reset, or classical "
                     "condition — statevector is not deterministic. Use run() for trajectory simulation.");
    } else if (!terminal_only(st)) {
        fail(QHIP_CAPABILITY,
             "QarpSimulator::run: mid-circuit measurement, Reset, classical conditions and branches need the "
             "per-shot trajectory path, which this backend does not provide");
    }
    const int kmax = cfg ? cfg->fusion_max_qubits : 3;
    const int kmin = cfg ? cfg->fusion_min_qubits : 12;
    const auto fused = fuse_for_dispatch(st, n, kmax, kmin);
    std::vector<HostOp> ops;
    ops.reserve(fused.size());
    for (const auto& it : fused) lower_item(st, it, n, ops, terminal);
    if (fused_ops) *fused_ops = static_cast<int64_t>(fused.size());
    return schedule(ops, n, precision, options(cfg, false));
}

static void fill_stats(qhip_apply_stats* stats, const Plan& plan, int64_t fused) {
    if (!stats) return;
    stats->fused_ops = fused;
    stats->passes = static_cast<int64_t>(plan.passes.size());
    stats->diagonal_ops = plan.diag_ops;
    stats->mfma_ops = plan.mfma_ops;
    stats->flops = plan.flops;
    stats->stream_passes = plan.stream_passes;
}

void apply_stream(const qhip_config* cfg, const qhip_stream* stream, qhip_state* s, qhip_apply_stats* stats,
                  bool terminal) {
    int64_t fused = 0;
    const Plan plan = plan_stream(cfg, stream, s->n, s->precision, terminal, &fused);
    launch_plan(plan, s, nullptr, nullptr, 0);
    fill_stats(stats, plan, fused);
}

// The adjoint's second state: up to 1 GiB it lives in a grow-only context
// buffer reused across calls (no allocation per gradient); larger states
// are allocated for the call and freed after it.
struct StateHolder {
    static constexpr size_t kKeepBytes = size_t{1} << 30;
    qhip_state st;
    bool own = false;
    explicit StateHolder(const qhip_state* like) {
        st.n = like->n;
        st.precision = like->precision;
        st.owned = false;   // never freed through a qhip_state path
        const size_t bytes = st.dim() * st.amp_bytes();
        if (bytes <= kKeepBytes) {
            st.d = ctx().phi.ensure(bytes);
            return;
        }
        own = true;
        if (hipMalloc(&st.d, bytes) != hipSuccess) {
            (void)hipGetLastError();
            st.d = nullptr;
            fail(QHIP_NOMEM, "adjoint: allocation of the phi state failed");
        }
    }
    ~StateHolder() {
        if (own && st.d) {
            (void)hipStreamSynchronize(ctx().stream);
            (void)hipFree(st.d);
        }
    }
};

void build_backward(const Stream& st, int n, std::vector<GradSlot>& slots, std::vector<HostOp>& back) {
    validate(st, n);
    for (int64_t i = 0; i < st.size(); ++i)
        if (needs_trajectory(st.op(i)))
            fail(QHIP_RUNTIME, "QarpSimulator::run_gradient: trajectory commands (Measure / Reset / classical "
                               "condition) are not differentiable.");
    // Backward op list (reverse order): generator inner product, then dagger.
    for (int64_t i = st.size() - 1; i >= 0; --i) {
        const qhip_op& op = st.op(i);
        if (op.n_deps > 0) {
            HostOp G;
            cplx c;
            if (!generator(st, i, G, c))
                fail(QHIP_CAPABILITY, "run_gradient: gate id " + std::to_string(op.gate) +
                                          " carries a symbolic parameter but has no adjoint generator rule "
                                          "(Rx/Ry/Rz/P/GPhase/CRx/CRy/CRz/CP/RXX/RYY/RZZ); rebase it to those "
                                          "gates before differentiating.");
            G.slot = static_cast<int32_t>(slots.size());
            slots.push_back({c, i});
            back.push_back(std::move(G));
        }
        lower_dagger(st, i, back);
    }
    merge_adjoint_runs(st, n, slots, back);
}

void accumulate_gradient(const qhip_stream* stream, const std::vector<GradSlot>& slots, const cplx* ip,
                         int n_params, double* grad) {
    const Stream st{stream};
    for (int p = 0; p < n_params; ++p) grad[p] = 0.0;
    for (size_t s = 0; s < slots.size(); ++s) {
        if (slots[s].param == kMergedSlot) continue;   // taken by a merged run's slot
        const double base = 2.0 * (slots[s].c * ip[s]).real();
        if (slots[s].param >= 0) {
            if (slots[s].param < n_params) grad[slots[s].param] += base;
            continue;
        }
        const qhip_op& op = st.op(slots[s].op);
        for (int d = 0; d < op.n_deps; ++d) {
            const int p = stream->dep_index[op.dep_offset + d];
            if (p >= 0 && p < n_params) grad[p] += stream->dep_alpha[op.dep_offset + d] * base;
        }
    }
}

ScheduleOptions schedule_options(const qhip_config* cfg, bool two) { return options(cfg, two); }

void run_gradient(const qhip_config* cfg, const qhip_stream* stream, int n_params, qhip_state* psi,
                  const qhip_observable* obs, const qhip_state* phi_in, double* grad, double* energy) {
    const Stream st{stream};
    const int n = psi->n;
    std::vector<GradSlot> slots;
    std::vector<HostOp> back;
    build_backward(st, n, slots, back);
    // Host work first while the device is idle (the forward schedule, the
    // observable's groups and their upload), then the forward sweep psi <-
    // U psi (fused; the gradient sweep itself is unfused) and phi = H psi are
    // queued, and the reverse schedule is built while the device runs them.
    int64_t fused = 0;
    const Plan fwd = plan_stream(cfg, stream, n, psi->precision, false, &fused);
    const auto prepared = obs ? prepare_observable(obs, n) : nullptr;
    launch_plan(fwd, psi, nullptr, nullptr, 0);
    StateHolder phi(psi);
    if (obs) {
        launch_prepared_observable(*prepared, psi, &phi.st);
    } else {
        QHIP_CHECK(hipMemcpyAsync(phi.st.d, phi_in->d, psi->dim() * psi->amp_bytes(), hipMemcpyDeviceToDevice,
                                  ctx().stream));
    }
    const Plan plan = schedule(back, n, psi->precision, options(cfg, true));
    // One device buffer: the inner-product totals, the energy <psi|phi>, and
    // the energy's partials.  The energy is reduced on the stream before the
    // reverse sweep changes psi and phi, and read back with the totals, so
    // the host does not wait between the two.
    auto& C = ctx();
    const size_t ns = slots.size();
    const bool want_energy = energy && obs;
    const int ib = want_energy ? inner_blocks(psi) : 0;
    double2* d_tot = static_cast<double2*>(C.misc2.ensure(sizeof(double2) * (ns + 2 + static_cast<size_t>(ib))));
    if (want_energy) inner_async(psi, &phi.st, d_tot + ns + 1, d_tot + ns);
    launch_plan(plan, psi, &phi.st, ns ? d_tot : nullptr, static_cast<int>(ns));
    std::vector<double2> h(ns + 1);
    if (ns || want_energy) {
        QHIP_CHECK(hipMemcpyAsync(h.data(), d_tot, sizeof(double2) * (ns + 1), hipMemcpyDeviceToHost, C.stream));
        QHIP_CHECK(hipStreamSynchronize(C.stream));
    }
    if (energy) *energy = want_energy ? h[ns].x : 0.0;
    std::vector<cplx> ip(ns);
    for (size_t k = 0; k < ns; ++k) ip[k] = cplx(h[k].x, h[k].y);
    accumulate_gradient(stream, slots, ip.data(), n_params, grad);
    // psi is back at the root state after the reverse sweep; leave U|root>
    // as documented: re-run the forward sweep only if the caller needs it —
    // the energy is already computed, so the root state is returned.
}

}  // namespace qhip

// ── C ABI ───────────────────────────────────────────────────────────────────

using qhip::Error;

template <typename F>
static int guarded(F&& f) {
    try {
        std::lock_guard<std::mutex> lk(qhip::g_lock);
        f();
        qhip::t_code = QHIP_OK;
        return QHIP_OK;
    } catch (const Error& e) {
        qhip::t_err = e.what();
        qhip::t_code = e.code;
        return e.code;
    } catch (const std::bad_alloc&) {
        qhip::t_err = "host allocation failed";
        qhip::t_code = QHIP_NOMEM;
        return QHIP_NOMEM;
    } catch (const std::exception& e) {
        qhip::t_err = e.what();
        qhip::t_code = QHIP_RUNTIME;
        return QHIP_RUNTIME;
    }
}

static void need(bool cond, const char* msg) {
    if (!cond) qhip::fail(QHIP_INVALID, msg);
}

struct qhip_fused {
    std::vector<qhip::FusedItem> items;
    std::vector<qhip_op> src_ops;
    std::vector<std::vector<uint32_t>> src_qubits;
    std::vector<std::vector<qhip::cplx>> src_mats;   // Custom ops re-emitted verbatim
};

extern "C" {

int qhip_last_error(char* buf, size_t len) {
    if (buf && len) {
        const size_t n = std::min(len - 1, qhip::t_err.size());
        std::memcpy(buf, qhip::t_err.data(), n);
        buf[n] = '\0';
    }
    return qhip::t_code;
}

int qhip_device_count(int32_t* out) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *out = n;
    return QHIP_OK;
}

int qhip_synchronize(void) {
    return guarded([&] { QHIP_CHECK(hipStreamSynchronize(qhip::ctx().stream)); });
}

int qhip_state_create(int32_t n_qubits, int32_t precision, qhip_state** out) {
    return guarded([&] {
        need(out != nullptr, "qhip_state_create: out is NULL");
        need(n_qubits >= 0 && n_qubits <= 40, "qhip_state_create: n_qubits must be in [0, 40]");
        need(precision == QHIP_C128 || precision == QHIP_C64, "qhip_state_create: bad precision");
        qhip::ctx();
        auto s = std::make_unique<qhip_state>();
        s->n = n_qubits;
        s->precision = precision;
        const size_t bytes = s->dim() * s->amp_bytes();
        if (hipMalloc(&s->d, bytes) != hipSuccess) {
            (void)hipGetLastError();
            qhip::fail(QHIP_NOMEM, "qhip_state_create: device allocation of " + std::to_string(bytes) +
                                       " bytes failed");
        }
        qhip::state_reset(s.get());
        QHIP_CHECK(hipStreamSynchronize(qhip::ctx().stream));
        *out = s.release();
    });
}

int qhip_state_destroy(qhip_state* s) {
    return guarded([&] {
        if (!s) return;
        QHIP_CHECK(hipStreamSynchronize(qhip::ctx().stream));
        if (s->owned && s->d) QHIP_CHECK(hipFree(s->d));
        delete s;
    });
}

int qhip_state_info(const qhip_state* s, int32_t* n_qubits, int32_t* precision) {
    return guarded([&] {
        need(s != nullptr, "qhip_state_info: NULL state");
        if (n_qubits) *n_qubits = s->n;
        if (precision) *precision = s->precision;
    });
}

int qhip_state_reset(qhip_state* s) {
    return guarded([&] {
        need(s != nullptr, "NULL state");
        qhip::state_reset(s);
        QHIP_CHECK(hipStreamSynchronize(qhip::ctx().stream));
    });
}

int qhip_state_upload(qhip_state* s, const void* host, uint64_t n_amps) {
    return guarded([&] {
        need(s && host, "qhip_state_upload: NULL argument");
        need(n_amps == s->dim(), "qhip_state_upload: length does not match 2^n_qubits");
        QHIP_CHECK(hipMemcpyAsync(s->d, host, n_amps * s->amp_bytes(), hipMemcpyHostToDevice, qhip::ctx().stream));
        QHIP_CHECK(hipStreamSynchronize(qhip::ctx().stream));
    });
}

int qhip_state_download(const qhip_state* s, void* host, uint64_t n_amps) {
    return guarded([&] {
        need(s && host, "qhip_state_download: NULL argument");
        need(n_amps == s->dim(), "qhip_state_download: length does not match 2^n_qubits");
        QHIP_CHECK(hipMemcpyAsync(host, s->d, n_amps * s->amp_bytes(), hipMemcpyDeviceToHost, qhip::ctx().stream));
        QHIP_CHECK(hipStreamSynchronize(qhip::ctx().stream));
    });
}

int qhip_state_upload_c128(qhip_state* s, const double* host, uint64_t n_amps) {
    return guarded([&] {
        need(s && host, "qhip_state_upload_c128: NULL argument");
        need(n_amps == s->dim(), "qhip_state_upload_c128: length does not match 2^n_qubits");
        qhip::state_convert_upload(s, host, n_amps);
    });
}

int qhip_state_download_c128(const qhip_state* s, double* host, uint64_t n_amps) {
    return guarded([&] {
        need(s && host, "qhip_state_download_c128: NULL argument");
        need(n_amps == s->dim(), "qhip_state_download_c128: length does not match 2^n_qubits");
        qhip::state_convert_download(s, host, n_amps);
    });
}

int qhip_state_copy(qhip_state* dst, const qhip_state* src) {
    return guarded([&] {
        need(dst && src, "qhip_state_copy: NULL argument");
        need(dst->n == src->n && dst->precision == src->precision, "qhip_state_copy: shape mismatch");
        QHIP_CHECK(hipMemcpyAsync(dst->d, src->d, src->dim() * src->amp_bytes(), hipMemcpyDeviceToDevice,
                                  qhip::ctx().stream));
        QHIP_CHECK(hipStreamSynchronize(qhip::ctx().stream));
    });
}

int qhip_state_norm2(const qhip_state* s, double* out) {
    return guarded([&] {
        need(s && out, "qhip_state_norm2: NULL argument");
        *out = qhip::state_norm2(s);
    });
}

int qhip_state_device_ptr(const qhip_state* s, void** out) {
    return guarded([&] {
        need(s && out, "qhip_state_device_ptr: NULL argument");
        *out = s->d;
    });
}

int qhip_fuse(const qhip_config* cfg, const qhip_stream* in, int32_t n_qubits, qhip_fused** out) {
    return guarded([&] {
        need(cfg && in && out, "qhip_fuse: NULL argument");
        const qhip::Stream st{in};
        qhip::validate(st, n_qubits);
        auto f = std::make_unique<qhip_fused>();
        f->items = qhip::fuse_for_dispatch(st, n_qubits, cfg->fusion_max_qubits, cfg->fusion_min_qubits);
        for (const auto& it : f->items) {
            if (it.src >= 0) {
                const qhip_op& op = st.op(it.src);
                f->src_ops.push_back(op);
                f->src_qubits.emplace_back(st.qubits(it.src), st.qubits(it.src) + op.n_qubits);
                std::vector<qhip::cplx> m;
                if (op.gate == QHIP_CUSTOM && op.unitary_offset >= 0 && op.n_qubits <= QHIP_MAX_FUSION_QUBITS) {
                    const size_t D = size_t{1} << op.n_qubits;
                    m.resize(D * D);
                    std::memcpy(m.data(), in->unitaries + 2 * op.unitary_offset, sizeof(qhip::cplx) * D * D);
                }
                f->src_mats.push_back(std::move(m));
            } else {
                f->src_ops.push_back(qhip_op{});
                f->src_qubits.emplace_back();
                f->src_mats.emplace_back();
            }
        }
        *out = f.release();
    });
}

int qhip_schedule_inspect(const qhip_config* cfg, const qhip_stream* in, int32_t n_qubits, int32_t precision,
                          qhip_fused** out) {
    return guarded([&] {
        need(in && out, "qhip_schedule_inspect: NULL argument");
        need(n_qubits >= 0 && n_qubits <= 40, "qhip_schedule_inspect: n_qubits must be in [0, 40]");
        need(precision == QHIP_C128 || precision == QHIP_C64, "qhip_schedule_inspect: bad precision");
        const qhip::Plan plan = qhip::plan_stream(cfg, in, n_qubits, precision, false, nullptr);
        auto f = std::make_unique<qhip_fused>();
        for (size_t pi = 0; pi < plan.passes.size(); ++pi) {
            const qhip::PassDesc& P = plan.passes[pi];
            auto global = [&](int pos) -> uint32_t {
                if (P.T == 0) return static_cast<uint32_t>(pos);
                return pos < 32 ? static_cast<uint32_t>(P.qpos[pos]) : static_cast<uint32_t>(pos - 32);
            };
            for (int o = P.op_begin; o < P.op_end; ++o) {
                const qhip::DevOp& d = plan.ops[o];
                qhip::FusedItem it;
                it.pass = static_cast<int32_t>(pi);
                switch (d.kind) {
                    case qhip::OP_BLOCK:
                        continue;   // register-block header: its OP_SUB entries follow
                    case qhip::OP_MFMA:
                        it.k = d.k;
                        for (int b = 0; b < d.k; ++b) it.q[b] = global(d.pos[b]);
                        it.m = qhip::mfma_matrix(plan.coefs.data() + d.coef, d.k, precision != QHIP_C128);
                        break;
                    case qhip::OP_SUB:   // pos[] holds the member op's own operand positions
                    case qhip::OP_DENSE: {
                        const bool sub = d.kind == qhip::OP_SUB;
                        const bool real2 = sub && ((d.k >= qhip::kReal2 && d.k < qhip::kReal2 + 3) ||
                                                   (d.k >= qhip::kImag2 && d.k < qhip::kImag2 + 3));
                        const bool neg = sub && d.k >= qhip::kNegPair && d.k < qhip::kNegPair + 3;
                        const bool diag = sub && ((d.k >= 6 && d.k <= 12) || neg);
                        if (diag) {   // local operands (ascending), then the ones outside the tile
                            const int kl = neg ? 2 : d.k == 12 ? 0 : d.k <= 8 ? 1 : 2, kg = neg ? 0 : d.spos[7];
                            it.k = kl + kg;
                            for (int b = 0; b < kl; ++b) it.q[b] = global(d.pos[b]);
                            for (int j = 0; j < kg; ++j) it.q[kl + j] = static_cast<uint32_t>(d.spos[j]);
                            const size_t D = size_t{1} << it.k;
                            it.m.assign(D * D, qhip::cplx(0.0));
                            for (size_t y = 0; y < D; ++y) it.m[y * D + y] = plan.coefs[d.coef + y];
                            break;
                        }
                        it.k = sub ? (d.k <= 2 || real2 ? 1 : 2) : d.k;
                        const size_t D = size_t{1} << it.k;
                        if (sub && it.k == 2 && d.pos[0] > d.pos[1]) {
                            // stored in block-bit order: report it on ascending positions
                            it.q[0] = global(d.pos[1]);
                            it.q[1] = global(d.pos[0]);
                        } else {
                            for (int b = 0; b < it.k; ++b) it.q[b] = global(d.pos[b]);
                        }
                        it.m.assign(plan.coefs.begin() + d.coef, plan.coefs.begin() + d.coef + D * D);
                        break;
                    }
                    case qhip::OP_DIAG: {
                        it.k = d.k;
                        for (int b = 0; b < d.k; ++b) it.q[b] = global(d.pos[b]);
                        const size_t D = size_t{1} << d.k;
                        it.m.assign(D * D, qhip::cplx(0.0));
                        for (size_t y = 0; y < D; ++y) it.m[y * D + y] = plan.coefs[d.coef + y];
                        break;
                    }
                    case qhip::OP_DIAGG: {   // local operands, then the global ones
                        const int kg = d.spos[7];
                        it.k = d.k + kg;
                        for (int b = 0; b < d.k; ++b) it.q[b] = global(d.pos[b]);
                        for (int j = 0; j < kg; ++j) it.q[d.k + j] = static_cast<uint32_t>(d.spos[j]);
                        const size_t D = size_t{1} << it.k;
                        it.m.assign(D * D, qhip::cplx(0.0));
                        for (size_t y = 0; y < D; ++y) it.m[y * D + y] = plan.coefs[d.coef + y];
                        break;
                    }
                    case qhip::OP_MASKPHASE: {
                        it.gate = QHIP_MCZ;
                        for (int l = 0; l < 64; ++l)
                            if ((d.lmask >> l) & 1) it.wide_q.push_back(global(l));
                        for (int q = 0; q < 64; ++q)
                            if ((d.gmask >> q) & 1) it.wide_q.push_back(static_cast<uint32_t>(q));
                        it.k = static_cast<int32_t>(it.wide_q.size());
                        it.phase = plan.coefs[d.coef];
                        break;
                    }
                    case qhip::OP_SCALAR:
                        it.gate = QHIP_GPHASE;
                        it.phase = plan.coefs[d.coef];
                        break;
                    case qhip::OP_PHASE: {
                        // One GPhase for the constant, then each term as the
                        // diagonal exp(i w z_a z_b) (or exp(i w z_a)).
                        auto push = [&](qhip::FusedItem&& x) {
                            f->items.push_back(std::move(x));
                            f->src_ops.push_back(qhip_op{});
                            f->src_qubits.emplace_back();
                            f->src_mats.emplace_back();
                        };
                        const qhip::cplx* rec = plan.coefs.data() + d.coef;
                        it.gate = QHIP_GPHASE;
                        it.phase = rec[0];
                        push(std::move(it));
                        auto key_qubit = [&](int kb) { return kb >= 56 ? static_cast<int>(global(kb - 56)) : kb == 55 ? -1 : kb; };
                        // A term: factor e where its operand bits agree, conj(e) where they differ.
                        auto term = [&](int qa, int qb, qhip::cplx e) {
                            qhip::FusedItem t;
                            t.pass = static_cast<int32_t>(pi);
                            if (qa < 0) std::swap(qa, qb);
                            t.k = qb < 0 ? 1 : 2;
                            t.q[0] = static_cast<uint32_t>(qa);
                            if (qb >= 0) t.q[1] = static_cast<uint32_t>(qb);
                            const size_t D = size_t{1} << t.k;
                            t.m.assign(D * D, qhip::cplx(0.0));
                            for (size_t y = 0; y < D; ++y) t.m[y * D + y] = std::popcount(y) & 1 ? std::conj(e) : e;
                            push(std::move(t));
                        };
                        const uint64_t lm = d.lmask, gm = d.gmask;
                        const size_t n_none = lm & 0xffffu;
                        const bool pair_table = ((lm >> 16) & 0xffffu) != 0;
                        size_t o = 1;
                        for (size_t i = 0; i < n_none; ++i, o += 2) {
                            const int code = static_cast<int>(rec[o + 1].real());
                            term(key_qubit(code & 0xff), key_qubit(code >> 8), rec[o]);
                        }
                        for (int m = 0; m < 6; ++m) {
                            const size_t cnt = (m < 2 ? lm >> (32 + 16 * m) : gm >> (16 * (m - 2))) & 0xffffu;
                            for (size_t i = 0; i < cnt; ++i, o += 2)
                                term(static_cast<int>(global(8 + m)), key_qubit(static_cast<int>(rec[o + 1].real())),
                                     rec[o]);
                        }
                        if (pair_table) {   // a diagonal over the row bits
                            qhip::FusedItem t;
                            t.pass = static_cast<int32_t>(pi);
                            t.k = P.T - 8;
                            for (int m = 0; m < t.k; ++m) t.q[m] = global(8 + m);
                            const size_t D = size_t{1} << t.k;
                            t.m.assign(D * D, qhip::cplx(0.0));
                            for (size_t y = 0; y < D; ++y) t.m[y * D + y] = rec[o + y];
                            push(std::move(t));
                        }
                        continue;
                    }
                    default:
                        qhip::fail(QHIP_RUNTIME, "qhip_schedule_inspect: unexpected op kind");
                }
                f->items.push_back(std::move(it));
                f->src_ops.push_back(qhip_op{});
                f->src_qubits.emplace_back();
                f->src_mats.emplace_back();
            }
            // The store permutation as SWAPs of the pass's global bits:
            // position i ends up holding what position sperm[i] held.
            int held[16];
            for (int i = 0; i < 16; ++i) held[i] = i;
            for (int i = 0; i < P.T && P.T <= 16; ++i) {
                int j = i;
                while (j < P.T && held[j] != P.sperm[i]) ++j;
                if (j == P.T) qhip::fail(QHIP_RUNTIME, "qhip_schedule_inspect: store permutation is not a permutation");
                if (j == i) continue;
                std::swap(held[i], held[j]);
                qhip::FusedItem it;
                it.pass = static_cast<int32_t>(pi);
                it.k = 2;
                it.q[0] = static_cast<uint32_t>(P.qpos[i]);
                it.q[1] = static_cast<uint32_t>(P.qpos[j]);
                it.m.assign(16, qhip::cplx(0.0));
                for (int x = 0; x < 4; ++x) it.m[(((x & 1) << 1) | (x >> 1)) * 4 + x] = 1.0;
                f->items.push_back(std::move(it));
                f->src_ops.push_back(qhip_op{});
                f->src_qubits.emplace_back();
                f->src_mats.emplace_back();
            }
        }
        *out = f.release();
    });
}

int qhip_fused_pass(const qhip_fused* f, int64_t i, int32_t* pass) {
    return guarded([&] {
        need(f && pass, "qhip_fused_pass: NULL argument");
        need(i >= 0 && i < static_cast<int64_t>(f->items.size()), "qhip_fused_pass: index out of range");
        *pass = f->items[i].pass;
    });
}

int qhip_fused_count(const qhip_fused* f, int64_t* n_ops) {
    return guarded([&] {
        need(f && n_ops, "qhip_fused_count: NULL argument");
        *n_ops = static_cast<int64_t>(f->items.size());
    });
}

int qhip_fused_get(const qhip_fused* f, int64_t i, int32_t* gate, int32_t* n_qubits, uint32_t* qubits,
                   double* params, double* matrix, int64_t capacity) {
    return guarded([&] {
        need(f && gate && n_qubits && qubits && params, "qhip_fused_get: NULL argument");
        need(i >= 0 && i < static_cast<int64_t>(f->items.size()), "qhip_fused_get: index out of range");
        const auto& it = f->items[i];
        if (it.src >= 0) {
            const qhip_op& op = f->src_ops[i];
            *gate = op.gate;
            *n_qubits = op.n_qubits;
            for (int b = 0; b < op.n_qubits && b < 64; ++b) qubits[b] = f->src_qubits[i][b];
            for (int p = 0; p < 4; ++p) params[p] = p < op.n_params ? op.params[p] : 0.0;
            const auto& m = f->src_mats[i];
            if (matrix && !m.empty()) {
                need(capacity >= static_cast<int64_t>(m.size()), "qhip_fused_get: matrix capacity too small");
                std::memcpy(matrix, m.data(), sizeof(qhip::cplx) * m.size());
            }
        } else if (it.gate == QHIP_MCZ || it.gate == QHIP_GPHASE) {
            *gate = it.gate;
            *n_qubits = it.k;
            for (int b = 0; b < it.k && b < 64; ++b) qubits[b] = it.wide_q[b];
            params[0] = it.phase.real();
            params[1] = it.phase.imag();
            params[2] = params[3] = 0.0;
        } else {
            *gate = QHIP_CUSTOM;
            *n_qubits = it.k;
            for (int b = 0; b < it.k; ++b) qubits[b] = it.q[b];
            for (int p = 0; p < 4; ++p) params[p] = 0.0;
            const int64_t D = int64_t{1} << it.k;
            if (matrix) {
                need(capacity >= D * D, "qhip_fused_get: matrix capacity too small");
                std::memcpy(matrix, it.m.data(), sizeof(qhip::cplx) * D * D);
            }
        }
    });
}

int qhip_fused_destroy(qhip_fused* f) {
    return guarded([&] { delete f; });
}

int qhip_apply(const qhip_config* cfg, const qhip_stream* stream, qhip_state* s, qhip_apply_stats* stats) {
    return guarded([&] {
        need(stream && s, "qhip_apply: NULL argument");
        qhip::apply_stream(cfg, stream, s, stats, false);
    });
}

int qhip_state_permute(qhip_state* s, const uint32_t* qubits, int32_t k, const uint64_t* table) {
    return guarded([&] {
        need(s && (k <= 0 || (qubits && table)), "qhip_state_permute: NULL argument");
        qhip::permute_state(s, qubits, k, table);
    });
}

int qhip_state_controlled_powers(const qhip_config* cfg, qhip_state* s, const uint32_t* controls,
                                 const uint64_t* exponents, int32_t n_controls, const uint32_t* targets,
                                 int32_t m, const double* matrix) {
    return guarded([&] {
        need(s && (n_controls <= 0 || (controls && exponents)) && (m <= 0 || (targets && matrix)),
             "qhip_state_controlled_powers: NULL argument");
        qhip::controlled_powers(cfg, s, controls, exponents, n_controls, targets, m, matrix);
    });
}

int qhip_state_set_basis(qhip_state* s, const uint64_t* labels, int64_t n_labels, uint64_t stride) {
    return guarded([&] {
        need(s && (n_labels <= 0 || labels), "qhip_state_set_basis: NULL argument");
        qhip::set_basis_states(s, labels, n_labels, stride);
    });
}

int qhip_compact(const qhip_state* s, uint64_t n_amps, double tolerance, int64_t* keys, double* values,
                 int64_t capacity, int64_t* n_out, double* max_pruned) {
    return guarded([&] {
        need(s && n_out, "qhip_compact: NULL argument");
        *n_out = qhip::compact_values(s, n_amps, tolerance, keys, values, capacity, max_pruned);
    });
}

int qhip_shots_apply1q(qhip_state* s, int32_t qubit, uint64_t stride, const double* mats, const uint8_t* active) {
    return guarded([&] {
        need(s && mats && active, "qhip_shots_apply1q: NULL argument");
        qhip::shots_apply1q(s, qubit, stride, mats, active);
    });
}

int qhip_shots_rho1(const qhip_state* s, int32_t qubit, uint64_t stride, double* out) {
    return guarded([&] {
        need(s && out, "qhip_shots_rho1: NULL argument");
        qhip::shots_rho1(s, qubit, stride, out);
    });
}

int qhip_shots_broadcast(qhip_state* dst, const qhip_state* src) {
    return guarded([&] {
        need(dst && src, "qhip_shots_broadcast: NULL argument");
        qhip::shots_broadcast(dst, src);
    });
}

int qhip_shots_restore(qhip_state* dst, const qhip_state* saved, const uint8_t* active, uint64_t stride) {
    return guarded([&] {
        need(dst && saved && active, "qhip_shots_restore: NULL argument");
        qhip::shots_restore(dst, saved, active, stride);
    });
}

int qhip_shots_prob0(const qhip_state* s, int32_t qubit, uint64_t stride, double* p0) {
    return guarded([&] {
        need(s && p0, "qhip_shots_prob0: NULL argument");
        qhip::shots_prob0(s, qubit, stride, p0);
    });
}

int qhip_shots_project(qhip_state* s, int32_t qubit, uint64_t stride, const int8_t* action, const double* scale) {
    return guarded([&] {
        need(s && action && scale, "qhip_shots_project: NULL argument");
        qhip::shots_project(s, qubit, stride, action, scale);
    });
}

int qhip_shots_sample(const qhip_state* s, uint64_t stride, const double* u, int64_t n_shots, uint64_t* out) {
    return guarded([&] {
        need(s && (n_shots <= 0 || (u && out)), "qhip_shots_sample: NULL argument");
        qhip::shots_sample(s, stride, u, n_shots, out);
    });
}

int qhip_mt_uniforms(uint32_t seed, int64_t n, double* out) {
    return guarded([&] {
        need(n <= 0 || out, "qhip_mt_uniforms: NULL argument");
        qhip::mt_uniforms(seed, n, out);
    });
}

int qhip_plan_create(const qhip_config* cfg, const qhip_stream* stream, int32_t n_qubits, int32_t precision,
                     qhip_plan** out) {
    return guarded([&] {
        need(stream && out, "qhip_plan_create: NULL argument");
        need(n_qubits >= 0 && n_qubits <= 40, "qhip_plan_create: n_qubits must be in [0, 40]");
        need(precision == QHIP_C128 || precision == QHIP_C64, "qhip_plan_create: bad precision");
        qhip::ctx();
        int64_t fused = 0;
        qhip::Plan plan = qhip::plan_stream(cfg, stream, n_qubits, precision, false, &fused);
        qhip::CompiledPlan* cp = qhip::compile_plan(std::move(plan), n_qubits, precision);
        cp->fused_ops = fused;
        *out = reinterpret_cast<qhip_plan*>(cp);
    });
}

int qhip_plan_apply(const qhip_plan* plan, qhip_state* s, qhip_apply_stats* stats) {
    return guarded([&] {
        need(plan && s, "qhip_plan_apply: NULL argument");
        const auto* cp = reinterpret_cast<const qhip::CompiledPlan*>(plan);
        need(cp->n == s->n && cp->precision == s->precision, "qhip_plan_apply: state shape differs from the plan");
        qhip::run_compiled(*cp, s, nullptr);
        qhip::fill_stats(stats, cp->plan, cp->fused_ops);
    });
}

int qhip_plan_apply_graph(qhip_plan* plan, qhip_state* s, qhip_apply_stats* stats) {
    return guarded([&] {
        need(plan && s, "qhip_plan_apply_graph: NULL argument");
        auto* cp = reinterpret_cast<qhip::CompiledPlan*>(plan);
        need(cp->n == s->n && cp->precision == s->precision, "qhip_plan_apply_graph: state shape differs from the plan");
        qhip::run_compiled_graph(*cp, s);
        qhip::fill_stats(stats, cp->plan, cp->fused_ops);
    });
}

int qhip_plan_update(qhip_plan* plan, const qhip_config* cfg, const qhip_stream* stream, int32_t* rebuilt) {
    return guarded([&] {
        need(plan && stream, "qhip_plan_update: NULL argument");
        auto* cp = reinterpret_cast<qhip::CompiledPlan*>(plan);
        int64_t fused = 0;
        qhip::Plan p = qhip::plan_stream(cfg, stream, cp->n, cp->precision, false, &fused);
        const bool r = qhip::update_compiled(*cp, std::move(p), fused);
        if (rebuilt) *rebuilt = r ? 1 : 0;
    });
}

int qhip_plan_apply_timed(const qhip_plan* plan, qhip_state* s, double* pass_ms, int64_t capacity) {
    return guarded([&] {
        need(plan && s && (pass_ms || capacity == 0), "qhip_plan_apply_timed: NULL argument");
        const auto* cp = reinterpret_cast<const qhip::CompiledPlan*>(plan);
        need(cp->n == s->n && cp->precision == s->precision, "qhip_plan_apply_timed: state shape differs");
        std::vector<double> ms;
        qhip::run_compiled(*cp, s, &ms);
        for (int64_t i = 0; i < capacity && i < static_cast<int64_t>(ms.size()); ++i) pass_ms[i] = ms[i];
    });
}

int qhip_plan_destroy(qhip_plan* plan) {
    return guarded([&] { qhip::free_compiled(reinterpret_cast<qhip::CompiledPlan*>(plan)); });
}

int qhip_apply_terminal(const qhip_config* cfg, const qhip_stream* stream, qhip_state* s,
                        qhip_apply_stats* stats) {
    return guarded([&] {
        need(stream && s, "qhip_apply_terminal: NULL argument");
        qhip::apply_stream(cfg, stream, s, stats, true);
    });
}

int qhip_transition(const qhip_state* bra, const qhip_state* ket, const qhip_observable* obs, double out[2]) {
    return guarded([&] {
        need(bra && ket && obs && out, "qhip_transition: NULL argument");
        if (ket->n < 0 || ket->n > 62)
            qhip::fail(QHIP_INVALID, "pauli_transition: n_qubits must be in [0, 62], got " + std::to_string(ket->n));
        need(bra->n == ket->n && bra->precision == ket->precision,
             "QarpSimulator::transition: bra/ket lengths do not match");
        const qhip::cplx v = qhip::transition(bra, ket, obs);
        out[0] = v.real();
        out[1] = v.imag();
    });
}

int qhip_transition_noalloc(const qhip_state* bra, const qhip_state* ket, const qhip_observable* obs,
                            double out[2]) {
    return guarded([&] {
        need(bra && ket && obs && out, "qhip_transition_noalloc: NULL argument");
        if (ket->n < 0 || ket->n > 62)
            qhip::fail(QHIP_INVALID, "pauli_transition: n_qubits must be in [0, 62], got " + std::to_string(ket->n));
        need(bra->n == ket->n && bra->precision == ket->precision,
             "QarpSimulator::transition: bra/ket lengths do not match");
        const qhip::cplx v = qhip::transition_noalloc(bra, ket, obs);
        out[0] = v.real();
        out[1] = v.imag();
    });
}

int qhip_apply_observable(const qhip_state* in, const qhip_observable* obs, qhip_state* out) {
    return guarded([&] {
        need(in && obs && out, "qhip_apply_observable: NULL argument");
        need(in->n == out->n && in->precision == out->precision && in != out,
             "qhip_apply_observable: in/out must be distinct states of one shape");
        qhip::apply_observable(in, obs, out);
    });
}

int qhip_sample(const qhip_state* s, int64_t n_shots, uint32_t seed, uint64_t* outcomes) {
    return guarded([&] {
        need(s && (outcomes || n_shots == 0), "qhip_sample: NULL argument");
        need(n_shots >= 0, "qhip_sample: n_shots must be >= 0");
        std::vector<double> u(static_cast<size_t>(n_shots));
        qhip::mt_uniforms(seed, n_shots, u.data());
        qhip::sample(s, u.data(), n_shots, outcomes);
    });
}

int qhip_sample_uniforms(const qhip_state* s, const double* uniforms, int64_t n, uint64_t* outcomes) {
    return guarded([&] {
        need(s && (n == 0 || (uniforms && outcomes)), "qhip_sample_uniforms: NULL argument");
        need(n >= 0, "qhip_sample_uniforms: n must be >= 0");
        for (int64_t i = 0; i < n; ++i)
            need(uniforms[i] >= 0.0 && uniforms[i] < 1.0, "qhip_sample_uniforms: uniforms must lie in [0, 1)");
        qhip::sample(s, uniforms, n, outcomes);
    });
}

int qhip_probabilities(const qhip_state* s, double* out, uint64_t n) {
    return guarded([&] {
        need(s && out, "qhip_probabilities: NULL argument");
        need(n == s->dim(), "qhip_probabilities: length does not match 2^n_qubits");
        qhip::probabilities(s, out);
    });
}

int qhip_exact(const qhip_state* s, double threshold, int64_t* keys, double* probs, int64_t capacity,
               int64_t* count) {
    return guarded([&] {
        need(s && count, "qhip_exact: NULL argument");
        need(capacity >= 0, "qhip_exact: capacity must be >= 0");
        need(threshold >= 0.0, "qhip_exact: threshold must be >= 0");
        *count = qhip::exact(s, threshold, keys, probs, capacity);
    });
}

int qhip_run_gradient(const qhip_config* cfg, const qhip_stream* stream, int32_t n_params, qhip_state* psi,
                      const qhip_observable* obs, double* grad, double* energy) {
    return guarded([&] {
        need(stream && psi && obs && (grad || n_params == 0), "qhip_run_gradient: NULL argument");
        need(n_params >= 0, "qhip_run_gradient: n_params must be >= 0");
        qhip::run_gradient(cfg, stream, n_params, psi, obs, nullptr, grad, energy);
    });
}

int qhip_run_gradient_phi(const qhip_config* cfg, const qhip_stream* stream, int32_t n_params, qhip_state* psi,
                          const qhip_state* phi, double* grad) {
    return guarded([&] {
        need(stream && psi && phi && (grad || n_params == 0), "qhip_run_gradient_phi: NULL argument");
        need(phi->n == psi->n && phi->precision == psi->precision, "run_gradient: |phi> size mismatch");
        qhip::run_gradient(cfg, stream, n_params, psi, nullptr, phi, grad, nullptr);
    });
}

int qhip_batch_energy_grad(const qhip_config* cfg, int32_t n_instances, const int32_t* n_qubits,
                           const qhip_stream* streams, const qhip_observable* observables, const int32_t* n_params,
                           int32_t precision, double* energies, double* grads, const int64_t* grad_offsets) {
    return guarded([&] {
        need(n_instances >= 0 && n_qubits && streams && observables && energies, "qhip_batch_energy_grad: NULL argument");
        need(!grads || (grad_offsets && n_params), "qhip_batch_energy_grad: grads need offsets and counts");
        qhip::batch_energy_grad(cfg, n_instances, n_qubits, streams, observables, n_params, precision, energies,
                                grads, grad_offsets);
    });
}

int qhip_batch_create(const qhip_config* cfg, int32_t n_instances, const int32_t* n_qubits,
                      const qhip_stream* streams, const qhip_observable* observables, const int32_t* n_params,
                      int32_t precision, int32_t gradients, qhip_batch** out) {
    return guarded([&] {
        need(out && n_instances >= 0 && (n_instances == 0 || (n_qubits && streams && observables)),
             "qhip_batch_create: NULL argument");
        need(!gradients || n_instances == 0 || n_params, "qhip_batch_create: gradients need n_params");
        *out = reinterpret_cast<qhip_batch*>(qhip::batch_create(cfg, n_instances, n_qubits, streams, observables,
                                                                n_params, precision, gradients != 0));
    });
}

int qhip_batch_run(qhip_batch* b, double* energies, double* grads, const int64_t* grad_offsets, int32_t use_graph) {
    return guarded([&] {
        need(b && energies, "qhip_batch_run: NULL argument");
        need(!grads || grad_offsets, "qhip_batch_run: grads need offsets");
        qhip::batch_run(reinterpret_cast<qhip::BatchSession*>(b), energies, grads, grad_offsets, use_graph != 0);
    });
}

int qhip_batch_update(qhip_batch* b, const qhip_stream* streams, int32_t* rebuilt) {
    return guarded([&] {
        need(b && streams, "qhip_batch_update: NULL argument");
        const bool r = qhip::batch_update(reinterpret_cast<qhip::BatchSession*>(b), streams);
        if (rebuilt) *rebuilt = r ? 1 : 0;
    });
}

int qhip_permutation_table(const qhip_stream* local, int32_t k, uint32_t max_rest, uint32_t max_work_log2,
                           int32_t* found, uint64_t* table) {
    return guarded([&] {
        need(local && found && table, "qhip_permutation_table: NULL argument");
        *found = qhip::permutation_table(local, k, max_rest, max_work_log2, table) ? 1 : 0;
    });
}

int qhip_fnv1a_words(const uint64_t* words, int64_t n, uint64_t seed, uint64_t* out) {
    return guarded([&] {
        need(out && n >= 0 && (words || n == 0), "qhip_fnv1a_words: NULL argument");
        *out = qhip::fnv1a_words(words, n, seed);
    });
}

int qhip_batch_stats(const qhip_batch* b, qhip_batch_info* out) {
    return guarded([&] {
        need(b && out, "qhip_batch_stats: NULL argument");
        qhip::batch_stats(reinterpret_cast<const qhip::BatchSession*>(b), out);
    });
}

int qhip_batch_destroy(qhip_batch* b) {
    return guarded([&] { qhip::batch_destroy(reinterpret_cast<qhip::BatchSession*>(b)); });
}

}  // extern "C"
