// batch.cpp — many independent instances (own width, circuit, observable)
// in batched launches.
//
// Stands in for the sequential parameter-set loops of batch_expectation
// (src/qarpx/src/simulator/qarp_simulator.cpp:971-984) and the per-point
// sweep of QarpEngine._sweep (src/qarp/engines/_qarp_engine.py:450-503),
// with every instance differentiated by the adjoint method of run_gradient
// (qarp_simulator.cpp:1519-1624).
//
// Every instance is planned on the host (fused forward schedule, unfused
// reverse schedule with its generator slots), all op tables, observables,
// launch descriptors and states go to the device in one arena each, and
// then:
//   * forward step k: ONE tile launch runs pass k of every instance that has
//     one (tile_batch_kernel: workgroup -> (instance, tile) by bisection of a
//     prefix table), so the number of launches is the longest schedule, not
//     the sum of all schedules;
//   * per instance, phi = H psi and <psi|phi> stay on the device;
//   * reverse step k: one two-state tile launch over all instances, then one
//     launch reducing every inner-product op of that step (fixed tile order);
//   * one copy brings energies and inner products back; gradients follow
//     on the host as in run_gradient.
// Instances whose schedules contain a pass the batched kernel does not run
// (a dense op wider than five qubits) go through run_gradient one by one.
#include <cmath>
#include "common.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <exception>
#include <memory>
#include <thread>

namespace qhip {

namespace {

struct Inst {
    int n = 0;
    bool batched = true;
    Plan fwd, back;
    std::vector<GradSlot> slots;
    size_t fwd_ops = 0, fwd_coefs = 0, back_ops = 0, back_coefs = 0;   // arena offsets (elements)
    size_t psi = 0, phi = 0;                                           // state arena offsets (bytes)
    size_t tot = 0;                                                    // first slot in the totals arena
};

struct DevArena {
    void* p = nullptr;
    DevArena() = default;
    DevArena(const DevArena&) = delete;
    DevArena& operator=(const DevArena&) = delete;
    ~DevArena() { reset(); }
    void reset() {
        if (p) {
            (void)hipStreamSynchronize(ctx().stream);
            (void)hipFree(p);
            p = nullptr;
        }
    }
    void* alloc(size_t bytes, const char* what) {
        reset();
        if (hipMalloc(&p, std::max<size_t>(bytes, 256)) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            fail(QHIP_NOMEM, std::string("batch: ") + what + " allocation failed");
        }
        return p;
    }
};

struct Step {   // one batched launch
    std::vector<BatchItem> items;
    std::vector<int64_t> first;
    int64_t total = 0;
    int tmax = 0, kmax = 0;
    std::vector<IpReduce> jobs;        // two-state steps: reductions after the launch
    std::vector<size_t> item_ip;       // byte offset of each item's partials (npos: none)
    std::vector<size_t> job_in;        // byte offset of each job's partial row
    size_t items_off = 0, first_off = 0, jobs_off = 0;   // byte offsets in the descriptor arena
};

constexpr size_t npos = ~size_t{0};

size_t round_up(size_t b) { return (b + 255) / 256 * 256; }

bool batchable(const Plan& p) {
    for (const PassDesc& P : p.passes)
        if (P.T <= 0) return false;
    return true;
}

}  // namespace

// One set of instances: plans, device arenas and launch descriptors built
// once; a sweep (states reset, forward steps, H psi and energies, reverse
// steps and reductions) is a fixed sequence of kernel launches that can be
// captured as one hipGraph and replayed after the coefficients change.
struct BatchSession {
    const qhip_config* cfg = nullptr;
    qhip_config cfg_copy{};
    int n_inst = 0;
    int precision = QHIP_C128;
    bool want_grad = true;
    size_t amp = 16;
    std::vector<int32_t> n_qubits;
    std::vector<int32_t> n_params;
    const qhip_stream* streams = nullptr;          // caller-owned, valid until the next update
    const qhip_observable* observables = nullptr;
    std::vector<Inst> inst;
    DevArena tables, states, results, desc, ip_area;
    std::unique_ptr<ObsArena, void (*)(ObsArena*)> obs{nullptr, obs_arena_destroy};
    char* d_tab = nullptr;
    size_t ops_b = 0;
    double2* d_energy = nullptr;
    double2* d_tot = nullptr;
    double2* d_scratch = nullptr;
    char* d_states = nullptr;
    char* d_desc = nullptr;
    size_t n_tot = 0;
    std::vector<Step> fsteps, bsteps;
    hipGraphExec_t graph = nullptr;

    ~BatchSession() {
        if (graph) {
            (void)hipStreamSynchronize(ctx().stream);
            (void)hipGraphExecDestroy(graph);
        }
    }

    void plan_one(Inst& I, int i) const {
        I.n = n_qubits[i];
        if (I.n < 0 || I.n > 34) fail(QHIP_INVALID, "batch: n_qubits out of range");
        int64_t fused = 0;
        I.fwd = plan_stream(cfg, &streams[i], I.n, precision, false, &fused);
        if (want_grad) {
            std::vector<HostOp> back;
            build_backward(Stream{&streams[i]}, I.n, I.slots, back);
            if (!I.slots.empty()) I.back = schedule(back, I.n, precision, schedule_options(cfg, true));
        }
        I.batched = batchable(I.fwd) && batchable(I.back);
    }

    // Host planning of every instance (fusion, lowering, pass selection,
    // backward schedule) on a pool of host threads; instances are independent
    // and the planning code shares no mutable state.  The first failure (in
    // instance order) is rethrown.
    void plan_all(std::vector<Inst>& out) const {
        out.assign(n_inst, Inst{});
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const int workers = static_cast<int>(std::min<unsigned>(std::min(hw, 16u), static_cast<unsigned>(n_inst)));
        std::vector<std::exception_ptr> err(n_inst);
        std::atomic<int> next{0};
        auto work = [&] {
            for (int i = next++; i < n_inst; i = next++) {
                try {
                    plan_one(out[i], i);
                } catch (...) {
                    err[i] = std::current_exception();
                }
            }
        };
        if (workers <= 1) {
            work();
        } else {
            std::vector<std::thread> pool;
            pool.reserve(workers - 1);
            for (int w = 1; w < workers; ++w) pool.emplace_back(work);
            work();
            for (auto& t : pool) t.join();
        }
        for (const auto& e : err)
            if (e) std::rethrow_exception(e);
    }

    void upload_coefs(const std::vector<cplx>& coefs) {
        if (coefs.empty()) return;
        if (precision == QHIP_C128) {
            QHIP_CHECK(hipMemcpy(d_tab + ops_b, coefs.data(), coefs.size() * 16, hipMemcpyHostToDevice));
        } else {
            std::vector<float2> cf(coefs.size());
            for (size_t k = 0; k < cf.size(); ++k)
                cf[k] = make_float2(static_cast<float>(coefs[k].real()), static_cast<float>(coefs[k].imag()));
            QHIP_CHECK(hipMemcpy(d_tab + ops_b, cf.data(), cf.size() * 8, hipMemcpyHostToDevice));
        }
    }

    void build();
    void enqueue();
    void run(double* energies, double* grads, const int64_t* grad_offsets, bool use_graph);
    bool update(const qhip_stream* new_streams);
};

static bool same_structure(const Plan& a, const Plan& b) {
    if (a.passes.size() != b.passes.size() || a.ops.size() != b.ops.size() || a.coefs.size() != b.coefs.size())
        return false;
    if (!a.passes.empty() && std::memcmp(a.passes.data(), b.passes.data(), a.passes.size() * sizeof(PassDesc)))
        return false;
    return a.ops.empty() || std::memcmp(a.ops.data(), b.ops.data(), a.ops.size() * sizeof(DevOp)) == 0;
}

void BatchSession::build() {
    plan_all(inst);
    // ── arenas: op tables, coefficients, states, totals ──
    std::vector<DevOp> ops;
    std::vector<cplx> coefs;
    size_t state_bytes = 0;
    n_tot = 0;
    for (Inst& I : inst) {
        if (!I.batched) continue;
        I.fwd_ops = ops.size();
        I.fwd_coefs = coefs.size();
        ops.insert(ops.end(), I.fwd.ops.begin(), I.fwd.ops.end());
        coefs.insert(coefs.end(), I.fwd.coefs.begin(), I.fwd.coefs.end());
        I.back_ops = ops.size();
        I.back_coefs = coefs.size();
        ops.insert(ops.end(), I.back.ops.begin(), I.back.ops.end());
        coefs.insert(coefs.end(), I.back.coefs.begin(), I.back.coefs.end());
        const size_t sb = round_up((size_t{1} << I.n) * amp);
        I.psi = state_bytes;
        I.phi = state_bytes + sb;
        state_bytes += 2 * sb;
        I.tot = n_tot;
        n_tot += I.slots.size();
    }
    ops_b = round_up(ops.size() * sizeof(DevOp));
    const size_t coef_b = round_up(coefs.size() * amp);
    d_tab = static_cast<char*>(tables.alloc(ops_b + coef_b, "op table"));
    QHIP_CHECK(hipMemcpy(d_tab, ops.data(), ops.size() * sizeof(DevOp), hipMemcpyHostToDevice));
    upload_coefs(coefs);
    const DevOp* d_ops = reinterpret_cast<const DevOp*>(d_tab);
    const char* d_coefs = d_tab + ops_b;
    d_states = static_cast<char*>(states.alloc(state_bytes, "state"));
    // results: energies (double2 per instance), slot totals, inner scratch
    int scratch_n = 1;
    for (const Inst& I : inst) {
        if (!I.batched) continue;
        qhip_state tmp;
        tmp.n = I.n;
        tmp.precision = precision;
        scratch_n = std::max(scratch_n, inner_blocks(&tmp) + 1);
    }
    const size_t en_b = round_up(sizeof(double2) * n_inst), tot_b = round_up(sizeof(double2) * (n_tot + 1));
    char* d_res = static_cast<char*>(results.alloc(en_b + tot_b + sizeof(double2) * scratch_n, "result"));
    d_energy = reinterpret_cast<double2*>(d_res);
    d_tot = reinterpret_cast<double2*>(d_res + en_b);
    d_scratch = reinterpret_cast<double2*>(d_res + en_b + tot_b);

    // ── launch descriptors, per step ──
    size_t max_fwd = 0, max_back = 0;
    for (const Inst& I : inst)
        if (I.batched) {
            max_fwd = std::max(max_fwd, I.fwd.passes.size());
            max_back = std::max(max_back, I.back.passes.size());
        }
    fsteps.assign(max_fwd, Step{});
    bsteps.assign(max_back, Step{});
    size_t ip_bytes = 0;   // largest per-step inner-product partial area
    auto add_item = [&](Step& S, const Inst& I, const Plan& plan, size_t p, size_t ops_off, size_t coef_off,
                        bool two, size_t& ip_off) {
        const PassDesc& P = plan.passes[p];
        BatchItem b{};
        b.P = P;
        b.psi = d_states + I.psi;
        b.phi = two ? d_states + I.phi : nullptr;
        b.ops = d_ops + ops_off;
        b.coefs = d_coefs + coef_off * amp;
        b.n_tiles = int64_t{1} << (P.n - P.T);
        b.ip_out = nullptr;
        size_t my_ip = npos;
        if (two) {
            int n_ip = 0;
            for (int o = P.op_begin; o < P.op_end; ++o) {
                const DevOp& d = plan.ops[o];
                if (!is_ip_kind(d.kind)) continue;
                S.jobs.push_back({nullptr, b.n_tiles, d_tot + I.tot + static_cast<size_t>(d.lmask)});
                S.job_in.push_back(ip_off + sizeof(double2) * static_cast<size_t>(n_ip) * b.n_tiles);
                ++n_ip;
            }
            if (n_ip > 0) {
                my_ip = ip_off;
                ip_off += sizeof(double2) * static_cast<size_t>(n_ip) * b.n_tiles;
            }
        }
        S.item_ip.push_back(my_ip);
        S.first.push_back(S.total);
        S.total += b.n_tiles;
        S.tmax = std::max(S.tmax, P.T);
        S.kmax = std::max(S.kmax & ~kPhaseKernel, P.kmax & ~kPhaseKernel) | ((S.kmax | P.kmax) & kPhaseKernel);
        S.items.push_back(b);
    };
    for (size_t k = 0; k < max_fwd; ++k) {
        size_t dummy = 0;
        for (const Inst& I : inst)
            if (I.batched && k < I.fwd.passes.size()) add_item(fsteps[k], I, I.fwd, k, I.fwd_ops, I.fwd_coefs, false, dummy);
    }
    for (size_t k = 0; k < max_back; ++k) {
        size_t ip_off = 0;
        for (const Inst& I : inst)
            if (I.batched && k < I.back.passes.size())
                add_item(bsteps[k], I, I.back, k, I.back_ops, I.back_coefs, true, ip_off);
        ip_bytes = std::max(ip_bytes, ip_off);
    }
    char* d_ip = static_cast<char*>(ip_area.alloc(ip_bytes, "inner-product partial"));
    // Partial rows live in the partial area, at the offsets planned above.
    for (Step& S : bsteps) {
        for (size_t k = 0; k < S.items.size(); ++k)
            if (S.item_ip[k] != npos) S.items[k].ip_out = reinterpret_cast<double2*>(d_ip + S.item_ip[k]);
        for (size_t k = 0; k < S.jobs.size(); ++k) S.jobs[k].in = reinterpret_cast<const double2*>(d_ip + S.job_in[k]);
    }
    size_t desc_bytes = 0;
    for (auto* steps : {&fsteps, &bsteps})
        for (Step& S : *steps) {
            S.items_off = desc_bytes;
            desc_bytes += round_up(S.items.size() * sizeof(BatchItem));
            S.first_off = desc_bytes;
            desc_bytes += round_up(S.first.size() * sizeof(int64_t));
            S.jobs_off = desc_bytes;
            desc_bytes += round_up(S.jobs.size() * sizeof(IpReduce));
        }
    std::vector<char> hdesc(std::max<size_t>(desc_bytes, 1));
    for (auto* steps : {&fsteps, &bsteps})
        for (const Step& S : *steps) {
            std::memcpy(hdesc.data() + S.items_off, S.items.data(), S.items.size() * sizeof(BatchItem));
            std::memcpy(hdesc.data() + S.first_off, S.first.data(), S.first.size() * sizeof(int64_t));
            std::memcpy(hdesc.data() + S.jobs_off, S.jobs.data(), S.jobs.size() * sizeof(IpReduce));
        }
    d_desc = static_cast<char*>(desc.alloc(hdesc.size(), "descriptor"));
    QHIP_CHECK(hipMemcpy(d_desc, hdesc.data(), hdesc.size(), hipMemcpyHostToDevice));
    obs.reset(obs_arena_create(n_inst, n_qubits.data(), observables));
}

void BatchSession::enqueue() {
    auto& C = ctx();
    auto state_of = [&](const Inst& I, size_t off) {
        qhip_state s;
        s.n = I.n;
        s.precision = precision;
        s.d = d_states + off;
        s.owned = false;
        return s;
    };
    for (Inst& I : inst) {
        if (!I.batched) continue;
        qhip_state s = state_of(I, I.psi);
        state_reset(&s);
    }
    for (const Step& S : fsteps)
        launch_tile_batch(precision, false, reinterpret_cast<const BatchItem*>(d_desc + S.items_off),
                          reinterpret_cast<const int64_t*>(d_desc + S.first_off), static_cast<int>(S.items.size()),
                          S.total, S.tmax, S.kmax, C.stream);
    for (int i = 0; i < n_inst; ++i) {
        const Inst& I = inst[i];
        if (!I.batched) continue;
        qhip_state psi = state_of(I, I.psi), phi = state_of(I, I.phi);
        obs_apply_async(obs.get(), i, &psi, &phi);
        inner_async(&psi, &phi, d_scratch, d_energy + i);
    }
    for (const Step& S : bsteps) {
        launch_tile_batch(precision, true, reinterpret_cast<const BatchItem*>(d_desc + S.items_off),
                          reinterpret_cast<const int64_t*>(d_desc + S.first_off), static_cast<int>(S.items.size()),
                          S.total, S.tmax, S.kmax, C.stream);
        launch_ip_reduce(reinterpret_cast<const IpReduce*>(d_desc + S.jobs_off), static_cast<int>(S.jobs.size()),
                         C.stream);
    }
}

void BatchSession::run(double* energies, double* grads, const int64_t* grad_offsets, bool use_graph) {
    auto& C = ctx();
    if (use_graph) {
        if (!graph) {
            hipGraph_t g = nullptr;
            QHIP_CHECK(hipStreamBeginCapture(C.stream, hipStreamCaptureModeThreadLocal));
            try {
                enqueue();
            } catch (...) {
                (void)hipStreamEndCapture(C.stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            QHIP_CHECK(hipStreamEndCapture(C.stream, &g));
            hipGraphExec_t exec = nullptr;
            const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            QHIP_CHECK(e);
            graph = exec;
        }
        QHIP_CHECK(hipGraphLaunch(graph, C.stream));
    } else {
        enqueue();
    }
    std::vector<double2> h_energy(n_inst), h_tot(n_tot + 1);
    QHIP_CHECK(hipMemcpyAsync(h_energy.data(), d_energy, sizeof(double2) * n_inst, hipMemcpyDeviceToHost,
                              C.stream));
    QHIP_CHECK(hipMemcpyAsync(h_tot.data(), d_tot, sizeof(double2) * (n_tot + 1), hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));
    for (int i = 0; i < n_inst; ++i) {
        const Inst& I = inst[i];
        if (!I.batched) continue;
        energies[i] = h_energy[i].x;
        if (want_grad && grads) {
            std::vector<cplx> ip(I.slots.size());
            for (size_t s = 0; s < ip.size(); ++s) ip[s] = cplx(h_tot[I.tot + s].x, h_tot[I.tot + s].y);
            accumulate_gradient(&streams[i], I.slots, ip.data(), n_params[i], grads + grad_offsets[i]);
        }
    }

    // ── instances outside the batched kernels, one by one ──
    for (int i = 0; i < n_inst; ++i) {
        if (inst[i].batched) continue;
        qhip_state st;
        st.n = inst[i].n;
        st.precision = precision;
        if (hipMalloc(&st.d, st.dim() * st.amp_bytes()) != hipSuccess) {
            (void)hipGetLastError();
            fail(QHIP_NOMEM, "batch: state allocation failed");
        }
        try {
            state_reset(&st);
            if (want_grad && grads) {
                run_gradient(cfg, &streams[i], n_params[i], &st, &observables[i], nullptr,
                             grads + grad_offsets[i], &energies[i]);
            } else {
                apply_stream(cfg, &streams[i], &st, nullptr, false);
                energies[i] = transition(&st, &st, &observables[i]).real();
            }
        } catch (...) {
            (void)hipStreamSynchronize(C.stream);
            (void)hipFree(st.d);
            throw;
        }
        QHIP_CHECK(hipStreamSynchronize(C.stream));
        QHIP_CHECK(hipFree(st.d));
    }
}

bool BatchSession::update(const qhip_stream* new_streams) {
    streams = new_streams;
    std::vector<Inst> fresh;
    plan_all(fresh);
    bool same = fresh.size() == inst.size();
    for (size_t i = 0; same && i < fresh.size(); ++i)
        same = fresh[i].batched == inst[i].batched && fresh[i].slots.size() == inst[i].slots.size() &&
               same_structure(fresh[i].fwd, inst[i].fwd) && same_structure(fresh[i].back, inst[i].back);
    if (same) {
        std::vector<cplx> coefs;
        for (size_t i = 0; i < fresh.size(); ++i) {
            if (!fresh[i].batched) continue;
            coefs.insert(coefs.end(), fresh[i].fwd.coefs.begin(), fresh[i].fwd.coefs.end());
            coefs.insert(coefs.end(), fresh[i].back.coefs.begin(), fresh[i].back.coefs.end());
            inst[i].fwd.coefs = std::move(fresh[i].fwd.coefs);
            inst[i].back.coefs = std::move(fresh[i].back.coefs);
            inst[i].slots = std::move(fresh[i].slots);
        }
        QHIP_CHECK(hipStreamSynchronize(ctx().stream));
        upload_coefs(coefs);
        return false;
    }
    // New structure: rebuild everything (the graph, if any, is captured again).
    if (graph) {
        QHIP_CHECK(hipStreamSynchronize(ctx().stream));
        (void)hipGraphExecDestroy(graph);
        graph = nullptr;
    }
    for (DevArena* a : {&tables, &states, &results, &desc, &ip_area}) a->reset();
    obs.reset();
    build();
    return true;
}

BatchSession* batch_create(const qhip_config* cfg, int n_instances, const int32_t* n_qubits,
                           const qhip_stream* streams, const qhip_observable* observables, const int32_t* n_params,
                           int precision, bool gradients) {
    if (precision != QHIP_C128 && precision != QHIP_C64) fail(QHIP_INVALID, "batch: bad precision");
    if (n_instances < 0) fail(QHIP_INVALID, "batch: n_instances must be >= 0");
    auto S = std::make_unique<BatchSession>();
    if (cfg) {
        S->cfg_copy = *cfg;
        S->cfg = &S->cfg_copy;
    }
    S->n_inst = n_instances;
    S->precision = precision;
    S->want_grad = gradients;
    S->amp = precision == QHIP_C128 ? 16 : 8;
    S->n_qubits.assign(n_qubits, n_qubits + n_instances);
    S->n_params.assign(n_params ? n_params : n_qubits, (n_params ? n_params : n_qubits) + n_instances);
    if (!n_params) std::fill(S->n_params.begin(), S->n_params.end(), 0);
    S->streams = streams;
    S->observables = observables;
    if (n_instances > 0) S->build();
    return S.release();
}

void batch_run(BatchSession* S, double* energies, double* grads, const int64_t* grad_offsets, bool use_graph) {
    if (S->n_inst == 0) return;
    if (grads && !S->want_grad) fail(QHIP_INVALID, "batch: the session was created without gradients");
    S->run(energies, grads, grad_offsets, use_graph);
}

bool batch_update(BatchSession* S, const qhip_stream* streams) { return S->n_inst ? S->update(streams) : false; }

void batch_destroy(BatchSession* S) { delete S; }

void batch_stats(const BatchSession* S, qhip_batch_info* out) {
    *out = qhip_batch_info{};
    out->fwd_launches = static_cast<int64_t>(S->fsteps.size());
    out->rev_launches = static_cast<int64_t>(S->bsteps.size());
    for (int i = 0; i < S->n_inst; ++i) {
        const Inst& I = S->inst[static_cast<size_t>(i)];
        if (!I.batched) {
            ++out->instances_single;
            continue;
        }
        ++out->instances_batched;
        const double state = std::ldexp(static_cast<double>(S->amp), I.n);
        out->fwd_bytes += 2.0 * state * static_cast<double>(I.fwd.passes.size());
        out->obs_bytes += (static_cast<double>(obs_arena_groups(S->obs.get(), i)) + 1.0) * state + 2.0 * state;
        if (S->want_grad) out->rev_bytes += 4.0 * state * static_cast<double>(I.back.passes.size());
    }
}

void batch_energy_grad(const qhip_config* cfg, int n_instances, const int32_t* n_qubits,
                       const qhip_stream* streams, const qhip_observable* observables,
                       const int32_t* n_params, int precision, double* energies, double* grads,
                       const int64_t* grad_offsets) {
    std::unique_ptr<BatchSession> S(
        batch_create(cfg, n_instances, n_qubits, streams, observables, n_params, precision, grads != nullptr));
    batch_run(S.get(), energies, grads, grad_offsets, false);
}

}  // namespace qhip
