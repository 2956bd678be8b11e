// kernels_apply.hip — state management, the wide-op kernel and the pass
// launcher.  The tile-pass kernel itself is in tile_kernel.h.
#include "tile_kernel.h"

#include <cstdlib>
#include <cstring>

namespace qhip {

using tile::cmad;
using tile::kThreads;

template <typename A> struct Real;
template <> struct Real<double2> { using T = double; };
template <> struct Real<float2> { using T = float; };

// Dense op wider than five qubits, straight from HBM: one group of 2^k
// amplitudes at a time per workgroup, staged through LDS, one output row per
// thread.  Operand positions are global qubits.
template <typename A>
__global__ __launch_bounds__(kThreads) void wide_dense_kernel(A* __restrict__ s, int n, const DevOp* __restrict__ op,
                                                              const A* __restrict__ U) {
    __shared__ A buf[256];
    const int K = op->k, D = 1 << K;
    const uint64_t groups = 1ull << (n - K);
    for (uint64_t g = blockIdx.x; g < groups; g += gridDim.x) {
        uint64_t base = g;
        for (int b = 0; b < K; ++b) {
            const uint64_t p = static_cast<uint64_t>(op->spos[b]);
            base = ((base >> p) << (p + 1)) | (base & ((1ull << p) - 1ull));
        }
        const int x = threadIdx.x;
        uint64_t ox = 0;
        if (x < D)
            for (int b = 0; b < K; ++b) ox |= static_cast<uint64_t>((x >> b) & 1) << op->pos[b];
        if (x < D) buf[x] = s[base | ox];
        __syncthreads();
        if (x < D) {
            A acc;
            acc.x = 0;
            acc.y = 0;
            for (int j = 0; j < D; ++j) acc = cmad(acc, U[x * D + j], buf[j]);
            s[base | ox] = acc;
        }
        __syncthreads();
    }
}

// ── state utilities ─────────────────────────────────────────────────────────

template <typename A>
__global__ void fill_zero_one(A* s, uint64_t dim) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < dim; i += stride) {
        A v;
        v.x = i == 0 ? 1 : 0;
        v.y = 0;
        s[i] = v;
    }
}

// Per-block FP64 sums of |a|^2 (fixed grid-stride order, then a block tree).
template <typename A>
__global__ __launch_bounds__(kThreads) void norm2_partials(const A* __restrict__ s, uint64_t dim,
                                                           double* __restrict__ out) {
    __shared__ double red[kThreads / 64];
    double acc = 0.0;
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kThreads;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < dim; i += stride) {
        const A a = s[i];
        acc += static_cast<double>(a.x) * a.x + static_cast<double>(a.y) * a.y;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) v += red[w];
        out[blockIdx.x] = v;
    }
}

// Sum of n doubles by one block, fixed order.
__global__ __launch_bounds__(kThreads) void sum_doubles(const double* __restrict__ in, int64_t n,
                                                        double* __restrict__ out) {
    __shared__ double red[kThreads];
    double acc = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kThreads) acc += in[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// Partials of the k-th inner-product op of a pass summed over tiles, in tile
// order, into totals[slot] (one block per op).
__global__ __launch_bounds__(kThreads) void reduce_pass_ip(const double2* __restrict__ in, int64_t n_tiles,
                                                           const int32_t* __restrict__ slots,
                                                           double2* __restrict__ totals) {
    __shared__ double rx[kThreads], ry[kThreads];
    const double2* row = in + static_cast<int64_t>(blockIdx.x) * n_tiles;
    double ax = 0.0, ay = 0.0;
    for (int64_t i = threadIdx.x; i < n_tiles; i += kThreads) { ax += row[i].x; ay += row[i].y; }
    rx[threadIdx.x] = ax;
    ry[threadIdx.x] = ay;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) { rx[threadIdx.x] += rx[threadIdx.x + w]; ry[threadIdx.x] += ry[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[slots[blockIdx.x]] = make_double2(rx[0], ry[0]);
}

// Sums of many inner-product partial rows at once: block b adds
// jobs[b].n_tiles partials in tile order (the tree of reduce_pass_ip) into
// *jobs[b].out.
__global__ __launch_bounds__(kThreads) void reduce_ip_batch(const IpReduce* __restrict__ jobs) {
    __shared__ double rx[kThreads], ry[kThreads];
    const IpReduce J = jobs[blockIdx.x];
    double ax = 0.0, ay = 0.0;
    for (int64_t i = threadIdx.x; i < J.n_tiles; i += kThreads) { ax += J.in[i].x; ay += J.in[i].y; }
    rx[threadIdx.x] = ax;
    ry[threadIdx.x] = ay;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) { rx[threadIdx.x] += rx[threadIdx.x + w]; ry[threadIdx.x] += ry[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) *J.out = make_double2(rx[0], ry[0]);
}

template <typename A>
__global__ void convert_c128_to(const double2* __restrict__ in, A* __restrict__ out, uint64_t n) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        A v;
        v.x = static_cast<typename Real<A>::T>(in[i].x);
        v.y = static_cast<typename Real<A>::T>(in[i].y);
        out[i] = v;
    }
}

template <typename A>
__global__ void convert_to_c128(const A* __restrict__ in, double2* __restrict__ out, uint64_t n) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = make_double2(in[i].x, in[i].y);
}

// ── host launchers ──────────────────────────────────────────────────────────

static int grid_for(uint64_t n, int per_block = kThreads) {
    const uint64_t blocks = (n + per_block - 1) / per_block;
    const uint64_t cap = static_cast<uint64_t>(std::max(1, ctx().n_cu)) * 32;
    return static_cast<int>(std::max<uint64_t>(1, std::min(blocks, cap)));
}

void state_reset(qhip_state* s) {
    auto& C = ctx();
    if (s->precision == QHIP_C128)
        hipLaunchKernelGGL(fill_zero_one<double2>, dim3(grid_for(s->dim())), dim3(kThreads), 0, C.stream,
                           static_cast<double2*>(s->d), s->dim());
    else
        hipLaunchKernelGGL(fill_zero_one<float2>, dim3(grid_for(s->dim())), dim3(kThreads), 0, C.stream,
                           static_cast<float2*>(s->d), s->dim());
    QHIP_CHECK(hipGetLastError());
}

double state_norm2(const qhip_state* s) {
    auto& C = ctx();
    const int blocks = grid_for(s->dim());
    double* part = static_cast<double*>(C.red.ensure(sizeof(double) * (blocks + 1)));
    if (s->precision == QHIP_C128)
        hipLaunchKernelGGL(norm2_partials<double2>, dim3(blocks), dim3(kThreads), 0, C.stream,
                           static_cast<const double2*>(s->d), s->dim(), part);
    else
        hipLaunchKernelGGL(norm2_partials<float2>, dim3(blocks), dim3(kThreads), 0, C.stream,
                           static_cast<const float2*>(s->d), s->dim(), part);
    hipLaunchKernelGGL(sum_doubles, dim3(1), dim3(kThreads), 0, C.stream, part, static_cast<int64_t>(blocks),
                       part + blocks);
    QHIP_CHECK(hipGetLastError());
    double out = 0.0;
    QHIP_CHECK(hipMemcpyAsync(&out, part + blocks, sizeof(double), hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));
    return out;
}

void state_convert_upload(qhip_state* s, const double* host, uint64_t n) {
    auto& C = ctx();
    if (s->precision == QHIP_C128) {
        QHIP_CHECK(hipMemcpyAsync(s->d, host, n * 16, hipMemcpyHostToDevice, C.stream));
    } else {
        // Stage in chunks through the misc buffer.
        const uint64_t chunk = uint64_t{1} << 24;
        double2* stage = static_cast<double2*>(C.misc.ensure(chunk * 16));
        for (uint64_t off = 0; off < n; off += chunk) {
            const uint64_t m = std::min(chunk, n - off);
            QHIP_CHECK(hipMemcpyAsync(stage, host + 2 * off, m * 16, hipMemcpyHostToDevice, C.stream));
            hipLaunchKernelGGL(convert_c128_to<float2>, dim3(grid_for(m)), dim3(kThreads), 0, C.stream, stage,
                               static_cast<float2*>(s->d) + off, m);
            QHIP_CHECK(hipGetLastError());
        }
    }
    QHIP_CHECK(hipStreamSynchronize(C.stream));
}

void state_convert_download(const qhip_state* s, double* host, uint64_t n) {
    auto& C = ctx();
    if (s->precision == QHIP_C128) {
        QHIP_CHECK(hipMemcpyAsync(host, s->d, n * 16, hipMemcpyDeviceToHost, C.stream));
    } else {
        const uint64_t chunk = uint64_t{1} << 24;
        double2* stage = static_cast<double2*>(C.misc.ensure(chunk * 16));
        for (uint64_t off = 0; off < n; off += chunk) {
            const uint64_t m = std::min(chunk, n - off);
            hipLaunchKernelGGL(convert_to_c128<float2>, dim3(grid_for(m)), dim3(kThreads), 0, C.stream,
                               static_cast<const float2*>(s->d) + off, stage, m);
            QHIP_CHECK(hipGetLastError());
            QHIP_CHECK(hipMemcpyAsync(host + 2 * off, stage, m * 16, hipMemcpyDeviceToHost, C.stream));
        }
    }
    QHIP_CHECK(hipStreamSynchronize(C.stream));
}

template <typename A>
static void launch_passes(const Plan& plan, A* psi, A* phi, double2* ip_part, double2* ip_tot,
                          const DevOp* d_ops, const A* d_coefs, const int32_t* d_slots,
                          std::vector<hipEvent_t>* events) {
    auto& C = ctx();
    int slot_base = 0;
    size_t ev = 0;
    for (const PassDesc& P : plan.passes) {
        if (events) QHIP_CHECK(hipEventRecord((*events)[ev++], C.stream));
        if (P.T == 0) {   // wide dense op, straight from HBM
            const uint64_t groups = uint64_t{1} << (P.n - P.kmax);
            const int blocks = static_cast<int>(std::min<uint64_t>(groups, static_cast<uint64_t>(C.n_cu) * 8));
            for (A* st : {psi, phi}) {
                if (!st) continue;
                hipLaunchKernelGGL(wide_dense_kernel<A>, dim3(blocks), dim3(kThreads), 0, C.stream, st, P.n,
                                   d_ops + P.op_begin, d_coefs + plan.ops[P.op_begin].coef);
                QHIP_CHECK(hipGetLastError());
            }
        } else {
            const int64_t n_tiles = int64_t{1} << (P.n - P.T);
            launch_tile_pass(P, psi, phi, ip_part, d_ops, d_coefs, n_tiles, C.stream);
            if (phi) {
                int n_ip = 0;
                for (int o = P.op_begin; o < P.op_end; ++o)
                    n_ip += is_ip_kind(plan.ops[o].kind);
                if (n_ip > 0) {
                    hipLaunchKernelGGL(reduce_pass_ip, dim3(n_ip), dim3(kThreads), 0, C.stream, ip_part, n_tiles,
                                       d_slots + slot_base, ip_tot);
                    QHIP_CHECK(hipGetLastError());
                    slot_base += n_ip;
                }
            }
        }
        if (events) QHIP_CHECK(hipEventRecord((*events)[ev++], C.stream));
    }
}

// Device copies of a plan's op tables: either in the shared scratch buffers
// (one-shot applies) or in a plan-owned allocation (compiled plans).
struct DeviceTables {
    DevOp* ops = nullptr;
    void* coefs = nullptr;
    int32_t* slots = nullptr;
    void* owned = nullptr;
    int max_ip = 0;
};

static DeviceTables upload_tables(const Plan& plan, int precision, bool own) {
    auto& C = ctx();
    DeviceTables t;
    std::vector<int32_t> slots;
    for (const PassDesc& P : plan.passes) {
        int n_ip = 0;
        for (int o = P.op_begin; o < P.op_end; ++o)
            if (is_ip_kind(plan.ops[o].kind)) {
                slots.push_back(static_cast<int32_t>(plan.ops[o].lmask));
                ++n_ip;
            }
        t.max_ip = std::max(t.max_ip, n_ip);
    }
    const size_t align = 256;
    auto up = [&](size_t b) { return (b + align - 1) / align * align; };
    const size_t ops_b = plan.ops.size() * sizeof(DevOp);
    const size_t slot_b = slots.size() * sizeof(int32_t);
    const size_t coef_b = plan.coefs.size() * (precision == QHIP_C128 ? sizeof(double2) : sizeof(float2));
    const size_t total = up(ops_b) + up(slot_b) + up(coef_b) + align;
    char* base;
    if (own) {
        if (hipMalloc(reinterpret_cast<void**>(&base), total) != hipSuccess) {
            (void)hipGetLastError();
            fail(QHIP_NOMEM, "plan: device table allocation failed");
        }
        t.owned = base;
    } else {
        base = static_cast<char*>(C.ops.ensure(total));
    }
    t.ops = reinterpret_cast<DevOp*>(base);
    t.slots = reinterpret_cast<int32_t*>(base + up(ops_b));
    t.coefs = base + up(ops_b) + up(slot_b);
    QHIP_CHECK(hipMemcpyAsync(t.ops, plan.ops.data(), ops_b, hipMemcpyHostToDevice, C.stream));
    if (slot_b) QHIP_CHECK(hipMemcpyAsync(t.slots, slots.data(), slot_b, hipMemcpyHostToDevice, C.stream));
    if (precision == QHIP_C128) {
        QHIP_CHECK(hipMemcpyAsync(t.coefs, plan.coefs.data(), coef_b, hipMemcpyHostToDevice, C.stream));
        QHIP_CHECK(hipStreamSynchronize(C.stream));
    } else {
        std::vector<float2> cf(plan.coefs.size());
        for (size_t i = 0; i < cf.size(); ++i)
            cf[i] = make_float2(static_cast<float>(plan.coefs[i].real()), static_cast<float>(plan.coefs[i].imag()));
        QHIP_CHECK(hipMemcpyAsync(t.coefs, cf.data(), coef_b, hipMemcpyHostToDevice, C.stream));
        QHIP_CHECK(hipStreamSynchronize(C.stream));   // cf must outlive the copy
    }
    return t;
}

static void run_tables(const Plan& plan, const DeviceTables& t, qhip_state* s, qhip_state* phi, void* d_ip_totals,
                       std::vector<double>* pass_ms) {
    auto& C = ctx();
    double2* ip_part = nullptr;
    if (phi && t.max_ip > 0) {
        int T0 = 0;
        for (const PassDesc& P : plan.passes) T0 = P.T > 0 ? P.T : T0;
        const int64_t n_tiles = int64_t{1} << (plan.passes[0].n - T0);
        ip_part = static_cast<double2*>(C.red.ensure(sizeof(double2) * static_cast<size_t>(t.max_ip) * n_tiles));
    }
    std::vector<hipEvent_t> events;
    if (pass_ms) {
        events.resize(2 * plan.passes.size());
        for (auto& e : events) QHIP_CHECK(hipEventCreate(&e));
    }
    if (s->precision == QHIP_C128)
        launch_passes<double2>(plan, static_cast<double2*>(s->d), phi ? static_cast<double2*>(phi->d) : nullptr,
                               ip_part, static_cast<double2*>(d_ip_totals), t.ops,
                               static_cast<const double2*>(t.coefs), t.slots, pass_ms ? &events : nullptr);
    else
        launch_passes<float2>(plan, static_cast<float2*>(s->d), phi ? static_cast<float2*>(phi->d) : nullptr,
                              ip_part, static_cast<double2*>(d_ip_totals), t.ops,
                              static_cast<const float2*>(t.coefs), t.slots, pass_ms ? &events : nullptr);
    QHIP_CHECK(hipStreamSynchronize(C.stream));
    if (pass_ms) {
        pass_ms->assign(plan.passes.size(), 0.0);
        for (size_t i = 0; i < plan.passes.size(); ++i) {
            float ms = 0.0f;
            QHIP_CHECK(hipEventElapsedTime(&ms, events[2 * i], events[2 * i + 1]));
            (*pass_ms)[i] = ms;
        }
        for (auto& e : events) (void)hipEventDestroy(e);
    }
}

void launch_ip_reduce(const IpReduce* d_jobs, int n_jobs, hipStream_t stream) {
    if (n_jobs <= 0) return;
    hipLaunchKernelGGL(reduce_ip_batch, dim3(static_cast<uint32_t>(n_jobs)), dim3(kThreads), 0, stream, d_jobs);
    QHIP_CHECK(hipGetLastError());
}

void launch_tile_batch_c128(bool two, const BatchItem* items, const int64_t* first, int n_items, int64_t total,
                            int tmax, int kmax, hipStream_t stream);
void launch_tile_batch_c64(bool two, const BatchItem* items, const int64_t* first, int n_items, int64_t total,
                           int tmax, int kmax, hipStream_t stream);

void launch_tile_batch(int precision, bool two, const BatchItem* d_items, const int64_t* d_first, int n_items,
                       int64_t total_tiles, int tmax, int kmax, hipStream_t stream) {
    if (precision == QHIP_C128)
        launch_tile_batch_c128(two, d_items, d_first, n_items, total_tiles, tmax, kmax, stream);
    else
        launch_tile_batch_c64(two, d_items, d_first, n_items, total_tiles, tmax, kmax, stream);
}

void launch_plan(const Plan& plan, qhip_state* s, qhip_state* phi, void* d_ip_totals, int n_slots) {
    (void)n_slots;
    if (plan.passes.empty()) return;
    const DeviceTables t = upload_tables(plan, s->precision, false);
    run_tables(plan, t, s, phi, d_ip_totals, nullptr);
}

CompiledPlan* compile_plan(Plan&& plan, int n, int precision) {
    auto* cp = new CompiledPlan();
    cp->plan = std::move(plan);
    cp->n = n;
    cp->precision = precision;
    if (!cp->plan.passes.empty()) {
        try {
            const DeviceTables t = upload_tables(cp->plan, precision, true);
            cp->d_tables = t.owned;
            cp->d_ops = t.ops;
            cp->d_coefs = t.coefs;
            cp->d_slots = t.slots;
            cp->max_ip = t.max_ip;
        } catch (...) {
            delete cp;
            throw;
        }
    }
    return cp;
}

void run_compiled(const CompiledPlan& cp, qhip_state* s, std::vector<double>* pass_ms) {
    if (cp.plan.passes.empty()) {
        if (pass_ms) pass_ms->clear();
        return;
    }
    DeviceTables t;
    t.ops = static_cast<DevOp*>(cp.d_ops);
    t.coefs = cp.d_coefs;
    t.slots = static_cast<int32_t*>(cp.d_slots);
    t.max_ip = cp.max_ip;
    run_tables(cp.plan, t, s, nullptr, nullptr, pass_ms);
}

static void drop_graph(CompiledPlan& cp) {
    if (cp.graph) {
        (void)hipStreamSynchronize(ctx().stream);
        (void)hipGraphExecDestroy(cp.graph);
        cp.graph = nullptr;
        cp.graph_state = nullptr;
    }
}

// Replays the plan's pass launches as one hipGraph: captured on the first
// call for a state (stream capture of exactly the launches run_compiled
// makes: tile / wide kernels only, no copies, fills or events), relaunched
// afterwards.  A different state re-captures.
void run_compiled_graph(CompiledPlan& cp, qhip_state* s) {
    if (cp.plan.passes.empty()) return;
    auto& C = ctx();
    if (cp.graph && cp.graph_state != s->d) drop_graph(cp);
    if (!cp.graph) {
        hipGraph_t g = nullptr;
        QHIP_CHECK(hipStreamBeginCapture(C.stream, hipStreamCaptureModeThreadLocal));
        try {
            if (s->precision == QHIP_C128)
                launch_passes<double2>(cp.plan, static_cast<double2*>(s->d), nullptr, nullptr, nullptr,
                                       static_cast<const DevOp*>(cp.d_ops), static_cast<const double2*>(cp.d_coefs),
                                       static_cast<const int32_t*>(cp.d_slots), nullptr);
            else
                launch_passes<float2>(cp.plan, static_cast<float2*>(s->d), nullptr, nullptr, nullptr,
                                      static_cast<const DevOp*>(cp.d_ops), static_cast<const float2*>(cp.d_coefs),
                                      static_cast<const int32_t*>(cp.d_slots), nullptr);
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
        cp.graph = exec;
        cp.graph_state = s->d;
    }
    QHIP_CHECK(hipGraphLaunch(cp.graph, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));
}

static bool same_structure(const Plan& a, const Plan& b) {
    if (a.passes.size() != b.passes.size() || a.ops.size() != b.ops.size() || a.coefs.size() != b.coefs.size())
        return false;
    if (std::memcmp(a.passes.data(), b.passes.data(), a.passes.size() * sizeof(PassDesc)) != 0) return false;
    return std::memcmp(a.ops.data(), b.ops.data(), a.ops.size() * sizeof(DevOp)) == 0;
}

bool update_compiled(CompiledPlan& cp, Plan&& plan, int64_t fused_ops) {
    auto& C = ctx();
    if (cp.d_tables && same_structure(cp.plan, plan)) {
        if (cp.precision == QHIP_C128) {
            QHIP_CHECK(hipMemcpyAsync(cp.d_coefs, plan.coefs.data(), plan.coefs.size() * sizeof(double2),
                                      hipMemcpyHostToDevice, C.stream));
            QHIP_CHECK(hipStreamSynchronize(C.stream));
        } else {
            std::vector<float2> cf(plan.coefs.size());
            for (size_t i = 0; i < cf.size(); ++i)
                cf[i] = make_float2(static_cast<float>(plan.coefs[i].real()),
                                    static_cast<float>(plan.coefs[i].imag()));
            QHIP_CHECK(hipMemcpyAsync(cp.d_coefs, cf.data(), cf.size() * sizeof(float2), hipMemcpyHostToDevice,
                                      C.stream));
            QHIP_CHECK(hipStreamSynchronize(C.stream));
        }
        cp.plan.coefs = std::move(plan.coefs);
        cp.plan.flops = plan.flops;
        cp.fused_ops = fused_ops;
        return false;
    }
    drop_graph(cp);
    if (cp.d_tables) {
        QHIP_CHECK(hipStreamSynchronize(C.stream));
        QHIP_CHECK(hipFree(cp.d_tables));
        cp.d_tables = cp.d_ops = cp.d_coefs = cp.d_slots = nullptr;
    }
    cp.plan = std::move(plan);
    cp.fused_ops = fused_ops;
    if (!cp.plan.passes.empty()) {
        const DeviceTables t = upload_tables(cp.plan, cp.precision, true);
        cp.d_tables = t.owned;
        cp.d_ops = t.ops;
        cp.d_coefs = t.coefs;
        cp.d_slots = t.slots;
        cp.max_ip = t.max_ip;
    }
    return true;
}

void free_compiled(CompiledPlan* cp) {
    if (!cp) return;
    drop_graph(*cp);
    if (cp->d_tables) {
        (void)hipStreamSynchronize(ctx().stream);
        (void)hipFree(cp->d_tables);
    }
    delete cp;
}

}  // namespace qhip
