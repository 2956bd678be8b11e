// kernels_pauli.hip — Pauli-sum expectation <bra|H|ket>, H|psi> and <a|b>.
//
// The observable is grouped on the host by X/Y flip mask x (terms sharing x
// read the same partner amplitude), with i^{n_Y} folded into each term's
// coefficient, as build_groups does (src/qarpx/src/simulator/
// pauli_expectation.cpp:34-71):
//     <bra|H|ket> = sum_g sum_i conj(bra[i ^ x_g]) ket[i] w_g(i),
//     w_g(i)      = sum_{t in g} c_t (-1)^popcount(i & z_t).
// Device layout: each workgroup owns a contiguous chunk of 2^12 ket indices
// (16 per thread, kept in registers) and sweeps every group over it; for a
// group the partner chunk is the chunk at (chunk ^ x_high), read as one
// contiguous block with the low bits of x permuting lanes.  Sums are FP64
// (complex64 states are widened on load): per-thread sums, a lane-shuffle
// wave reduction, a fixed-order block reduction, then one block sums the
// per-chunk partials in index order — the result does not depend on timing.
#include "common.h"

#include <algorithm>
#include <cstring>
#include <memory>
#include <type_traits>

#include <unordered_map>

namespace qhip {

constexpr int kT = 256;
constexpr int kPer = 16;                 // ket amplitudes per thread
constexpr uint64_t kChunk = kT * kPer;   // per workgroup

struct GroupDev {
    uint64_t x;
    int32_t t_begin, t_end;
};

struct Grouped {
    std::vector<GroupDev> groups;
    std::vector<uint64_t> z;
    std::vector<double2> c;
};

static Grouped group_terms(const qhip_observable* obs, int n, const char* who) {
    Grouped out;
    std::unordered_map<uint64_t, size_t> index;
    std::vector<std::vector<std::pair<uint64_t, cplx>>> per;
    for (int64_t t = 0; t < obs->n_terms; ++t) {
        const qhip_pauli_term& term = obs->terms[t];
        uint64_t x = 0, z = 0, seen = 0;
        int ny = 0;
        for (int l = 0; l < term.n_letters; ++l) {
            const uint32_t q = obs->qubits[term.offset + l];
            const char ch = obs->letters[term.offset + l];
            if (q >= static_cast<uint32_t>(n))
                fail(QHIP_INVALID, std::string(who) + ": observable qubit " + std::to_string(q) +
                                       " out of range for " + std::to_string(n) + " qubits");
            const uint64_t bit = uint64_t{1} << q;
            if (seen & bit)
                fail(QHIP_INVALID, std::string(who) + ": qubit " + std::to_string(q) +
                                       " repeated within one Pauli string");
            seen |= bit;
            switch (ch) {
                case 'X': x |= bit; break;
                case 'Z': z |= bit; break;
                case 'Y': x |= bit; z |= bit; ++ny; break;
                case 'I': break;
                default: fail(QHIP_INVALID, std::string(who) + ": invalid Pauli '" + std::string(1, ch) + "'");
            }
        }
        static const cplx ipow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
        const cplx c = cplx(term.coeff_re, term.coeff_im) * ipow[ny & 3];
        auto [it, fresh] = index.try_emplace(x, per.size());
        if (fresh) per.emplace_back();
        per[it->second].emplace_back(z, c);
    }
    std::vector<uint64_t> xs(per.size());
    for (const auto& [x, i] : index) xs[i] = x;
    for (size_t g = 0; g < per.size(); ++g) {
        GroupDev d;
        d.x = xs[g];
        d.t_begin = static_cast<int32_t>(out.z.size());
        for (const auto& [z, c] : per[g]) {
            out.z.push_back(z);
            out.c.push_back(make_double2(c.real(), c.imag()));
        }
        d.t_end = static_cast<int32_t>(out.z.size());
        out.groups.push_back(d);
    }
    return out;
}

__device__ __forceinline__ double2 load_amp(const double2* p, uint64_t i) { return p[i]; }
__device__ __forceinline__ double2 load_amp(const float2* p, uint64_t i) {
    const float2 v = p[i];
    return make_double2(v.x, v.y);
}

__device__ __forceinline__ double2 block_sum(double2 v, double2* red) {
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off);
        v.y += __shfl_xor(v.y, off);
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double2 r = make_double2(0.0, 0.0);
    if (threadIdx.x == 0)
        for (int w = 0; w < kT / 64; ++w) { r.x += red[w].x; r.y += red[w].y; }
    return r;
}

// Weight of group g at index i.
__device__ __forceinline__ double2 weight(uint64_t i, const GroupDev& G, const uint64_t* __restrict__ z,
                                          const double2* __restrict__ c) {
    double2 w = make_double2(0.0, 0.0);
    for (int t = G.t_begin; t < G.t_end; ++t) {
        const double s = (__popcll(i & z[t]) & 1) ? -1.0 : 1.0;
        w.x += s * c[t].x;
        w.y += s * c[t].y;
    }
    return w;
}

// HERM (bra == ket, every coefficient real, so every group operator H_x is
// Hermitian): with h the top bit of x != 0, the terms at i and i ^ x are
// complex conjugates, so the real part is twice the sum over i with bit h
// clear — half the partner reads.  Only the real part of the result is
// meaningful then (the exact imaginary part is zero).
template <typename A, bool HERM>
__global__ __launch_bounds__(kT) void transition_kernel(const A* __restrict__ bra, const A* __restrict__ ket,
                                                        uint64_t dim, const GroupDev* __restrict__ groups,
                                                        int n_groups, const uint64_t* __restrict__ z,
                                                        const double2* __restrict__ c,
                                                        double2* __restrict__ partials, uint64_t chunk0) {
    __shared__ double2 red[kT / 64];
    const uint64_t base = (chunk0 + blockIdx.x) * kChunk;
    double2 k[kPer];
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const uint64_t i = base + threadIdx.x + static_cast<uint64_t>(r) * kT;
        k[r] = i < dim ? load_amp(ket, i) : make_double2(0.0, 0.0);
    }
    for (int g = 0; g < n_groups; ++g) {
        const GroupDev G = groups[g];
        uint64_t hb = 0;   // HERM: top bit of x (indices with it set are skipped)
        double sc = 1.0;
        if (HERM && G.x) {
            hb = uint64_t{1} << (63 - __clzll(G.x));
            sc = 2.0;
            if (base & hb) continue;   // a bit above the chunk: uniform over the workgroup
        }
#pragma unroll
        for (int r = 0; r < kPer; ++r) {
            const uint64_t i = base + threadIdx.x + static_cast<uint64_t>(r) * kT;
            if (i >= dim || (HERM && (i & hb))) continue;
            const double2 b = load_amp(bra, i ^ G.x);
            double2 w = weight(i, G, z, c);
            if (HERM) {
                w.x *= sc;
                w.y *= sc;
            }
            // conj(b) * w * ket[i]
            const double vx = b.x * w.x + b.y * w.y, vy = b.x * w.y - b.y * w.x;
            acc.x += vx * k[r].x - vy * k[r].y;
            acc.y += vx * k[r].y + vy * k[r].x;
        }
    }
    const double2 tot = block_sum(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

template <typename A>
__global__ __launch_bounds__(kT) void observable_kernel(const A* __restrict__ in, A* __restrict__ out,
                                                        uint64_t dim, const GroupDev* __restrict__ groups,
                                                        int n_groups, const uint64_t* __restrict__ z,
                                                        const double2* __restrict__ c) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kT;
    for (uint64_t j = static_cast<uint64_t>(blockIdx.x) * kT + threadIdx.x; j < dim; j += stride) {
        double2 acc = make_double2(0.0, 0.0);
        for (int g = 0; g < n_groups; ++g) {
            const GroupDev G = groups[g];
            const uint64_t src = j ^ G.x;
            const double2 v = load_amp(in, src);
            const double2 w = weight(src, G, z, c);
            acc.x += w.x * v.x - w.y * v.y;
            acc.y += w.x * v.y + w.y * v.x;
        }
        A o;
        o.x = acc.x;
        o.y = acc.y;
        out[j] = o;
    }
}

template <typename A>
__global__ __launch_bounds__(kT) void inner_kernel(const A* __restrict__ a, const A* __restrict__ b, uint64_t dim,
                                                   double2* __restrict__ partials) {
    __shared__ double2 red[kT / 64];
    double2 acc = make_double2(0.0, 0.0);
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kT;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kT + threadIdx.x; i < dim; i += stride) {
        const double2 x = load_amp(a, i), y = load_amp(b, i);
        acc.x += x.x * y.x + x.y * y.y;
        acc.y += x.x * y.y - x.y * y.x;
    }
    const double2 tot = block_sum(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kT) void sum_partials(const double2* __restrict__ in, int64_t n,
                                                   double2* __restrict__ out) {
    __shared__ double rx[kT], ry[kT];
    double ax = 0.0, ay = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kT) { ax += in[i].x; ay += in[i].y; }
    rx[threadIdx.x] = ax;
    ry[threadIdx.x] = ay;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) { rx[threadIdx.x] += rx[threadIdx.x + w]; ry[threadIdx.x] += ry[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = make_double2(rx[0], ry[0]);
}

struct DeviceGroups {
    const GroupDev* g;
    const uint64_t* z;
    const double2* c;
    int n_groups;
};

static DeviceGroups upload_groups(const Grouped& gr) {
    auto& C = ctx();
    const size_t gb = gr.groups.size() * sizeof(GroupDev);
    const size_t zb = gr.z.size() * sizeof(uint64_t);
    const size_t cb = gr.c.size() * sizeof(double2);
    char* base = static_cast<char*>(C.ops.ensure(gb + zb + cb + 64));
    QHIP_CHECK(hipMemcpyAsync(base, gr.groups.data(), gb, hipMemcpyHostToDevice, C.stream));
    QHIP_CHECK(hipMemcpyAsync(base + gb, gr.z.data(), zb, hipMemcpyHostToDevice, C.stream));
    QHIP_CHECK(hipMemcpyAsync(base + gb + zb, gr.c.data(), cb, hipMemcpyHostToDevice, C.stream));
    return {reinterpret_cast<const GroupDev*>(base), reinterpret_cast<const uint64_t*>(base + gb),
            reinterpret_cast<const double2*>(base + gb + zb), static_cast<int>(gr.groups.size())};
}

static cplx finish(double2* part, int64_t n_part) {
    auto& C = ctx();
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(kT), 0, C.stream, part, n_part, part + n_part);
    QHIP_CHECK(hipGetLastError());
    double2 out;
    QHIP_CHECK(hipMemcpyAsync(&out, part + n_part, sizeof(double2), hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));
    return cplx(out.x, out.y);
}

cplx transition(const qhip_state* bra, const qhip_state* ket, const qhip_observable* obs) {
    const int n = ket->n;
    const Grouped gr = group_terms(obs, n, "pauli_transition");
    if (gr.groups.empty()) return cplx(0.0, 0.0);
    auto& C = ctx();
    const DeviceGroups dg = upload_groups(gr);
    const uint64_t dim = ket->dim();
    const int64_t blocks = static_cast<int64_t>((dim + kChunk - 1) / kChunk);
    double2* part = static_cast<double2*>(C.red.ensure(sizeof(double2) * (blocks + 1)));
    bool herm = bra->d == ket->d;
    for (int64_t t = 0; herm && t < obs->n_terms; ++t) herm = obs->terms[t].coeff_im == 0.0;
    auto go = [&](auto herm_tag) {
        constexpr bool H = decltype(herm_tag)::value;
        if (ket->precision == QHIP_C128)
            hipLaunchKernelGGL((transition_kernel<double2, H>), dim3(static_cast<uint32_t>(blocks)), dim3(kT), 0,
                               C.stream, static_cast<const double2*>(bra->d), static_cast<const double2*>(ket->d), dim,
                               dg.g, dg.n_groups, dg.z, dg.c, part, uint64_t{0});
        else
            hipLaunchKernelGGL((transition_kernel<float2, H>), dim3(static_cast<uint32_t>(blocks)), dim3(kT), 0,
                               C.stream, static_cast<const float2*>(bra->d), static_cast<const float2*>(ket->d), dim,
                               dg.g, dg.n_groups, dg.z, dg.c, part, uint64_t{0});
    };
    if (herm) go(std::true_type{});
    else go(std::false_type{});
    QHIP_CHECK(hipGetLastError());
    const cplx r = finish(part, blocks);
    return herm ? cplx(r.real(), 0.0) : r;
}

// ── <bra|H|ket> with no device allocation ───────────────────────────────────
// For callers that cannot allocate (graph capture, serving loops): the group
// table and the reduction partials live in a fixed module-scope device array.
// The ket runs in slices of kSliceChunks chunks; after each slice
// accumulate_partials adds its chunk partials into 256 running sums, thread t
// taking chunks t, t + 256, ... in increasing order — the additions
// sum_partials performs over all chunks at once — so the result is bit-equal
// to transition().  Every call in the process shares the array (the library
// lock serialises them); callers that bypass that lock on one GPU race.

constexpr int64_t kSliceChunks = 16384;              // multiple of kT
constexpr int kNoAllocMaxGroups = 4096;
constexpr int kNoAllocMaxTerms = 16384;
constexpr size_t kNaGroupsOff = 0;
constexpr size_t kNaZOff = kNaGroupsOff + sizeof(GroupDev) * kNoAllocMaxGroups;
constexpr size_t kNaCOff = kNaZOff + sizeof(uint64_t) * kNoAllocMaxTerms;
constexpr size_t kNaPartOff = kNaCOff + sizeof(double2) * kNoAllocMaxTerms;
constexpr size_t kNaAccOff = kNaPartOff + sizeof(double2) * kSliceChunks;
constexpr size_t kNaOutOff = kNaAccOff + sizeof(double2) * kT;
constexpr size_t kNoAllocBytes = kNaOutOff + 256;    // 725,248 bytes
static_assert(kSliceChunks % kT == 0, "slices must keep chunk residues mod kT");
static_assert(kNaPartOff % 256 == 0 && kNaAccOff % 256 == 0 && kNaOutOff % 256 == 0, "aligned sections");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// The array's device address per device, from hipGetSymbolAddress on the
// first call with that device current (the name itself is a host-side stand-in).
constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices] = {};

static unsigned char* noalloc_scratch() {
    const int dev = ctx().device;
    if (dev < 0 || dev >= kMaxDevices) fail(QHIP_DEVICE, "qhip_transition_noalloc: device index out of range");
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            fail(QHIP_DEVICE, std::string("qhip_transition_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                                  hipGetErrorString(e));
        }
        g_scratch_ptr[dev] = p;
    }
    return static_cast<unsigned char*>(g_scratch_ptr[dev]);
}

__global__ __launch_bounds__(kT) void accumulate_partials(const double2* __restrict__ in, int64_t n,
                                                          double2* __restrict__ acc, int first) {
    double ax = 0.0, ay = 0.0;
    if (!first) {
        ax = acc[threadIdx.x].x;
        ay = acc[threadIdx.x].y;
    }
    for (int64_t i = threadIdx.x; i < n; i += kT) { ax += in[i].x; ay += in[i].y; }
    acc[threadIdx.x] = make_double2(ax, ay);
}

cplx transition_noalloc(const qhip_state* bra, const qhip_state* ket, const qhip_observable* obs) {
    const int n = ket->n;
    const Grouped gr = group_terms(obs, n, "pauli_transition");
    if (gr.groups.empty()) return cplx(0.0, 0.0);
    if (gr.groups.size() > static_cast<size_t>(kNoAllocMaxGroups) || gr.z.size() > static_cast<size_t>(kNoAllocMaxTerms))
        fail(QHIP_INVALID, "qhip_transition_noalloc: the observable has " + std::to_string(gr.groups.size()) +
                               " flip-mask groups and " + std::to_string(gr.z.size()) +
                               " terms; the fixed scratch holds at most " + std::to_string(kNoAllocMaxGroups) +
                               " groups and " + std::to_string(kNoAllocMaxTerms) + " terms");
    auto& C = ctx();
    unsigned char* base = noalloc_scratch();
    auto* d_groups = reinterpret_cast<GroupDev*>(base + kNaGroupsOff);
    auto* d_z = reinterpret_cast<uint64_t*>(base + kNaZOff);
    auto* d_c = reinterpret_cast<double2*>(base + kNaCOff);
    auto* d_part = reinterpret_cast<double2*>(base + kNaPartOff);
    auto* d_acc = reinterpret_cast<double2*>(base + kNaAccOff);
    auto* d_out = reinterpret_cast<double2*>(base + kNaOutOff);
    QHIP_CHECK(hipMemcpyAsync(d_groups, gr.groups.data(), gr.groups.size() * sizeof(GroupDev),
                              hipMemcpyHostToDevice, C.stream));
    QHIP_CHECK(hipMemcpyAsync(d_z, gr.z.data(), gr.z.size() * sizeof(uint64_t), hipMemcpyHostToDevice, C.stream));
    QHIP_CHECK(hipMemcpyAsync(d_c, gr.c.data(), gr.c.size() * sizeof(double2), hipMemcpyHostToDevice, C.stream));
    const uint64_t dim = ket->dim();
    const int64_t chunks = static_cast<int64_t>((dim + kChunk - 1) / kChunk);
    bool herm = bra->d == ket->d;
    for (int64_t t = 0; herm && t < obs->n_terms; ++t) herm = obs->terms[t].coeff_im == 0.0;
    const int ng = static_cast<int>(gr.groups.size());
    for (int64_t c0 = 0; c0 < chunks; c0 += kSliceChunks) {
        const int64_t nb = std::min<int64_t>(kSliceChunks, chunks - c0);
        const dim3 grid(static_cast<uint32_t>(nb));
        const uint64_t off = static_cast<uint64_t>(c0);
        if (ket->precision == QHIP_C128) {
            const auto* b = static_cast<const double2*>(bra->d);
            const auto* k = static_cast<const double2*>(ket->d);
            if (herm)
                hipLaunchKernelGGL((transition_kernel<double2, true>), grid, dim3(kT), 0, C.stream, b, k, dim, d_groups,
                                   ng, d_z, d_c, d_part, off);
            else
                hipLaunchKernelGGL((transition_kernel<double2, false>), grid, dim3(kT), 0, C.stream, b, k, dim,
                                   d_groups, ng, d_z, d_c, d_part, off);
        } else {
            const auto* b = static_cast<const float2*>(bra->d);
            const auto* k = static_cast<const float2*>(ket->d);
            if (herm)
                hipLaunchKernelGGL((transition_kernel<float2, true>), grid, dim3(kT), 0, C.stream, b, k, dim, d_groups,
                                   ng, d_z, d_c, d_part, off);
            else
                hipLaunchKernelGGL((transition_kernel<float2, false>), grid, dim3(kT), 0, C.stream, b, k, dim,
                                   d_groups, ng, d_z, d_c, d_part, off);
        }
        hipLaunchKernelGGL(accumulate_partials, dim3(1), dim3(kT), 0, C.stream, d_part, nb, d_acc,
                           c0 == 0 ? 1 : 0);
        QHIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(kT), 0, C.stream, d_acc, static_cast<int64_t>(kT), d_out);
    QHIP_CHECK(hipGetLastError());
    double2 r;
    QHIP_CHECK(hipMemcpyAsync(&r, d_out, sizeof(double2), hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));
    return herm ? cplx(r.x, 0.0) : cplx(r.x, r.y);
}

void apply_observable(const qhip_state* in, const qhip_observable* obs, qhip_state* out) {
    const int n = in->n;
    const Grouped gr = group_terms(obs, n, "run_gradient");
    auto& C = ctx();
    if (gr.groups.empty()) {
        QHIP_CHECK(hipMemsetAsync(out->d, 0, out->dim() * out->amp_bytes(), C.stream));
        QHIP_CHECK(hipStreamSynchronize(C.stream));
        return;
    }
    const DeviceGroups dg = upload_groups(gr);
    const uint64_t dim = in->dim();
    const int blocks = static_cast<int>(std::min<uint64_t>((dim + kT - 1) / kT,
                                                           static_cast<uint64_t>(std::max(1, C.n_cu)) * 64));
    if (in->precision == QHIP_C128)
        hipLaunchKernelGGL(observable_kernel<double2>, dim3(blocks), dim3(kT), 0, C.stream,
                           static_cast<const double2*>(in->d), static_cast<double2*>(out->d), dim, dg.g,
                           dg.n_groups, dg.z, dg.c);
    else
        hipLaunchKernelGGL(observable_kernel<float2>, dim3(blocks), dim3(kT), 0, C.stream,
                           static_cast<const float2*>(in->d), static_cast<float2*>(out->d), dim, dg.g,
                           dg.n_groups, dg.z, dg.c);
    QHIP_CHECK(hipGetLastError());
    QHIP_CHECK(hipStreamSynchronize(C.stream));
}

// ── many observables at once (batched instances) ───────────────────────────

struct ObsArena {
    void* d = nullptr;
    std::vector<DeviceGroups> per;
    ~ObsArena() {
        if (d) {
            (void)hipStreamSynchronize(ctx().stream);
            (void)hipFree(d);
        }
    }
};

ObsArena* obs_arena_create(int n_obs, const int32_t* n_qubits, const qhip_observable* obs) {
    std::vector<Grouped> gs;
    gs.reserve(n_obs);
    size_t gb = 0, zb = 0, cb = 0;
    for (int i = 0; i < n_obs; ++i) {
        gs.push_back(group_terms(&obs[i], n_qubits[i], "run_gradient"));
        gb += gs.back().groups.size() * sizeof(GroupDev);
        zb += gs.back().z.size() * sizeof(uint64_t);
        cb += gs.back().c.size() * sizeof(double2);
    }
    auto a = std::make_unique<ObsArena>();
    std::vector<char> host(gb + zb + cb + 64);
    if (hipMalloc(&a->d, host.size()) != hipSuccess) {
        (void)hipGetLastError();
        a->d = nullptr;
        fail(QHIP_NOMEM, "batch: observable table allocation failed");
    }
    char* dev = static_cast<char*>(a->d);
    size_t og = 0, oz = gb, oc = gb + zb;
    for (const Grouped& g : gs) {
        DeviceGroups dg{reinterpret_cast<const GroupDev*>(dev + og), reinterpret_cast<const uint64_t*>(dev + oz),
                        reinterpret_cast<const double2*>(dev + oc), static_cast<int>(g.groups.size())};
        std::memcpy(host.data() + og, g.groups.data(), g.groups.size() * sizeof(GroupDev));
        std::memcpy(host.data() + oz, g.z.data(), g.z.size() * sizeof(uint64_t));
        std::memcpy(host.data() + oc, g.c.data(), g.c.size() * sizeof(double2));
        og += g.groups.size() * sizeof(GroupDev);
        oz += g.z.size() * sizeof(uint64_t);
        oc += g.c.size() * sizeof(double2);
        a->per.push_back(dg);
    }
    QHIP_CHECK(hipMemcpy(a->d, host.data(), host.size(), hipMemcpyHostToDevice));
    return a.release();
}

void obs_arena_destroy(ObsArena* a) { delete a; }

// out = H_i in, stream-ordered (no host synchronisation).
void obs_apply_async(const ObsArena* a, int i, const qhip_state* in, qhip_state* out) {
    auto& C = ctx();
    const DeviceGroups& dg = a->per[i];
    const uint64_t dim = in->dim();
    const int blocks = static_cast<int>(std::min<uint64_t>((dim + kT - 1) / kT,
                                                           static_cast<uint64_t>(std::max(1, C.n_cu)) * 64));
    if (in->precision == QHIP_C128)
        hipLaunchKernelGGL(observable_kernel<double2>, dim3(blocks), dim3(kT), 0, C.stream,
                           static_cast<const double2*>(in->d), static_cast<double2*>(out->d), dim, dg.g,
                           dg.n_groups, dg.z, dg.c);
    else
        hipLaunchKernelGGL(observable_kernel<float2>, dim3(blocks), dim3(kT), 0, C.stream,
                           static_cast<const float2*>(in->d), static_cast<float2*>(out->d), dim, dg.g,
                           dg.n_groups, dg.z, dg.c);
    QHIP_CHECK(hipGetLastError());
}

// *d_out = <a|b> in FP64, stream-ordered; scratch holds inner_blocks(a) + 1 entries.
int inner_blocks(const qhip_state* a) {
    return static_cast<int>(std::min<uint64_t>((a->dim() + kT - 1) / kT,
                                               static_cast<uint64_t>(std::max(1, ctx().n_cu)) * 32));
}

void inner_async(const qhip_state* a, const qhip_state* b, double2* scratch, double2* d_out) {
    auto& C = ctx();
    const uint64_t dim = a->dim();
    const int blocks = inner_blocks(a);
    if (a->precision == QHIP_C128)
        hipLaunchKernelGGL(inner_kernel<double2>, dim3(blocks), dim3(kT), 0, C.stream,
                           static_cast<const double2*>(a->d), static_cast<const double2*>(b->d), dim, scratch);
    else
        hipLaunchKernelGGL(inner_kernel<float2>, dim3(blocks), dim3(kT), 0, C.stream,
                           static_cast<const float2*>(a->d), static_cast<const float2*>(b->d), dim, scratch);
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(kT), 0, C.stream, scratch, static_cast<int64_t>(blocks), d_out);
    QHIP_CHECK(hipGetLastError());
}

cplx inner(const qhip_state* a, const qhip_state* b) {
    auto& C = ctx();
    const uint64_t dim = a->dim();
    const int blocks = static_cast<int>(std::min<uint64_t>((dim + kT - 1) / kT,
                                                           static_cast<uint64_t>(std::max(1, C.n_cu)) * 32));
    double2* part = static_cast<double2*>(C.red.ensure(sizeof(double2) * (blocks + 1)));
    if (a->precision == QHIP_C128)
        hipLaunchKernelGGL(inner_kernel<double2>, dim3(blocks), dim3(kT), 0, C.stream,
                           static_cast<const double2*>(a->d), static_cast<const double2*>(b->d), dim, part);
    else
        hipLaunchKernelGGL(inner_kernel<float2>, dim3(blocks), dim3(kT), 0, C.stream,
                           static_cast<const float2*>(a->d), static_cast<const float2*>(b->d), dim, part);
    QHIP_CHECK(hipGetLastError());
    return finish(part, blocks);
}

}  // namespace qhip
