// Shot-batched trajectories (the per-shot loop of QarpSimulator::run,
// src/qarpx/src/simulator/qarp_simulator.cpp:1038-1211, and
// apply_command_trajectory, :600-689).  B shots live in one (n + b)-qubit
// state, shot j in amplitudes [j * 2^n, (j + 1) * 2^n): unitary runs act on
// every shot at once through the ordinary apply, and these kernels do the
// per-shot parts — broadcasting the shared prefix state, restoring shots a
// conditional run must not touch, the probability of |0> on a qubit per
// shot (M0_prob), the per-shot projection + renormalisation of Measure and
// Reset (P0/P1_gate, normalize, X), and each shot's final draw.
#include "common.h"

#include <algorithm>
#include <bit>
#include <string>
#include <vector>

namespace qhip {
namespace {

constexpr int kBlk = 256;
constexpr uint64_t kChunk = uint64_t{1} << 14;   // amplitude pairs per partial of prob0
// Shots of at least this many amplitudes take the chunk-scan walk for their
// final draw; smaller ones a thread each.
constexpr uint64_t kWalkBlock = uint64_t{1} << 20;

int blocks_for(uint64_t n, int cap_per_cu = 32) {
    return static_cast<int>(std::max<uint64_t>(
        1, std::min<uint64_t>((n + kBlk - 1) / kBlk, static_cast<uint64_t>(ctx().n_cu) * cap_per_cu)));
}

template <typename A>
__global__ __launch_bounds__(kBlk) void broadcast_kernel(A* __restrict__ dst, const A* __restrict__ src, uint64_t total,
                                                         uint64_t low_mask) {
    const uint64_t step = static_cast<uint64_t>(gridDim.x) * kBlk;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlk + threadIdx.x; i < total; i += step)
        dst[i] = src[i & low_mask];
}

template <typename A>
__global__ __launch_bounds__(kBlk) void restore_kernel(A* __restrict__ dst, const A* __restrict__ saved, uint64_t total,
                                                       int log_stride, const uint8_t* __restrict__ active) {
    const uint64_t step = static_cast<uint64_t>(gridDim.x) * kBlk;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlk + threadIdx.x; i < total; i += step)
        if (!active[i >> log_stride]) dst[i] = saved[i];
}

// partial[j * chunks + c] = sum of |a|^2 over pairs c*kChunk .. of shot j with
// the qubit's bit clear (fixed order inside a workgroup: grid-stride sums,
// then a tree over the workgroup).
template <typename A>
__global__ __launch_bounds__(kBlk) void prob0_kernel(const A* __restrict__ s, int q, int log_stride, uint64_t pairs,
                                                     uint64_t chunks, double* __restrict__ partial) {
    __shared__ double red[kBlk];
    const uint64_t j = blockIdx.x / chunks, c = blockIdx.x % chunks;
    const uint64_t lo = c * kChunk, hi = std::min(pairs, lo + kChunk);
    const uint64_t low = (uint64_t{1} << q) - 1;
    double acc = 0.0;
    for (uint64_t p = lo + threadIdx.x; p < hi; p += kBlk) {
        const uint64_t i = (j << log_stride) | ((p & ~low) << 1) | (p & low);
        const double x = s[i].x, y = s[i].y;
        acc += x * x + y * y;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlk / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Per shot: action -1 leaves it alone; 0 keeps bit 0 (scaled), 1 keeps bit 1
// (scaled), 2 moves the bit-1 amplitude to bit 0 (Reset after outcome 1).
template <typename A>
__global__ __launch_bounds__(kBlk) void project_kernel(A* __restrict__ s, int q, int log_stride, uint64_t pairs_total,
                                                       const int8_t* __restrict__ action,
                                                       const double* __restrict__ scale) {
    const uint64_t step = static_cast<uint64_t>(gridDim.x) * kBlk;
    const uint64_t low = (uint64_t{1} << q) - 1, bit = uint64_t{1} << q;
    for (uint64_t p = static_cast<uint64_t>(blockIdx.x) * kBlk + threadIdx.x; p < pairs_total; p += step) {
        const uint64_t i0 = ((p & ~low) << 1) | (p & low), i1 = i0 | bit;
        const uint64_t j = i0 >> log_stride;
        const int act = action[j];
        if (act < 0) continue;
        const double f = scale[j];
        A a0 = s[i0], a1 = s[i1], z;
        z.x = 0;
        z.y = 0;
        A keep = act == 0 ? a0 : a1;
        keep.x = keep.x * f;
        keep.y = keep.y * f;
        if (act == 1) {
            s[i0] = z;
            s[i1] = keep;
        } else {
            s[i0] = keep;
            s[i1] = z;
        }
    }
}

// One thread per shot: the reference's final draw, sequentially — total =
// sum of p_i = re*re + im*im in index order, threshold = u * total, first
// index whose running sum reaches it (else the last).  The products and sums
// are rounded one by one (no fused multiply-add), as the host code does.
template <typename A>
__global__ __launch_bounds__(64) void sample_blocks_kernel(const A* __restrict__ s, uint64_t stride,
                                                           const double* __restrict__ u, int64_t n_shots,
                                                           uint64_t* __restrict__ out) {
    // Loads run kAhead elements ahead of the sequential sum, so a thread waits
    // one memory latency per kAhead terms; the additions stay in index order.
    constexpr int kAhead = 16;
    const int64_t j = static_cast<int64_t>(blockIdx.x) * 64 + threadIdx.x;
    if (j >= n_shots) return;
    const A* b = s + static_cast<uint64_t>(j) * stride;
    auto terms = [&](uint64_t i0, double* p) {
#pragma unroll
        for (int t = 0; t < kAhead; ++t) {
            double x = 0.0, y = 0.0;
            if (i0 + t < stride) {
                x = b[i0 + t].x;
                y = b[i0 + t].y;
            }
            p[t] = __dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y));
        }
    };
    double total = 0.0;
    for (uint64_t i0 = 0; i0 < stride; i0 += kAhead) {
        double p[kAhead];
        terms(i0, p);
#pragma unroll
        for (int t = 0; t < kAhead; ++t) total = __dadd_rn(total, p[t]);   // + 0 past the end: exact
    }
    const double threshold = __dmul_rn(u[j], total);
    double acc = 0.0;
    uint64_t pick = stride - 1;
    bool found = false;
    for (uint64_t i0 = 0; i0 < stride && !found; i0 += kAhead) {
        double p[kAhead];
        terms(i0, p);
#pragma unroll
        for (int t = 0; t < kAhead; ++t) {
            if (!found && i0 + t < stride) {
                acc = __dadd_rn(acc, p[t]);
                if (acc >= threshold) {
                    pick = i0 + t;
                    found = true;
                }
            }
        }
    }
    out[j] = pick;
}

// Per block j with active[j]: the 2x2 complex matrix mats[8j .. 8j + 8)
// (row-major (re, im)) on `q` — the noise channels' Pauli and Kraus
// operators, chosen per shot.
template <typename A>
__global__ __launch_bounds__(kBlk) void apply1q_kernel(A* __restrict__ s, int q, int log_stride, uint64_t pairs_total,
                                                       const double* __restrict__ mats,
                                                       const uint8_t* __restrict__ active) {
    const uint64_t step = static_cast<uint64_t>(gridDim.x) * kBlk;
    const uint64_t low = (uint64_t{1} << q) - 1, bit = uint64_t{1} << q;
    for (uint64_t p = static_cast<uint64_t>(blockIdx.x) * kBlk + threadIdx.x; p < pairs_total; p += step) {
        const uint64_t i0 = ((p & ~low) << 1) | (p & low), i1 = i0 | bit;
        const uint64_t j = i0 >> log_stride;
        if (!active[j]) continue;
        const double* m = mats + 8 * j;
        const A a0 = s[i0], a1 = s[i1];
        const double x0 = a0.x, y0 = a0.y, x1 = a1.x, y1 = a1.y;
        A o0, o1;
        o0.x = (m[0] * x0 - m[1] * y0) + (m[2] * x1 - m[3] * y1);
        o0.y = (m[0] * y0 + m[1] * x0) + (m[2] * y1 + m[3] * x1);
        o1.x = (m[4] * x0 - m[5] * y0) + (m[6] * x1 - m[7] * y1);
        o1.y = (m[4] * y0 + m[5] * x0) + (m[6] * y1 + m[7] * x1);
        s[i0] = o0;
        s[i1] = o1;
    }
}

// partial[(j * chunks + c) * 4 + {0..3}] = sums over pairs c*kChunk.. of shot
// j of |a|^2, |b|^2, Re(a conj b), Im(a conj b), with a the amplitude with
// the qubit's bit clear and b its partner (fixed-order reductions).
template <typename A>
__global__ __launch_bounds__(kBlk) void rho1_kernel(const A* __restrict__ s, int q, int log_stride, uint64_t pairs,
                                                    uint64_t chunks, double* __restrict__ partial) {
    __shared__ double red[4][kBlk];
    const uint64_t j = blockIdx.x / chunks, c = blockIdx.x % chunks;
    const uint64_t lo = c * kChunk, hi = std::min(pairs, lo + kChunk);
    const uint64_t low = (uint64_t{1} << q) - 1, bit = uint64_t{1} << q;
    double r00 = 0.0, r11 = 0.0, rx = 0.0, ry = 0.0;
    for (uint64_t p = lo + threadIdx.x; p < hi; p += kBlk) {
        const uint64_t i0 = (j << log_stride) | ((p & ~low) << 1) | (p & low);
        const double ax = s[i0].x, ay = s[i0].y, bx = s[i0 | bit].x, by = s[i0 | bit].y;
        r00 += ax * ax + ay * ay;
        r11 += bx * bx + by * by;
        rx += ax * bx + ay * by;
        ry += ay * bx - ax * by;
    }
    red[0][threadIdx.x] = r00;
    red[1][threadIdx.x] = r11;
    red[2][threadIdx.x] = rx;
    red[3][threadIdx.x] = ry;
    __syncthreads();
    for (int w = kBlk / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w)
            for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) partial[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

void check_blocks(const qhip_state* s, uint64_t stride, const char* what) {
    if (stride == 0 || stride > s->dim() || (stride & (stride - 1)) != 0)
        fail(QHIP_INVALID, std::string(what) + ": shot stride must be a power of two <= 2^n");
}

}  // namespace

void shots_broadcast(qhip_state* dst, const qhip_state* src) {
    if (dst->precision != src->precision || dst->n < src->n)
        fail(QHIP_INVALID, "qhip_shots_broadcast: destination narrower than the source or of another precision");
    auto& C = ctx();
    const uint64_t total = dst->dim(), low = src->dim() - 1;
    if (dst->precision == QHIP_C128)
        hipLaunchKernelGGL(broadcast_kernel<double2>, dim3(blocks_for(total)), dim3(kBlk), 0, C.stream,
                           static_cast<double2*>(dst->d), static_cast<const double2*>(src->d), total, low);
    else
        hipLaunchKernelGGL(broadcast_kernel<float2>, dim3(blocks_for(total)), dim3(kBlk), 0, C.stream,
                           static_cast<float2*>(dst->d), static_cast<const float2*>(src->d), total, low);
    QHIP_CHECK(hipGetLastError());
    QHIP_CHECK(hipStreamSynchronize(C.stream));
}

void shots_restore(qhip_state* dst, const qhip_state* saved, const uint8_t* active, uint64_t stride) {
    if (dst->precision != saved->precision || dst->n != saved->n)
        fail(QHIP_INVALID, "qhip_shots_restore: states differ in shape");
    check_blocks(dst, stride, "qhip_shots_restore");
    auto& C = ctx();
    const uint64_t nb = dst->dim() / stride;
    auto* d_act = static_cast<uint8_t*>(C.misc2.ensure(nb));
    QHIP_CHECK(hipMemcpyAsync(d_act, active, nb, hipMemcpyHostToDevice, C.stream));
    const int ls = std::countr_zero(stride);
    if (dst->precision == QHIP_C128)
        hipLaunchKernelGGL(restore_kernel<double2>, dim3(blocks_for(dst->dim())), dim3(kBlk), 0, C.stream,
                           static_cast<double2*>(dst->d), static_cast<const double2*>(saved->d), dst->dim(), ls, d_act);
    else
        hipLaunchKernelGGL(restore_kernel<float2>, dim3(blocks_for(dst->dim())), dim3(kBlk), 0, C.stream,
                           static_cast<float2*>(dst->d), static_cast<const float2*>(saved->d), dst->dim(), ls, d_act);
    QHIP_CHECK(hipGetLastError());
    QHIP_CHECK(hipStreamSynchronize(C.stream));
}

void shots_prob0(const qhip_state* s, int q, uint64_t stride, double* p0) {
    check_blocks(s, stride, "qhip_shots_prob0");
    if (q < 0 || (uint64_t{1} << q) >= stride) fail(QHIP_INVALID, "qhip_shots_prob0: qubit outside a shot");
    auto& C = ctx();
    const uint64_t nb = s->dim() / stride, pairs = stride / 2;
    const uint64_t chunks = (pairs + kChunk - 1) / kChunk;
    auto* d_part = static_cast<double*>(C.misc2.ensure(sizeof(double) * nb * chunks));
    const int ls = std::countr_zero(stride);
    const dim3 grid(static_cast<uint32_t>(nb * chunks));
    if (s->precision == QHIP_C128)
        hipLaunchKernelGGL(prob0_kernel<double2>, grid, dim3(kBlk), 0, C.stream, static_cast<const double2*>(s->d), q,
                           ls, pairs, chunks, d_part);
    else
        hipLaunchKernelGGL(prob0_kernel<float2>, grid, dim3(kBlk), 0, C.stream, static_cast<const float2*>(s->d), q,
                           ls, pairs, chunks, d_part);
    QHIP_CHECK(hipGetLastError());
    std::vector<double> part(nb * chunks);
    QHIP_CHECK(hipMemcpyAsync(part.data(), d_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));
    for (uint64_t j = 0; j < nb; ++j) {
        double t = 0.0;
        for (uint64_t c = 0; c < chunks; ++c) t += part[j * chunks + c];
        p0[j] = t;
    }
}

void shots_sample(const qhip_state* s, uint64_t stride, const double* u, int64_t n_shots, uint64_t* out) {
    check_blocks(s, stride, "qhip_shots_sample");
    if (n_shots < 0 || static_cast<uint64_t>(n_shots) > s->dim() / stride)
        fail(QHIP_INVALID, "qhip_shots_sample: more shots than blocks of the state");
    if (n_shots == 0) return;
    if (stride >= kWalkBlock) {
        // Wide shots: each block's walk through the exact chunk scan of the
        // terminal sampler (one view of the block per shot).
        for (int64_t j = 0; j < n_shots; ++j) {
            qhip_state view{};
            view.n = std::countr_zero(stride);
            view.precision = s->precision;
            view.d = static_cast<char*>(s->d) + static_cast<uint64_t>(j) * stride * s->amp_bytes();
            view.owned = false;
            out[j] = walk_draw(&view, u[j]);
        }
        return;
    }
    auto& C = ctx();
    char* base = static_cast<char*>(C.misc2.ensure(16 * static_cast<size_t>(n_shots) + 64));
    auto* d_u = reinterpret_cast<double*>(base);
    auto* d_out = reinterpret_cast<uint64_t*>(base + 8 * n_shots);
    QHIP_CHECK(hipMemcpyAsync(d_u, u, sizeof(double) * n_shots, hipMemcpyHostToDevice, C.stream));
    const dim3 grid(static_cast<uint32_t>((n_shots + 63) / 64));
    if (s->precision == QHIP_C128)
        hipLaunchKernelGGL(sample_blocks_kernel<double2>, grid, dim3(64), 0, C.stream,
                           static_cast<const double2*>(s->d), stride, d_u, n_shots, d_out);
    else
        hipLaunchKernelGGL(sample_blocks_kernel<float2>, grid, dim3(64), 0, C.stream,
                           static_cast<const float2*>(s->d), stride, d_u, n_shots, d_out);
    QHIP_CHECK(hipGetLastError());
    QHIP_CHECK(hipMemcpyAsync(out, d_out, sizeof(uint64_t) * n_shots, hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));
}

void shots_apply1q(qhip_state* s, int q, uint64_t stride, const double* mats, const uint8_t* active) {
    check_blocks(s, stride, "qhip_shots_apply1q");
    if (q < 0 || (uint64_t{1} << q) >= stride) fail(QHIP_INVALID, "qhip_shots_apply1q: qubit outside a shot");
    auto& C = ctx();
    const uint64_t nb = s->dim() / stride;
    char* base = static_cast<char*>(C.misc2.ensure(nb * (8 * sizeof(double) + 1) + 64));
    auto* d_mats = reinterpret_cast<double*>(base);
    auto* d_act = reinterpret_cast<uint8_t*>(base + nb * 8 * sizeof(double));
    QHIP_CHECK(hipMemcpyAsync(d_mats, mats, 8 * sizeof(double) * nb, hipMemcpyHostToDevice, C.stream));
    QHIP_CHECK(hipMemcpyAsync(d_act, active, nb, hipMemcpyHostToDevice, C.stream));
    const int ls = std::countr_zero(stride);
    const uint64_t pairs = s->dim() / 2;
    if (s->precision == QHIP_C128)
        hipLaunchKernelGGL(apply1q_kernel<double2>, dim3(blocks_for(pairs)), dim3(kBlk), 0, C.stream,
                           static_cast<double2*>(s->d), q, ls, pairs, d_mats, d_act);
    else
        hipLaunchKernelGGL(apply1q_kernel<float2>, dim3(blocks_for(pairs)), dim3(kBlk), 0, C.stream,
                           static_cast<float2*>(s->d), q, ls, pairs, d_mats, d_act);
    QHIP_CHECK(hipGetLastError());
    QHIP_CHECK(hipStreamSynchronize(C.stream));
}

void shots_rho1(const qhip_state* s, int q, uint64_t stride, double* out) {
    check_blocks(s, stride, "qhip_shots_rho1");
    if (q < 0 || (uint64_t{1} << q) >= stride) fail(QHIP_INVALID, "qhip_shots_rho1: qubit outside a shot");
    auto& C = ctx();
    const uint64_t nb = s->dim() / stride, pairs = stride / 2;
    const uint64_t chunks = (pairs + kChunk - 1) / kChunk;
    auto* d_part = static_cast<double*>(C.misc2.ensure(4 * sizeof(double) * nb * chunks));
    const int ls = std::countr_zero(stride);
    const dim3 grid(static_cast<uint32_t>(nb * chunks));
    if (s->precision == QHIP_C128)
        hipLaunchKernelGGL(rho1_kernel<double2>, grid, dim3(kBlk), 0, C.stream, static_cast<const double2*>(s->d), q,
                           ls, pairs, chunks, d_part);
    else
        hipLaunchKernelGGL(rho1_kernel<float2>, grid, dim3(kBlk), 0, C.stream, static_cast<const float2*>(s->d), q,
                           ls, pairs, chunks, d_part);
    QHIP_CHECK(hipGetLastError());
    std::vector<double> part(4 * nb * chunks);
    QHIP_CHECK(hipMemcpyAsync(part.data(), d_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));
    for (uint64_t j = 0; j < nb; ++j)
        for (int k = 0; k < 4; ++k) {
            double t = 0.0;
            for (uint64_t c = 0; c < chunks; ++c) t += part[(j * chunks + c) * 4 + k];
            out[4 * j + k] = t;
        }
}

void shots_project(qhip_state* s, int q, uint64_t stride, const int8_t* action, const double* scale) {
    check_blocks(s, stride, "qhip_shots_project");
    if (q < 0 || (uint64_t{1} << q) >= stride) fail(QHIP_INVALID, "qhip_shots_project: qubit outside a shot");
    auto& C = ctx();
    const uint64_t nb = s->dim() / stride;
    char* base = static_cast<char*>(C.misc2.ensure(nb * (sizeof(double) + 1) + 64));
    auto* d_scale = reinterpret_cast<double*>(base);
    auto* d_act = reinterpret_cast<int8_t*>(base + nb * sizeof(double));
    QHIP_CHECK(hipMemcpyAsync(d_scale, scale, sizeof(double) * nb, hipMemcpyHostToDevice, C.stream));
    QHIP_CHECK(hipMemcpyAsync(d_act, action, nb, hipMemcpyHostToDevice, C.stream));
    const int ls = std::countr_zero(stride);
    const uint64_t pairs = s->dim() / 2;
    if (s->precision == QHIP_C128)
        hipLaunchKernelGGL(project_kernel<double2>, dim3(blocks_for(pairs)), dim3(kBlk), 0, C.stream,
                           static_cast<double2*>(s->d), q, ls, pairs, d_act, d_scale);
    else
        hipLaunchKernelGGL(project_kernel<float2>, dim3(blocks_for(pairs)), dim3(kBlk), 0, C.stream,
                           static_cast<float2*>(s->d), q, ls, pairs, d_act, d_scale);
    QHIP_CHECK(hipGetLastError());
    QHIP_CHECK(hipStreamSynchronize(C.stream));
}

}  // namespace qhip
