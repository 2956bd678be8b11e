// kernels_apply.hip — state management and the tile-pass gate kernel (gfx950).
//
// Tile pass: one workgroup of 256 threads per tile.  The tile holds 2^T
// amplitudes in LDS (T = 12 for complex128, 13 for complex64: 64 KiB, two
// workgroups per CU); local index l maps to the global index by depositing
// l's bits onto the pass's tile qubits and the tile number's bits onto the
// remaining qubits.  Load and store move 16-B (c128) or 8-B (c64) elements
// with consecutive threads on consecutive low qubits, so every wavefront
// reads whole 2^low-amplitude runs.  In between, each op of the pass runs
// out of LDS with a barrier after it:
//   dense   — every thread takes groups of 2^k amplitudes (the op's operands
//             inserted as zero bits into the group number), multiplies the
//             2^k x 2^k matrix (read through the scalar cache: the indices
//             are wave-uniform) and writes the group back;
//   diag    — per amplitude, phase d[bits of the full index at the operands];
//   mask    — per amplitude, phase where every mask bit is set;
//   scalar  — global phase;
//   ip_*    — (adjoint only) <phi| G |psi> over the tile, wave-reduced with
//             lane shuffles into a per-wave LDS slot.
// In two-state mode (adjoint) the tile of psi and the tile of phi share the
// workgroup and every unitary op acts on both.
#include "common.h"

#include <cstdlib>

namespace qhip {

template <typename A> struct Real;
template <> struct Real<double2> { using T = double; };
template <> struct Real<float2> { using T = float; };

constexpr int kThreads = 256;

// Deposit of the tile number onto the non-tile qubits.
__device__ __forceinline__ uint64_t tile_base(uint64_t t, uint64_t outer_mask) {
    uint64_t g = 0;
    while (outer_mask) {
        const uint64_t low = outer_mask & (~outer_mask + 1);
        if (t & 1ull) g |= low;
        t >>= 1;
        outer_mask ^= low;
    }
    return g;
}

// Tile loads / stores touch every amplitude exactly once per pass: marked
// non-temporal so the streamed state does not evict reused data from L2.
constexpr int kBatch = 8;
__device__ __forceinline__ double2 stream_load(const double2* p) {
    double2 v;
    v.x = __builtin_nontemporal_load(&p->x);
    v.y = __builtin_nontemporal_load(&p->y);
    return v;
}
__device__ __forceinline__ float2 stream_load(const float2* p) {
    float2 v;
    v.x = __builtin_nontemporal_load(&p->x);
    v.y = __builtin_nontemporal_load(&p->y);
    return v;
}
__device__ __forceinline__ void stream_store(double2* p, double2 v) {
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}
__device__ __forceinline__ void stream_store(float2* p, float2 v) {
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}

template <typename A>
__device__ __forceinline__ A cmad(A acc, A u, A x) {
    acc.x += u.x * x.x - u.y * x.y;
    acc.y += u.x * x.y + u.y * x.x;
    return acc;
}

template <typename A>
__device__ __forceinline__ A cmul(A a, A b) {
    A r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

__device__ __forceinline__ double2 wave_sum(double2 v) {
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_xor(v.x, off);
        v.y += __shfl_xor(v.y, off);
    }
    return v;
}

template <int K>
__device__ __forceinline__ uint32_t insert_zeros(uint32_t g, const int8_t* spos) {
#pragma unroll
    for (int b = 0; b < K; ++b) {
        const uint32_t p = static_cast<uint32_t>(spos[b]);
        g = ((g >> p) << (p + 1)) | (g & ((1u << p) - 1u));
    }
    return g;
}

template <int K>
__device__ __forceinline__ uint32_t local_offset(uint32_t x, const int8_t* pos) {
    uint32_t o = 0;
#pragma unroll
    for (int b = 0; b < K; ++b) o |= ((x >> b) & 1u) << pos[b];
    return o;
}

// One dense k-qubit op over the LDS tile: groups of 2^K amplitudes.
template <typename A, int K>
__device__ __forceinline__ void dense_groups(A* s, int T, const DevOp* op, const A* __restrict__ U) {
    constexpr int D = 1 << K;
    int8_t pos[K], spos[K];
#pragma unroll
    for (int b = 0; b < K; ++b) { pos[b] = op->pos[b]; spos[b] = op->spos[b]; }
    const uint32_t groups = 1u << (T - K);
    for (uint32_t g = threadIdx.x; g < groups; g += kThreads) {
        const uint32_t base = insert_zeros<K>(g, spos);
        A in[D];
#pragma unroll
        for (int x = 0; x < D; ++x) in[x] = s[base | local_offset<K>(x, pos)];
        if constexpr (K <= 3) {
#pragma unroll
            for (int y = 0; y < D; ++y) {
                A acc;
                acc.x = 0;
                acc.y = 0;
#pragma unroll
                for (int x = 0; x < D; ++x) acc = cmad(acc, U[y * D + x], in[x]);
                s[base | local_offset<K>(y, pos)] = acc;
            }
        } else {
#pragma unroll 1
            for (int y = 0; y < D; ++y) {
                A acc;
                acc.x = 0;
                acc.y = 0;
#pragma unroll
                for (int x = 0; x < D; ++x) acc = cmad(acc, U[y * D + x], in[x]);
                s[base | local_offset<K>(y, pos)] = acc;
            }
        }
    }
}

template <typename A, int KMAX>
__device__ __forceinline__ void dense_op(A* s, int T, const DevOp* op, const A* __restrict__ U) {
    switch (op->k) {
        case 1: dense_groups<A, 1>(s, T, op, U); break;
        case 2: if constexpr (KMAX >= 2) dense_groups<A, 2>(s, T, op, U); break;
        case 3: if constexpr (KMAX >= 3) dense_groups<A, 3>(s, T, op, U); break;
        case 4: if constexpr (KMAX >= 4) dense_groups<A, 4>(s, T, op, U); break;
        case 5: if constexpr (KMAX >= 5) dense_groups<A, 5>(s, T, op, U); break;
        default: break;
    }
}

__device__ __forceinline__ uint32_t diag_index(uint32_t l, uint64_t gbase, const DevOp* op, int k) {
    uint32_t idx = 0;
    for (int b = 0; b < k; ++b) {
        const int p = op->pos[b];
        const uint32_t bit = p < 32 ? (l >> p) & 1u : static_cast<uint32_t>((gbase >> (p - 32)) & 1ull);
        idx |= bit << b;
    }
    return idx;
}

// <phi| G |psi> over the tile for a dense generator on K <= 2 qubits.
template <typename A, int K>
__device__ __forceinline__ double2 ip_dense(const A* s0, const A* s1, int T, const DevOp* op,
                                            const A* __restrict__ G) {
    constexpr int D = 1 << K;
    int8_t pos[K], spos[K];
#pragma unroll
    for (int b = 0; b < K; ++b) { pos[b] = op->pos[b]; spos[b] = op->spos[b]; }
    double2 acc = make_double2(0.0, 0.0);
    const uint32_t groups = 1u << (T - K);
    for (uint32_t g = threadIdx.x; g < groups; g += kThreads) {
        const uint32_t base = insert_zeros<K>(g, spos);
        A in[D];
#pragma unroll
        for (int x = 0; x < D; ++x) in[x] = s0[base | local_offset<K>(x, pos)];
#pragma unroll
        for (int y = 0; y < D; ++y) {
            A gy;
            gy.x = 0;
            gy.y = 0;
#pragma unroll
            for (int x = 0; x < D; ++x) gy = cmad(gy, G[y * D + x], in[x]);
            const A f = s1[base | local_offset<K>(y, pos)];
            const double fr = f.x, fi = f.y, gr = gy.x, gi = gy.y;   // conj(f) * gy
            acc.x += fr * gr + fi * gi;
            acc.y += fr * gi - fi * gr;
        }
    }
    return acc;
}

// The ops of pass P over the tile in LDS (s0; s1 too in two-state mode),
// a barrier after each.  Returns the number of inner-product ops, whose
// per-wave partials are left in wave_ip.
template <typename A, bool TWO, int KMAX>
__device__ __forceinline__ int run_ops(A* s0, A* s1, const PassDesc& P, uint64_t gbase,
                                       const DevOp* __restrict__ ops, const A* __restrict__ coefs,
                                       double2* wave_ip) {
    const uint32_t L = 1u << P.T;
    int n_ip = 0;
    for (int o = P.op_begin; o < P.op_end; ++o) {
        const DevOp* op = ops + o;
        const int kind = op->kind, k = op->k;
        const A* U = coefs + op->coef;
        switch (kind) {
            case OP_DENSE:
                dense_op<A, KMAX>(s0, P.T, op, U);
                if (TWO) dense_op<A, KMAX>(s1, P.T, op, U);
                break;
            case OP_DIAG:
                for (uint32_t l = threadIdx.x; l < L; l += kThreads) {
                    const A d = U[diag_index(l, gbase, op, k)];
                    s0[l] = cmul(s0[l], d);
                    if (TWO) s1[l] = cmul(s1[l], d);
                }
                break;
            case OP_MASKPHASE: {
                const A ph = U[0];
                const uint64_t lm = op->lmask, gm = op->gmask;
                if ((gbase & gm) == gm)
                    for (uint32_t l = threadIdx.x; l < L; l += kThreads)
                        if ((l & lm) == lm) {
                            s0[l] = cmul(s0[l], ph);
                            if (TWO) s1[l] = cmul(s1[l], ph);
                        }
                break;
            }
            case OP_SCALAR: {
                const A ph = U[0];
                for (uint32_t l = threadIdx.x; l < L; l += kThreads) {
                    s0[l] = cmul(s0[l], ph);
                    if (TWO) s1[l] = cmul(s1[l], ph);
                }
                break;
            }
            case OP_IP_DENSE:
            case OP_IP_DIAG:
                if constexpr (TWO) {
                    double2 acc = make_double2(0.0, 0.0);
                    if (kind == OP_IP_DENSE) {
                        acc = k == 1 ? ip_dense<A, 1>(s0, s1, P.T, op, U) : ip_dense<A, 2>(s0, s1, P.T, op, U);
                    } else {
                        for (uint32_t l = threadIdx.x; l < L; l += kThreads) {
                            const A d = U[k == 0 ? 0 : diag_index(l, gbase, op, k)];
                            const A x = s0[l], f = s1[l];
                            const double gr = static_cast<double>(d.x) * x.x - static_cast<double>(d.y) * x.y;
                            const double gi = static_cast<double>(d.x) * x.y + static_cast<double>(d.y) * x.x;
                            acc.x += static_cast<double>(f.x) * gr + static_cast<double>(f.y) * gi;
                            acc.y += static_cast<double>(f.x) * gi - static_cast<double>(f.y) * gr;
                        }
                    }
                    acc = wave_sum(acc);
                    if ((threadIdx.x & 63) == 0) wave_ip[n_ip * (kThreads / 64) + (threadIdx.x >> 6)] = acc;
                    ++n_ip;
                }
                break;
        }
        __syncthreads();
    }

    return n_ip;
}

// Tile of pass P: load, run the pass's ops in LDS, store.  KMAX bounds the
// widest dense op of the pass (register budget of the group vectors).
template <typename A, bool TWO, int KMAX>
__global__ __launch_bounds__(kThreads) void tile_pass_kernel(
    A* __restrict__ psi, A* __restrict__ phi, const PassDesc P, const DevOp* __restrict__ ops,
    const A* __restrict__ coefs, double2* __restrict__ ip_out, int64_t n_tiles) {
    extern __shared__ unsigned char smem_raw[];
    A* s0 = reinterpret_cast<A*>(smem_raw);
    A* s1 = s0 + (1u << P.T);
    double2* wave_ip = reinterpret_cast<double2*>(s0 + (TWO ? 2u : 1u) * (1u << P.T));
    const uint32_t L = 1u << P.T;
    const uint64_t t = blockIdx.x;
    const uint64_t full = P.n >= 64 ? ~0ull : ((1ull << P.n) - 1ull);
    const uint64_t gbase = tile_base(t, full & ~P.tmask);

    // Global index of local l = threadIdx.x + 256 r: the thread's own low
    // eight local bits are deposited once; the subsets of hmask enumerate
    // the deposits of r in increasing order.
    uint64_t dj = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b)
        if (b < P.T) dj |= static_cast<uint64_t>((threadIdx.x >> b) & 1u) << P.qpos[b];
    // Loads are issued kBatch at a time before any LDS write, so each thread
    // keeps kBatch HBM requests in flight (tiles of >= 2^11 amplitudes; the
    // R = 2^(T-8) rows per thread are a multiple of kBatch there).
    const uint32_t R = L / kThreads;
    if (R >= kBatch) {
        uint64_t dr = 0;
        for (uint32_t r0 = 0; r0 < R; r0 += kBatch) {
            A v0[kBatch], v1[kBatch];
#pragma unroll
            for (int i = 0; i < kBatch; ++i) {
                const uint64_t g = gbase | dj | dr;
                v0[i] = stream_load(psi + g);
                if (TWO) v1[i] = stream_load(phi + g);
                dr = (dr - P.hmask) & P.hmask;
            }
#pragma unroll
            for (int i = 0; i < kBatch; ++i) {
                const uint32_t l = threadIdx.x + (r0 + i) * kThreads;
                s0[l] = v0[i];
                if (TWO) s1[l] = v1[i];
            }
        }
    } else {
        uint64_t dr = 0;
        for (uint32_t l = threadIdx.x; l < L; l += kThreads) {
            const uint64_t g = gbase | dj | dr;
            s0[l] = psi[g];
            if (TWO) s1[l] = phi[g];
            dr = (dr - P.hmask) & P.hmask;
        }
    }
    __syncthreads();

    const int n_ip = run_ops<A, TWO, KMAX>(s0, s1, P, gbase, ops, coefs, wave_ip);

    if (TWO && static_cast<int>(threadIdx.x) < n_ip) {
        // Sum the per-wave values in fixed order: partial k of this tile.
        const int k = threadIdx.x;
        double2 v = wave_ip[k * (kThreads / 64)];
        for (int w = 1; w < kThreads / 64; ++w) {
            v.x += wave_ip[k * (kThreads / 64) + w].x;
            v.y += wave_ip[k * (kThreads / 64) + w].y;
        }
        ip_out[static_cast<int64_t>(k) * n_tiles + static_cast<int64_t>(t)] = v;
    }

    if (R >= kBatch) {
        uint64_t dr = 0;
        for (uint32_t r0 = 0; r0 < R; r0 += kBatch) {
            A v0[kBatch], v1[kBatch];
#pragma unroll
            for (int i = 0; i < kBatch; ++i) {
                const uint32_t l = threadIdx.x + (r0 + i) * kThreads;
                v0[i] = s0[l];
                if (TWO) v1[i] = s1[l];
            }
#pragma unroll
            for (int i = 0; i < kBatch; ++i) {
                const uint64_t g = gbase | dj | dr;
                stream_store(psi + g, v0[i]);
                if (TWO) stream_store(phi + g, v1[i]);
                dr = (dr - P.hmask) & P.hmask;
            }
        }
    } else {
        uint64_t dr = 0;
        for (uint32_t l = threadIdx.x; l < L; l += kThreads) {
            const uint64_t g = gbase | dj | dr;
            psi[g] = s0[l];
            if (TWO) phi[g] = s1[l];
            dr = (dr - P.hmask) & P.hmask;
        }
    }
}

// Persistent, software-pipelined variant for single-state passes with
// R = 2^(T-8) rows per thread: each workgroup walks tiles blockIdx.x,
// blockIdx.x + gridDim.x, ...; the next tile's HBM loads are issued into
// registers before the current tile's ops run in LDS, so load latency hides
// behind the compute and the store of the current tile.
template <typename A, int KMAX, int R>
__global__ __launch_bounds__(kThreads) void stream_pass_kernel(
    A* __restrict__ psi, const PassDesc P, const DevOp* __restrict__ ops, const A* __restrict__ coefs,
    int64_t n_tiles) {
    extern __shared__ unsigned char smem_raw[];
    A* s0 = reinterpret_cast<A*>(smem_raw);
    const uint64_t full = P.n >= 64 ? ~0ull : ((1ull << P.n) - 1ull);
    const uint64_t outer = full & ~P.tmask;
    uint64_t dj = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) dj |= static_cast<uint64_t>((threadIdx.x >> b) & 1u) << P.qpos[b];
    uint64_t dr_tab[R];
    {
        uint64_t dr = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            dr_tab[r] = dj | dr;
            dr = (dr - P.hmask) & P.hmask;
        }
    }
    int64_t t = blockIdx.x;
    if (t >= n_tiles) return;
    A pre[R];
    uint64_t gbase = tile_base(static_cast<uint64_t>(t), outer);
#pragma unroll
    for (int r = 0; r < R; ++r) pre[r] = stream_load(psi + (gbase | dr_tab[r]));
    while (true) {
#pragma unroll
        for (int r = 0; r < R; ++r) s0[threadIdx.x + r * kThreads] = pre[r];
        __syncthreads();
        const int64_t tn = t + gridDim.x;
        const uint64_t gnext = tn < n_tiles ? tile_base(static_cast<uint64_t>(tn), outer) : 0;
        if (tn < n_tiles) {
#pragma unroll
            for (int r = 0; r < R; ++r) pre[r] = stream_load(psi + (gnext | dr_tab[r]));
        }
        (void)run_ops<A, false, KMAX>(s0, nullptr, P, gbase, ops, coefs, nullptr);
#pragma unroll
        for (int r = 0; r < R; ++r) stream_store(psi + (gbase | dr_tab[r]), s0[threadIdx.x + r * kThreads]);
        if (tn >= n_tiles) break;
        __syncthreads();   // LDS reads of this tile done before the next tile lands
        t = tn;
        gbase = gnext;
    }
}

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

// QHIP_PIPELINE=1 selects the persistent pipelined kernel for single-state
// passes of 2^11..2^13 tiles; the default one-tile-per-workgroup kernel keeps
// more workgroups resident per CU, which measured faster.
static bool tile_pipelining() {
    static const bool on = [] {
        const char* e = std::getenv("QHIP_PIPELINE");
        return e && e[0] == '1';
    }();
    return on;
}

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

template <typename A, bool TWO, int KMAX>
static const void* tile_fn() {
    return reinterpret_cast<const void*>(&tile_pass_kernel<A, TWO, KMAX>);
}

template <typename A, int KMAX, int R>
static void launch_stream(const PassDesc& P, A* psi, const DevOp* d_ops, const A* d_coefs, int64_t n_tiles,
                          size_t lds) {
    auto& C = ctx();
    static int per_cu = 0;   // resident workgroups per CU for this instantiation
    if (per_cu == 0) {
        QHIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(&stream_pass_kernel<A, KMAX, R>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        int blocks = 0;
        QHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &blocks, reinterpret_cast<const void*>(&stream_pass_kernel<A, KMAX, R>), kThreads, lds));
        per_cu = std::max(1, blocks);
    }
    const int64_t grid = std::min<int64_t>(n_tiles, static_cast<int64_t>(C.n_cu) * per_cu);
    hipLaunchKernelGGL((stream_pass_kernel<A, KMAX, R>), dim3(static_cast<uint32_t>(grid)), dim3(kThreads), lds,
                       C.stream, psi, P, d_ops, d_coefs, n_tiles);
    QHIP_CHECK(hipGetLastError());
}

template <typename A, int KMAX>
static bool try_stream(const PassDesc& P, A* psi, const DevOp* d_ops, const A* d_coefs, int64_t n_tiles,
                       size_t lds) {
    switch (P.T) {
        case 11: launch_stream<A, KMAX, 8>(P, psi, d_ops, d_coefs, n_tiles, lds); return true;
        case 12: launch_stream<A, KMAX, 16>(P, psi, d_ops, d_coefs, n_tiles, lds); return true;
        case 13: launch_stream<A, KMAX, 32>(P, psi, d_ops, d_coefs, n_tiles, lds); return true;
        default: return false;
    }
}

template <typename A, bool TWO>
static void launch_tile(const PassDesc& P, A* psi, A* phi, double2* ip_part, const DevOp* d_ops, const A* d_coefs,
                        int64_t n_tiles, size_t lds) {
    auto& C = ctx();
    const int km = P.kmax <= 2 ? 2 : std::min(P.kmax, 5);
    if (!TWO && tile_pipelining()) {
        bool done = false;
        switch (km) {
            case 2: done = try_stream<A, 2>(P, psi, d_ops, d_coefs, n_tiles, lds); break;
            case 3: done = try_stream<A, 3>(P, psi, d_ops, d_coefs, n_tiles, lds); break;
            case 4: done = try_stream<A, 4>(P, psi, d_ops, d_coefs, n_tiles, lds); break;
            default: done = try_stream<A, 5>(P, psi, d_ops, d_coefs, n_tiles, lds); break;
        }
        if (done) return;
    }
    const dim3 grid(static_cast<uint32_t>(n_tiles)), block(kThreads);
    // Above 64 KiB of dynamic LDS a kernel must opt in (160 KiB per CU).
    static bool opted[4] = {false, false, false, false};
    const int kk = km - 2;
    if (lds > 65536 && !opted[kk]) {
        const void* fn = kk == 0 ? tile_fn<A, TWO, 2>() : kk == 1 ? tile_fn<A, TWO, 3>()
                       : kk == 2 ? tile_fn<A, TWO, 4>() : tile_fn<A, TWO, 5>();
        QHIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
        opted[kk] = true;
    }
    switch (km) {
        case 2: hipLaunchKernelGGL((tile_pass_kernel<A, TWO, 2>), grid, block, lds, C.stream, psi, phi, P, d_ops,
                                   d_coefs, ip_part, n_tiles); break;
        case 3: hipLaunchKernelGGL((tile_pass_kernel<A, TWO, 3>), grid, block, lds, C.stream, psi, phi, P, d_ops,
                                   d_coefs, ip_part, n_tiles); break;
        case 4: hipLaunchKernelGGL((tile_pass_kernel<A, TWO, 4>), grid, block, lds, C.stream, psi, phi, P, d_ops,
                                   d_coefs, ip_part, n_tiles); break;
        default: hipLaunchKernelGGL((tile_pass_kernel<A, TWO, 5>), grid, block, lds, C.stream, psi, phi, P, d_ops,
                                    d_coefs, ip_part, n_tiles); break;
    }
    QHIP_CHECK(hipGetLastError());
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
            const size_t lds = (size_t{1} << P.T) * sizeof(A) * (phi ? 2 : 1) +
                               (phi ? 64 * (kThreads / 64) * sizeof(double2) : 0);
            if (phi) launch_tile<A, true>(P, psi, phi, ip_part, d_ops, d_coefs, n_tiles, lds);
            else launch_tile<A, false>(P, psi, phi, ip_part, d_ops, d_coefs, n_tiles, lds);
            if (phi) {
                int n_ip = 0;
                for (int o = P.op_begin; o < P.op_end; ++o)
                    n_ip += plan.ops[o].kind == OP_IP_DENSE || plan.ops[o].kind == OP_IP_DIAG;
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
            if (plan.ops[o].kind == OP_IP_DENSE || plan.ops[o].kind == OP_IP_DIAG) {
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

void free_compiled(CompiledPlan* cp) {
    if (!cp) return;
    if (cp->d_tables) {
        (void)hipStreamSynchronize(ctx().stream);
        (void)hipFree(cp->d_tables);
    }
    delete cp;
}

}  // namespace qhip
