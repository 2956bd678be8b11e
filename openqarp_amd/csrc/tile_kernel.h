// tile_kernel.h — the tile-pass gate kernel (gfx950), instantiated once per
// amplitude type by tile_c128.hip and tile_c64.hip.
//
// Tile pass: one workgroup of 256 threads per tile.  The tile holds 2^T
// amplitudes in a static LDS array sized for 2^TB >= 2^T (TB is a template
// parameter, so the array's address is a link-time constant and every LDS
// access is one XOR of a per-thread and a wave-uniform byte offset); local
// index l maps to the global index by depositing l's bits onto the pass's
// tile qubits and the tile number's bits onto the remaining qubits.  Load and
// store move 16-B (c128) or 8-B (c64) elements with consecutive threads on
// consecutive low qubits, so every wavefront reads whole 2^low-amplitude
// runs.  In between, each op of the pass runs out of LDS with a barrier
// after it:
//   dense   — every thread takes groups of 2^k amplitudes (the group number's
//             bits deposited on the non-operand positions, in an order chosen
//             on the host so that 16 neighbouring lanes hit 16 distinct LDS
//             slots modulo 16), multiplies the 2^k x 2^k matrix (read
//             through the scalar cache: the indices are wave-uniform) and
//             writes the group back; for k <= 2 two groups are in flight;
//   diag    — per amplitude, phase d[bits of the full index at the operands];
//   mask    — per amplitude, phase where every mask bit is set;
//   scalar  — global phase;
//   ip_*    — (adjoint only) <phi| G |psi> over the tile, wave-reduced with
//             lane shuffles into a per-wave LDS slot.
// In two-state mode (adjoint) the tile of psi and the tile of phi share the
// workgroup and every unitary op acts on both.
//
// LDS layout: amplitude l of state z sits at element z * 2^TB + swz(l) with
// swz(l) = l ^ (((l >> 4) ^ (l >> 8)) & 15).  swz is linear over XOR and
// only changes bits 0-3, so the slot of (thread bits) | (uniform bits) is
// swz(thread bits) ^ swz(uniform bits).
#pragma once

#include <cstddef>
#include <type_traits>

#include "common.h"

namespace qhip {
namespace tile {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kBatch = 8;

template <typename A> struct Amp;
template <> struct Amp<double2> { using R = double; static constexpr int kShift = 4; };
template <> struct Amp<float2> { using R = float; static constexpr int kShift = 3; };

__host__ __device__ constexpr uint32_t swz(uint32_t l) { return l ^ (((l >> 4) ^ (l >> 8)) & 15u); }

template <typename A>
__device__ __forceinline__ A& at(A* s, uint32_t byte_offset) {
    return *reinterpret_cast<A*>(reinterpret_cast<char*>(s) + byte_offset);
}

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

// Complex multiply-accumulate as four explicit FMAs (the compiler otherwise
// splits it into separate multiplies and adds).
__device__ __forceinline__ double2 cmad(double2 acc, double2 u, double2 x) {
    acc.x = __fma_rn(u.x, x.x, acc.x);
    acc.x = __fma_rn(-u.y, x.y, acc.x);
    acc.y = __fma_rn(u.x, x.y, acc.y);
    acc.y = __fma_rn(u.y, x.x, acc.y);
    return acc;
}
__device__ __forceinline__ float2 cmad(float2 acc, float2 u, float2 x) {
    acc.x = __fmaf_rn(u.x, x.x, acc.x);
    acc.x = __fmaf_rn(-u.y, x.y, acc.x);
    acc.y = __fmaf_rn(u.x, x.y, acc.y);
    acc.y = __fmaf_rn(u.y, x.x, acc.y);
    return acc;
}
template <typename A>
__device__ __forceinline__ A cmul(A a, A b) {
    A z;
    z.x = 0;
    z.y = 0;
    return cmad(z, a, b);
}

// A double moved across lanes by a DPP control (both halves).
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double wave_sum1(double v) {
    v += dpp_move<0xb1>(v);    // quad_perm [1,0,3,2]: lane ^ 1
    v += dpp_move<0x4e>(v);    // quad_perm [2,3,0,1]: lane ^ 2 (every lane: its quad's sum)
    v += dpp_move<0x141>(v);   // row_half_mirror: the other quad of the half-row
    v += dpp_move<0x140>(v);   // row_mirror: the other half of the row (every lane: its row's sum)
    const double r1 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 16),
                                       __builtin_amdgcn_readlane(__double2loint(v), 16));
    const double r2 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 32),
                                       __builtin_amdgcn_readlane(__double2loint(v), 32));
    const double r3 = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 48),
                                       __builtin_amdgcn_readlane(__double2loint(v), 48));
    return ((v + r1) + r2) + r3;
}

// Sum over the 64 lanes of a wave, valid in lanes 0..15 (every lane must be
// active).  Lane moves within rows of 16 by DPP and three lane reads across
// rows: no LDS instructions (a __shfl_xor chain is six dependent
// ds_bpermute round trips per component).
__device__ __forceinline__ double2 wave_sum(double2 v) { return make_double2(wave_sum1(v.x), wave_sum1(v.y)); }

// One DevOp as read through the scalar cache: eight 8-B loads from a
// wave-uniform address, byte fields unpacked with shifts (sub-dword loads
// would go through the vector memory path, one round trip each).
struct OpRegs {
    int kind, k;
    uint64_t pos, spos, gpos0, gpos1;   // packed int8 fields of DevOp
    int64_t coef;
    uint64_t lmask, gmask;
    __device__ int p(int b) const { return static_cast<int>((pos >> (8 * b)) & 0xffu); }
    __device__ int g(int b) const {
        return static_cast<int>(((b < 8 ? gpos0 : gpos1) >> (8 * (b & 7))) & 0xffu);
    }
};
static_assert(sizeof(DevOp) == 64 && offsetof(DevOp, pos) == 8 && offsetof(DevOp, spos) == 16 &&
                  offsetof(DevOp, coef) == 24 && offsetof(DevOp, lmask) == 32 && offsetof(DevOp, gmask) == 40 &&
                  offsetof(DevOp, gpos) == 48,
              "DevOp layout read by load_op");

__device__ __forceinline__ OpRegs load_op(const DevOp* __restrict__ op) {
    const uint64_t* q = reinterpret_cast<const uint64_t*>(op);
    OpRegs r;
    const uint64_t w0 = q[0];
    r.kind = static_cast<int>(static_cast<uint32_t>(w0));
    r.k = static_cast<int>(w0 >> 32);
    r.pos = q[1];
    r.spos = q[2];
    r.coef = static_cast<int64_t>(q[3]);
    r.lmask = q[4];
    r.gmask = q[5];
    r.gpos0 = q[6];
    r.gpos1 = q[7];
    return r;
}

template <int K>
__device__ __forceinline__ uint32_t local_offset(uint32_t x, const OpRegs& op) {
    uint32_t o = 0;
#pragma unroll
    for (int b = 0; b < K; ++b) o |= ((x >> b) & 1u) << op.p(b);
    return o;
}

// Sub-table of a grouped diagonal for a tile: the bits of its operands
// outside the tile (global qubits in spos bytes 0..5, their count in byte 7).
__device__ __forceinline__ uint32_t diag_sub_table(uint64_t gbase, const OpRegs& op) {
    const int kg = static_cast<int>((op.spos >> 56) & 0xffu);
    uint32_t gsel = 0;
    for (int j = 0; j < kg; ++j) gsel |= static_cast<uint32_t>((gbase >> ((op.spos >> (8 * j)) & 0xffu)) & 1u) << j;
    return gsel;
}

// Entries of a register-block member: a diagonal member (pattern >= 6) with
// operands outside the tile uses the sub-table its tile selects.
__device__ __forceinline__ int member_local_bits(int pattern) { return pattern == 12 ? 0 : pattern <= 8 ? 1 : 2; }

// Byte offset (swizzled) of this thread's part of a dense op's group number:
// thread bit b of the group number sits at local position gpos[b].
template <typename A>
__device__ __forceinline__ uint32_t thread_part(const OpRegs& op, int G) {
    uint32_t tp = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b)
        if (b < G) tp |= ((threadIdx.x >> b) & 1u) << op.g(b);
    return swz(tp) << Amp<A>::kShift;
}

// y = U x for one group held in registers, written back to LDS: row y goes
// to byte offset a[y] (K <= 3), or to tp ^ h ^ slot(y) computed per row
// (K >= 4, whose rows are not unrolled).
template <typename A, int K>
__device__ __forceinline__ void matvec_store(A* s, const uint32_t* a, uint32_t tp, uint32_t h, const OpRegs& op,
                                             const A* in, const A* __restrict__ U) {
    constexpr int D = 1 << K;
    if constexpr (K <= 3) {
#pragma unroll
        for (int y = 0; y < D; ++y) {
            A acc;
            acc.x = 0;
            acc.y = 0;
#pragma unroll
            for (int x = 0; x < D; ++x) acc = cmad(acc, U[y * D + x], in[x]);
            at(s, a[y]) = acc;
        }
    } else {
#pragma unroll 1
        for (int y = 0; y < D; ++y) {
            A acc;
            acc.x = 0;
            acc.y = 0;
#pragma unroll
            for (int x = 0; x < D; ++x) acc = cmad(acc, U[y * D + x], in[x]);
            const uint32_t o = swz(local_offset<K>(static_cast<uint32_t>(y), op)) << Amp<A>::kShift;
            at(s, tp ^ (h ^ o)) = acc;
        }
    }
}

// One dense k-qubit op over one state's LDS tile.  tp: this thread's byte
// offset (thread_part | state offset); G = T - K group-number bits, whose
// bits >= 8 sit on the local positions of op.gmask.
template <typename A, int K>
__device__ __forceinline__ void dense_groups(A* s, uint32_t tp, int G, const OpRegs& op, const A* __restrict__ U) {
    constexpr int D = 1 << K;
    constexpr int sh = Amp<A>::kShift;
    if (G < 8 && threadIdx.x >= (1u << G)) return;
    uint32_t so[D];
#pragma unroll
    for (int x = 0; x < D; ++x) so[x] = swz(local_offset<K>(x, op)) << sh;
    const uint32_t hm = static_cast<uint32_t>(op.gmask);
    const uint32_t iters = G > 8 ? 1u << (G - 8) : 1u;
    uint32_t hi = 0;   // deposit of the current group number's bits >= 8
    uint32_t i = 0;
    if constexpr (K <= 2) {
        for (; i + 1 < iters; i += 2) {
            const uint32_t h0 = swz(hi) << sh;
            hi = (hi - hm) & hm;
            const uint32_t h1 = swz(hi) << sh;
            hi = (hi - hm) & hm;
            uint32_t a0[D], a1[D];
            A in0[D], in1[D];
#pragma unroll
            for (int x = 0; x < D; ++x) {
                a0[x] = tp ^ (h0 ^ so[x]);
                a1[x] = tp ^ (h1 ^ so[x]);
                in0[x] = at(s, a0[x]);
                in1[x] = at(s, a1[x]);
            }
            matvec_store<A, K>(s, a0, tp, h0, op, in0, U);
            matvec_store<A, K>(s, a1, tp, h1, op, in1, U);
        }
    }
    for (; i < iters; ++i) {
        const uint32_t h0 = swz(hi) << sh;
        hi = (hi - hm) & hm;
        uint32_t a0[D];
        A in0[D];
#pragma unroll
        for (int x = 0; x < D; ++x) {
            a0[x] = tp ^ (h0 ^ so[x]);
            in0[x] = at(s, a0[x]);
        }
        matvec_store<A, K>(s, a0, tp, h0, op, in0, U);
    }
}

template <typename A, int KMAX>
__device__ __forceinline__ void dense_op(A* s, uint32_t tp, int G, const OpRegs& op, const A* __restrict__ U) {
    switch (op.k) {
        case 1: dense_groups<A, 1>(s, tp, G, op, U); break;
        case 2: if constexpr (KMAX >= 2) dense_groups<A, 2>(s, tp, G, op, U); break;
        case 3: if constexpr (KMAX >= 3) dense_groups<A, 3>(s, tp, G, op, U); break;
        case 4: if constexpr (KMAX >= 4) dense_groups<A, 4>(s, tp, G, op, U); break;
        case 5: if constexpr (KMAX >= 5) dense_groups<A, 5>(s, tp, G, op, U); break;
        default: break;
    }
}

// Dense k-qubit op (k = 3..5) on the matrix cores.  Groups go
// 16 to a chunk (the MFMA's 16 columns, lane % 16); wave w takes chunks w,
// w + 4, ....  Lane group c = lane / 16 loads input amplitudes c, c + 4, ...
// of its group (re then im feed consecutive K-steps), the D/8 M-tiles of 16
// real rows are accumulated over D/2 K-steps from per-lane fragments F,
// and lane group g writes output amplitudes 8mt + 2g and 8mt + 2g + 1 of
// its group.  A wave reads its whole chunk before writing any of it.
// FP64 (complex128) and FP32 (complex64) matrix-core steps.  Both read A
// at lane (i = lane % 16, k = lane / 16) and B at lane (j = lane % 16,
// k = lane / 16); D row i of the 16x16 result lands in register r of lane
// group g = lane / 16 as i = 4r + g (f64) or i = 4g + r (f32), measured on
// the device.
template <typename A> struct Mfma;
template <> struct Mfma<double2> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 step(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static constexpr bool kGroupRows = false;   // row = 4r + g
};
template <> struct Mfma<float2> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 step(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static constexpr bool kGroupRows = true;    // row = 4g + r
};

// k = 3, realified 16x16 matrix.  The host orders the fragment rows so that
// lane group g receives output amplitudes 8mt + 2g and 8mt + 2g + 1 in
// registers (0, 1) and (2, 3) for either row layout.
template <typename A, int K>
__device__ __forceinline__ void dense_mfma(A* s, uint32_t zoff, int G, const OpRegs& op, const A* __restrict__ F) {
    using M = Mfma<A>;
    constexpr int sh = Amp<A>::kShift;
    constexpr int D = 1 << K, NT = D / 8, H = D / 4;   // M-tiles; input amplitudes (= fragment pairs) per lane
    const uint32_t lane = threadIdx.x & 63u, j = lane & 15u, c = lane >> 4;
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform: chunk loop in SALU
    uint32_t lp = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) lp |= ((j >> b) & 1u) << op.g(b);
    lp = (swz(lp) << sh) | zoff;
    uint32_t so_in[H];
#pragma unroll
    for (int t = 0; t < H; ++t) so_in[t] = swz(local_offset<K>(c + 4u * t, op)) << sh;
    const uint32_t chunks = 1u << (G - 4);
    for (uint32_t ch = w; ch < chunks; ch += kWaves) {
        uint32_t up = 0;   // deposit of the group-number bits >= 4 (wave-uniform)
        for (int b = 4; b < G; ++b) up |= ((ch >> (b - 4)) & 1u) << op.g(b);
        const uint32_t base = lp ^ (swz(up) << sh);
        A x[H];
#pragma unroll
        for (int t = 0; t < H; ++t) x[t] = at(s, base ^ so_in[t]);
#pragma unroll
        for (int mt = 0; mt < NT; ++mt) {
            typename M::v4 acc = {0, 0, 0, 0};
#pragma unroll
            for (int t = 0; t < H; ++t) {
                const A f = F[(mt * H + t) * 64 + lane];
                acc = M::step(f.x, x[t].x, acc);
                acc = M::step(f.y, x[t].y, acc);
            }
            const uint32_t b0 = 8u * static_cast<uint32_t>(mt) + 2u * c;
            A y0, y1;
            y0.x = acc[0];
            y0.y = acc[1];
            y1.x = acc[2];
            y1.y = acc[3];
            at(s, base ^ (swz(local_offset<K>(b0, op)) << sh)) = y0;
            at(s, base ^ (swz(local_offset<K>(b0 + 1u, op)) << sh)) = y1;
        }
    }
}

// k = 4, 5: the complex product as three real ones (Gauss):
// P1 = Ur·xr, P2 = Ui·xi, P3 = (Ur + Ui)·(xr + xi); y = (P1 − P2) + i(P3 − P1 − P2).
// Real 16-row M-tiles over output amplitudes (row i of tile mt is amplitude
// 16mt + i), K-steps over input amplitudes (step s, column c is amplitude
// 4s + c, so lane group c feeds amplitudes c, c + 4, ...): D/16 tiles x D/4
// steps x 3 products = 3/4 of the MFMAs of the realified form.  Fragments:
// per tile and pair of steps, (Ur, Ui, Ur + Ui) as three pairs per lane at
// [((mt * D/8 + t) * 3 + p) * 64 + lane]; with PAIRS2 (complex128 fragments
// staged in LDS) only (Ur, Ui) at [((mt * D/8 + t) * 2 + p) * 64 + lane], the
// sum formed here with the additions the host makes.
template <typename A, int K, bool PAIRS2>
__device__ __forceinline__ void dense_mfma3(A* s, uint32_t zoff, int G, const OpRegs& op, const A* __restrict__ F) {
    using M = Mfma<A>;
    constexpr int sh = Amp<A>::kShift;
    constexpr int D = 1 << K, NT = D / 16, H = D / 4, HP = H / 2;
    const uint32_t lane = threadIdx.x & 63u, j = lane & 15u, c = lane >> 4;
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform: chunk loop in SALU
    uint32_t lp = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) lp |= ((j >> b) & 1u) << op.g(b);
    lp = (swz(lp) << sh) | zoff;
    uint32_t so_in[H];
#pragma unroll
    for (int t = 0; t < H; ++t) so_in[t] = swz(local_offset<K>(c + 4u * t, op)) << sh;
    const uint32_t chunks = 1u << (G - 4);
    for (uint32_t ch = w; ch < chunks; ch += kWaves) {
        uint32_t up = 0;
        for (int b = 4; b < G; ++b) up |= ((ch >> (b - 4)) & 1u) << op.g(b);
        const uint32_t base = lp ^ (swz(up) << sh);
        A x[H];
#pragma unroll
        for (int t = 0; t < H; ++t) x[t] = at(s, base ^ so_in[t]);
#pragma unroll(K < 5 ? NT : 1)
        for (int mt = 0; mt < NT; ++mt) {
            typename M::v4 p1 = {0, 0, 0, 0}, p2 = p1, p3 = p1;
#pragma unroll
            for (int t = 0; t < HP; ++t) {
                A fr, fi, fs;
                if constexpr (PAIRS2) {
                    const A* f = F + ((mt * HP + t) * 2) * 64 + lane;
                    fr = f[0];
                    fi = f[64];
                    fs.x = fr.x + fi.x;
                    fs.y = fr.y + fi.y;
                } else {
                    const A* f = F + ((mt * HP + t) * 3) * 64 + lane;
                    fr = f[0];
                    fi = f[64];
                    fs = f[128];
                }
                const A x0 = x[2 * t], x1 = x[2 * t + 1];
                p1 = M::step(fr.x, x0.x, p1);
                p2 = M::step(fi.x, x0.y, p2);
                p3 = M::step(fs.x, x0.x + x0.y, p3);
                p1 = M::step(fr.y, x1.x, p1);
                p2 = M::step(fi.y, x1.y, p2);
                p3 = M::step(fs.y, x1.x + x1.y, p3);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t b = 16u * static_cast<uint32_t>(mt) + (M::kGroupRows ? 4u * c + r : 4u * r + c);
                A y;
                y.x = p1[r] - p2[r];
                y.y = p3[r] - p1[r] - p2[r];
                at(s, base ^ (swz(local_offset<K>(b, op)) << sh)) = y;
            }
        }
    }
}

// Chunk walk of a matrix-core op (chunks of 16 groups, lane % 16 = group-
// number bits 0-3): wave w takes the chunks whose group-number bits 4 and 5
// spell w (with G < 6 the waves beyond the chunk count sit out) and visits
// the deposits of the remaining bits (local positions in m) as
// hi -> (hi - m) & m from 0, which meets every subset of m once.  base is
// this lane's byte offset at hi = 0; n the wave's chunk count.
struct ChunkWalk {
    uint32_t base, m, n;
};
template <typename A>
__device__ __forceinline__ ChunkWalk chunk_walk(uint32_t zoff, int G, const OpRegs& op) {
    constexpr int sh = Amp<A>::kShift;
    const uint32_t j = threadIdx.x & 15u;
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int cb = G - 4, wb = cb < 2 ? cb : 2;
    ChunkWalk cw;
    cw.n = w < (1u << wb) ? 1u << (cb - wb) : 0u;
    uint32_t lp = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) lp |= ((j >> b) & 1u) << op.g(b);
    for (int b = 0; b < wb; ++b) lp |= ((w >> b) & 1u) << op.g(4 + b);
    cw.base = (swz(lp) << sh) | zoff;
    uint32_t m = 0;
    for (int b = 4 + wb; b < G; ++b) m |= 1u << op.g(b);
    cw.m = m;
    return cw;
}

// k = 3 (realified) and k = 4 (three products) software-pipelined over a
// wave's chunks: the products of chunk i are issued before the results of
// chunk i - 1 are combined and written, so the matrix cores work while the
// VALU and LDS finish the previous chunk, and the reads of chunk i overlap
// the products of chunk i - 1.  Chunk i's reads come before chunk i - 1's
// writes; the chunks hold disjoint groups.  Same products, same order of
// accumulation per output as dense_mfma / dense_mfma3.
template <typename A>
__device__ __forceinline__ void dense_mfma_k3p(A* s, uint32_t zoff, int G, const OpRegs& op, const A* __restrict__ F) {
    using M = Mfma<A>;
    using V = typename M::v4;
    constexpr int sh = Amp<A>::kShift;
    constexpr int K = 3, H = 2;
    const uint32_t lane = threadIdx.x & 63u, c = lane >> 4;
    const ChunkWalk cw = chunk_walk<A>(zoff, G, op);
    if (cw.n == 0) return;
    uint32_t so_in[H];
#pragma unroll
    for (int t = 0; t < H; ++t) so_in[t] = swz(local_offset<K>(c + 4u * t, op)) << sh;
    const uint32_t so0 = swz(local_offset<K>(2u * c, op)) << sh, so1 = swz(local_offset<K>(2u * c + 1u, op)) << sh;
    A f[H];
#pragma unroll
    for (int t = 0; t < H; ++t) f[t] = F[t * 64 + lane];
    auto issue = [&](const A* x) {
        V acc = {0, 0, 0, 0};
#pragma unroll
        for (int t = 0; t < H; ++t) {
            acc = M::step(f[t].x, x[t].x, acc);
            acc = M::step(f[t].y, x[t].y, acc);
        }
        return acc;
    };
    auto finish = [&](uint32_t base, const V& acc) {
        A y0, y1;
        y0.x = acc[0];
        y0.y = acc[1];
        y1.x = acc[2];
        y1.y = acc[3];
        at(s, base ^ so0) = y0;
        at(s, base ^ so1) = y1;
    };
    // Two accumulator sets in turn (a, b): no copy of a result that is
    // still in flight.
    uint32_t hi = 0, ba = cw.base, bb = 0;
    A x[H];
    auto load = [&](uint32_t base) {
#pragma unroll
        for (int t = 0; t < H; ++t) x[t] = at(s, base ^ so_in[t]);
    };
    auto next = [&]() {
        hi = (hi - cw.m) & cw.m;
        return cw.base ^ (swz(hi) << sh);
    };
    load(ba);
    V a = issue(x), b;
    uint32_t i = 1;
    for (; i + 1 < cw.n; i += 2) {
        bb = next();
        load(bb);
        b = issue(x);
        finish(ba, a);
        ba = next();
        load(ba);
        a = issue(x);
        finish(bb, b);
    }
    if (i < cw.n) {
        bb = next();
        load(bb);
        b = issue(x);
        finish(ba, a);
        finish(bb, b);
    } else {
        finish(ba, a);
    }
}

template <typename A, bool PAIRS2>
__device__ __forceinline__ void dense_mfma_k4p(A* s, uint32_t zoff, int G, const OpRegs& op, const A* __restrict__ F) {
    using M = Mfma<A>;
    using V = typename M::v4;
    constexpr int sh = Amp<A>::kShift;
    constexpr int K = 4, H = 4, HP = 2;
    const uint32_t lane = threadIdx.x & 63u, c = lane >> 4;
    const ChunkWalk cw = chunk_walk<A>(zoff, G, op);
    if (cw.n == 0) return;
    uint32_t so_in[H], so_out[4];
#pragma unroll
    for (int t = 0; t < H; ++t) so_in[t] = swz(local_offset<K>(c + 4u * t, op)) << sh;
#pragma unroll
    for (int r = 0; r < 4; ++r)
        so_out[r] = swz(local_offset<K>(M::kGroupRows ? 4u * c + r : 4u * r + c, op)) << sh;
    A fr[HP], fi[HP], fs[HP];
#pragma unroll
    for (int t = 0; t < HP; ++t) {
        if constexpr (PAIRS2) {
            const A* f = F + (t * 2) * 64 + lane;
            fr[t] = f[0];
            fi[t] = f[64];
            fs[t].x = fr[t].x + fi[t].x;
            fs[t].y = fr[t].y + fi[t].y;
        } else {
            const A* f = F + (t * 3) * 64 + lane;
            fr[t] = f[0];
            fi[t] = f[64];
            fs[t] = f[128];
        }
    }
    struct P3 {
        V p1, p2, p3;
    };
    auto issue = [&](const A* x) {
        P3 a;
        a.p1 = V{0, 0, 0, 0};
        a.p2 = a.p1;
        a.p3 = a.p1;
#pragma unroll
        for (int t = 0; t < HP; ++t) {
            const A x0 = x[2 * t], x1 = x[2 * t + 1];
            a.p1 = M::step(fr[t].x, x0.x, a.p1);
            a.p2 = M::step(fi[t].x, x0.y, a.p2);
            a.p3 = M::step(fs[t].x, x0.x + x0.y, a.p3);
            a.p1 = M::step(fr[t].y, x1.x, a.p1);
            a.p2 = M::step(fi[t].y, x1.y, a.p2);
            a.p3 = M::step(fs[t].y, x1.x + x1.y, a.p3);
        }
        return a;
    };
    auto finish = [&](uint32_t base, const P3& a) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            A y;
            y.x = a.p1[r] - a.p2[r];
            y.y = a.p3[r] - a.p1[r] - a.p2[r];
            at(s, base ^ so_out[r]) = y;
        }
    };
    uint32_t hi = 0, ba = cw.base, bb = 0;
    A x[H];
    auto load = [&](uint32_t base) {
#pragma unroll
        for (int t = 0; t < H; ++t) x[t] = at(s, base ^ so_in[t]);
    };
    auto next = [&]() {
        hi = (hi - cw.m) & cw.m;
        return cw.base ^ (swz(hi) << sh);
    };
    load(ba);
    P3 a = issue(x), b;
    uint32_t i = 1;
    for (; i + 1 < cw.n; i += 2) {
        bb = next();
        load(bb);
        b = issue(x);
        finish(ba, a);
        ba = next();
        load(ba);
        a = issue(x);
        finish(bb, b);
    }
    if (i < cw.n) {
        bb = next();
        load(bb);
        b = issue(x);
        finish(ba, a);
        finish(bb, b);
    } else {
        finish(ba, a);
    }
}

// STAGED: the fragments come from the streaming kernel's LDS copy.
template <typename A, bool STAGED>
__device__ __forceinline__ void mfma_op(A* s, uint32_t zoff, int G, const OpRegs& op, const A* __restrict__ F) {
    constexpr bool kPairs2 = STAGED && std::is_same<A, double2>::value;
    switch (op.k) {
        case 3: dense_mfma_k3p<A>(s, zoff, G, op, F); break;
        case 4: dense_mfma_k4p<A, kPairs2>(s, zoff, G, op, F); break;
        case 5: dense_mfma3<A, 5, kPairs2>(s, zoff, G, op, F); break;
        default: break;
    }
}

// Register blocks: a thread holds the 8 amplitudes of one group of the
// block's three tile positions (v[x], x = block bits) while the member ops
// act on them.
template <typename A, int B>   // 2x2 matrix on block bit B
__device__ __forceinline__ void sub_single(A* v, const A* __restrict__ U) {
    const A u00 = U[0], u01 = U[1], u10 = U[2], u11 = U[3];
#pragma unroll
    for (int x = 0; x < 8; ++x)
        if (!((x >> B) & 1)) {
            const A a = v[x], b = v[x | (1 << B)];
            v[x] = cmad(cmul(u00, a), u01, b);
            v[x | (1 << B)] = cmad(cmul(u10, a), u11, b);
        }
}

template <typename A, int B0, int B1>   // 4x4 matrix, operand 0 on block bit B0 < B1
__device__ __forceinline__ void sub_pair(A* v, const A* __restrict__ U) {
    constexpr int C = 3 - B0 - B1;   // the block bit the op does not touch
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int i[4] = {r << C, (r << C) | (1 << B0), (r << C) | (1 << B1), (r << C) | (1 << B0) | (1 << B1)};
        const A in[4] = {v[i[0]], v[i[1]], v[i[2]], v[i[3]]};
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            A acc;
            acc.x = 0;
            acc.y = 0;
#pragma unroll
            for (int x = 0; x < 4; ++x) acc = cmad(acc, U[y * 4 + x], in[x]);
            v[i[y]] = acc;
        }
    }
}

template <typename A, int B0, int B1>   // diagonal on block bits B0 < B1 (B1 < 0: one bit)
__device__ __forceinline__ void sub_diag(A* v, const A* __restrict__ Dg) {
    constexpr int N = B1 < 0 ? 2 : 4;
    A d[N];
#pragma unroll
    for (int e = 0; e < N; ++e) d[e] = Dg[e];
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        const int e = ((x >> B0) & 1) | (B1 < 0 ? 0 : ((x >> B1) & 1) << 1);
        v[x] = cmul(v[x], d[e]);
    }
}

template <typename A, int B>   // real 2x2 matrix (zero imaginary parts) on block bit B
__device__ __forceinline__ void sub_real(A* v, const A* __restrict__ U) {
    const auto u00 = U[0].x, u01 = U[1].x, u10 = U[2].x, u11 = U[3].x;
#pragma unroll
    for (int x = 0; x < 8; ++x)
        if (!((x >> B) & 1)) {
            const A a = v[x], b = v[x | (1 << B)];
            v[x].x = fma(u01, b.x, u00 * a.x);
            v[x].y = fma(u01, b.y, u00 * a.y);
            v[x | (1 << B)].x = fma(u11, b.x, u10 * a.x);
            v[x | (1 << B)].y = fma(u11, b.y, u10 * a.y);
        }
}

template <typename A, int B>   // [[a, i b], [i c, d]], a..d real, on block bit B
__device__ __forceinline__ void sub_imag(A* v, const A* __restrict__ U) {
    const auto a = U[0].x, b = U[1].y, c = U[2].y, d = U[3].x;
#pragma unroll
    for (int x = 0; x < 8; ++x)
        if (!((x >> B) & 1)) {
            const A p = v[x], q = v[x | (1 << B)];
            v[x].x = fma(-b, q.y, a * p.x);
            v[x].y = fma(b, q.x, a * p.y);
            v[x | (1 << B)].x = fma(-c, p.y, d * q.x);
            v[x | (1 << B)].y = fma(c, p.x, d * q.y);
        }
}

template <typename A, int B0, int B1>   // -1 where block bits B0 and B1 are both set
__device__ __forceinline__ void sub_negate(A* v) {
#pragma unroll
    for (int x = 0; x < 8; ++x)
        if (((x >> B0) & 1) && ((x >> B1) & 1)) {
            v[x].x = -v[x].x;
            v[x].y = -v[x].y;
        }
}

template <typename A>
__device__ __forceinline__ void sub_apply(A* v, int pattern, const A* __restrict__ U) {
    switch (pattern) {
        case kReal2 + 0: sub_real<A, 0>(v, U); break;
        case kReal2 + 1: sub_real<A, 1>(v, U); break;
        case kReal2 + 2: sub_real<A, 2>(v, U); break;
        case kImag2 + 0: sub_imag<A, 0>(v, U); break;
        case kImag2 + 1: sub_imag<A, 1>(v, U); break;
        case kImag2 + 2: sub_imag<A, 2>(v, U); break;
        case kNegPair + 0: sub_negate<A, 0, 1>(v); break;
        case kNegPair + 1: sub_negate<A, 0, 2>(v); break;
        case kNegPair + 2: sub_negate<A, 1, 2>(v); break;
        case 0: sub_single<A, 0>(v, U); break;
        case 1: sub_single<A, 1>(v, U); break;
        case 2: sub_single<A, 2>(v, U); break;
        case 3: sub_pair<A, 0, 1>(v, U); break;
        case 4: sub_pair<A, 0, 2>(v, U); break;
        case 5: sub_pair<A, 1, 2>(v, U); break;
        case 6: sub_diag<A, 0, -1>(v, U); break;
        case 7: sub_diag<A, 1, -1>(v, U); break;
        case 8: sub_diag<A, 2, -1>(v, U); break;
        case 9: sub_diag<A, 0, 1>(v, U); break;
        case 10: sub_diag<A, 0, 2>(v, U); break;
        case 11: sub_diag<A, 1, 2>(v, U); break;
        default: {   // 12: no operand on the block
            const A d = U[0];
#pragma unroll
            for (int x = 0; x < 8; ++x) v[x] = cmul(v[x], d);
            break;
        }
    }
}

// One member on both states of a two-state block (psi in v, phi in f) under
// one dispatch: the member's coefficients are read once, and the two states'
// products are independent chains the compiler can interleave.
template <typename A>
__device__ __forceinline__ void sub_apply2(A* v, A* f, int pattern, const A* __restrict__ U) {
    switch (pattern) {
        case kReal2 + 0: sub_real<A, 0>(v, U); sub_real<A, 0>(f, U); break;
        case kReal2 + 1: sub_real<A, 1>(v, U); sub_real<A, 1>(f, U); break;
        case kReal2 + 2: sub_real<A, 2>(v, U); sub_real<A, 2>(f, U); break;
        case kImag2 + 0: sub_imag<A, 0>(v, U); sub_imag<A, 0>(f, U); break;
        case kImag2 + 1: sub_imag<A, 1>(v, U); sub_imag<A, 1>(f, U); break;
        case kImag2 + 2: sub_imag<A, 2>(v, U); sub_imag<A, 2>(f, U); break;
        case kNegPair + 0: sub_negate<A, 0, 1>(v); sub_negate<A, 0, 1>(f); break;
        case kNegPair + 1: sub_negate<A, 0, 2>(v); sub_negate<A, 0, 2>(f); break;
        case kNegPair + 2: sub_negate<A, 1, 2>(v); sub_negate<A, 1, 2>(f); break;
        case 0: sub_single<A, 0>(v, U); sub_single<A, 0>(f, U); break;
        case 1: sub_single<A, 1>(v, U); sub_single<A, 1>(f, U); break;
        case 2: sub_single<A, 2>(v, U); sub_single<A, 2>(f, U); break;
        case 3: sub_pair<A, 0, 1>(v, U); sub_pair<A, 0, 1>(f, U); break;
        case 4: sub_pair<A, 0, 2>(v, U); sub_pair<A, 0, 2>(f, U); break;
        case 5: sub_pair<A, 1, 2>(v, U); sub_pair<A, 1, 2>(f, U); break;
        case 6: sub_diag<A, 0, -1>(v, U); sub_diag<A, 0, -1>(f, U); break;
        case 7: sub_diag<A, 1, -1>(v, U); sub_diag<A, 1, -1>(f, U); break;
        case 8: sub_diag<A, 2, -1>(v, U); sub_diag<A, 2, -1>(f, U); break;
        case 9: sub_diag<A, 0, 1>(v, U); sub_diag<A, 0, 1>(f, U); break;
        case 10: sub_diag<A, 0, 2>(v, U); sub_diag<A, 0, 2>(f, U); break;
        case 11: sub_diag<A, 1, 2>(v, U); sub_diag<A, 1, 2>(f, U); break;
        default: {   // 12: no operand on the block
            const A d = U[0];
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                v[x] = cmul(v[x], d);
                f[x] = cmul(f[x], d);
            }
            break;
        }
    }
}


// One OP_BLOCK over one state's tile: load each group's 8 amplitudes, run
// the nsub OP_SUB entries that follow the header, store.  PAIRS: a thread
// with two or more groups takes them two at a time.
template <typename A, bool PAIRS>
__device__ __forceinline__ void block_apply(A* s, uint32_t tp, int G, const OpRegs& hdr,
                                            const DevOp* __restrict__ subs, int nsub, const A* __restrict__ coefs,
                                            uint64_t gbase) {
    constexpr int sh = Amp<A>::kShift;
    if (G < 8 && threadIdx.x >= (1u << G)) return;
    uint32_t so[8];
#pragma unroll
    for (int x = 0; x < 8; ++x) so[x] = swz(local_offset<3>(x, hdr)) << sh;
    const uint32_t hm = static_cast<uint32_t>(hdr.gmask);
    const uint32_t iters = G > 8 ? 1u << (G - 8) : 1u;
    uint32_t hi = 0;
    if (PAIRS && iters >= 2) {
        // Two groups at a time: each member is dispatched once for both, and
        // the two groups' products are independent chains.
        for (uint32_t i = 0; i < iters; i += 2) {
            const uint32_t h0 = swz(hi) << sh;
            hi = (hi - hm) & hm;
            const uint32_t h1 = swz(hi) << sh;
            hi = (hi - hm) & hm;
            uint32_t a[8], b[8];
            A v[8], w[8];
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                a[x] = tp ^ (h0 ^ so[x]);
                b[x] = tp ^ (h1 ^ so[x]);
                v[x] = at(s, a[x]);
                w[x] = at(s, b[x]);
            }
            OpRegs nxt = load_op(subs);
            for (int j = 0; j < nsub; ++j) {
                const OpRegs sub = nxt;
                if (j + 1 < nsub) nxt = load_op(subs + j + 1);
                const A* U = coefs + sub.coef;
                if (sub.k >= 6 && sub.k <= 12) U += diag_sub_table(gbase, sub) << member_local_bits(sub.k);
                sub_apply2(v, w, sub.k, U);
            }
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                at(s, a[x]) = v[x];
                at(s, b[x]) = w[x];
            }
        }
        return;
    }
    for (uint32_t i = 0; i < iters; ++i) {
        const uint32_t h = swz(hi) << sh;
        hi = (hi - hm) & hm;
        uint32_t a[8];
        A v[8];
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            a[x] = tp ^ (h ^ so[x]);
            v[x] = at(s, a[x]);
        }
        OpRegs nxt = load_op(subs);   // descriptors one ahead, as in block_apply2
        for (int j = 0; j < nsub; ++j) {
            const OpRegs sub = nxt;
            if (j + 1 < nsub) nxt = load_op(subs + j + 1);
            const A* U = coefs + sub.coef;
            if (sub.k >= 6 && sub.k <= 12) U += diag_sub_table(gbase, sub) << member_local_bits(sub.k);
            sub_apply(v, sub.k, U);
        }
#pragma unroll
        for (int x = 0; x < 8; ++x) at(s, a[x]) = v[x];
    }
}

// Whether a one-circuit kernel's single-state blocks take their groups two
// at a time: only tiles of 2^12 amplitudes or more give a thread two groups,
// and the batched and streaming kernels have no registers to spare for the
// second group's eight amplitudes.
constexpr bool block_pairs(int tb) { return tb >= 12; }

// Grouped diagonal over one state's tile: each thread scales its groups'
// 2^K amplitudes by the entries D[x] (wave-uniform, through the scalar cache).
template <typename A, int K>
__device__ __forceinline__ void diag_groups(A* s, uint32_t tp, int G, const OpRegs& op, const A* __restrict__ Dg) {
    constexpr int D = 1 << K;
    constexpr int sh = Amp<A>::kShift;
    if (G < 8 && threadIdx.x >= (1u << G)) return;
    uint32_t so[D];
    A d[D];
#pragma unroll
    for (int x = 0; x < D; ++x) {
        so[x] = swz(local_offset<K>(x, op)) << sh;
        d[x] = Dg[x];
    }
    const uint32_t hm = static_cast<uint32_t>(op.gmask);
    const uint32_t iters = G > 8 ? 1u << (G - 8) : 1u;
    uint32_t hi = 0;
    for (uint32_t i = 0; i < iters; ++i) {
        const uint32_t h = swz(hi) << sh;
        hi = (hi - hm) & hm;
        uint32_t a[D];
        A v[D];
#pragma unroll
        for (int x = 0; x < D; ++x) {
            a[x] = tp ^ (h ^ so[x]);
            v[x] = at(s, a[x]);
        }
#pragma unroll
        for (int x = 0; x < D; ++x) at(s, a[x]) = cmul(v[x], d[x]);
    }
}

// <phi| D |psi> over one state pair's tile for a grouped diagonal (entries
// D[x] for local operand bits x); tp0 / tp1 address psi's and phi's tiles.
template <typename A, int K>
__device__ __forceinline__ double2 ip_diag_groups(A* s, uint32_t tp0, uint32_t tp1, int G, const OpRegs& op,
                                                  const A* __restrict__ Dg) {
    constexpr int D = 1 << K;
    constexpr int sh = Amp<A>::kShift;
    double2 acc = make_double2(0.0, 0.0);
    if (G < 8 && threadIdx.x >= (1u << G)) return acc;
    uint32_t so[D];
    double2 d[D];
#pragma unroll
    for (int x = 0; x < D; ++x) {
        so[x] = swz(local_offset<K>(x, op)) << sh;
        d[x] = make_double2(Dg[x].x, Dg[x].y);
    }
    const uint32_t hm = static_cast<uint32_t>(op.gmask);
    const uint32_t iters = G > 8 ? 1u << (G - 8) : 1u;
    uint32_t hi = 0;
    for (uint32_t i = 0; i < iters; ++i) {
        const uint32_t h = swz(hi) << sh;
        hi = (hi - hm) & hm;
#pragma unroll
        for (int x = 0; x < D; ++x) {
            const A xv = at(s, tp0 ^ (h ^ so[x])), f = at(s, tp1 ^ (h ^ so[x]));
            const double gr = d[x].x * xv.x - d[x].y * xv.y, gi = d[x].x * xv.y + d[x].y * xv.x;
            acc.x += static_cast<double>(f.x) * gr + static_cast<double>(f.y) * gi;
            acc.y += static_cast<double>(f.x) * gi - static_cast<double>(f.y) * gr;
        }
    }
    return acc;
}

// <f| G |v> over a block's 8 amplitudes for a generator on block bit B
// (2x2) or on block bits B0 < B1 (4x4, operand 0 on B0); FP64 sums.
template <typename A>
__device__ __forceinline__ void ip_acc(double2& acc, const A& f, const A& g) {
    acc.x += static_cast<double>(f.x) * g.x + static_cast<double>(f.y) * g.y;
    acc.y += static_cast<double>(f.x) * g.y - static_cast<double>(f.y) * g.x;
}

template <typename A, int B>
__device__ __forceinline__ double2 ip_single(const A* v, const A* f, const A* __restrict__ U) {
    const A u00 = U[0], u01 = U[1], u10 = U[2], u11 = U[3];
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int x = 0; x < 8; ++x)
        if (!((x >> B) & 1)) {
            const A a = v[x], b = v[x | (1 << B)];
            ip_acc(acc, f[x], cmad(cmul(u00, a), u01, b));
            ip_acc(acc, f[x | (1 << B)], cmad(cmul(u10, a), u11, b));
        }
    return acc;
}

template <typename A, int B0, int B1>
__device__ __forceinline__ double2 ip_pair(const A* v, const A* f, const A* __restrict__ U) {
    constexpr int C = 3 - B0 - B1;
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int i[4] = {r << C, (r << C) | (1 << B0), (r << C) | (1 << B1), (r << C) | (1 << B0) | (1 << B1)};
#pragma unroll
        for (int y = 0; y < 4; ++y) {
            A g;
            g.x = 0;
            g.y = 0;
#pragma unroll
            for (int x = 0; x < 4; ++x) g = cmad(g, U[y * 4 + x], v[i[x]]);
            ip_acc(acc, f[i[y]], g);
        }
    }
    return acc;
}

template <typename A, int B0, int B1>   // diagonal generator on block bits B0 < B1 (B1 < 0: one bit)
__device__ __forceinline__ double2 ip_diag(const A* v, const A* f, const A* __restrict__ Dg) {
    constexpr int N = B1 < 0 ? 2 : 4;
    A d[N];
#pragma unroll
    for (int e = 0; e < N; ++e) d[e] = Dg[e];
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        const int e = ((x >> B0) & 1) | (B1 < 0 ? 0 : ((x >> B1) & 1) << 1);
        ip_acc(acc, f[x], cmul(d[e], v[x]));
    }
    return acc;
}

// <f| P |v> for a Pauli generator P on block bit B: X swaps the pair, Y
// swaps it with -i (upper) / +i (lower), Z negates the upper amplitude.
template <typename A, int B>
__device__ __forceinline__ double2 ip_pauli_x(const A* v, const A* f) {
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int x = 0; x < 8; ++x) ip_acc(acc, f[x], v[x ^ (1 << B)]);
    return acc;
}
template <typename A, int B>
__device__ __forceinline__ double2 ip_pauli_y(const A* v, const A* f) {
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        const A w = v[x ^ (1 << B)];
        A g;
        if ((x >> B) & 1) {   // i w
            g.x = -w.y;
            g.y = w.x;
        } else {              // -i w
            g.x = w.y;
            g.y = -w.x;
        }
        ip_acc(acc, f[x], g);
    }
    return acc;
}
template <typename A, int B>
__device__ __forceinline__ double2 ip_pauli_z(const A* v, const A* f) {
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        A g = v[x];
        if ((x >> B) & 1) {
            g.x = -g.x;
            g.y = -g.y;
        }
        ip_acc(acc, f[x], g);
    }
    return acc;
}

// One OP_BLOCK over both states of a two-state pass: a thread holds a
// group's 8 amplitudes of psi (v) and phi (f); OP_SUB members act on both,
// OP_IPSUB members add <phi| G |psi> to their accumulator.  The header's
// spos[7] counts the OP_IPSUB members; their per-wave sums go to wave_ip
// slots n_ip, n_ip + 1, ... in member order.  Returns that count.
template <typename A, bool JUMP>
__device__ __forceinline__ int block_apply2(A* s, uint32_t tp, uint32_t kPhi, int G, const OpRegs& hdr,
                                            const DevOp* __restrict__ subs, int nsub, const A* __restrict__ coefs,
                                            double2* wave_ip, int n_ip, uint64_t gbase) {
    constexpr int sh = Amp<A>::kShift;
    const int nipb = static_cast<int>((hdr.spos >> 56) & 0xffu);
    double2 acc[kMaxBlockIp];
#pragma unroll
    for (int t = 0; t < kMaxBlockIp; ++t) acc[t] = make_double2(0.0, 0.0);
    if (!(G < 8 && threadIdx.x >= (1u << G))) {
        uint32_t so[8];
#pragma unroll
        for (int x = 0; x < 8; ++x) so[x] = swz(local_offset<3>(x, hdr)) << sh;
        const uint32_t hm = static_cast<uint32_t>(hdr.gmask);
        const uint32_t iters = G > 8 ? 1u << (G - 8) : 1u;
        uint32_t hi = 0;
        for (uint32_t i = 0; i < iters; ++i) {
            const uint32_t h = swz(hi) << sh;
            hi = (hi - hm) & hm;
            uint32_t a[8];
            A v[8], f[8];
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                a[x] = tp ^ (h ^ so[x]);
                v[x] = at(s, a[x]);
                f[x] = at(s, a[x] | kPhi);
            }
            int j_ip = 0;
            // Member descriptors one ahead: the next one's scalar load is in
            // flight with this one's coefficient loads (one round trip each).
            OpRegs nxt = load_op(subs);
            for (int j = 0; j < nsub; ++j) {
                const OpRegs sub = nxt;
                if (j + 1 < nsub) nxt = load_op(subs + j + 1);
                const A* U = coefs + sub.coef;
                if (sub.k >= 6 && sub.k <= 12) U += diag_sub_table(gbase, sub) << member_local_bits(sub.k);
                if (sub.kind == OP_IPSUB) {
                    double2 r;
                    switch (sub.k) {
                        case kXGen + 0: r = ip_pauli_x<A, 0>(v, f); break;
                        case kXGen + 1: r = ip_pauli_x<A, 1>(v, f); break;
                        case kXGen + 2: r = ip_pauli_x<A, 2>(v, f); break;
                        case kYGen + 0: r = ip_pauli_y<A, 0>(v, f); break;
                        case kYGen + 1: r = ip_pauli_y<A, 1>(v, f); break;
                        case kYGen + 2: r = ip_pauli_y<A, 2>(v, f); break;
                        case kZGen + 0: r = ip_pauli_z<A, 0>(v, f); break;
                        case kZGen + 1: r = ip_pauli_z<A, 1>(v, f); break;
                        case kZGen + 2: r = ip_pauli_z<A, 2>(v, f); break;
                        case 0: r = ip_single<A, 0>(v, f, U); break;
                        case 1: r = ip_single<A, 1>(v, f, U); break;
                        case 2: r = ip_single<A, 2>(v, f, U); break;
                        case 3: r = ip_pair<A, 0, 1>(v, f, U); break;
                        case 4: r = ip_pair<A, 0, 2>(v, f, U); break;
                        case 5: r = ip_pair<A, 1, 2>(v, f, U); break;
                        case 6: r = ip_diag<A, 0, -1>(v, f, U); break;
                        case 7: r = ip_diag<A, 1, -1>(v, f, U); break;
                        case 8: r = ip_diag<A, 2, -1>(v, f, U); break;
                        case 9: r = ip_diag<A, 0, 1>(v, f, U); break;
                        case 10: r = ip_diag<A, 0, 2>(v, f, U); break;
                        case 11: r = ip_diag<A, 1, 2>(v, f, U); break;
                        default: {   // 12: no operand on the block
                            const A d = U[0];
                            r = make_double2(0.0, 0.0);
#pragma unroll
                            for (int x = 0; x < 8; ++x) ip_acc(r, f[x], cmul(d, v[x]));
                            break;
                        }
                    }
                    // j_ip is wave-uniform: a scalar branch, no register
                    // indexing.  JUMP: one switch (a three-level branch
                    // tree); otherwise a compare per accumulator, which the
                    // batched kernel, short of registers, keeps (the switch
                    // cost it 9 % there while saving 4 % in the one-circuit
                    // kernel).
                    static_assert(kMaxBlockIp == 8, "one case per accumulator");
                    if constexpr (!JUMP) {
#pragma unroll
                        for (int t = 0; t < kMaxBlockIp; ++t)
                            if (t == j_ip) {
                                acc[t].x += r.x;
                                acc[t].y += r.y;
                            }
                    } else switch (j_ip) {
#define QHIP_ACC(T)        \
    case T:                \
        acc[T].x += r.x;   \
        acc[T].y += r.y;   \
        break;
                        QHIP_ACC(0) QHIP_ACC(1) QHIP_ACC(2) QHIP_ACC(3) QHIP_ACC(4) QHIP_ACC(5) QHIP_ACC(6) QHIP_ACC(7)
#undef QHIP_ACC
                        default: break;
                    }
                    ++j_ip;
                    continue;
                }
                sub_apply2(v, f, sub.k, U);
            }
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                at(s, a[x]) = v[x];
                at(s, a[x] | kPhi) = f[x];
            }
        }
    }
#pragma unroll
    for (int t = 0; t < kMaxBlockIp; ++t)
        if (t < nipb) {
            const double2 r = wave_sum(acc[t]);
            if ((threadIdx.x & 63) == 0) wave_ip[(n_ip + t) * kWaves + (threadIdx.x >> 6)] = r;
        }
    return nipb;
}

__device__ __forceinline__ uint32_t diag_index(uint32_t l, uint64_t gbase, const OpRegs& op, int k) {
    uint32_t idx = 0;
    for (int b = 0; b < k; ++b) {
        const int p = op.p(b);
        const uint32_t bit = p < 32 ? (l >> p) & 1u : static_cast<uint32_t>((gbase >> (p - 32)) & 1ull);
        idx |= bit << b;
    }
    return idx;
}

// <phi| G |psi> over the tile for a dense generator on K <= 2 qubits; tp0 /
// tp1 are the thread's byte offsets into psi's and phi's tiles.
template <typename A, int K>
__device__ __forceinline__ double2 ip_dense(A* s, uint32_t tp0, uint32_t tp1, int G, const OpRegs& op,
                                            const A* __restrict__ Gm) {
    constexpr int D = 1 << K;
    constexpr int sh = Amp<A>::kShift;
    double2 acc = make_double2(0.0, 0.0);
    if (G < 8 && threadIdx.x >= (1u << G)) return acc;
    uint32_t so[D];
#pragma unroll
    for (int x = 0; x < D; ++x) so[x] = swz(local_offset<K>(x, op)) << sh;
    const uint32_t hm = static_cast<uint32_t>(op.gmask);
    const uint32_t iters = G > 8 ? 1u << (G - 8) : 1u;
    uint32_t hi = 0;
    for (uint32_t i = 0; i < iters; ++i) {
        const uint32_t h = swz(hi) << sh;
        hi = (hi - hm) & hm;
        A in[D];
#pragma unroll
        for (int x = 0; x < D; ++x) in[x] = at(s, tp0 ^ (h ^ so[x]));
#pragma unroll
        for (int y = 0; y < D; ++y) {
            A gy;
            gy.x = 0;
            gy.y = 0;
#pragma unroll
            for (int x = 0; x < D; ++x) gy = cmad(gy, Gm[y * D + x], in[x]);
            const A f = at(s, tp1 ^ (h ^ so[y]));
            const double fr = f.x, fi = f.y, gr = gy.x, gi = gy.y;   // conj(f) * gy
            acc.x += fr * gr + fi * gi;
            acc.y += fr * gi - fi * gr;
        }
    }
    return acc;
}

// e or conj(e) (FP64), by a parity bit.
__device__ __forceinline__ double2 unit_or_conj(double2 e, uint64_t odd) {
    return make_double2(e.x, odd ? -e.y : e.y);
}
__device__ __forceinline__ double2 mul_unit(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// f times, over cnt terms of a merged phase op (two records each: exp(i w),
// then {operand code, 0}), exp(i w) or its conjugate as odd(code) says.
// Four terms' records are loaded before their products, so the loads are in
// flight together, and the products run in four independent chains joined
// at the end (a single chain of dependent FP64 products is latency-bound).
template <typename A, typename Odd>
__device__ __forceinline__ double2 term_product(double2 f, const A*& r, uint32_t cnt, Odd odd) {
    double2 p[4] = {f, make_double2(1.0, 0.0), make_double2(1.0, 0.0), make_double2(1.0, 0.0)};
    uint32_t i = 0;
    for (; i + 4 <= cnt; i += 4, r += 8) {
        A e[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) e[t] = r[t];
#pragma unroll
        for (int t = 0; t < 4; ++t)
            p[t] = mul_unit(p[t], unit_or_conj(make_double2(static_cast<double>(e[2 * t].x),
                                                            static_cast<double>(e[2 * t].y)),
                                               odd(static_cast<uint32_t>(e[2 * t + 1].x))));
    }
    for (; i < cnt; ++i, r += 2) {
        const A e0 = r[0], e1 = r[1];
        p[0] = mul_unit(p[0], unit_or_conj(make_double2(static_cast<double>(e0.x), static_cast<double>(e0.y)),
                                           odd(static_cast<uint32_t>(e1.x))));
    }
    return mul_unit(mul_unit(p[0], p[1]), mul_unit(p[2], p[3]));
}

// f plus, over cnt terms of an inner-product phase op (two records each:
// the weight k, then {operand code, 0}), k or -k as odd(code) says.
template <typename A, typename Odd>
__device__ __forceinline__ double2 term_sum(double2 f, const A*& r, uint32_t cnt, Odd odd) {
    double2 p[4] = {f, make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
    uint32_t i = 0;
    for (; i + 4 <= cnt; i += 4, r += 8) {
        A e[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) e[t] = r[t];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const double sg = odd(static_cast<uint32_t>(e[2 * t + 1].x)) ? -1.0 : 1.0;
            p[t].x = fma(sg, static_cast<double>(e[2 * t].x), p[t].x);
            p[t].y = fma(sg, static_cast<double>(e[2 * t].y), p[t].y);
        }
    }
    for (; i < cnt; ++i, r += 2) {
        const A e0 = r[0], e1 = r[1];
        const double sg = odd(static_cast<uint32_t>(e1.x)) ? -1.0 : 1.0;
        p[0].x = fma(sg, static_cast<double>(e0.x), p[0].x);
        p[0].y = fma(sg, static_cast<double>(e0.y), p[0].y);
    }
    return make_double2((p[0].x + p[1].x) + (p[2].x + p[3].x), (p[0].y + p[1].y) + (p[2].y + p[3].y));
}

// <phi| D |psi> over this thread's amplitudes for an inner-product phase op
// (OP_IP_PHASE): D at amplitude r = k0 plus the terms with no row operand
// (a0, per thread) plus, per row bit m, +-g_m (the terms with one operand
// there) plus the row-pair table's entry; the thread key as for OP_PHASE.
// Sums and products in FP64.
template <typename A, int MAXROWS>
__device__ __forceinline__ double2 ip_phase_op(A* s, int T, uint64_t gbase, uint32_t te, uint32_t R, bool active,
                                               uint32_t phi_off, const OpRegs& op, const A* __restrict__ rec) {
    constexpr int sh = Amp<A>::kShift;
    const uint64_t key = gbase | (static_cast<uint64_t>(threadIdx.x) << 56);
    const A* r = rec;
    double2 a0 = make_double2(static_cast<double>(r[0].x), static_cast<double>(r[0].y));
    ++r;
    a0 = term_sum(a0, r, static_cast<uint32_t>(op.gmask & 0xffffu),
                  [&](uint32_t c) { return ((key >> (c & 0xffu)) ^ (key >> (c >> 8))) & 1u; });
    const int rows = T > 8 ? T - 8 : 0;
    double2 g[MAXROWS > 0 ? MAXROWS : 1];
#pragma unroll
    for (int m = 0; m < MAXROWS; ++m) {
        g[m] = make_double2(0.0, 0.0);
        const uint32_t cnt = static_cast<uint32_t>(
            (m < 2 ? op.gmask >> (32 + 16 * m) : op.gpos0 >> (16 * (m - 2))) & 0xffffu);
        if (m < rows) g[m] = term_sum(g[m], r, cnt, [&](uint32_t c) { return (key >> c) & 1u; });
    }
    const bool pair_table = ((op.gmask >> 16) & 0xffffu) != 0;
    double2 acc = make_double2(0.0, 0.0);
    if (!active) return acc;
    for (uint32_t i0 = 0; i0 < R; i0 += 4) {
        A v[4], f[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i0 + u < R) {
                const uint32_t a = te ^ (swz((i0 + u) << 8) << sh);
                v[u] = at(s, a);
                f[u] = at(s, a | phi_off);
            }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i0 + u < R) {
                const uint32_t rr = i0 + u;
                double2 d = a0;
#pragma unroll
                for (int m = 0; m < MAXROWS; ++m)
                    if (m < rows) {
                        const double sg = ((rr >> m) & 1u) ? -1.0 : 1.0;
                        d.x = fma(sg, g[m].x, d.x);
                        d.y = fma(sg, g[m].y, d.y);
                    }
                if (pair_table) {
                    const A e = r[rr];
                    d.x += static_cast<double>(e.x);
                    d.y += static_cast<double>(e.y);
                }
                const double xr = v[u].x, xi = v[u].y, fr = f[u].x, fi = f[u].y;
                const double gr = d.x * xr - d.y * xi, gi = d.x * xi + d.y * xr;   // (D psi)_r
                acc.x += fr * gr + fi * gi;                                          // conj(phi_r) (D psi)_r
                acc.y += fr * gi - fi * gr;
            }
    }
    return acc;
}

// A merged phase op (OP_PHASE) on this thread's amplitudes l = threadIdx.x +
// 256 r, r < R.  Each term is stored as its factor exp(i w) and its operand
// code; it contributes exp(i w) where its operands' bits agree and the
// conjugate where they differ.  The thread's key holds the tile's global
// bits and, at bits 56..63, the thread bits (tile positions 0..7); bit 55 is
// clear.  The factor of amplitude r = f0 (the constant and the terms with no
// operand on a row bit, per thread) x, per row bit m (tile position 8 + m),
// g_m or its conjugate (the terms with one operand there) x the row-pair
// table's entry for row code r.  Products in FP64 in both precisions.
template <typename A, bool TWO, int MAXROWS>
__device__ __forceinline__ void phase_op(A* s, int T, uint64_t gbase, uint32_t te, uint32_t R, bool active,
                                         uint32_t phi_off, const OpRegs& op, const A* __restrict__ rec) {
    constexpr int sh = Amp<A>::kShift;
    const uint64_t key = gbase | (static_cast<uint64_t>(threadIdx.x) << 56);
    const A* r = rec;
    double2 f0 = make_double2(static_cast<double>(r[0].x), static_cast<double>(r[0].y));
    ++r;
    f0 = term_product(f0, r, static_cast<uint32_t>(op.lmask & 0xffffu),
                      [&](uint32_t c) { return ((key >> (c & 0xffu)) ^ (key >> (c >> 8))) & 1u; });
    const int rows = T > 8 ? T - 8 : 0;
    double2 g[MAXROWS > 0 ? MAXROWS : 1];
#pragma unroll
    for (int m = 0; m < MAXROWS; ++m) {
        g[m] = make_double2(1.0, 0.0);
        const uint32_t cnt = static_cast<uint32_t>(
            (m < 2 ? op.lmask >> (32 + 16 * m) : op.gmask >> (16 * (m - 2))) & 0xffffu);
        if (m < rows) g[m] = term_product(g[m], r, cnt, [&](uint32_t c) { return (key >> c) & 1u; });
    }
    const bool pair_table = ((op.lmask >> 16) & 0xffffu) != 0;   // row-pair terms: a table over row codes
    // Rows in Gray-code order: consecutive codes differ in one row bit m, so
    // each factor is the previous one times conj(g_m)^2 (bit m set) or
    // g_m^2 (bit m cleared).  Four amplitudes per step: their LDS reads are
    // issued together, then the products, then the writes.
    double2 sq[MAXROWS > 0 ? MAXROWS : 1];
    double2 base = f0;
#pragma unroll
    for (int m = 0; m < MAXROWS; ++m)
        if (m < rows) {
            base = mul_unit(base, g[m]);
            sq[m] = mul_unit(g[m], g[m]);
        }
    if (!active) return;
    for (uint32_t i0 = 0; i0 < R; i0 += 4) {
        A v[4], w[4];
        uint32_t ad[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i0 + u < R) {
                const uint32_t i = i0 + u, rr = i ^ (i >> 1);
                ad[u] = te ^ (swz(rr << 8) << sh);
                v[u] = at(s, ad[u]);
                if (TWO) w[u] = at(s, ad[u] | phi_off);
            }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i0 + u < R) {
                const uint32_t i = i0 + u, rr = i ^ (i >> 1);
                if (i > 0) {
                    const int mb = __builtin_ctz(i);   // wave-uniform
                    double2 q = make_double2(1.0, 0.0);
#pragma unroll
                    for (int m = 0; m < MAXROWS; ++m)
                        if (m == mb) q = sq[m];
                    base = mul_unit(base, unit_or_conj(q, (rr >> mb) & 1u));
                }
                double2 f = base;
                if (pair_table) {
                    const A e = r[rr];
                    f = mul_unit(f, make_double2(static_cast<double>(e.x), static_cast<double>(e.y)));
                }
                A fa;
                fa.x = f.x;
                fa.y = f.y;
                v[u] = cmul(v[u], fa);
                if (TWO) w[u] = cmul(w[u], fa);
            }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i0 + u < R) {
                at(s, ad[u]) = v[u];
                if (TWO) at(s, ad[u] | phi_off) = w[u];
            }
    }
}

// The ops of pass P over the tile in LDS (psi at element 0; phi at element
// 2^TB in two-state mode), a barrier after each.  te: this thread's byte
// offset for amplitudes l = threadIdx.x + 256 r.  Returns the number of
// inner-product ops, whose per-wave partials are left in wave_ip.  STAGED
// (single state): the matrix-core ops read their fragments from frag, the
// LDS copy the streaming kernel made, in op order.
template <typename A, bool TWO, int KMAX, int TB, bool STAGED, bool PAIRS, bool BATCH>
__device__ __forceinline__ int run_ops(A* s, const PassDesc& P, uint64_t gbase, uint32_t te,
                                       const DevOp* __restrict__ ops, const A* __restrict__ coefs,
                                       double2* wave_ip, const A* frag) {
    static_assert(!(STAGED && TWO), "staged fragments are a single-state form");
    constexpr int KD = KMAX & ~kPhaseKernel;                      // widest VALU dense op
    constexpr bool PH = (KMAX & kPhaseKernel) != 0;              // merged phase ops compiled in
    constexpr int sh = Amp<A>::kShift;
    constexpr uint32_t kPhi = (1u << TB) << sh;   // byte offset of phi's tile
    const uint32_t L = 1u << P.T;
    const uint32_t R = L >= kThreads ? L / kThreads : 1u;
    const bool active = threadIdx.x < L;
    int n_ip = 0;
    uint32_t foff = 0;   // staged fragments consumed so far
    for (int o = P.op_begin; o < P.op_end; ++o) {
        const OpRegs op = load_op(ops + o);
        const int kind = op.kind, k = op.k;
        const A* U = coefs + op.coef;
        switch (kind) {
            case OP_DENSE: {
                const int G = P.T - k;
                const uint32_t tp = thread_part<A>(op, G);
                dense_op<A, KD>(s, tp, G, op, U);
                if (TWO) dense_op<A, KD>(s, tp | kPhi, G, op, U);
                break;
            }
            case OP_MFMA: {
                const int G = P.T - k;
                if constexpr (STAGED) {
                    // The staging rule of tile_stream_kernel: an op's fragments
                    // are in LDS when they fitted after the earlier ops'.
                    const uint32_t nf = staged_fragment_elems(k, std::is_same<A, float2>::value);
                    if (foff + nf <= kStreamFragBytes / sizeof(A)) {
                        mfma_op<A, true>(s, 0u, G, op, frag + foff);
                        foff += nf;
                    } else {
                        mfma_op<A, false>(s, 0u, G, op, U);
                    }
                } else {
                    mfma_op<A, false>(s, 0u, G, op, U);
                    if (TWO) mfma_op<A, false>(s, kPhi, G, op, U);
                }
                break;
            }
            case OP_BLOCK: {
                const int nsub = static_cast<int>(op.lmask);
                const int G = P.T - 3;
                const uint32_t tp = thread_part<A>(op, G);
                if constexpr (TWO) {
                    // Both states in registers, each member dispatched once
                    // for the two (and inner-product members, if any).
                    n_ip += block_apply2<A, !BATCH>(s, tp, kPhi, G, op, ops + o + 1, nsub, coefs, wave_ip, n_ip, gbase);
                } else {
                    block_apply<A, PAIRS>(s, tp, G, op, ops + o + 1, nsub, coefs, gbase);
                }
                o += nsub;   // the member entries are consumed here
                break;
            }
            case OP_DIAG:
                if (active)
                    for (uint32_t r = 0; r < R; ++r) {
                        const uint32_t l = threadIdx.x + r * kThreads;
                        const uint32_t a = te ^ (swz(r << 8) << sh);
                        const A d = U[diag_index(l, gbase, op, k)];
                        at(s, a) = cmul(at(s, a), d);
                        if (TWO) at(s, a | kPhi) = cmul(at(s, a | kPhi), d);
                    }
                break;
            case OP_DIAGG: {
                const A* Dg = U + (diag_sub_table(gbase, op) << k);
                if (k == 0) {
                    const A ph = Dg[0];
                    if (active)
                        for (uint32_t r = 0; r < R; ++r) {
                            const uint32_t a = te ^ (swz(r << 8) << sh);
                            at(s, a) = cmul(at(s, a), ph);
                            if (TWO) at(s, a | kPhi) = cmul(at(s, a | kPhi), ph);
                        }
                } else {
                    const int G = P.T - k;
                    const uint32_t tp = thread_part<A>(op, G);
                    if (k == 1) {
                        diag_groups<A, 1>(s, tp, G, op, Dg);
                        if (TWO) diag_groups<A, 1>(s, tp | kPhi, G, op, Dg);
                    } else {
                        diag_groups<A, 2>(s, tp, G, op, Dg);
                        if (TWO) diag_groups<A, 2>(s, tp | kPhi, G, op, Dg);
                    }
                }
                break;
            }
            case OP_MASKPHASE: {
                const A ph = U[0];
                const uint64_t lm = op.lmask, gm = op.gmask;
                if ((gbase & gm) == gm && active)
                    for (uint32_t r = 0; r < R; ++r) {
                        const uint32_t l = threadIdx.x + r * kThreads;
                        if ((l & lm) == lm) {
                            const uint32_t a = te ^ (swz(r << 8) << sh);
                            at(s, a) = cmul(at(s, a), ph);
                            if (TWO) at(s, a | kPhi) = cmul(at(s, a | kPhi), ph);
                        }
                    }
                break;
            }
            case OP_SCALAR: {
                const A ph = U[0];
                if (active)
                    for (uint32_t r = 0; r < R; ++r) {
                        const uint32_t a = te ^ (swz(r << 8) << sh);
                        at(s, a) = cmul(at(s, a), ph);
                        if (TWO) at(s, a | kPhi) = cmul(at(s, a | kPhi), ph);
                    }
                break;
            }
            case OP_PHASE:
                if constexpr (PH)
                    phase_op<A, TWO, (TB > 14 ? 6 : TB > 8 ? TB - 8 : 0)>(s, P.T, gbase, te, R, active, kPhi, op, U);
                break;
            case OP_IP_PHASE:   // two-state passes only
                if constexpr (TWO && PH) {
                    double2 acc =
                        ip_phase_op<A, (TB > 14 ? 6 : TB > 8 ? TB - 8 : 0)>(s, P.T, gbase, te, R, active, kPhi, op, U);
                    acc = wave_sum(acc);
                    if ((threadIdx.x & 63) == 0) wave_ip[n_ip * kWaves + (threadIdx.x >> 6)] = acc;
                    ++n_ip;
                }
                break;
            case OP_IP_DENSE:
            case OP_IP_DIAG:
            case OP_IP_DIAGG:
                if constexpr (TWO) {
                    double2 acc = make_double2(0.0, 0.0);
                    if (kind == OP_IP_DENSE) {
                        const int G = P.T - k;
                        const uint32_t tp = thread_part<A>(op, G);
                        acc = k == 1 ? ip_dense<A, 1>(s, tp, tp | kPhi, G, op, U)
                                     : ip_dense<A, 2>(s, tp, tp | kPhi, G, op, U);
                    } else if (kind == OP_IP_DIAGG && k > 0) {
                        const int G = P.T - k;
                        const uint32_t tp = thread_part<A>(op, G);
                        const A* Dg = U + (diag_sub_table(gbase, op) << k);
                        acc = k == 1 ? ip_diag_groups<A, 1>(s, tp, tp | kPhi, G, op, Dg)
                                     : ip_diag_groups<A, 2>(s, tp, tp | kPhi, G, op, Dg);
                    } else if (kind == OP_IP_DIAGG && active) {   // no local operand: one entry per tile
                        const A dv = U[diag_sub_table(gbase, op)];
                        const double dx = dv.x, dy = dv.y;
                        for (uint32_t r = 0; r < R; ++r) {
                            const uint32_t a = te ^ (swz(r << 8) << sh);
                            const A x = at(s, a), f = at(s, a | kPhi);
                            const double gr = dx * x.x - dy * x.y, gi = dx * x.y + dy * x.x;
                            acc.x += static_cast<double>(f.x) * gr + static_cast<double>(f.y) * gi;
                            acc.y += static_cast<double>(f.x) * gi - static_cast<double>(f.y) * gr;
                        }
                    } else if (kind == OP_IP_DIAG && active) {
                        for (uint32_t r = 0; r < R; ++r) {
                            const uint32_t l = threadIdx.x + r * kThreads;
                            const uint32_t a = te ^ (swz(r << 8) << sh);
                            const A d = U[k == 0 ? 0 : diag_index(l, gbase, op, k)];
                            const A x = at(s, a), f = at(s, a | kPhi);
                            const double gr = static_cast<double>(d.x) * x.x - static_cast<double>(d.y) * x.y;
                            const double gi = static_cast<double>(d.x) * x.y + static_cast<double>(d.y) * x.x;
                            acc.x += static_cast<double>(f.x) * gr + static_cast<double>(f.y) * gi;
                            acc.y += static_cast<double>(f.x) * gi - static_cast<double>(f.y) * gr;
                        }
                    }
                    acc = wave_sum(acc);
                    if ((threadIdx.x & 63) == 0) wave_ip[n_ip * kWaves + (threadIdx.x >> 6)] = acc;
                    ++n_ip;
                }
                break;
        }
        __syncthreads();
    }
    return n_ip;
}

// Tile of pass P: load, run the pass's ops in LDS, store.  KMAX bounds the
// widest dense op of the pass (register budget of the group vectors); the
// tile array holds 2^TB >= 2^T amplitudes per state.
template <typename A, bool TWO, int TB>
struct TileLds {
    static constexpr uint32_t kTile = 1u << TB;
    static constexpr uint32_t kIpElems = TWO ? kMaxIpPerPass * kWaves * sizeof(double2) / sizeof(A) : 0;
    static constexpr uint32_t kElems = (TWO ? 2 : 1) * kTile + kIpElems;
};

// One tile t of pass P in the LDS array s (TileLds<A, TWO, TB>::kElems
// elements).  PAIRS: single-state blocks take a thread's groups two at a time.
// BATCH: the batched kernel (tighter on registers).
template <typename A, bool TWO, int KMAX, int TB, bool PAIRS, bool BATCH>
__device__ __forceinline__ void tile_body(A* s, A* __restrict__ psi, A* __restrict__ phi, const PassDesc& P,
                                          const DevOp* __restrict__ ops, const A* __restrict__ coefs,
                                          double2* __restrict__ ip_out, int64_t n_tiles, uint64_t t) {
    constexpr int sh = Amp<A>::kShift;
    constexpr uint32_t kTile = 1u << TB;
    constexpr uint32_t kPhi = kTile << sh;
    double2* wave_ip = reinterpret_cast<double2*>(s + (TWO ? 2 : 1) * kTile);

    const uint32_t L = 1u << P.T;
    const uint32_t R = L >= kThreads ? L / kThreads : 1u;
    const bool active = threadIdx.x < L;
    const uint64_t full = P.n >= 64 ? ~0ull : ((1ull << P.n) - 1ull);
    const uint64_t gbase = tile_base(t, full & ~P.tmask);
    const uint32_t te = swz(threadIdx.x) << sh;

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
                const uint32_t a = te ^ (swz((r0 + i) << 8) << sh);
                at(s, a) = v0[i];
                if (TWO) at(s, a | kPhi) = v1[i];
            }
        }
    } else if (active) {
        uint64_t dr = 0;
        for (uint32_t r = 0; r < R; ++r) {
            const uint64_t g = gbase | dj | dr;
            const uint32_t a = te ^ (swz(r << 8) << sh);
            at(s, a) = psi[g];
            if (TWO) at(s, a | kPhi) = phi[g];
            dr = (dr - P.hmask) & P.hmask;
        }
    }
    __syncthreads();

    const int n_ip = run_ops<A, TWO, KMAX, TB, false, PAIRS, BATCH>(s, P, gbase, te, ops, coefs, wave_ip, nullptr);

    if (TWO && static_cast<int>(threadIdx.x) < n_ip) {
        // Sum the per-wave values in fixed order: partial k of this tile.
        const int k = threadIdx.x;
        double2 v = wave_ip[k * kWaves];
        for (int w = 1; w < kWaves; ++w) {
            v.x += wave_ip[k * kWaves + w].x;
            v.y += wave_ip[k * kWaves + w].y;
        }
        ip_out[static_cast<int64_t>(k) * n_tiles + static_cast<int64_t>(t)] = v;
    }

    // Store through the pass's permutation of tile positions (the identity
    // unless the pass relabels qubits): output local index m reads input
    // index perm(m) = sum_i bit_i(m) << sperm[i].  perm and the swizzle are
    // linear over XOR, so the LDS offset is the thread's part XOR one
    // uniform offset per set row bit.
    static_assert(TB <= 13, "tile rows carry at most five permuted bits");
    uint32_t tperm = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b)
        if (b < P.T) tperm |= ((threadIdx.x >> b) & 1u) << P.sperm[b];
    const uint32_t to = swz(tperm) << sh;
    uint32_t rbit[5];
#pragma unroll
    for (int b = 0; b < 5; ++b) rbit[b] = 8 + b < P.T ? swz(1u << P.sperm[8 + b]) << sh : 0u;
    auto row_perm = [&](uint32_t r) {
        uint32_t p = 0;
#pragma unroll
        for (int b = 0; b < 5; ++b) p ^= ((r >> b) & 1u) ? rbit[b] : 0u;
        return p;
    };
    if (R >= kBatch) {
        uint64_t dr = 0;
        for (uint32_t r0 = 0; r0 < R; r0 += kBatch) {
            A v0[kBatch], v1[kBatch];
#pragma unroll
            for (int i = 0; i < kBatch; ++i) {
                const uint32_t a = to ^ row_perm(r0 + i);
                v0[i] = at(s, a);
                if (TWO) v1[i] = at(s, a | kPhi);
            }
#pragma unroll
            for (int i = 0; i < kBatch; ++i) {
                const uint64_t g = gbase | dj | dr;
                stream_store(psi + g, v0[i]);
                if (TWO) stream_store(phi + g, v1[i]);
                dr = (dr - P.hmask) & P.hmask;
            }
        }
    } else if (active) {
        uint64_t dr = 0;
        for (uint32_t r = 0; r < R; ++r) {
            const uint64_t g = gbase | dj | dr;
            const uint32_t a = to ^ row_perm(r);
            psi[g] = at(s, a);
            if (TWO) phi[g] = at(s, a | kPhi);
            dr = (dr - P.hmask) & P.hmask;
        }
    }
}

// Two-state tiles of the common kind (dense ops of at most two qubits) ask
// the register allocator for two waves per SIMD (VGPRs + AGPRs <= 256), the
// occupancy their LDS allows (VGPRs and AGPRs share one register file).
constexpr int two_state_waves(bool two, int kmax) { return two && kmax == 2 ? 2 : 1; }
// Single-state 32-KiB tiles of the common kind in the one-circuit kernel:
// four waves per SIMD (VGPRs + AGPRs <= 128); their LDS would allow five.
// (The batched kernel's instance lookup does not fit that budget.)
template <typename A, bool TWO, int KMAX, int TB, bool BATCH>
constexpr int tile_waves() {
    if (TWO) return two_state_waves(TWO, KMAX & ~kPhaseKernel);
    return (!BATCH && (KMAX & ~kPhaseKernel) == 2 && TB + Amp<A>::kShift == 15) ? 4 : 1;
}

template <typename A, bool TWO, int KMAX, int TB>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(tile_waves<A, TWO, KMAX, TB, false>()))) void tile_pass_kernel(
    A* __restrict__ psi, A* __restrict__ phi, const PassDesc P, const DevOp* __restrict__ ops,
    const A* __restrict__ coefs, double2* __restrict__ ip_out, int64_t n_tiles) {
    __shared__ __align__(16) A s[TileLds<A, TWO, TB>::kElems];
    tile_body<A, TWO, KMAX, TB, block_pairs(TB), false>(s, psi, phi, P, ops, coefs, ip_out, n_tiles, blockIdx.x);
}

// Many instances' passes in one launch: workgroup b runs tile b - first[j]
// of item j, where first[j] <= b < first[j + 1] (first[] is the exclusive
// prefix sum of the items' tile counts, found by bisection; wave-uniform).
template <typename A, bool TWO, int KMAX, int TB>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(tile_waves<A, TWO, KMAX, TB, true>()))) void tile_batch_kernel(const BatchItem* __restrict__ items,
                                                              const int64_t* __restrict__ first, int n_items) {
    __shared__ __align__(16) A s[TileLds<A, TWO, TB>::kElems];
    const int64_t b = blockIdx.x;
    int lo = 0, hi = n_items - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const BatchItem& it = items[lo];
    tile_body<A, TWO, KMAX, TB, false, true>(s, static_cast<A*>(it.psi), static_cast<A*>(it.phi), it.P,
                                static_cast<const DevOp*>(it.ops), static_cast<const A*>(it.coefs), it.ip_out,
                                it.n_tiles, static_cast<uint64_t>(b - first[lo]));
}

// Streaming single-state tile pass.  Persistent workgroups walk the tiles
// t = blockIdx.x, blockIdx.x + gridDim.x, ... of a pass whose tile fills the
// LDS array (T = TB, so every thread holds R = 2^(TB-8) rows).  The loads of
// a workgroup's next tile are issued into registers before the ops of the
// current tile run, so the HBM reads overlap the pass's compute; stores leave
// asynchronously as in the one-tile kernel.  The matrix-core fragments of the
// pass are copied into LDS once per workgroup: a vector load from global
// memory inside the op loop would wait for the prefetch in flight, since
// vector memory loads complete in issue order.  Same arithmetic, same order
// of operations per amplitude as tile_pass_kernel.
template <typename A, int TB>
struct StreamLds {
    static constexpr uint32_t kTile = 1u << TB;
    static constexpr uint32_t kFrag = kStreamFragBytes / sizeof(A);
    static constexpr uint32_t kBytes = (kTile + kFrag) * sizeof(A);
};
// Waves per SIMD the LDS allows (160 KiB per CU, four SIMDs, one wave of
// each workgroup per SIMD), at most four.
template <typename A, int TB>
constexpr int stream_waves() {
    constexpr uint32_t w = 163840u / StreamLds<A, TB>::kBytes;
    return w < 1 ? 1 : w > 4 ? 4 : static_cast<int>(w);
}

template <typename A, int TB>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(stream_waves<A, TB>()))) void tile_stream_kernel(
    A* __restrict__ psi, const PassDesc P, const DevOp* __restrict__ ops, const A* __restrict__ coefs,
    int64_t n_tiles) {
    using L = StreamLds<A, TB>;
    constexpr int sh = Amp<A>::kShift;
    constexpr uint32_t R = L::kTile / kThreads;
    constexpr uint32_t RB = R < kBatch ? R : kBatch;
    constexpr bool f32 = std::is_same<A, float2>::value;
    constexpr int kRowBits = TB - 8;
    static_assert(TB >= 9 && TB <= 14, "rows of 256 amplitudes, at most six permuted row bits");
    __shared__ __align__(16) A s[L::kTile + L::kFrag];
    A* frag = s + L::kTile;

    // Fragments of the pass's matrix-core ops, in op order, while they fit
    // (run_ops applies the same rule).
    uint32_t off = 0;
    for (int o = P.op_begin; o < P.op_end; ++o) {
        const OpRegs op = load_op(ops + o);
        if (op.kind != OP_MFMA) continue;
        const uint32_t n = static_cast<uint32_t>(staged_fragment_elems(op.k, f32));
        if (off + n > L::kFrag) continue;
        const A* src = coefs + op.coef;
        for (uint32_t e = threadIdx.x; e < n; e += kThreads) {
            const uint32_t j = e >> 6, lane = e & 63u;
            frag[off + e] = src[(op.k >= 4 && !f32) ? ((j >> 1) * 3 + (j & 1)) * 64 + lane : e];
        }
        off += n;
    }

    const uint64_t full = P.n >= 64 ? ~0ull : ((1ull << P.n) - 1ull);
    const uint64_t outer = full & ~P.tmask;
    const uint32_t te = swz(threadIdx.x) << sh;
    uint64_t dj = 0;   // the thread's eight low local bits, deposited
#pragma unroll
    for (int b = 0; b < 8; ++b) dj |= static_cast<uint64_t>((threadIdx.x >> b) & 1u) << P.qpos[b];
    // Store permutation (as in tile_body): output row r of this thread reads
    // LDS offset to ^ row_perm(r).
    uint32_t tperm = 0;
#pragma unroll
    for (int b = 0; b < 8; ++b) tperm |= ((threadIdx.x >> b) & 1u) << P.sperm[b];
    const uint32_t to = swz(tperm) << sh;
    uint32_t rbit[kRowBits];
#pragma unroll
    for (int b = 0; b < kRowBits; ++b) rbit[b] = swz(1u << P.sperm[8 + b]) << sh;

    A pre[R];
    auto load_tile = [&](int64_t tt) {
        const uint64_t g0 = tile_base(static_cast<uint64_t>(tt), outer) | dj;
        uint64_t dr = 0;
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            pre[r] = stream_load(psi + (g0 | dr));
            dr = (dr - P.hmask) & P.hmask;
        }
    };
    auto put_tile = [&]() {
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) at(s, te ^ (swz(r << 8) << sh)) = pre[r];
    };
    int64_t t = blockIdx.x;
    if (t >= n_tiles) return;   // the host sizes the grid to at most n_tiles
    load_tile(t);
    put_tile();
    __syncthreads();   // the first tile and the fragments in LDS
    for (;;) {
        const int64_t tn = t + gridDim.x;
        // The next tile's loads go out before this tile's ops run.  They are
        // unconditional so that the wait before put_tile counts only this
        // iteration's loads and stores: past the end the last tile is read
        // again and never used.
        load_tile(tn < n_tiles ? tn : n_tiles - 1);
        const uint64_t gbase = tile_base(static_cast<uint64_t>(t), outer);
        run_ops<A, false, 2 | kPhaseKernel, TB, true, false, false>(s, P, gbase, te, ops, coefs, nullptr, frag);
        const uint64_t g0 = gbase | dj;
        uint64_t dr = 0;
#pragma unroll
        for (uint32_t r0 = 0; r0 < R; r0 += RB) {
            A v[RB];
#pragma unroll
            for (uint32_t i = 0; i < RB; ++i) {
                const uint32_t r = r0 + i;
                uint32_t pr = 0;
#pragma unroll
                for (int b = 0; b < kRowBits; ++b) pr ^= ((r >> b) & 1u) ? rbit[b] : 0u;
                v[i] = at(s, to ^ pr);
            }
#pragma unroll
            for (uint32_t i = 0; i < RB; ++i) {
                stream_store(psi + (g0 | dr), v[i]);
                dr = (dr - P.hmask) & P.hmask;
            }
        }
        if (tn >= n_tiles) break;
        __syncthreads();   // every wave has read this tile out of LDS
        put_tile();
        __syncthreads();
        t = tn;
    }
}

// ── launch ──────────────────────────────────────────────────────────────────
// TB0 .. TB0 + 2 are the instantiated tile-array sizes of one amplitude type:
// a pass runs with the smallest array that holds its 2^T amplitudes.

template <typename A, bool TWO, int TB>
static void launch_k(int km, const PassDesc& P, A* psi, A* phi, double2* ip, const DevOp* ops, const A* coefs,
                     int64_t n_tiles, hipStream_t st) {
    const dim3 grid(static_cast<uint32_t>(n_tiles)), block(kThreads);
    switch (km) {
        case 2: hipLaunchKernelGGL((tile_pass_kernel<A, TWO, 2, TB>), grid, block, 0, st, psi, phi, P, ops, coefs, ip,
                                   n_tiles); break;
        case 3: hipLaunchKernelGGL((tile_pass_kernel<A, TWO, 3, TB>), grid, block, 0, st, psi, phi, P, ops, coefs, ip,
                                   n_tiles); break;
        case 4: hipLaunchKernelGGL((tile_pass_kernel<A, TWO, 4, TB>), grid, block, 0, st, psi, phi, P, ops, coefs, ip,
                                   n_tiles); break;
        case 2 | kPhaseKernel:
            hipLaunchKernelGGL((tile_pass_kernel<A, TWO, 2 | kPhaseKernel, TB>), grid, block, 0, st, psi, phi, P, ops,
                               coefs, ip, n_tiles);
            break;
        case 5 | kPhaseKernel:
            hipLaunchKernelGGL((tile_pass_kernel<A, TWO, 5 | kPhaseKernel, TB>), grid, block, 0, st, psi, phi, P, ops,
                               coefs, ip, n_tiles);
            break;
        default: hipLaunchKernelGGL((tile_pass_kernel<A, TWO, 5, TB>), grid, block, 0, st, psi, phi, P, ops, coefs,
                                    ip, n_tiles); break;
    }
}

// Kernel variant of a pass: its VALU dense width (2..5), and for passes
// with merged phase ops the phase variant (width 2 or 5).
inline int kernel_variant(int kmax) {
    const int w = kmax & ~kPhaseKernel;
    const int km = w <= 2 ? 2 : std::min(w, 5);
    if (kmax & kPhaseKernel) return (km == 2 ? 2 : 5) | kPhaseKernel;
    return km;
}

template <typename A, bool TWO, int TB0>
static void launch_tb(const PassDesc& P, A* psi, A* phi, double2* ip, const DevOp* ops, const A* coefs,
                      int64_t n_tiles, hipStream_t st) {
    const int km = kernel_variant(P.kmax);
    const int tb = std::max(P.T, TB0);
    switch (tb - TB0) {
        case 0: launch_k<A, TWO, TB0>(km, P, psi, phi, ip, ops, coefs, n_tiles, st); break;
        case 1: launch_k<A, TWO, TB0 + 1>(km, P, psi, phi, ip, ops, coefs, n_tiles, st); break;
        case 2: launch_k<A, TWO, TB0 + 2>(km, P, psi, phi, ip, ops, coefs, n_tiles, st); break;
        default: fail(QHIP_INVALID, "tile pass: " + std::to_string(P.T) + " tile qubits exceed the LDS tile");
    }
}

inline int cu_count() {
    static int cus[64] = {};
    int dev = 0;
    QHIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) fail(QHIP_RUNTIME, "device index out of range");
    if (!cus[dev]) QHIP_CHECK(hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev));
    return cus[dev];
}

// Persistent grid: as many workgroups as fit on the device at once (LDS
// bound), never more than there are tiles.
template <typename A, int TB>
static void launch_stream(const PassDesc& P, A* psi, const DevOp* ops, const A* coefs, int64_t n_tiles,
                          hipStream_t st) {
    static int per_cu = 0;
    if (!per_cu) {
        int nb = 0;
        QHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tile_stream_kernel<A, TB>, kThreads, 0));
        per_cu = nb > 0 ? nb : 1;
    }
    const int64_t grid = std::min<int64_t>(n_tiles, static_cast<int64_t>(per_cu) * cu_count());
    hipLaunchKernelGGL((tile_stream_kernel<A, TB>), dim3(static_cast<uint32_t>(grid)), dim3(kThreads), 0, st, psi,
                       P, ops, coefs, n_tiles);
}

template <typename A, int TB0>
static bool streams(const PassDesc& P, const A* phi) {
    constexpr bool c128 = std::is_same<A, double2>::value;
    static_assert(TB0 == (c128 ? kTileArrayQubitsC128 : kTileArrayQubitsC64), "tile arrays of streaming_pass");
    return !phi && streaming_pass(P, c128 ? QHIP_C128 : QHIP_C64);
}

template <typename A, int TB0>
static void launch(const PassDesc& P, A* psi, A* phi, double2* ip, const DevOp* ops, const A* coefs, int64_t n_tiles,
                   hipStream_t st) {
    if (P.T < 1 || P.T > P.n) fail(QHIP_RUNTIME, "tile pass: bad tile width");
    if (streams<A, TB0>(P, phi)) {
        switch (P.T - TB0) {
            case 0: launch_stream<A, TB0>(P, psi, ops, coefs, n_tiles, st); break;
            case 1: launch_stream<A, TB0 + 1>(P, psi, ops, coefs, n_tiles, st); break;
            case 2: launch_stream<A, TB0 + 2>(P, psi, ops, coefs, n_tiles, st); break;
            default: launch_stream<A, TB0 + 3>(P, psi, ops, coefs, n_tiles, st); break;
        }
    } else if (P.T > TB0 + 2) {
        fail(QHIP_INVALID, "tile pass: " + std::to_string(P.T) + " tile qubits run only on the streaming kernel");
    } else if (phi) launch_tb<A, true, TB0>(P, psi, phi, ip, ops, coefs, n_tiles, st);
    else launch_tb<A, false, TB0>(P, psi, phi, ip, ops, coefs, n_tiles, st);
    QHIP_CHECK(hipGetLastError());
}

template <typename A, bool TWO, int TB>
static void launch_batch_k(int km, const BatchItem* items, const int64_t* first, int n_items, int64_t total,
                           hipStream_t st) {
    const dim3 grid(static_cast<uint32_t>(total)), block(kThreads);
    switch (km) {
        case 2: hipLaunchKernelGGL((tile_batch_kernel<A, TWO, 2, TB>), grid, block, 0, st, items, first, n_items); break;
        case 3: hipLaunchKernelGGL((tile_batch_kernel<A, TWO, 3, TB>), grid, block, 0, st, items, first, n_items); break;
        case 4: hipLaunchKernelGGL((tile_batch_kernel<A, TWO, 4, TB>), grid, block, 0, st, items, first, n_items); break;
        case 2 | kPhaseKernel:
            hipLaunchKernelGGL((tile_batch_kernel<A, TWO, 2 | kPhaseKernel, TB>), grid, block, 0, st, items, first,
                               n_items);
            break;
        case 5 | kPhaseKernel:
            hipLaunchKernelGGL((tile_batch_kernel<A, TWO, 5 | kPhaseKernel, TB>), grid, block, 0, st, items, first,
                               n_items);
            break;
        default: hipLaunchKernelGGL((tile_batch_kernel<A, TWO, 5, TB>), grid, block, 0, st, items, first, n_items); break;
    }
}

template <typename A, int TB0>
static void launch_batch(bool two, const BatchItem* items, const int64_t* first, int n_items, int64_t total, int tmax,
                         int kmax, hipStream_t st) {
    if (n_items <= 0 || total <= 0) return;
    const int tb = std::max(tmax, TB0) - TB0;
    if (tb > 2) fail(QHIP_INVALID, "batched pass: " + std::to_string(tmax) + " tile qubits exceed the LDS tile");
    auto go = [&](auto two_tag) {
        constexpr bool TWO = decltype(two_tag)::value;
        const int km = kernel_variant(kmax);
        switch (tb) {
            case 0: launch_batch_k<A, TWO, TB0>(km, items, first, n_items, total, st); break;
            case 1: launch_batch_k<A, TWO, TB0 + 1>(km, items, first, n_items, total, st); break;
            default: launch_batch_k<A, TWO, TB0 + 2>(km, items, first, n_items, total, st); break;
        }
    };
    if (two) go(std::true_type{});
    else go(std::false_type{});
    QHIP_CHECK(hipGetLastError());
}

}  // namespace tile
}  // namespace qhip
