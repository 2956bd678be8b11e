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
// contiguous block with the low bits of x permuting lanes.  A thread's 16
// indices differ only in bits 8..11 (r = 0..15), so a group's terms are split
// on the host by their z bits there (the pattern p): per pattern the thread
// sums c_t (-1)^popcount(i_0 & z_t) once at its r = 0 index i_0, and the
// weight at index r is the sum over patterns of (-1)^popcount(r & p) times
// that sum — one popcount per term per thread instead of one per term per
// amplitude, with the r-dependent signs wave-uniform.  Sums are FP64
// (complex64 states are widened on load): per-thread sums, a lane-shuffle
// wave reduction, a fixed-order block reduction, then one block sums the
// per-chunk partials in index order — the result does not depend on timing.
#include "common.h"

#include <algorithm>
#include <bit>
#include <cstring>
#include <memory>
#include <type_traits>

#include <unordered_map>

namespace qhip {

constexpr int kT = 256;
constexpr int kPer = 16;                 // ket amplitudes per thread
constexpr uint64_t kChunk = kT * kPer;   // per workgroup
constexpr int kRShift = 8;               // a thread's amplitudes differ in bits 8..11
constexpr int kChunkBits = 12;           // log2(kChunk)
static_assert(kT == 1 << kRShift && kPer == 16, "r occupies index bits 8..11");

struct GroupDev {
    uint64_t x;
    int32_t s_begin, s_end;              // its patterns in the SubDev table
    int32_t t_begin, t_end;              // its terms (contiguous across its patterns)
};

struct SubDev {
    int32_t rpat;                        // z bits 8..11 shared by the terms; bits 16..31: sign mask
    int32_t t_begin, t_end;
    // Its terms run real coefficients first (n_real of them), then purely
    // imaginary ones (n_imag), then the rest; both counts are 0 when either
    // would not fit in 16 bits (every term is then treated as general, which
    // gives the same values).
    uint16_t n_real, n_imag;
};

// Coefficient kinds of a term (imaginary part 0, real part 0, or neither):
// in H|psi> a real or imaginary coefficient adds one multiply-add per
// amplitude to its group's weight instead of two.
enum TermKind { kTermReal = 0, kTermImag = 1, kTermGeneral = 2 };
inline TermKind term_kind(const cplx& c) {
    return c.imag() == 0.0 ? kTermReal : c.real() == 0.0 ? kTermImag : kTermGeneral;
}

struct Grouped {
    std::vector<GroupDev> groups;
    std::vector<SubDev> subs;
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
    // The diagonal group (x = 0) first: the expectation kernel treats it apart.
    if (auto it = index.find(0); it != index.end() && it->second != 0) {
        std::swap(xs[0], xs[it->second]);
        std::swap(per[0], per[it->second]);
    }
    // Then by the flip mask above bit kChunkBits (stable): H|psi> stages one
    // partner chunk per run of equal high parts (x = 0 stays first).
    std::vector<size_t> order(per.size());
    for (size_t g = 0; g < order.size(); ++g) order[g] = g;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t a, size_t b) { return (xs[a] >> kChunkBits) < (xs[b] >> kChunkBits); });
    for (size_t g : order) {
        GroupDev d;
        d.x = xs[g];
        d.s_begin = static_cast<int32_t>(out.subs.size());
        // Terms by pattern, then by coefficient kind; first appearance order
        // kept within a (pattern, kind) run.
        auto& terms = per[g];
        auto key = [](const std::pair<uint64_t, cplx>& t) {
            return static_cast<int>((t.first >> kRShift) & 15) * 3 + static_cast<int>(term_kind(t.second));
        };
        std::stable_sort(terms.begin(), terms.end(), [&](const auto& a, const auto& b) { return key(a) < key(b); });
        size_t n_kind[3] = {0, 0, 0};
        auto close_sub = [&] {
            SubDev& S = out.subs.back();
            const bool fits = n_kind[kTermReal] <= 0xffff && n_kind[kTermImag] <= 0xffff;
            S.n_real = fits ? static_cast<uint16_t>(n_kind[kTermReal]) : 0;
            S.n_imag = fits ? static_cast<uint16_t>(n_kind[kTermImag]) : 0;
            n_kind[0] = n_kind[1] = n_kind[2] = 0;
        };
        for (size_t t = 0; t < terms.size(); ++t) {
            const int32_t pat = static_cast<int32_t>((terms[t].first >> kRShift) & 15);
            if (t == 0 || pat != out.subs.back().rpat) {
                if (t > 0) close_sub();
                out.subs.push_back(SubDev{pat, static_cast<int32_t>(out.z.size()), 0, 0, 0});
            }
            const cplx ct = terms[t].second;
            ++n_kind[term_kind(ct)];
            out.z.push_back(terms[t].first);
            out.c.push_back(make_double2(ct.real(), ct.imag()));
            out.subs.back().t_end = static_cast<int32_t>(out.z.size());
        }
        if (!terms.empty()) close_sub();
        // Bits 16..31 of rpat: bit r = (-1)^popcount(r & pattern) is -1 (the
        // sign of the pattern's terms at the r-th output of a thread).
        for (int32_t si = d.s_begin; si < static_cast<int32_t>(out.subs.size()); ++si) {
            int32_t mask = 0;
            for (int r = 0; r < 16; ++r) mask |= (std::popcount(static_cast<unsigned>(r & out.subs[si].rpat)) & 1) << r;
            out.subs[si].rpat |= mask << 16;
        }
        d.s_end = static_cast<int32_t>(out.subs.size());
        d.t_begin = out.subs[d.s_begin].t_begin;
        d.t_end = out.subs.back().t_end;
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

// Sum over the terms of pattern S of c_t (-1)^popcount(i0 & z_t), i0 with
// bits 8..11 clear.  The weight of the group at index i0 | (r << 8) is then
// the sum over its patterns p of (-1)^popcount(r & p) times this sum.
__device__ __forceinline__ double2 pattern_sum(uint64_t i0, const SubDev& S, const uint64_t* __restrict__ z,
                                               const double2* __restrict__ c) {
    double ax = 0.0, ay = 0.0;
#pragma unroll 4
    for (int t = S.t_begin; t < S.t_end; ++t) {
        const double sg = (__popcll(i0 & z[t]) & 1) ? -1.0 : 1.0;
        ax += sg * c[t].x;
        ay += sg * c[t].y;
    }
    return make_double2(ax, ay);
}

// (-1)^popcount(r & p) for a compile-time r and a wave-uniform p.
__device__ __forceinline__ double rsign(int r, int p) { return (__popc(r & p) & 1) ? -1.0 : 1.0; }

// HERM (bra == ket, every coefficient real, so every group operator H_x is
// Hermitian): with h the top bit of x != 0, the terms at i and i ^ x are
// complex conjugates, so the real part is twice the sum over i with bit h
// clear — half the partner reads.  Only the real part of the result is
// meaningful then (the exact imaginary part is zero).
// FULL: the state has at least one whole chunk (n >= 12), so no index needs a
// bounds check.  Indices are formed from i_0 and the wave-uniform r bits, so
// no per-r 64-bit index stays live across the group loop.
template <typename A, bool HERM, bool FULL>
__global__ __launch_bounds__(kT) void transition_kernel(const A* __restrict__ bra, const A* __restrict__ ket,
                                                        uint64_t dim, const GroupDev* __restrict__ groups,
                                                        int n_groups, const SubDev* __restrict__ subs,
                                                        const uint64_t* __restrict__ z,
                                                        const double2* __restrict__ c,
                                                        double2* __restrict__ partials, uint64_t chunk0, int same) {
    __shared__ double2 red[kT / 64];
    const uint64_t base = (chunk0 + blockIdx.x) * kChunk;
    const uint64_t i0 = base + threadIdx.x;
    constexpr uint64_t rmask = uint64_t{15} << kRShift;
    double2 k[kPer];
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const uint32_t lid = threadIdx.x + static_cast<uint32_t>(r) * kT;
        k[r] = (FULL || lid < dim) ? load_amp(ket, i0 | (static_cast<uint64_t>(r) << kRShift))
                                   : make_double2(0.0, 0.0);
    }
    int g0 = 0;
    if (same && n_groups > 0 && groups[0].x == 0) {
        // The diagonal group (placed first by the host) with bra == ket: its
        // partner is the ket amplitude itself, u[r] = |ket[i_r]|^2 is real, and
        // every pattern sum is formed once.  Often the group with most terms.
        const GroupDev G = groups[0];
        double u[kPer];
#pragma unroll
        for (int r = 0; r < kPer; ++r) u[r] = k[r].x * k[r].x + k[r].y * k[r].y;
        for (int si = G.s_begin; si < G.s_end; ++si) {
            const SubDev S = subs[si];
            const double2 a = pattern_sum(i0, S, z, c);
            double sx = 0.0;
#pragma unroll
            for (int r = 0; r < kPer; ++r) sx += rsign(r, S.rpat) * u[r];
            acc.x += a.x * sx;
            acc.y += a.y * sx;
        }
        g0 = 1;
    }
    for (int g = g0; g < n_groups; ++g) {
        const GroupDev G = groups[g];
        uint64_t hb = 0;   // HERM: top bit of x (indices with it set are skipped)
        if (HERM && G.x) {
            hb = uint64_t{1} << (63 - __clzll(G.x));
            if (base & hb) continue;   // a bit above the chunk: uniform over the workgroup
        }
        const bool lane_skip = HERM && (i0 & hb);        // h among the thread bits
        const uint64_t p0 = (i0 ^ G.x) & ~rmask;         // partner of r, bits 8..11 cleared
        const int xr = static_cast<int>((G.x >> kRShift) & 15);
        const int hr = static_cast<int>((hb >> kRShift) & 15);   // h among the r bits
        // u[r] = conj(bra[i_r ^ x]) ket[i_r]; the group adds sum_r w(i_r) u[r],
        // in two halves of eight r so that few registers stay live (the
        // pattern sums are formed once per half).
        double2 gacc = make_double2(0.0, 0.0);
#pragma unroll
        for (int h = 0; h < kPer; h += kPer / 2) {
            double2 u[kPer / 2];
#pragma unroll
            for (int q = 0; q < kPer / 2; ++q) {
                const int r = h + q;
                const uint32_t lid = threadIdx.x + static_cast<uint32_t>(r) * kT;
                if ((!FULL && lid >= dim) || (HERM && (lane_skip || (r & hr)))) {
                    u[q] = make_double2(0.0, 0.0);
                    continue;
                }
                const double2 b = load_amp(bra, p0 | (static_cast<uint64_t>(r ^ xr) << kRShift));
                u[q] = make_double2(b.x * k[r].x + b.y * k[r].y, b.x * k[r].y - b.y * k[r].x);
            }
            for (int si = G.s_begin; si < G.s_end; ++si) {
                const SubDev S = subs[si];
                const double2 a = pattern_sum(i0, S, z, c);
                double sx = 0.0, sy = 0.0;
#pragma unroll
                for (int q = 0; q < kPer / 2; ++q) {
                    const double sg = rsign(h + q, S.rpat);
                    sx += sg * u[q].x;
                    sy += sg * u[q].y;
                }
                gacc.x += a.x * sx - a.y * sy;
                gacc.y += a.x * sy + a.y * sx;
            }
        }
        if (HERM && G.x) {   // the skipped half is the conjugate of the visited one
            gacc.x *= 2.0;
            gacc.y *= 2.0;
        }
        acc.x += gacc.x;
        acc.y += gacc.y;
    }
    const double2 tot = block_sum(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// out[j] = sum_g w_g(j ^ x_g) in[j ^ x_g] with w_g(i) = sum_{t in g} c_t
// (-1)^popcount(i & z_t).  The group and term tables are read through the
// scalar unit (one per CU), so the cost that bounds this kernel is the scalar
// work per group and term; each thread therefore carries kObsOut outputs
// (j, j + Q, ..., Q = grid threads) through one walk of the tables, and issues
// the partner loads of kObsAhead groups before it weighs the first.  A real or
// imaginary term costs an AND, a popcount, a shift-OR and one multiply-add
// per output.  I is uint32_t when the state has at most 2^32 amplitudes
// (every index and mask fits), uint64_t otherwise.
constexpr int kObsOut = 4;
constexpr int kObsAhead = 2;

// (-1)^popcount(src & z) as a double: the parity goes to the sign bit of 1.0.
template <typename I>
__device__ __forceinline__ double parity_sign(I src, uint64_t zt) {
    const uint32_t p = sizeof(I) == 4 ? __popc(static_cast<uint32_t>(src) & static_cast<uint32_t>(zt))
                                      : __popcll(static_cast<uint64_t>(src) & zt);
    return __hiloint2double(static_cast<int>((p << 31) | 0x3ff00000u), 0);
}

// acc[u] += w v[u] with w = sum over the group's terms of c_t (-1)^popcount(
// src[u] & z_t) (the weight summed first, as the reference's group sums are).
// A real or imaginary coefficient adds one multiply-add to w; the product
// with v is four multiply-adds (a zero part of w leaves acc exactly as is).
template <typename I>
__device__ __forceinline__ void add_group(double2 (&acc)[kObsOut], const I (&src)[kObsOut], const GroupDev& G,
                                          const SubDev* __restrict__ subs, const uint64_t* __restrict__ z,
                                          const double2* __restrict__ c, const double2 (&v)[kObsOut]) {
    double wr[kObsOut], wi[kObsOut];
#pragma unroll
    for (int u = 0; u < kObsOut; ++u) wr[u] = wi[u] = 0.0;
    for (int si = G.s_begin; si < G.s_end; ++si) {
        const SubDev S = subs[si];
        const int r_end = S.t_begin + S.n_real, i_end = r_end + S.n_imag;
        for (int t = S.t_begin; t < r_end; ++t) {
            const uint64_t zt = z[t];
            const double cr = c[t].x;
#pragma unroll
            for (int u = 0; u < kObsOut; ++u) wr[u] = __fma_rn(parity_sign(src[u], zt), cr, wr[u]);
        }
        for (int t = r_end; t < i_end; ++t) {
            const uint64_t zt = z[t];
            const double ci = c[t].y;
#pragma unroll
            for (int u = 0; u < kObsOut; ++u) wi[u] = __fma_rn(parity_sign(src[u], zt), ci, wi[u]);
        }
        for (int t = i_end; t < S.t_end; ++t) {
            const uint64_t zt = z[t];
            const double2 ct = c[t];
#pragma unroll
            for (int u = 0; u < kObsOut; ++u) {
                const double sg = parity_sign(src[u], zt);
                wr[u] = __fma_rn(sg, ct.x, wr[u]);
                wi[u] = __fma_rn(sg, ct.y, wi[u]);
            }
        }
    }
#pragma unroll
    for (int u = 0; u < kObsOut; ++u) {
        acc[u].x = __fma_rn(wr[u], v[u].x, acc[u].x);
        acc[u].x = __fma_rn(-wi[u], v[u].y, acc[u].x);
        acc[u].y = __fma_rn(wr[u], v[u].y, acc[u].y);
        acc[u].y = __fma_rn(wi[u], v[u].x, acc[u].y);
    }
}

template <typename A, typename I>
__global__ __launch_bounds__(kT) void observable_kernel(const A* __restrict__ in, A* __restrict__ out,
                                                        uint64_t dim, const GroupDev* __restrict__ groups,
                                                        int n_groups, const SubDev* __restrict__ subs,
                                                        const uint64_t* __restrict__ z,
                                                        const double2* __restrict__ c) {
    const uint64_t q = static_cast<uint64_t>(gridDim.x) * kT;
    const int full = n_groups - n_groups % kObsAhead;
    for (uint64_t b = static_cast<uint64_t>(blockIdx.x) * kT + threadIdx.x; b < dim; b += q * kObsOut) {
        // Outputs past the end read the first output's partners and are not stored.
        I j[kObsOut];
#pragma unroll
        for (int u = 0; u < kObsOut; ++u) {
            const uint64_t ju = b + static_cast<uint64_t>(u) * q;
            j[u] = static_cast<I>(ju < dim ? ju : b);
        }
        double2 acc[kObsOut];
#pragma unroll
        for (int u = 0; u < kObsOut; ++u) acc[u] = make_double2(0.0, 0.0);
        for (int g0 = 0; g0 < full; g0 += kObsAhead) {
            double2 v[kObsAhead][kObsOut];
#pragma unroll
            for (int a = 0; a < kObsAhead; ++a) {
                const I x = static_cast<I>(groups[g0 + a].x);
#pragma unroll
                for (int u = 0; u < kObsOut; ++u) v[a][u] = load_amp(in, j[u] ^ x);
            }
#pragma unroll
            for (int a = 0; a < kObsAhead; ++a) {
                const GroupDev G = groups[g0 + a];
                I src[kObsOut];
#pragma unroll
                for (int u = 0; u < kObsOut; ++u) src[u] = j[u] ^ static_cast<I>(G.x);
                add_group(acc, src, G, subs, z, c, v[a]);
            }
        }
        for (int g = full; g < n_groups; ++g) {
            const GroupDev G = groups[g];
            I src[kObsOut];
            double2 v[kObsOut];
#pragma unroll
            for (int u = 0; u < kObsOut; ++u) {
                src[u] = j[u] ^ static_cast<I>(G.x);
                v[u] = load_amp(in, src[u]);
            }
            add_group(acc, src, G, subs, z, c, v);
        }
#pragma unroll
        for (int u = 0; u < kObsOut; ++u) {
            if (b + static_cast<uint64_t>(u) * q < dim) {
                A o;
                o.x = acc[u].x;
                o.y = acc[u].y;
                out[j[u]] = o;
            }
        }
    }
}

// H|psi> for states of at least one chunk (2^12 amplitudes): workgroup c
// owns the outputs of chunk c, thread t those at j_r = c 2^12 + t + 256 r
// (r < 16, so the outputs differ in bits 8..11).  The partners of a group
// with flip mask x are chunk c ^ (x >> 12) with the local index XOR-ed by
// x's low bits, and the host orders groups by x >> 12: each run of equal
// high parts stages its partner chunk in LDS once (coalesced loads, widened
// to FP64), and every group of the run reads its partners there — the
// gathered loads of observable_kernel, which the texture-address path bound,
// become one chunk load per run.  Weights: per term one parity per thread at
// j_0 ^ x; the output r's sign of a term with pattern p (its z bits 8..11) is
// (-1)^popcount(r & p), read from the pattern's mask (SubDev::rpat bits
// 16..31); per pattern A = sum of its terms' signed coefficients, then
// acc_r += (+-A) psi[partner_r] — the group weight is summed per pattern
// first.
// acc[r] += (-1)^popcount(r & P) a psi[partner of output r] for a pattern P
// known at compile time (so every sign is a modifier of the multiply-add);
// vb: the byte offset in the staged chunk of output 0's partner, output r's
// partner at vb ^ (r << 12).
template <int P, bool REAL>
__device__ __forceinline__ void accumulate_pattern(double2 (&acc)[kPer], const char* S, uint32_t vb, double2 a) {
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        const double2 v = *reinterpret_cast<const double2*>(S + (vb ^ (static_cast<uint32_t>(r) << kChunkBits)));
        const bool neg = (__builtin_popcount(static_cast<unsigned>(r & P)) & 1) != 0;
        const double ax = neg ? -a.x : a.x;
        acc[r].x = __fma_rn(ax, v.x, acc[r].x);
        acc[r].y = __fma_rn(ax, v.y, acc[r].y);
        if (!REAL) {
            const double ay = neg ? -a.y : a.y;
            acc[r].x = __fma_rn(-ay, v.y, acc[r].x);
            acc[r].y = __fma_rn(ay, v.x, acc[r].y);
        }
    }
}

// Sum over a pattern's terms of c_t (-1)^popcount(src & z_t) (the pattern
// sum of the expectation kernel, with I-wide indices and masks).
template <typename I>
__device__ __forceinline__ double2 pattern_weight(I src, int t0, int t1, const uint64_t* __restrict__ z,
                                                  const double2* __restrict__ c) {
    double ax = 0.0, ay = 0.0;
    for (int t = t0; t < t1; ++t) {
        const uint32_t p = sizeof(I) == 4 ? __popc(static_cast<uint32_t>(src) & static_cast<uint32_t>(z[t]))
                                          : __popcll(static_cast<uint64_t>(src) & z[t]);
        const double sg = __hiloint2double(static_cast<int>((p << 31) | 0x3ff00000u), 0);
        ax = __fma_rn(sg, c[t].x, ax);
        ay = __fma_rn(sg, c[t].y, ay);
    }
    return make_double2(ax, ay);
}

template <typename A, typename I>
__global__ __launch_bounds__(kT) void observable_chunk_kernel(const A* __restrict__ in, A* __restrict__ out,
                                                              const GroupDev* __restrict__ groups, int n_groups,
                                                              const SubDev* __restrict__ subs,
                                                              const uint64_t* __restrict__ z,
                                                              const double2* __restrict__ c) {
    __shared__ double2 S[kChunk];
    const char* Sb = reinterpret_cast<const char*>(S);
    const uint64_t chunk = blockIdx.x;
    const uint32_t t = threadIdx.x;
    const uint64_t j0 = (chunk << kChunkBits) | t;
    double2 acc[kPer];
#pragma unroll
    for (int r = 0; r < kPer; ++r) acc[r] = make_double2(0.0, 0.0);
    // The next run's partner chunk is loaded into registers while the
    // current run computes (the runs are wave-uniform: every thread reads
    // the same group table), so a run's staging waits only for LDS.
    A pre[kPer];
    auto fetch = [&](uint64_t xh) {
        const A* src = in + ((chunk ^ xh) << kChunkBits);
#pragma unroll
        for (int r = 0; r < kPer; ++r) pre[r] = src[t + kT * r];
    };
    if (n_groups > 0) fetch(groups[0].x >> kChunkBits);
    for (int g = 0; g < n_groups;) {
        const uint64_t xh = groups[g].x >> kChunkBits;
        int g_end = g + 1;
        while (g_end < n_groups && (groups[g_end].x >> kChunkBits) == xh) ++g_end;
        __syncthreads();  // the previous run's reads are done
#pragma unroll
        for (int r = 0; r < kPer; ++r) S[t + kT * r] = make_double2(pre[r].x, pre[r].y);
        __syncthreads();
        if (g_end < n_groups) fetch(groups[g_end].x >> kChunkBits);
        for (; g < g_end; ++g) {
            const GroupDev G = groups[g];
            const uint32_t xlo = static_cast<uint32_t>(G.x) & (kChunk - 1);
            const uint32_t vb = (t << 4) ^ (xlo << 4);   // output 0's partner: local index t ^ xlo
            const I src = static_cast<I>(j0 ^ G.x);
            for (int si = G.s_begin; si < G.s_end; ++si) {
                // One 16-byte scalar load: rpat (pattern and sign mask), the
                // term range and the kind counts.
                const int4 w = reinterpret_cast<const int4*>(subs)[si];
                const double2 a = pattern_weight<I>(src, w.y, w.z, z, c);
                const bool real = w.y + (w.w & 0xffff) == w.z;
                const int pat = w.x & (kPer - 1);
#define QHIP_PAT(P)                                                                    \
    case P:                                                                            \
        if (real) accumulate_pattern<P, true>(acc, Sb, vb, a);                         \
        else accumulate_pattern<P, false>(acc, Sb, vb, a);                             \
        break;
                switch (pat) {
                    QHIP_PAT(0) QHIP_PAT(1) QHIP_PAT(2) QHIP_PAT(3) QHIP_PAT(4) QHIP_PAT(5) QHIP_PAT(6) QHIP_PAT(7)
                    QHIP_PAT(8) QHIP_PAT(9) QHIP_PAT(10) QHIP_PAT(11) QHIP_PAT(12) QHIP_PAT(13) QHIP_PAT(14)
                    QHIP_PAT(15)
                }
#undef QHIP_PAT
            }
        }
    }
#pragma unroll
    for (int r = 0; r < kPer; ++r) {
        A o;
        o.x = acc[r].x;
        o.y = acc[r].y;
        out[j0 + kT * r] = o;
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

// The chunk partials of <bra|H|ket> are summed in slices of kSliceChunks:
// block b adds chunks [b * kSliceChunks, (b + 1) * kSliceChunks) in the order
// of sum_partials (thread t takes t, t + 256, ... then a fixed tree), and
// sum_partials then adds the slice sums.  The no-allocation path reduces its
// slices one launch at a time with the same kernel, so both agree bit for bit.
constexpr int64_t kSliceChunks = 16384;
static_assert(kSliceChunks % kT == 0, "slices keep chunk residues mod kT");

__global__ __launch_bounds__(kT) void reduce_slices(const double2* __restrict__ in, int64_t n,
                                                    double2* __restrict__ out) {
    __shared__ double rx[kT], ry[kT];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kSliceChunks;
    const int64_t m = n - base < kSliceChunks ? n - base : kSliceChunks;
    double ax = 0.0, ay = 0.0;
    for (int64_t i = threadIdx.x; i < m; i += kT) { ax += in[base + i].x; ay += in[base + i].y; }
    rx[threadIdx.x] = ax;
    ry[threadIdx.x] = ay;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) { rx[threadIdx.x] += rx[threadIdx.x + w]; ry[threadIdx.x] += ry[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = make_double2(rx[0], ry[0]);
}

struct DeviceGroups {
    const GroupDev* g;
    const SubDev* s;
    const uint64_t* z;
    const double2* c;
    int n_groups;
};

// Byte offsets of the four tables packed in one buffer (16-byte aligned).
struct TableLayout {
    size_t gb, sb, zb, cb;
    size_t so() const { return gb; }
    size_t zo() const { return gb + sb; }
    size_t co() const { return (gb + sb + zb + 15) & ~size_t{15}; }
    size_t total() const { return co() + cb; }
};

static TableLayout layout_of(const Grouped& gr) {
    return {gr.groups.size() * sizeof(GroupDev), gr.subs.size() * sizeof(SubDev), gr.z.size() * sizeof(uint64_t),
            gr.c.size() * sizeof(double2)};
}

static DeviceGroups view_groups(char* base, const Grouped& gr) {
    const TableLayout L = layout_of(gr);
    return {reinterpret_cast<const GroupDev*>(base), reinterpret_cast<const SubDev*>(base + L.so()),
            reinterpret_cast<const uint64_t*>(base + L.zo()), reinterpret_cast<const double2*>(base + L.co()),
            static_cast<int>(gr.groups.size())};
}

static void pack_tables(char* host, const Grouped& gr) {
    const TableLayout L = layout_of(gr);
    std::memcpy(host, gr.groups.data(), L.gb);
    std::memcpy(host + L.so(), gr.subs.data(), L.sb);
    std::memcpy(host + L.zo(), gr.z.data(), L.zb);
    std::memcpy(host + L.co(), gr.c.data(), L.cb);
}

static void copy_tables(char* dev, const Grouped& gr) {
    auto& C = ctx();
    const TableLayout L = layout_of(gr);
    QHIP_CHECK(hipMemcpyAsync(dev, gr.groups.data(), L.gb, hipMemcpyHostToDevice, C.stream));
    QHIP_CHECK(hipMemcpyAsync(dev + L.so(), gr.subs.data(), L.sb, hipMemcpyHostToDevice, C.stream));
    QHIP_CHECK(hipMemcpyAsync(dev + L.zo(), gr.z.data(), L.zb, hipMemcpyHostToDevice, C.stream));
    QHIP_CHECK(hipMemcpyAsync(dev + L.co(), gr.c.data(), L.cb, hipMemcpyHostToDevice, C.stream));
}

static DeviceGroups upload_groups(const Grouped& gr) {
    char* base = static_cast<char*>(ctx().ops.ensure(layout_of(gr).total() + 64));
    copy_tables(base, gr);
    return view_groups(base, gr);
}

template <bool HERM>
static void launch_transition(const qhip_state* bra, const qhip_state* ket, const DeviceGroups& dg, int64_t blocks,
                              double2* part, uint64_t chunk0) {
    auto& C = ctx();
    const int same = bra->d == ket->d ? 1 : 0;
    const dim3 grid(static_cast<uint32_t>(blocks));
    auto go = [&](auto full_tag, const auto* b, const auto* k) {
        constexpr bool FULL = decltype(full_tag)::value;
        using A = std::remove_cv_t<std::remove_pointer_t<decltype(b)>>;
        hipLaunchKernelGGL((transition_kernel<A, HERM, FULL>), grid, dim3(kT), 0, C.stream, b, k, ket->dim(), dg.g,
                           dg.n_groups, dg.s, dg.z, dg.c, part, chunk0, same);
    };
    const bool full = ket->dim() >= kChunk;
    if (ket->precision == QHIP_C128) {
        const auto* b = static_cast<const double2*>(bra->d);
        const auto* k = static_cast<const double2*>(ket->d);
        if (full) go(std::true_type{}, b, k);
        else go(std::false_type{}, b, k);
    } else {
        const auto* b = static_cast<const float2*>(bra->d);
        const auto* k = static_cast<const float2*>(ket->d);
        if (full) go(std::true_type{}, b, k);
        else go(std::false_type{}, b, k);
    }
}

static void launch_observable(const qhip_state* in, qhip_state* out, const DeviceGroups& dg) {
    auto& C = ctx();
    const uint64_t dim = in->dim();
    const int blocks = static_cast<int>(std::min<uint64_t>((dim + kT * kObsOut - 1) / (kT * kObsOut),
                                                           static_cast<uint64_t>(std::max(1, C.n_cu)) * 64));
    auto go = [&](auto amp_tag, auto index_tag) {
        using A = decltype(amp_tag);
        using I = decltype(index_tag);
        hipLaunchKernelGGL((observable_kernel<A, I>), dim3(blocks), dim3(kT), 0, C.stream,
                           static_cast<const A*>(in->d), static_cast<A*>(out->d), dim, dg.g, dg.n_groups, dg.s, dg.z,
                           dg.c);
    };
    if (dim >= kChunk) {
        if (dim / kChunk > 0x7fffffffu) fail(QHIP_CAPABILITY, "apply_observable: state too large for one grid");
        const dim3 grid(static_cast<uint32_t>(dim / kChunk));
        auto chunked = [&](auto amp_tag, auto index_tag) {
            using A = decltype(amp_tag);
            using I = decltype(index_tag);
            hipLaunchKernelGGL((observable_chunk_kernel<A, I>), grid, dim3(kT), 0, C.stream,
                               static_cast<const A*>(in->d), static_cast<A*>(out->d), dg.g, dg.n_groups, dg.s, dg.z,
                               dg.c);
        };
        const bool small32 = dim <= (uint64_t{1} << 32);
        if (in->precision == QHIP_C128) {
            if (small32) chunked(double2{}, uint32_t{});
            else chunked(double2{}, uint64_t{});
        } else {
            if (small32) chunked(float2{}, uint32_t{});
            else chunked(float2{}, uint64_t{});
        }
        QHIP_CHECK(hipGetLastError());
        return;
    }
    const bool small = dim <= (uint64_t{1} << 32);
    if (in->precision == QHIP_C128) {
        if (small) go(double2{}, uint32_t{});
        else go(double2{}, uint64_t{});
    } else {
        if (small) go(float2{}, uint32_t{});
        else go(float2{}, uint64_t{});
    }
    QHIP_CHECK(hipGetLastError());
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
    const int64_t slices = (blocks + kSliceChunks - 1) / kSliceChunks;
    double2* part = static_cast<double2*>(C.red.ensure(sizeof(double2) * (blocks + slices + 1)));
    bool herm = bra->d == ket->d;
    for (int64_t t = 0; herm && t < obs->n_terms; ++t) herm = obs->terms[t].coeff_im == 0.0;
    if (herm) launch_transition<true>(bra, ket, dg, blocks, part, 0);
    else launch_transition<false>(bra, ket, dg, blocks, part, 0);
    QHIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(reduce_slices, dim3(static_cast<uint32_t>(slices)), dim3(kT), 0, C.stream, part, blocks,
                       part + blocks);
    QHIP_CHECK(hipGetLastError());
    const cplx r = finish(part + blocks, slices);
    return herm ? cplx(r.real(), 0.0) : r;
}

// ── <bra|H|ket> with no device allocation ───────────────────────────────────
// For callers that cannot allocate (graph capture, serving loops): the group
// tables, the chunk partials of one slice and the slice sums live in a fixed
// module-scope device array.  The ket runs one slice of kSliceChunks chunks
// per launch, each slice reduced by reduce_slices into its own sum, and the
// slice sums are added as transition() adds them, so the result is bit-equal
// to transition().  Every call in the process shares the array (the library
// lock serialises them); callers that bypass that lock on one GPU race.

constexpr int kNoAllocMaxGroups = 4096;
constexpr int kNoAllocMaxTerms = 16384;              // patterns never outnumber terms
constexpr int64_t kNoAllocMaxSlices = 4096;          // 2^38 amplitudes
constexpr size_t kNaTablesBytes = sizeof(GroupDev) * kNoAllocMaxGroups + sizeof(SubDev) * kNoAllocMaxTerms +
                                  sizeof(uint64_t) * kNoAllocMaxTerms + sizeof(double2) * kNoAllocMaxTerms + 256;
constexpr size_t kNaPartOff = kNaTablesBytes;
constexpr size_t kNaSliceOff = kNaPartOff + sizeof(double2) * kSliceChunks;
constexpr size_t kNaOutOff = kNaSliceOff + sizeof(double2) * kNoAllocMaxSlices;
constexpr size_t kNoAllocBytes = kNaOutOff + 256;    // 1,081,856 bytes
static_assert(kNoAllocBytes == 1081856, "documented size");
static_assert(kNaPartOff % 256 == 0 && kNaSliceOff % 256 == 0 && kNaOutOff % 256 == 0, "aligned sections");

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

cplx transition_noalloc(const qhip_state* bra, const qhip_state* ket, const qhip_observable* obs) {
    const int n = ket->n;
    const Grouped gr = group_terms(obs, n, "pauli_transition");
    if (gr.groups.empty()) return cplx(0.0, 0.0);
    if (gr.groups.size() > static_cast<size_t>(kNoAllocMaxGroups) || gr.z.size() > static_cast<size_t>(kNoAllocMaxTerms))
        fail(QHIP_INVALID, "qhip_transition_noalloc: the observable has " + std::to_string(gr.groups.size()) +
                               " flip-mask groups and " + std::to_string(gr.z.size()) +
                               " terms; the fixed scratch holds at most " + std::to_string(kNoAllocMaxGroups) +
                               " groups and " + std::to_string(kNoAllocMaxTerms) + " terms");
    if (layout_of(gr).total() > kNaTablesBytes) fail(QHIP_RUNTIME, "qhip_transition_noalloc: table layout overflow");
    auto& C = ctx();
    unsigned char* base = noalloc_scratch();
    auto* d_part = reinterpret_cast<double2*>(base + kNaPartOff);
    auto* d_slices = reinterpret_cast<double2*>(base + kNaSliceOff);
    auto* d_out = reinterpret_cast<double2*>(base + kNaOutOff);
    const uint64_t dim = ket->dim();
    const int64_t chunks = static_cast<int64_t>((dim + kChunk - 1) / kChunk);
    const int64_t slices = (chunks + kSliceChunks - 1) / kSliceChunks;
    if (slices > kNoAllocMaxSlices)
        fail(QHIP_INVALID, "qhip_transition_noalloc: " + std::to_string(ket->n) +
                               " qubits exceed the fixed scratch's 2^38 amplitudes");
    copy_tables(reinterpret_cast<char*>(base), gr);
    const DeviceGroups dg = view_groups(reinterpret_cast<char*>(base), gr);
    bool herm = bra->d == ket->d;
    for (int64_t t = 0; herm && t < obs->n_terms; ++t) herm = obs->terms[t].coeff_im == 0.0;
    for (int64_t sl = 0; sl < slices; ++sl) {
        const int64_t c0 = sl * kSliceChunks;
        const int64_t nb = std::min<int64_t>(kSliceChunks, chunks - c0);
        if (herm) launch_transition<true>(bra, ket, dg, nb, d_part, static_cast<uint64_t>(c0));
        else launch_transition<false>(bra, ket, dg, nb, d_part, static_cast<uint64_t>(c0));
        hipLaunchKernelGGL(reduce_slices, dim3(1), dim3(kT), 0, C.stream, d_part, nb, d_slices + sl);
        QHIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(sum_partials, dim3(1), dim3(kT), 0, C.stream, d_slices, slices, d_out);
    QHIP_CHECK(hipGetLastError());
    double2 r;
    QHIP_CHECK(hipMemcpyAsync(&r, d_out, sizeof(double2), hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));
    return herm ? cplx(r.x, 0.0) : cplx(r.x, r.y);
}

// An observable grouped on the host with its tables already on the device
// (in ctx().obs), so that launching H|in> needs no host work or copy.
struct PreparedObservable {
    Grouped gr;         // host tables (alive while the copies may run)
    DeviceGroups dg{};
};

std::shared_ptr<PreparedObservable> prepare_observable(const qhip_observable* obs, int n) {
    auto po = std::make_shared<PreparedObservable>();
    po->gr = group_terms(obs, n, "run_gradient");
    if (!po->gr.groups.empty()) {
        char* base = static_cast<char*>(ctx().obs.ensure(layout_of(po->gr).total() + 64));
        copy_tables(base, po->gr);
        po->dg = view_groups(base, po->gr);
    }
    return po;
}

void launch_prepared_observable(const PreparedObservable& po, const qhip_state* in, qhip_state* out) {
    if (po.gr.groups.empty()) {
        QHIP_CHECK(hipMemsetAsync(out->d, 0, out->dim() * out->amp_bytes(), ctx().stream));
        return;
    }
    launch_observable(in, out, po.dg);
}

void apply_observable(const qhip_state* in, const qhip_observable* obs, qhip_state* out) {
    const int n = in->n;
    const Grouped gr = group_terms(obs, n, "apply_observable");
    auto& C = ctx();
    if (gr.groups.empty()) {
        QHIP_CHECK(hipMemsetAsync(out->d, 0, out->dim() * out->amp_bytes(), C.stream));
        QHIP_CHECK(hipStreamSynchronize(C.stream));
        return;
    }
    launch_observable(in, out, upload_groups(gr));
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
    std::vector<size_t> off(static_cast<size_t>(n_obs) + 1, 0);
    for (int i = 0; i < n_obs; ++i) {
        gs.push_back(group_terms(&obs[i], n_qubits[i], "run_gradient"));
        off[i + 1] = off[i] + ((layout_of(gs.back()).total() + 255) & ~size_t{255});
    }
    auto a = std::make_unique<ObsArena>();
    std::vector<char> host(off[n_obs] + 64);
    if (hipMalloc(&a->d, host.size()) != hipSuccess) {
        (void)hipGetLastError();
        a->d = nullptr;
        fail(QHIP_NOMEM, "batch: observable table allocation failed");
    }
    char* dev = static_cast<char*>(a->d);
    for (int i = 0; i < n_obs; ++i) {
        pack_tables(host.data() + off[i], gs[i]);
        a->per.push_back(view_groups(dev + off[i], gs[i]));
    }
    QHIP_CHECK(hipMemcpy(a->d, host.data(), host.size(), hipMemcpyHostToDevice));
    return a.release();
}

void obs_arena_destroy(ObsArena* a) { delete a; }

int obs_arena_groups(const ObsArena* a, int i) { return a->per[static_cast<size_t>(i)].n_groups; }

// out = H_i in, stream-ordered (no host synchronisation).
void obs_apply_async(const ObsArena* a, int i, const qhip_state* in, qhip_state* out) {
    launch_observable(in, out, a->per[i]);
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
