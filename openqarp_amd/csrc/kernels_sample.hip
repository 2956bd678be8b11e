// kernels_sample.hip — terminal sampling, identical to libstdc++'s
// std::discrete_distribution over std::mt19937 as sample_terminal calls it
// (src/qarpx/src/simulator/qarp_simulator.cpp:751-783).
//
// libstdc++ builds the CDF sequentially: S = (((0 + p_0) + p_1) + ...),
// q_i = p_i / S, cp_i = cp_{i-1} + q_i, last entry forced to 1.0, and a draw
// u returns lower_bound(cp, u).  A parallel scan rounds differently, so the
// device reproduces the sequential recurrence exactly:
//
//   While a running sum s stays inside one binade [2^e, 2^(e+1)), every
//   double there is an integer multiple of ulp = 2^(e-52), so
//   s_i = RN(s_{i-1} + v_i) = s_{i-1} + ulp * r_i with r_i = round(v_i / ulp)
//   — exact integer arithmetic, independent of s_{i-1} except on an exact
//   rounding tie (fraction 1/2), where ties-to-even looks at s's parity.
//
// Pass A (device): approximate FP64 chunk sums -> a guess of each chunk's
//   binade.  Pass B (device): per chunk, the exact integer sum of r_i on the
//   guessed grid plus tie / overflow flags.  Resolve (host, one step per
//   chunk): start from s = 0 and advance s by ulp * R_c when s really lies in
//   the guessed binade, no flag is set and the end stays below 2^(e+1);
//   otherwise download that chunk and run the float recurrence serially.
//   The resolve is exact either way; the guesses only decide how often the
//   serial path runs (chunks that cross a power of two, about one per
//   binade).  It runs once over p (giving S) and once over q = p / S (giving
//   the exact CDF at every chunk boundary).
// Search (device, one thread per draw): binary search over the exact chunk
//   ends, then the float recurrence from the chunk's exact start.
//
// Every sum and product here is separately rounded: contraction to FMA is
// off for this file.
#pragma clang fp contract(off)

#include "common.h"

#include <cmath>
#include <random>

namespace qhip {

constexpr int kST = 256;
constexpr int kPerS = 32;
constexpr uint64_t kSChunk = kST * kPerS;   // 8192 amplitudes per chunk
constexpr int kSerial = -100000;

__device__ __forceinline__ double prob(const double2* s, uint64_t i) {
    const double2 a = s[i];
    const double xx = a.x * a.x;
    const double yy = a.y * a.y;
    return xx + yy;                       // std::norm: x*x + y*y
}
__device__ __forceinline__ double prob(const float2* s, uint64_t i) {
    const float2 a = s[i];
    const double x = a.x, y = a.y;
    const double xx = x * x;
    const double yy = y * y;
    return xx + yy;
}

template <typename A>
__global__ __launch_bounds__(kST) void chunk_sums(const A* __restrict__ s, uint64_t dim, double* __restrict__ out) {
    __shared__ double red[kST];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSChunk;
    double acc = 0.0;
    for (int r = 0; r < kPerS; ++r) {
        const uint64_t i = base + static_cast<uint64_t>(r) * kST + threadIdx.x;
        if (i < dim) acc += prob(s, i);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kST / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// Exact integer increments on the guessed grid.  flags: bit0 tie, bit1 overflow.
template <typename A>
__global__ __launch_bounds__(kST) void chunk_increments(const A* __restrict__ s, uint64_t dim,
                                                        const int32_t* __restrict__ grid_e, double divisor,
                                                        int use_div, unsigned long long* __restrict__ R,
                                                        int32_t* __restrict__ flags) {
    __shared__ unsigned long long rr[kST];
    __shared__ int rf[kST];
    const int e = grid_e[blockIdx.x];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kSChunk;
    unsigned long long acc = 0;
    int flag = 0;
    if (e != kSerial) {
        for (int r = 0; r < kPerS; ++r) {
            const uint64_t i = base + static_cast<uint64_t>(r) * kST + threadIdx.x;
            if (i >= dim) continue;
            double v = prob(s, i);
            if (use_div) v = v / divisor;
            const double x = ldexp(v, 52 - e);          // exact power-of-two scaling
            // A chunk that stays in the binade has every term below 2^e, i.e.
            // x < 2^52; anything larger (or NaN) sends the chunk to the serial
            // path.  Its whole sum is below 2^52 too, so a thread whose
            // partial reaches 2^53 flags the chunk as well: the 256 partials
            // then stay below 2^61 and the block sum cannot wrap.
            if (!(x < 4503599627370496.0)) { flag |= 2; continue; }
            const double fl = floor(x);
            const double fr = x - fl;                   // exact (Sterbenz)
            acc += static_cast<unsigned long long>(fl) + (fr > 0.5 ? 1u : 0u);
            if (fr == 0.5) flag |= 1;
            if (acc >= (1ULL << 53)) {
                flag |= 2;
                acc = 0;
            }
        }
    }
    rr[threadIdx.x] = acc;
    rf[threadIdx.x] = flag;
    __syncthreads();
    for (int w = kST / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) { rr[threadIdx.x] += rr[threadIdx.x + w]; rf[threadIdx.x] |= rf[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        R[blockIdx.x] = rr[0];
        flags[blockIdx.x] = e == kSerial ? 4 : rf[0];
    }
}

template <typename A>
__global__ __launch_bounds__(kST) void search_draws(const A* __restrict__ s, uint64_t dim, double S,
                                                    const double* __restrict__ starts,
                                                    const double* __restrict__ ends, int64_t n_chunks,
                                                    const double* __restrict__ u, int64_t n_draws,
                                                    unsigned long long* __restrict__ out) {
    const int64_t d = static_cast<int64_t>(blockIdx.x) * kST + threadIdx.x;
    if (d >= n_draws) return;
    const double x = u[d];
    // lower_bound over chunk ends (the last one is 1.0 > x).
    int64_t lo = 0, count = n_chunks;
    while (count > 0) {
        const int64_t step = count / 2, mid = lo + step;
        if (ends[mid] < x) { lo = mid + 1; count -= step + 1; }
        else count = step;
    }
    if (lo >= n_chunks) lo = n_chunks - 1;
    const uint64_t b = static_cast<uint64_t>(lo) * kSChunk;
    const uint64_t e = b + kSChunk < dim ? b + kSChunk : dim;
    double run = starts[lo];
    uint64_t found = e - 1;
    for (uint64_t i = b; i < e; ++i) {
        const double q = prob(s, i) / S;
        run = (i == 0) ? q : run + q;
        if (i == dim - 1) run = 1.0;
        if (!(run < x)) { found = i; break; }
    }
    out[d] = found;
}

template <typename A>
__global__ void probs_kernel(const A* __restrict__ s, uint64_t dim, double* __restrict__ out) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < dim; i += stride)
        out[i] = prob(s, i);
}

// EXACT readout: indices with |amp|^2 above a threshold, compacted in
// ascending index order (the np.nonzero of _exact_result).  Block b owns
// amplitudes [4096 b, 4096 b + 4096): a count pass, an exclusive scan of
// the block counts, then a scatter pass that ranks kept lanes with wave
// ballots in index order.
constexpr uint64_t kXChunk = kST * 16;

// Two magnitudes, each the one its reference path computes.
//
// hypot_libm: std::hypot of the C library, which the column compaction of
// transition_columns.cu:271-276 calls (and numpy's hypot).  It follows the
// published algorithm that library uses (Borges, "An improved algorithm for
// hypot(a, b)", 2019, the corrected form without fused multiply-add):
// h = sqrt(x^2 + y^2) rounded, then one Newton step h -= (h^2 - x^2 - y^2)/(2h)
// with the residual rearranged about h - y or h - x so that nothing cancels.
// Arguments below 2^-450 are scaled by 2^600 (exact, and every step commutes
// with it) so that no square underflows.  This file compiles with
// contraction off, so none of these products is fused.
__device__ __forceinline__ double hypot_libm(double x, double y) {
    double ax = fabs(x), ay = fabs(y);
    if (ax < ay) {
        const double t = ax;
        ax = ay;
        ay = t;
    }
    if (ay == 0.0) return ax;
    double scale = 1.0;
    if (ax < 0x1p-450) {
        ax *= 0x1p600;
        ay *= 0x1p600;
        scale = 0x1p-600;
    }
    double h = sqrt(ax * ax + ay * ay);
    double t1, t2;
    if (h <= 2.0 * ay) {
        const double d = h - ay;
        t1 = ax * (2.0 * d - ax);
        t2 = (d - 2.0 * (ax - ay)) * d;
    } else {
        const double d = h - ax;
        t1 = 2.0 * d * (ax - 2.0 * ay);
        t2 = (4.0 * d - ay) * ay + d * d;
    }
    h -= (t1 + t2) / (2.0 * h);
    return h * scale;
}

// abs_numpy: numpy's absolute value of a complex128 array, which the EXACT
// readout squares (probs = np.abs(sv) ** 2, src/qarp/engines/_engine.py:
// 100-102): with L = max(|x|, |y|) and S = min, L * sqrt(fma(S/L, S/L, 1))
// (numpy's vectorised complex absolute value; 0 when L is 0).
__device__ __forceinline__ double abs_numpy(double x, double y) {
    const double ax = fabs(x), ay = fabs(y);
    const double l = ax < ay ? ay : ax, sm = ax < ay ? ax : ay;
    if (l == 0.0) return 0.0;
    const double r = sm / l;
    return l * sqrt(fma(r, r, 1.0));
}

__device__ __forceinline__ double born_prob(const double2* s, uint64_t i) {
    const double2 a = s[i];
    const double m = abs_numpy(a.x, a.y);
    return m * m;
}
__device__ __forceinline__ double born_prob(const float2* s, uint64_t i) {
    const float2 a = s[i];
    const double m = abs_numpy(static_cast<double>(a.x), static_cast<double>(a.y));
    return m * m;
}

template <typename A>
__global__ __launch_bounds__(kST) void exact_count(const A* __restrict__ s, uint64_t dim, double thr,
                                                   uint64_t* __restrict__ counts) {
    __shared__ uint32_t wsum[kST / 64];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kXChunk;
    uint32_t n = 0;
    for (int r = 0; r < 16; ++r) {
        const uint64_t i = base + static_cast<uint64_t>(r) * kST + threadIdx.x;
        if (i < dim && born_prob(s, i) > thr) ++n;
    }
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < kST / 64; ++w) t += wsum[w];
        counts[blockIdx.x] = t;
    }
}

// out[i] = in[0] + ... + in[i-1] for i <= n (one block; thread t scans a
// contiguous segment).
__global__ __launch_bounds__(kST) void exclusive_scan_u64(const uint64_t* __restrict__ in, uint64_t n,
                                                          uint64_t* __restrict__ out) {
    __shared__ uint64_t part[kST];
    const uint64_t seg = (n + kST - 1) / kST;
    const uint64_t lo = threadIdx.x * seg, hi = lo + seg < n ? lo + seg : n;
    uint64_t sum = 0;
    for (uint64_t i = lo; i < hi; ++i) sum += in[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t run = 0;
        for (int t = 0; t < kST; ++t) {
            const uint64_t v = part[t];
            part[t] = run;
            run += v;
        }
        out[n] = run;
    }
    __syncthreads();
    uint64_t run = part[threadIdx.x];
    for (uint64_t i = lo; i < hi; ++i) {
        out[i] = run;
        run += in[i];
    }
}

template <typename A>
__global__ __launch_bounds__(kST) void exact_scatter(const A* __restrict__ s, uint64_t dim, double thr,
                                                     const uint64_t* __restrict__ offsets,
                                                     int64_t* __restrict__ keys, double* __restrict__ probs) {
    __shared__ uint32_t wcount[kST / 64];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kXChunk;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t pos = offsets[blockIdx.x];
    for (int r = 0; r < 16; ++r) {
        const uint64_t i = base + static_cast<uint64_t>(r) * kST + threadIdx.x;
        const double p = i < dim ? born_prob(s, i) : 0.0;
        const bool keep = i < dim && p > thr;
        const uint64_t ballot = __ballot(keep);
        const uint32_t rank = __popcll(ballot & ((1ull << lane) - 1ull));
        if (lane == 0) wcount[w] = __popcll(ballot);
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int v = 0; v < kST / 64; ++v) {
            before += v < static_cast<int>(w) ? wcount[v] : 0u;
            total += wcount[v];
        }
        if (keep) {
            keys[pos + before + rank] = static_cast<int64_t>(i);
            probs[pos + before + rank] = p;
        }
        pos += total;
        __syncthreads();   // wcount is rewritten next round
    }
}

// Value compaction (transition columns): indices whose amplitude magnitude
// hypot(re, im) exceeds a tolerance, with the amplitudes, in ascending index
// order — the same count / scan / ballot-ranked scatter as the EXACT readout.
// The count pass also keeps, per block, the largest magnitude it prunes (the
// provider's max_pruned_amplitude diagnostic).
template <typename A>
__global__ __launch_bounds__(kST) void compact_count(const A* __restrict__ s, uint64_t dim, double tol,
                                                     uint64_t* __restrict__ counts, double* __restrict__ pmax) {
    __shared__ uint32_t wsum[kST / 64];
    __shared__ double wmax[kST / 64];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kXChunk;
    uint32_t n = 0;
    double mx = 0.0;
    for (int r = 0; r < 16; ++r) {
        const uint64_t i = base + static_cast<uint64_t>(r) * kST + threadIdx.x;
        if (i >= dim) continue;
        const double m = hypot_libm(static_cast<double>(s[i].x), static_cast<double>(s[i].y));
        if (m > tol) ++n;
        else mx = fmax(mx, m);
    }
    for (int off = 32; off > 0; off >>= 1) {
        n += __shfl_xor(n, off);
        mx = fmax(mx, __shfl_xor(mx, off));
    }
    if ((threadIdx.x & 63) == 0) {
        wsum[threadIdx.x >> 6] = n;
        wmax[threadIdx.x >> 6] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        double m = 0.0;
        for (int w = 0; w < kST / 64; ++w) {
            t += wsum[w];
            m = fmax(m, wmax[w]);
        }
        counts[blockIdx.x] = t;
        pmax[blockIdx.x] = m;
    }
}

__global__ __launch_bounds__(kST) void max_reduce(const double* __restrict__ in, uint64_t n,
                                                  double* __restrict__ out) {
    __shared__ double red[kST];
    double m = 0.0;
    for (uint64_t i = threadIdx.x; i < n; i += kST) m = fmax(m, in[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = kST / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

template <typename A>
__global__ __launch_bounds__(kST) void compact_scatter(const A* __restrict__ s, uint64_t dim, double tol,
                                                       const uint64_t* __restrict__ offsets,
                                                       int64_t* __restrict__ keys, double2* __restrict__ values) {
    __shared__ uint32_t wcount[kST / 64];
    const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kXChunk;
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t pos = offsets[blockIdx.x];
    for (int r = 0; r < 16; ++r) {
        const uint64_t i = base + static_cast<uint64_t>(r) * kST + threadIdx.x;
        double2 v = {0.0, 0.0};
        if (i < dim) v = {static_cast<double>(s[i].x), static_cast<double>(s[i].y)};
        const bool keep = i < dim && hypot_libm(v.x, v.y) > tol;
        const uint64_t ballot = __ballot(keep);
        const uint32_t rank = __popcll(ballot & ((1ull << lane) - 1ull));
        if (lane == 0) wcount[w] = __popcll(ballot);
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (int u = 0; u < kST / 64; ++u) {
            before += u < static_cast<int>(w) ? wcount[u] : 0u;
            total += wcount[u];
        }
        if (keep) {
            keys[pos + before + rank] = static_cast<int64_t>(i);
            values[pos + before + rank] = v;
        }
        pos += total;
        __syncthreads();
    }
}

// ── host side ───────────────────────────────────────────────────────────────

static int binade(double v) {      // e with v in [2^e, 2^(e+1)), v > 0 finite
    int E;
    std::frexp(v, &E);
    return E - 1;
}

static double host_prob(const qhip_state* st, const std::vector<char>& raw, uint64_t j) {
    if (st->precision == QHIP_C128) {
        const double* a = reinterpret_cast<const double*>(raw.data()) + 2 * j;
        const double xx = a[0] * a[0];
        const double yy = a[1] * a[1];
        return xx + yy;
    }
    const float* a = reinterpret_cast<const float*>(raw.data()) + 2 * j;
    const double x = a[0], y = a[1];
    const double xx = x * x;
    const double yy = y * y;
    return xx + yy;
}

// Serial recurrence over one chunk from `start` (exactly as libstdc++).
static double serial_chunk(const qhip_state* st, int64_t c, double start, double divisor, bool use_div,
                           bool first_is_init) {
    const uint64_t b = static_cast<uint64_t>(c) * kSChunk;
    const uint64_t e = std::min(b + kSChunk, st->dim());
    std::vector<char> raw((e - b) * st->amp_bytes());
    QHIP_CHECK(hipMemcpy(raw.data(), static_cast<const char*>(st->d) + b * st->amp_bytes(), raw.size(),
                         hipMemcpyDeviceToHost));
    double s = start;
    for (uint64_t i = b; i < e; ++i) {
        double v = host_prob(st, raw, i - b);
        if (use_div) v = v / divisor;
        s = (first_is_init && i == 0) ? v : s + v;
    }
    return s;
}

struct ChunkScan {
    std::vector<double> starts, ends;
    int64_t serial_chunks = 0;
};

template <typename A>
static ChunkScan resolve(const qhip_state* st, const std::vector<double>& approx, double divisor, bool use_div,
                         bool first_is_init) {
    auto& C = ctx();
    const int64_t nc = static_cast<int64_t>(approx.size());
    // Guess each chunk's binade from approximate running sums.
    std::vector<int32_t> grid(nc, kSerial);
    double run = 0.0;
    for (int64_t c = 0; c < nc; ++c) {
        const double a = use_div ? approx[c] / divisor : approx[c];
        const double lo = run, hi = run + a;
        run = hi;
        if (!(lo > 0.0) || !std::isfinite(hi)) continue;
        const int e0 = binade(lo), e1 = binade(hi);
        if (e0 != e1) continue;
        const double edge_lo = std::ldexp(1.0, e0), edge_hi = std::ldexp(1.0, e0 + 1);
        if ((lo - edge_lo) < 1e-6 * edge_lo || (edge_hi - hi) < 1e-6 * edge_hi) continue;
        grid[c] = e0;
    }
    int32_t* d_grid = static_cast<int32_t*>(C.misc.ensure(nc * (sizeof(int32_t) * 2 + sizeof(uint64_t)) + 64));
    unsigned long long* d_R = reinterpret_cast<unsigned long long*>(d_grid + 2 * nc + 2);
    int32_t* d_flags = d_grid + nc;
    QHIP_CHECK(hipMemcpyAsync(d_grid, grid.data(), nc * sizeof(int32_t), hipMemcpyHostToDevice, C.stream));
    hipLaunchKernelGGL(chunk_increments<A>, dim3(static_cast<uint32_t>(nc)), dim3(kST), 0, C.stream,
                       static_cast<const A*>(st->d), st->dim(), d_grid, divisor, use_div ? 1 : 0, d_R, d_flags);
    QHIP_CHECK(hipGetLastError());
    std::vector<unsigned long long> R(nc);
    std::vector<int32_t> flags(nc);
    QHIP_CHECK(hipMemcpyAsync(R.data(), d_R, nc * sizeof(uint64_t), hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipMemcpyAsync(flags.data(), d_flags, nc * sizeof(int32_t), hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));

    ChunkScan out;
    out.starts.resize(nc);
    out.ends.resize(nc);
    double s = 0.0;
    for (int64_t c = 0; c < nc; ++c) {
        out.starts[c] = s;
        bool fast = flags[c] == 0 && grid[c] != kSerial && s > 0.0 && !(first_is_init && c == 0);
        if (fast) fast = binade(s) == grid[c];
        if (fast) {
            const double ulp = std::ldexp(1.0, grid[c] - 52);
            const unsigned long long K = static_cast<unsigned long long>(s / ulp);   // exact: s is on the grid
            const unsigned long long Kend = K + R[c];
            fast = Kend < (1ULL << 53);
            if (fast) s = static_cast<double>(Kend) * ulp;
        }
        if (!fast) {
            s = serial_chunk(st, c, s, divisor, use_div, first_is_init);
            ++out.serial_chunks;
        }
        out.ends[c] = s;
    }
    return out;
}

template <typename A>
static void sample_impl(const qhip_state* st, const double* uniforms, int64_t n, uint64_t* outcomes) {
    auto& C = ctx();
    const uint64_t dim = st->dim();
    if (dim < 2) {     // discrete_distribution with < 2 weights always yields 0
        for (int64_t i = 0; i < n; ++i) outcomes[i] = 0;
        return;
    }
    const int64_t nc = static_cast<int64_t>((dim + kSChunk - 1) / kSChunk);
    double* d_approx = static_cast<double*>(C.red.ensure(sizeof(double) * nc));
    hipLaunchKernelGGL(chunk_sums<A>, dim3(static_cast<uint32_t>(nc)), dim3(kST), 0, C.stream,
                       static_cast<const A*>(st->d), dim, d_approx);
    QHIP_CHECK(hipGetLastError());
    std::vector<double> approx(nc);
    QHIP_CHECK(hipMemcpyAsync(approx.data(), d_approx, sizeof(double) * nc, hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));

    // S = accumulate(p, 0.0): the same recurrence started at 0.0.
    const ChunkScan total = resolve<A>(st, approx, 1.0, false, false);
    const double S = total.ends.back();
    if (!(S > 0.0) || !std::isfinite(S))
        fail(QHIP_INVALID, "sample: the state has zero (or non-finite) total probability");
    // cp = partial_sum(p / S), last entry 1.0.
    ChunkScan cdf = resolve<A>(st, approx, S, true, true);
    cdf.ends.back() = 1.0;
    if (n == 0) return;

    char* base = static_cast<char*>(C.misc2.ensure(sizeof(double) * (2 * nc + n) + sizeof(uint64_t) * n + 64));
    double* d_starts = reinterpret_cast<double*>(base);
    double* d_ends = d_starts + nc;
    double* d_u = d_ends + nc;
    unsigned long long* d_out = reinterpret_cast<unsigned long long*>(d_u + n);
    QHIP_CHECK(hipMemcpyAsync(d_starts, cdf.starts.data(), sizeof(double) * nc, hipMemcpyHostToDevice, C.stream));
    QHIP_CHECK(hipMemcpyAsync(d_ends, cdf.ends.data(), sizeof(double) * nc, hipMemcpyHostToDevice, C.stream));
    QHIP_CHECK(hipMemcpyAsync(d_u, uniforms, sizeof(double) * n, hipMemcpyHostToDevice, C.stream));
    const uint32_t blocks = static_cast<uint32_t>((n + kST - 1) / kST);
    hipLaunchKernelGGL(search_draws<A>, dim3(blocks), dim3(kST), 0, C.stream, static_cast<const A*>(st->d), dim, S,
                       d_starts, d_ends, nc, d_u, n, d_out);
    QHIP_CHECK(hipGetLastError());
    QHIP_CHECK(hipMemcpyAsync(outcomes, d_out, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));
}

// run()'s inline CDF walk over one state (qarp_simulator.cpp:1151-1163):
// total = the sequential sum of p_i from 0.0, threshold = u * total, the first
// index whose sequential running sum reaches it (else the last).  The exact
// chunk ends of that running sum come from resolve(); only the chunk that
// holds the threshold is walked term by term (on the host, from its exact
// start).
template <typename A>
static uint64_t walk_impl(const qhip_state* st, double u) {
    auto& C = ctx();
    const uint64_t dim = st->dim();
    const int64_t nc = static_cast<int64_t>((dim + kSChunk - 1) / kSChunk);
    double* d_approx = static_cast<double*>(C.red.ensure(sizeof(double) * nc));
    hipLaunchKernelGGL(chunk_sums<A>, dim3(static_cast<uint32_t>(nc)), dim3(kST), 0, C.stream,
                       static_cast<const A*>(st->d), dim, d_approx);
    QHIP_CHECK(hipGetLastError());
    std::vector<double> approx(nc);
    QHIP_CHECK(hipMemcpyAsync(approx.data(), d_approx, sizeof(double) * nc, hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));
    const ChunkScan run = resolve<A>(st, approx, 1.0, false, false);
    const double threshold = u * run.ends.back();
    int64_t c = 0;
    while (c < nc && run.ends[c] < threshold) ++c;
    if (c == nc) return dim - 1;
    const uint64_t b = static_cast<uint64_t>(c) * kSChunk;
    const uint64_t e = std::min(b + kSChunk, dim);
    std::vector<char> raw((e - b) * st->amp_bytes());
    QHIP_CHECK(hipMemcpy(raw.data(), static_cast<const char*>(st->d) + b * st->amp_bytes(), raw.size(),
                         hipMemcpyDeviceToHost));
    double acc = run.starts[c];
    for (uint64_t i = b; i < e; ++i) {
        acc = acc + host_prob(st, raw, i - b);
        if (acc >= threshold) return i;
    }
    return dim - 1;
}

uint64_t walk_draw(const qhip_state* s, double u) {
    return s->precision == QHIP_C128 ? walk_impl<double2>(s, u) : walk_impl<float2>(s, u);
}

void sample(const qhip_state* s, const double* uniforms, int64_t n, uint64_t* out) {
    if (s->precision == QHIP_C128) sample_impl<double2>(s, uniforms, n, out);
    else sample_impl<float2>(s, uniforms, n, out);
}

void probabilities(const qhip_state* s, double* out) {
    auto& C = ctx();
    const uint64_t dim = s->dim();
    double* d = static_cast<double*>(C.misc.ensure(sizeof(double) * dim));
    const int blocks = static_cast<int>(std::min<uint64_t>((dim + kST - 1) / kST, 65535));
    if (s->precision == QHIP_C128)
        hipLaunchKernelGGL(probs_kernel<double2>, dim3(blocks), dim3(kST), 0, C.stream,
                           static_cast<const double2*>(s->d), dim, d);
    else
        hipLaunchKernelGGL(probs_kernel<float2>, dim3(blocks), dim3(kST), 0, C.stream,
                           static_cast<const float2*>(s->d), dim, d);
    QHIP_CHECK(hipGetLastError());
    QHIP_CHECK(hipMemcpyAsync(out, d, sizeof(double) * dim, hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));
}

int64_t exact(const qhip_state* s, double threshold, int64_t* keys, double* probs, int64_t capacity) {
    auto& C = ctx();
    const uint64_t dim = s->dim();
    const uint64_t nb = (dim + kXChunk - 1) / kXChunk;
    uint64_t* counts = static_cast<uint64_t*>(C.red.ensure(sizeof(uint64_t) * (2 * nb + 1)));
    uint64_t* offsets = counts + nb;
    if (s->precision == QHIP_C128)
        hipLaunchKernelGGL(exact_count<double2>, dim3(static_cast<uint32_t>(nb)), dim3(kST), 0, C.stream,
                           static_cast<const double2*>(s->d), dim, threshold, counts);
    else
        hipLaunchKernelGGL(exact_count<float2>, dim3(static_cast<uint32_t>(nb)), dim3(kST), 0, C.stream,
                           static_cast<const float2*>(s->d), dim, threshold, counts);
    hipLaunchKernelGGL(exclusive_scan_u64, dim3(1), dim3(kST), 0, C.stream, counts, nb, offsets);
    QHIP_CHECK(hipGetLastError());
    uint64_t total = 0;
    QHIP_CHECK(hipMemcpyAsync(&total, offsets + nb, sizeof(uint64_t), hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));
    if (!keys || !probs) return static_cast<int64_t>(total);
    if (static_cast<uint64_t>(capacity) < total)
        fail(QHIP_INVALID, "qhip_exact: capacity " + std::to_string(capacity) + " is below the " +
                               std::to_string(total) + " retained outcomes");
    if (total == 0) return 0;
    char* out = static_cast<char*>(C.misc.ensure(16 * total));
    int64_t* d_keys = reinterpret_cast<int64_t*>(out);
    double* d_probs = reinterpret_cast<double*>(out + 8 * total);
    if (s->precision == QHIP_C128)
        hipLaunchKernelGGL(exact_scatter<double2>, dim3(static_cast<uint32_t>(nb)), dim3(kST), 0, C.stream,
                           static_cast<const double2*>(s->d), dim, threshold, offsets, d_keys, d_probs);
    else
        hipLaunchKernelGGL(exact_scatter<float2>, dim3(static_cast<uint32_t>(nb)), dim3(kST), 0, C.stream,
                           static_cast<const float2*>(s->d), dim, threshold, offsets, d_keys, d_probs);
    QHIP_CHECK(hipGetLastError());
    QHIP_CHECK(hipMemcpyAsync(keys, d_keys, 8 * total, hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipMemcpyAsync(probs, d_probs, 8 * total, hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));
    return static_cast<int64_t>(total);
}

int64_t compact_values(const qhip_state* s, uint64_t n_amps, double tol, int64_t* keys, double* values,
                       int64_t capacity, double* max_pruned) {
    auto& C = ctx();
    if (n_amps > s->dim()) fail(QHIP_INVALID, "qhip_compact: more amplitudes than the state holds");
    if (!(tol >= 0.0) || tol == HUGE_VAL) fail(QHIP_INVALID, "qhip_compact: tolerance must be finite and >= 0");
    const uint64_t dim = n_amps;
    if (max_pruned) *max_pruned = 0.0;
    if (dim == 0) return 0;
    const uint64_t nb = (dim + kXChunk - 1) / kXChunk;
    // counts[nb] | offsets[nb + 1] | block maxima[nb] | overall maximum
    uint64_t* counts = static_cast<uint64_t*>(C.red.ensure(sizeof(uint64_t) * (3 * nb + 2)));
    uint64_t* offsets = counts + nb;
    double* pmax = reinterpret_cast<double*>(offsets + nb + 1);
    if (s->precision == QHIP_C128)
        hipLaunchKernelGGL(compact_count<double2>, dim3(static_cast<uint32_t>(nb)), dim3(kST), 0, C.stream,
                           static_cast<const double2*>(s->d), dim, tol, counts, pmax);
    else
        hipLaunchKernelGGL(compact_count<float2>, dim3(static_cast<uint32_t>(nb)), dim3(kST), 0, C.stream,
                           static_cast<const float2*>(s->d), dim, tol, counts, pmax);
    hipLaunchKernelGGL(exclusive_scan_u64, dim3(1), dim3(kST), 0, C.stream, counts, nb, offsets);
    hipLaunchKernelGGL(max_reduce, dim3(1), dim3(kST), 0, C.stream, pmax, nb, pmax + nb);
    QHIP_CHECK(hipGetLastError());
    uint64_t total = 0;
    QHIP_CHECK(hipMemcpyAsync(&total, offsets + nb, sizeof(uint64_t), hipMemcpyDeviceToHost, C.stream));
    double mp = 0.0;
    QHIP_CHECK(hipMemcpyAsync(&mp, pmax + nb, sizeof(double), hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));
    if (max_pruned) *max_pruned = mp;
    if (!keys || !values) return static_cast<int64_t>(total);
    if (static_cast<uint64_t>(capacity) < total)
        fail(QHIP_INVALID, "qhip_compact: capacity " + std::to_string(capacity) + " is below the " +
                               std::to_string(total) + " retained entries");
    if (total == 0) return 0;
    char* out = static_cast<char*>(C.misc.ensure(24 * total));
    int64_t* d_keys = reinterpret_cast<int64_t*>(out);
    double2* d_vals = reinterpret_cast<double2*>(out + 8 * total);
    if (s->precision == QHIP_C128)
        hipLaunchKernelGGL(compact_scatter<double2>, dim3(static_cast<uint32_t>(nb)), dim3(kST), 0, C.stream,
                           static_cast<const double2*>(s->d), dim, tol, offsets, d_keys, d_vals);
    else
        hipLaunchKernelGGL(compact_scatter<float2>, dim3(static_cast<uint32_t>(nb)), dim3(kST), 0, C.stream,
                           static_cast<const float2*>(s->d), dim, tol, offsets, d_keys, d_vals);
    QHIP_CHECK(hipGetLastError());
    QHIP_CHECK(hipMemcpyAsync(keys, d_keys, 8 * total, hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipMemcpyAsync(values, d_vals, 16 * total, hipMemcpyDeviceToHost, C.stream));
    QHIP_CHECK(hipStreamSynchronize(C.stream));
    return static_cast<int64_t>(total);
}

// generate_canonical<double, 53> over std::mt19937: the toolchain's own
// libstdc++ implementation, the one the reference calls.
void mt_uniforms(uint32_t seed, int64_t n, double* u) {
    std::mt19937 g(seed);
    for (int64_t i = 0; i < n; ++i) u[i] = std::generate_canonical<double, 53>(g);
}

}  // namespace qhip
