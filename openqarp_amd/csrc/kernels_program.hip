// Structured-program kernels (the reference's Program kernels,
// src/qarpx/include/qarpx/simulator/program.h:17-46):
//  - a basis permutation |x> -> |table[x]> of a qubit subset, applied out of
//    place through a scratch state (apply_permutation, program.cpp:610-646);
//  - controlled matrix powers matrix^e_j on the targets wherever control j
//    is |1>, lowered to controlled dense ops of the ordinary tile pass on up
//    to 7 targets, and run as matrix-core GEMMs (squaring chain included) on
//    8..14 (apply_controlled_powers, program.cpp:648-712).
#include "common.h"

#include <algorithm>
#include <bit>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace qhip {
namespace {

constexpr int kMaxTableQubits = 26;       // program.cpp:35
constexpr int kBlock = 256;
constexpr int kMaxPowerTargets = QHIP_MAX_FUSION_QUBITS - 1;   // controlled op = targets + control

struct PermQubits {
    uint8_t q[32];   // local bit b <-> global bit q[b]
};

// Gather form of |x> -> |table[x]>: out[j] = in[(j & keep) | deposit(inv[x(j)])],
// so each thread's write is coalesced and its read is a gathered 8/16-B load.
// A thread takes kUnroll outputs kBlock apart per step and issues all their
// table loads, then all their amplitude loads, before any store: the two
// dependent loads of one output cost one latency per step, not per output.
constexpr int kUnroll = 8;

// KB: the table width rounded up (8, 16 or 26); positions past k hold 63,
// whose bit is 0 in every index and table entry, so the bit loops have a
// fixed trip count and the positions stay in scalar registers.
template <typename A, int KB>
__global__ __launch_bounds__(kBlock) void permute_kernel(const A* __restrict__ in, A* __restrict__ out,
                                                         uint64_t dim, const uint32_t* __restrict__ inv,
                                                         const PermQubits Q, uint64_t keep) {
    const uint64_t step = static_cast<uint64_t>(gridDim.x) * kBlock * kUnroll;
    for (uint64_t j0 = static_cast<uint64_t>(blockIdx.x) * kBlock * kUnroll + threadIdx.x; j0 < dim; j0 += step) {
        uint32_t s[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const uint64_t j = j0 + static_cast<uint64_t>(u) * kBlock;
            uint32_t x = 0;
#pragma unroll
            for (int b = 0; b < KB; ++b) x |= static_cast<uint32_t>((j >> Q.q[b]) & 1u) << b;
            s[u] = j < dim ? inv[x] : 0u;
        }
        A v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const uint64_t j = j0 + static_cast<uint64_t>(u) * kBlock;
            uint64_t i = j & keep;
#pragma unroll
            for (int b = 0; b < KB; ++b) i |= static_cast<uint64_t>((s[u] >> b) & 1u) << Q.q[b];
            if (j < dim) v[u] = in[i];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const uint64_t j = j0 + static_cast<uint64_t>(u) * kBlock;
            if (j < dim) {
                // Non-temporal: the output is not read again by this kernel
                // (measured: plain stores wrote 1.87x the bytes to memory).
                __builtin_nontemporal_store(v[u].x, &out[j].x);
                __builtin_nontemporal_store(v[u].y, &out[j].y);
            }
        }
    }
}

template <typename A>
void launch_permute(const qhip_state* s, void* scratch, const uint32_t* d_inv, const PermQubits& Q, int k,
                    uint64_t keep) {
    auto& C = ctx();
    const uint64_t dim = s->dim();
    const uint64_t want = (dim + kBlock * kUnroll - 1) / (kBlock * kUnroll);
    const int blocks = std::max(1, static_cast<int>(std::min<uint64_t>(want, static_cast<uint64_t>(C.n_cu) * 16)));
    const A* in = static_cast<const A*>(s->d);
    A* out = static_cast<A*>(scratch);
    if (k <= 8)
        hipLaunchKernelGGL((permute_kernel<A, 8>), dim3(blocks), dim3(kBlock), 0, C.stream, in, out, dim, d_inv, Q,
                           keep);
    else if (k <= 16)
        hipLaunchKernelGGL((permute_kernel<A, 16>), dim3(blocks), dim3(kBlock), 0, C.stream, in, out, dim, d_inv, Q,
                           keep);
    else
        hipLaunchKernelGGL((permute_kernel<A, kMaxTableQubits>), dim3(blocks), dim3(kBlock), 0, C.stream, in, out,
                           dim, d_inv, Q, keep);
    QHIP_CHECK(hipGetLastError());
}

// Zero the state, then amplitude b * stride + labels[b] = 1 for each b.
template <typename A>
__global__ __launch_bounds__(kBlock) void set_basis_kernel(A* s, uint64_t dim, const uint64_t* __restrict__ labels,
                                                           int64_t n_labels, int log_stride) {
    const uint64_t step = static_cast<uint64_t>(gridDim.x) * kBlock;
    const uint64_t low = (uint64_t{1} << log_stride) - 1;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < dim; i += step) {
        A v;
        v.x = 0;
        v.y = 0;
        const uint64_t b = i >> log_stride;
        if (b < static_cast<uint64_t>(n_labels) && labels[b] == (i & low)) v.x = 1;
        s[i] = v;
    }
}

using Mat = std::vector<cplx>;   // row-major d x d

// Row-by-row product on split real and imaginary planes, so the inner loop
// is plain vectorisable FP64 (twice as fast as std::complex products at
// d = 512, the same values).
Mat matmul(const Mat& a, const Mat& b, size_t d) {
    std::vector<double> cr(d * d, 0.0), ci(d * d, 0.0), br(d * d), bi(d * d);
    for (size_t i = 0; i < d * d; ++i) {
        br[i] = b[i].real();
        bi[i] = b[i].imag();
    }
    for (size_t y = 0; y < d; ++y) {
        double* r = &cr[y * d];
        double* m = &ci[y * d];
        for (size_t t = 0; t < d; ++t) {
            const double ar = a[y * d + t].real(), ai = a[y * d + t].imag();
            const double* pr = &br[t * d];
            const double* pi = &bi[t * d];
            for (size_t x = 0; x < d; ++x) {
                r[x] += ar * pr[x] - ai * pi[x];
                m[x] += ar * pi[x] + ai * pr[x];
            }
        }
    }
    Mat c(d * d);
    for (size_t i = 0; i < d * d; ++i) c[i] = cplx(cr[i], ci[i]);
    return c;
}

// ── controlled powers on 8..14 targets: a matrix-core GEMM ─────────────────
// Too wide for one tile-pass op (targets + control > 8 operands).  For a
// control bit c, the 2^(n-1-m) groups of 2^m amplitudes whose control bit is
// set are the columns of X (group g, amplitude k at base(g) | off(k): zero
// bits inserted at the sorted targets and control, then the control bit set;
// off(k) deposits k on the targets in operand order), and the kernel forms
// Y = P X on the matrix cores (apply_controlled_powers, program.cpp:676-706,
// one matrix-vector product per group).  A 256-thread workgroup computes a
// 64-row x 64-group tile of Y: chunks of 16 k of P and X are staged in LDS
// (the next chunk's loads are issued before the current one is multiplied),
// each wave owns 32 x 32 outputs as 2 x 2 blocks of 16 x 16 with real and
// imaginary accumulators (four real products per complex block), and Y is
// written row-major (Y[i * groups + g]) into a scratch buffer, from which one
// scatter pass puts it back into the state.  Tiles map to blocks so that the
// row tiles of one group tile run on one XCD (block b runs on XCD b mod 8),
// whose L2 then serves their shared X panel.  The same kernel squares
// matrices on the device: a row-major d x d matrix read as a 2m-qubit
// register (targets on the high m bits, no control) gives Y = P M row-major.
constexpr int kMaxGemmTargets = 14;    // P and its squares: 4 GiB each at 14
constexpr int kGemmT = 256;
constexpr int kTileI = 64, kTileG = 64, kChunkK = 16;

struct GemmGeom {
    uint8_t  tpos[kMaxTableQubits];      // target local bit b <-> register bit tpos[b]
    uint8_t  ins[kMaxTableQubits + 1];   // bits removed from the group number, ascending
    int32_t  n_ins;
    int32_t  m;
    uint64_t cbit;                       // control bit (0: none)
    uint64_t groups;                     // columns of X
    int32_t  row_tiles, group_tiles;
    int32_t  kfast;                      // loaders walk k (targets start at bit 0) or groups
};

// Matrix-core steps: A at lane (i = lane % 16, k = lane / 16), B at lane
// (j = lane % 16, k = lane / 16); D row i of the 16 x 16 result is in
// register r of lane group q = lane / 16 as i = 4r + q (f64) or 4q + r (f32).
template <typename A> struct GemmMfma;
template <> struct GemmMfma<double2> {
    typedef double R;
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 step(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int r, int q) { return 4 * r + q; }
};
template <> struct GemmMfma<float2> {
    typedef float R;
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 step(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int r, int q) { return 4 * q + r; }
};

// The geometry is read through a pointer (scalar loads with wave-uniform
// indices); a by-value kernel argument indexed by run-time bit numbers would
// be copied to scratch memory.
__device__ __forceinline__ uint64_t gemm_deposit(uint64_t k, const GemmGeom* __restrict__ geo, int from, int to) {
    uint64_t o = 0;
    for (int b = from; b < to; ++b) o |= ((k >> b) & 1u) << geo->tpos[b];
    return o;
}

__device__ __forceinline__ uint64_t gemm_base(uint64_t g, const GemmGeom* __restrict__ geo) {
    uint64_t base = g;
    const int n_ins = geo->n_ins;
    for (int b = 0; b < n_ins; ++b) {
        const uint64_t p = geo->ins[b];
        base = ((base >> p) << (p + 1)) | (base & ((uint64_t{1} << p) - 1u));
    }
    return base | geo->cbit;
}

template <typename A>
__global__ __launch_bounds__(kGemmT) void controlled_gemm_kernel(const A* __restrict__ in, const A* __restrict__ P,
                                                                 A* __restrict__ Y, const GemmGeom* __restrict__ gp) {
    using MF = GemmMfma<A>;
    using R = typename MF::R;
    using v4 = typename MF::v4;
    __shared__ A Ps[kChunkK][kTileI];
    __shared__ A Xs[kChunkK][kTileG];
    const uint32_t row_tiles = static_cast<uint32_t>(gp->row_tiles);
    const uint32_t b = blockIdx.x, xcd = b & 7u, local = b >> 3;
    const uint32_t gt = (local / row_tiles) * 8u + xcd;
    const uint32_t rt = local % row_tiles;
    if (gt >= static_cast<uint32_t>(gp->group_tiles)) return;   // the whole block
    const int m = gp->m;
    const uint64_t d = uint64_t{1} << m, groups = gp->groups;
    const bool kfast = gp->kfast != 0;
    const int t = threadIdx.x;

    // Loaders: four P and four X elements per thread and chunk.  P is read
    // along its rows (4 consecutive k per thread); X along k when the targets
    // start at register bit 0, else along the groups.
    const int pi_ = t >> 2, pk = (t & 3) * 4;
    int xg[4], xk[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        if (kfast) { xg[s] = (t >> 4) + 16 * s; xk[s] = t & 15; }
        else { xg[s] = t & 63; xk[s] = (t >> 6) + 4 * s; }
    }
    uint64_t xoff[4];
    bool xok[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const uint64_t g = static_cast<uint64_t>(gt) * kTileG + xg[s];
        xok[s] = g < groups;
        xoff[s] = (xok[s] ? gemm_base(g, gp) : 0) | gemm_deposit(static_cast<uint64_t>(xk[s]), gp, 0, 4);
    }
    const A* prow = P + (static_cast<uint64_t>(rt) * kTileI + pi_) * d + pk;
    A pv0, pv1, pv2, pv3, xv0, xv1, xv2, xv3;
#define QHIP_GEMM_LOAD(K0)                                                         \
    {                                                                              \
        const uint64_t hi_ = gemm_deposit((K0), gp, 4, m);                         \
        pv0 = prow[(K0) + 0];                                                      \
        pv1 = prow[(K0) + 1];                                                      \
        pv2 = prow[(K0) + 2];                                                      \
        pv3 = prow[(K0) + 3];                                                      \
        xv0 = xok[0] ? in[xoff[0] | hi_] : A{};                                    \
        xv1 = xok[1] ? in[xoff[1] | hi_] : A{};                                    \
        xv2 = xok[2] ? in[xoff[2] | hi_] : A{};                                    \
        xv3 = xok[3] ? in[xoff[3] | hi_] : A{};                                    \
    }

    const int w = t >> 6, lane = t & 63, li = lane & 15, lq = lane >> 4;
    const int wi = (w & 1) * 32, wg = (w >> 1) * 32;
    v4 accr[2][2], acci[2][2];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int gb = 0; gb < 2; ++gb)
#pragma unroll
            for (int r = 0; r < 4; ++r) { accr[rb][gb][r] = 0; acci[rb][gb][r] = 0; }

    QHIP_GEMM_LOAD(uint64_t{0})
    for (uint64_t k0 = 0; k0 < d; k0 += kChunkK) {
        Ps[pk + 0][pi_] = pv0;
        Ps[pk + 1][pi_] = pv1;
        Ps[pk + 2][pi_] = pv2;
        Ps[pk + 3][pi_] = pv3;
        Xs[xk[0]][xg[0]] = xv0;
        Xs[xk[1]][xg[1]] = xv1;
        Xs[xk[2]][xg[2]] = xv2;
        Xs[xk[3]][xg[3]] = xv3;
        __syncthreads();
        if (k0 + kChunkK < d) QHIP_GEMM_LOAD(k0 + kChunkK)
#pragma unroll
        for (int ks = 0; ks < kChunkK / 4; ++ks) {
            const int kq = ks * 4 + lq;
            R ar[2], ai[2], an[2], br[2], bi[2];
#pragma unroll
            for (int rb = 0; rb < 2; ++rb) {
                const A v = Ps[kq][wi + rb * 16 + li];
                ar[rb] = v.x;
                ai[rb] = v.y;
                an[rb] = -v.y;
            }
#pragma unroll
            for (int gb = 0; gb < 2; ++gb) {
                const A v = Xs[kq][wg + gb * 16 + li];
                br[gb] = v.x;
                bi[gb] = v.y;
            }
#pragma unroll
            for (int rb = 0; rb < 2; ++rb)
#pragma unroll
                for (int gb = 0; gb < 2; ++gb) {
                    accr[rb][gb] = MF::step(ar[rb], br[gb], accr[rb][gb]);
                    accr[rb][gb] = MF::step(an[rb], bi[gb], accr[rb][gb]);
                    acci[rb][gb] = MF::step(ar[rb], bi[gb], acci[rb][gb]);
                    acci[rb][gb] = MF::step(ai[rb], br[gb], acci[rb][gb]);
                }
        }
        __syncthreads();   // Ps / Xs are rewritten by the next chunk
    }
#undef QHIP_GEMM_LOAD
#pragma unroll
    for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int gb = 0; gb < 2; ++gb) {
            const uint64_t g = static_cast<uint64_t>(gt) * kTileG + wg + gb * 16 + li;
            if (g >= groups) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint64_t i = static_cast<uint64_t>(rt) * kTileI + wi + rb * 16 + MF::row(r, lq);
                A o;
                o.x = accr[rb][gb][r];
                o.y = acci[rb][gb][r];
                Y[i * groups + g] = o;
            }
        }
}

// state[base(g) | off(i)] = Y[i * groups + g] for every row i and group g.
template <typename A>
__global__ __launch_bounds__(kBlock) void gemm_scatter_kernel(A* __restrict__ s, const A* __restrict__ Y,
                                                              const GemmGeom* __restrict__ gp) {
    const int m = gp->m;
    const uint64_t groups = gp->groups, total = groups << m;
    const int gshift = __builtin_ctzll(groups);
    const uint64_t step = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; e < total; e += step) {
        const uint64_t i = e >> gshift, g = e & (groups - 1);
        const A v = Y[e];
        A* dst = s + (gemm_base(g, gp) | gemm_deposit(i, gp, 0, m));
        __builtin_nontemporal_store(v.x, &dst->x);   // plain stores wrote twice the bytes
        __builtin_nontemporal_store(v.y, &dst->y);
    }
}

// The same scatter when the targets' first six bits are register bits 0..5
// (off(i) contiguous over 64 consecutive i): a 64-row x 32-group tile of Y
// is read along its rows (contiguous groups), transposed through LDS, and
// written 64 contiguous amplitudes per group.
constexpr int kScRows = 64, kScGroups = 32;

template <typename A>
__global__ __launch_bounds__(kBlock) void gemm_scatter_rows_kernel(A* __restrict__ s, const A* __restrict__ Y,
                                                                   const GemmGeom* __restrict__ gp) {
    __shared__ A T[kScGroups][kScRows + 1];
    const int m = gp->m;
    const uint64_t groups = gp->groups;
    const uint64_t gtiles = (groups + kScGroups - 1) / kScGroups;
    const uint64_t i0 = (blockIdx.x / gtiles) * kScRows, g0 = (blockIdx.x % gtiles) * kScGroups;
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < kScRows * kScGroups / kBlock; ++q) {
        const int r = (t >> 5) + (kBlock / kScGroups) * q, c = t & (kScGroups - 1);
        if (g0 + c < groups) T[c][r] = Y[(i0 + r) * groups + g0 + c];
    }
    __syncthreads();
    const uint64_t hi = gemm_deposit(i0, gp, 6, m);
#pragma unroll
    for (int q = 0; q < kScRows * kScGroups / kBlock; ++q) {
        const int r = t & (kScRows - 1), c = (t >> 6) + (kBlock / kScRows) * q;
        if (g0 + c < groups) {
            const A v = T[c][r];
            A* dst = s + (gemm_base(g0 + c, gp) | hi | static_cast<uint64_t>(r));
            __builtin_nontemporal_store(v.x, &dst->x);
            __builtin_nontemporal_store(v.y, &dst->y);
        }
    }
}

template <typename A>
__global__ __launch_bounds__(kBlock) void narrow_kernel(const double2* __restrict__ in, A* __restrict__ out,
                                                        uint64_t n) {
    const uint64_t step = static_cast<uint64_t>(gridDim.x) * kBlock;
    for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; e < n; e += step) {
        A o;
        o.x = in[e].x;
        o.y = in[e].y;
        out[e] = o;
    }
}

GemmGeom gemm_geom(int m, const uint32_t* tpos, const std::vector<uint32_t>& removed, uint64_t cbit,
                   uint64_t groups) {
    GemmGeom geo{};
    geo.m = m;
    for (int b = 0; b < m; ++b) geo.tpos[b] = static_cast<uint8_t>(tpos[b]);
    std::vector<uint32_t> srt(removed);
    std::sort(srt.begin(), srt.end());
    geo.n_ins = static_cast<int32_t>(srt.size());
    for (size_t b = 0; b < srt.size(); ++b) geo.ins[b] = static_cast<uint8_t>(srt[b]);
    geo.cbit = cbit;
    geo.groups = groups;
    geo.row_tiles = static_cast<int32_t>((uint64_t{1} << m) / kTileI);
    geo.group_tiles = static_cast<int32_t>((groups + kTileG - 1) / kTileG);
    geo.kfast = tpos[0] == 0 ? 1 : 0;
    return geo;
}

// geo: the host copy (grid sizes); dgeo: the same record on the device.
template <typename A>
void launch_gemm(const A* in, const A* P, A* Y, const GemmGeom& geo, const GemmGeom* dgeo) {
    auto& C = ctx();
    const uint64_t gt8 = (static_cast<uint64_t>(geo.group_tiles) + 7) / 8 * 8;
    const uint64_t blocks = gt8 * static_cast<uint64_t>(geo.row_tiles);
    if (blocks > 0x7fffffffu) fail(QHIP_CAPABILITY, "controlled powers: GEMM grid too large");
    hipLaunchKernelGGL(controlled_gemm_kernel<A>, dim3(static_cast<uint32_t>(blocks)), dim3(kGemmT), 0, C.stream,
                       in, P, Y, dgeo);
    QHIP_CHECK(hipGetLastError());
}

template <typename A>
void launch_scatter(A* s, const A* Y, const GemmGeom& geo, const GemmGeom* dgeo) {
    auto& C = ctx();
    bool low6 = geo.m >= 6;
    for (int b = 0; b < 6 && low6; ++b) low6 = geo.tpos[b] == b;
    if (low6) {
        const uint64_t blocks = (uint64_t{1} << geo.m) / kScRows * ((geo.groups + kScGroups - 1) / kScGroups);
        hipLaunchKernelGGL(gemm_scatter_rows_kernel<A>, dim3(static_cast<uint32_t>(blocks)), dim3(kBlock), 0,
                           C.stream, s, Y, dgeo);
        QHIP_CHECK(hipGetLastError());
        return;
    }
    const uint64_t total = geo.groups << geo.m;
    const int blocks = static_cast<int>(std::min<uint64_t>((total + kBlock - 1) / kBlock,
                                                           static_cast<uint64_t>(C.n_cu) * 32));
    hipLaunchKernelGGL(gemm_scatter_kernel<A>, dim3(blocks), dim3(kBlock), 0, C.stream, s, Y, dgeo);
    QHIP_CHECK(hipGetLastError());
}

// Device matrices (row-major d x d, complex128) for the squaring chain.
struct DevMat {
    void* p = nullptr;
    explicit DevMat(size_t bytes) {
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            fail(QHIP_NOMEM, "controlled powers: device allocation of a " + std::to_string(bytes) +
                                 "-byte matrix failed");
        }
    }
    DevMat(const DevMat&) = delete;
    DevMat& operator=(const DevMat&) = delete;
    ~DevMat() {
        if (p) {
            (void)hipStreamSynchronize(ctx().stream);
            (void)hipFree(p);
        }
    }
    double2* d() const { return static_cast<double2*>(p); }
};

// Geometry of out = a b for row-major d x d matrices: b read as a 2m-qubit
// register whose high m bits are the targets (no control).
GemmGeom matmul_geom(int m) {
    std::vector<uint32_t> tpos(static_cast<size_t>(m)), removed(static_cast<size_t>(m));
    for (int q = 0; q < m; ++q) tpos[q] = removed[q] = static_cast<uint32_t>(m + q);
    return gemm_geom(m, tpos.data(), removed, 0, uint64_t{1} << m);
}

void check_distinct(std::vector<uint32_t> q, const char* what) {
    std::sort(q.begin(), q.end());
    if (std::adjacent_find(q.begin(), q.end()) != q.end())
        fail(QHIP_INVALID, std::string(what) + ": qubits must be distinct");
}

}  // namespace

void permute_state(qhip_state* s, const uint32_t* qubits, int k, const uint64_t* table) {
    if (k < 1 || k > kMaxTableQubits)
        fail(QHIP_INVALID, "add_permutation: 1.." + std::to_string(kMaxTableQubits) + " qubits required");
    std::vector<uint32_t> qs(qubits, qubits + k);
    check_distinct(qs, "add_permutation");
    for (uint32_t q : qs)
        if (q >= static_cast<uint32_t>(s->n))
            fail(QHIP_INVALID, "qhip_state_permute: qubit " + std::to_string(q) + " outside the " +
                                   std::to_string(s->n) + "-qubit register");
    const uint64_t n = uint64_t{1} << k;
    std::vector<uint32_t> inv(n, ~0u);
    for (uint64_t x = 0; x < n; ++x) {
        const uint64_t v = table[x];
        if (v >= n || inv[v] != ~0u) fail(QHIP_INVALID, "add_permutation: table is not a bijection");
        inv[v] = static_cast<uint32_t>(x);
    }
    auto& C = ctx();
    PermQubits Q{};
    for (int b = 0; b < 32; ++b) Q.q[b] = 63;   // padding: bit 63 is clear in every index and entry
    uint64_t keep = ~uint64_t{0};
    for (int b = 0; b < k; ++b) {
        Q.q[b] = static_cast<uint8_t>(qs[b]);
        keep &= ~(uint64_t{1} << qs[b]);
    }
    const size_t bytes = s->dim() * s->amp_bytes();
    void* scratch = C.perm.ensure(bytes);
    auto* d_inv = static_cast<uint32_t*>(C.misc2.ensure(sizeof(uint32_t) * n));
    QHIP_CHECK(hipMemcpyAsync(d_inv, inv.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, C.stream));
    if (s->precision == QHIP_C128) launch_permute<double2>(s, scratch, d_inv, Q, k, keep);
    else launch_permute<float2>(s, scratch, d_inv, Q, k, keep);
    if (s->owned) {
        // The state takes the permuted buffer; its old one becomes the scratch.
        QHIP_CHECK(hipStreamSynchronize(C.stream));
        C.perm.ptr = s->d;
        C.perm.bytes = bytes;
        s->d = scratch;
    } else {
        // A state wrapping caller memory keeps its buffer: copy back.
        QHIP_CHECK(hipMemcpyAsync(s->d, scratch, bytes, hipMemcpyDeviceToDevice, C.stream));
        QHIP_CHECK(hipStreamSynchronize(C.stream));
    }
}

void set_basis_states(qhip_state* s, const uint64_t* labels, int64_t n_labels, uint64_t stride) {
    const uint64_t dim = s->dim();
    if (stride == 0 || stride > dim || (stride & (stride - 1)) != 0)
        fail(QHIP_INVALID, "qhip_state_set_basis: stride must be a power of two <= 2^n");
    const int log_stride = std::countr_zero(stride);
    if (n_labels < 0 || static_cast<uint64_t>(n_labels) > dim / stride)
        fail(QHIP_INVALID, "qhip_state_set_basis: more labels than blocks of the state");
    for (int64_t b = 0; b < n_labels; ++b)
        if (labels[b] >= stride) fail(QHIP_INVALID, "qhip_state_set_basis: label outside its block");
    auto& C = ctx();
    auto* d_labels = static_cast<uint64_t*>(C.misc2.ensure(sizeof(uint64_t) * std::max<int64_t>(n_labels, 1)));
    if (n_labels > 0)
        QHIP_CHECK(hipMemcpyAsync(d_labels, labels, sizeof(uint64_t) * n_labels, hipMemcpyHostToDevice, C.stream));
    const int blocks = std::max(1, static_cast<int>(std::min<uint64_t>((dim + kBlock - 1) / kBlock,
                                                                       static_cast<uint64_t>(C.n_cu) * 32)));
    if (s->precision == QHIP_C128)
        hipLaunchKernelGGL(set_basis_kernel<double2>, dim3(blocks), dim3(kBlock), 0, C.stream,
                           static_cast<double2*>(s->d), dim, d_labels, n_labels, log_stride);
    else
        hipLaunchKernelGGL(set_basis_kernel<float2>, dim3(blocks), dim3(kBlock), 0, C.stream,
                           static_cast<float2*>(s->d), dim, d_labels, n_labels, log_stride);
    QHIP_CHECK(hipGetLastError());
    QHIP_CHECK(hipStreamSynchronize(C.stream));
}

void controlled_powers(const qhip_config* cfg, qhip_state* s, const uint32_t* controls, const uint64_t* exponents,
                       int n_controls, const uint32_t* targets, int m, const double* matrix) {
    if (n_controls < 1 || m < 1) fail(QHIP_INVALID, "add_controlled_powers: controls and targets required");
    std::vector<uint32_t> all(controls, controls + n_controls);
    all.insert(all.end(), targets, targets + m);
    check_distinct(all, "add_controlled_powers");
    for (uint32_t q : all)
        if (q >= static_cast<uint32_t>(s->n))
            fail(QHIP_INVALID, "qhip_state_controlled_powers: qubit " + std::to_string(q) + " outside the " +
                                   std::to_string(s->n) + "-qubit register");
    if (m > kMaxGemmTargets)
        fail(QHIP_CAPABILITY, "qhip_state_controlled_powers: at most " + std::to_string(kMaxGemmTargets) +
                                  " targets (the 2^m x 2^m matrix and its squares are held on the device)");
    const size_t d = size_t{1} << m;
    uint64_t any = 0;
    for (int j = 0; j < n_controls; ++j) any |= exponents[j];
    const int n_sq = static_cast<int>(std::bit_width(any));
    if (n_sq == 0) return;

    if (m > kMaxPowerTargets) {
        // squares[b] = M^(2^b) along one chain, and a product of squares from
        // the lowest bit up for an exponent with several bits
        // (program.cpp:656-673), all on the device; then per control one GEMM
        // into the scratch buffer and one scatter back into the state.
        auto& C = ctx();
        const size_t mat_bytes = d * d * sizeof(double2);
        const uint64_t groups = uint64_t{1} << (s->n - m - 1);
        // Geometries: [0] the chain's products, [1 + j] control j.
        std::vector<GemmGeom> geos(1 + static_cast<size_t>(n_controls));
        geos[0] = matmul_geom(m);
        for (int j = 0; j < n_controls; ++j) {
            std::vector<uint32_t> removed(targets, targets + m);
            removed.push_back(controls[j]);
            geos[1 + j] = gemm_geom(m, targets, removed, uint64_t{1} << controls[j], groups);
        }
        DevMat dgeo(geos.size() * sizeof(GemmGeom));
        QHIP_CHECK(hipMemcpyAsync(dgeo.p, geos.data(), geos.size() * sizeof(GemmGeom), hipMemcpyHostToDevice,
                                  C.stream));
        const GemmGeom* dg = static_cast<const GemmGeom*>(dgeo.p);
        auto matmul = [&](const double2* a, const double2* b, double2* out) {
            launch_gemm<double2>(b, a, out, geos[0], dg);
        };
        std::vector<std::unique_ptr<DevMat>> squares;
        squares.push_back(std::make_unique<DevMat>(mat_bytes));
        QHIP_CHECK(hipMemcpyAsync(squares[0]->p, matrix, mat_bytes, hipMemcpyHostToDevice, C.stream));
        for (int b = 1; b < n_sq; ++b) {
            squares.push_back(std::make_unique<DevMat>(mat_bytes));
            matmul(squares[b - 1]->d(), squares[b - 1]->d(), squares[b]->d());
        }
        std::map<uint64_t, std::unique_ptr<DevMat>> products;
        for (int j = 0; j < n_controls; ++j) {
            const uint64_t e = exponents[j];
            if (std::popcount(e) < 2 || products.count(e)) continue;
            auto r = std::make_unique<DevMat>(mat_bytes);
            auto tmp = std::make_unique<DevMat>(mat_bytes);
            QHIP_CHECK(hipMemcpyAsync(r->p, squares[static_cast<size_t>(std::countr_zero(e))]->p, mat_bytes,
                                      hipMemcpyDeviceToDevice, C.stream));
            for (uint64_t x = e & (e - 1); x != 0; x &= x - 1) {
                matmul(r->d(), squares[static_cast<size_t>(std::countr_zero(x))]->d(), tmp->d());
                std::swap(r, tmp);
            }
            products.emplace(e, std::move(r));
        }
        const bool c128 = s->precision == QHIP_C128;
        void* Y = C.perm.ensure(static_cast<size_t>(groups << m) * s->amp_bytes());
        std::unique_ptr<DevMat> narrow;
        if (!c128) narrow = std::make_unique<DevMat>(d * d * sizeof(float2));
        for (int j = 0; j < n_controls; ++j) {
            const uint64_t e = exponents[j];
            if (e == 0) continue;
            const double2* P = std::has_single_bit(e) ? squares[static_cast<size_t>(std::countr_zero(e))]->d()
                                                      : products.at(e)->d();
            const GemmGeom& geo = geos[1 + j];
            if (c128) {
                launch_gemm<double2>(static_cast<const double2*>(s->d), P, static_cast<double2*>(Y), geo, dg + 1 + j);
                launch_scatter<double2>(static_cast<double2*>(s->d), static_cast<const double2*>(Y), geo, dg + 1 + j);
            } else {
                const int blocks = static_cast<int>(std::min<uint64_t>((d * d + kBlock - 1) / kBlock,
                                                                       static_cast<uint64_t>(C.n_cu) * 32));
                hipLaunchKernelGGL(narrow_kernel<float2>, dim3(blocks), dim3(kBlock), 0, C.stream, P,
                                   static_cast<float2*>(narrow->p), static_cast<uint64_t>(d * d));
                QHIP_CHECK(hipGetLastError());
                launch_gemm<float2>(static_cast<const float2*>(s->d), static_cast<const float2*>(narrow->p),
                                    static_cast<float2*>(Y), geo, dg + 1 + j);
                launch_scatter<float2>(static_cast<float2*>(s->d), static_cast<const float2*>(Y), geo, dg + 1 + j);
            }
        }
        QHIP_CHECK(hipStreamSynchronize(C.stream));
        return;
    }

    Mat M(d * d);
    for (size_t i = 0; i < d * d; ++i) M[i] = cplx(matrix[2 * i], matrix[2 * i + 1]);
    // squares[b] = M^(2^b) along one chain; an exponent with several set bits
    // multiplies its squares from the lowest bit up (program.cpp:656-673).
    std::vector<Mat> squares;
    squares.push_back(M);
    for (int b = 1; b < n_sq; ++b) squares.push_back(matmul(squares.back(), squares.back(), d));
    std::map<uint64_t, Mat> products;
    for (int j = 0; j < n_controls; ++j) {
        const uint64_t e = exponents[j];
        if (std::popcount(e) < 2 || products.count(e)) continue;
        Mat r = squares[static_cast<size_t>(std::countr_zero(e))];
        for (uint64_t x = e & (e - 1); x != 0; x &= x - 1)
            r = matmul(r, squares[static_cast<size_t>(std::countr_zero(x))], d);
        products.emplace(e, std::move(r));
    }

    // One Custom op per control with a non-zero exponent: operands
    // targets..., control (local bit m), matrix diag(I, P).
    std::vector<qhip_op> ops;
    std::vector<uint32_t> pool;
    std::vector<double> unitaries;
    const size_t D = 2 * d;
    for (int j = 0; j < n_controls; ++j) {
        const uint64_t e = exponents[j];
        if (e == 0) continue;
        const Mat& P = std::has_single_bit(e) ? squares[static_cast<size_t>(std::countr_zero(e))] : products.at(e);
        qhip_op op{};
        op.gate = QHIP_CUSTOM;
        op.n_qubits = m + 1;
        op.qubit_offset = static_cast<int64_t>(pool.size());
        op.unitary_offset = static_cast<int64_t>(unitaries.size() / 2);
        op.dep_offset = -1;
        op.cbit = -1;
        pool.insert(pool.end(), targets, targets + m);
        pool.push_back(controls[j]);
        for (size_t y = 0; y < D; ++y)
            for (size_t x = 0; x < D; ++x) {
                cplx v(0.0);
                if (y < d && x < d) v = y == x ? cplx(1.0) : cplx(0.0);
                else if (y >= d && x >= d) v = P[(y - d) * d + (x - d)];
                unitaries.push_back(v.real());
                unitaries.push_back(v.imag());
            }
        ops.push_back(op);
    }
    if (ops.empty()) return;
    qhip_stream st{};
    st.ops = ops.data();
    st.n_ops = static_cast<int64_t>(ops.size());
    st.qubits = pool.data();
    st.unitaries = unitaries.data();
    apply_stream(cfg, &st, s, nullptr, false);
}

}  // namespace qhip
