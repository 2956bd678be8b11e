// Structured-program kernels (the reference's Program kernels,
// src/qarpx/include/qarpx/simulator/program.h:17-46):
//  - a basis permutation |x> -> |table[x]> of a qubit subset, applied out of
//    place through a scratch state (apply_permutation, program.cpp:610-646);
//  - controlled matrix powers matrix^e_j on the targets wherever control j
//    is |1>, lowered to controlled dense ops of the ordinary tile pass on up
//    to 7 targets, and run by a controlled dense kernel on 8..12
//    (apply_controlled_powers, program.cpp:648-712).
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
            if (j < dim) out[j] = v[u];
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

// ── controlled powers on 8..12 targets ─────────────────────────────────────
// Too wide for one tile-pass op (targets + control > 8 operands): per
// control, a kernel applies P on the targets of every group whose control
// bit is set (apply_controlled_powers, program.cpp:676-706).  A group is the
// 2^m amplitudes base | deposit(l); the groups with the control set are
// enumerated directly (zero bits inserted at the sorted targets and control,
// then the control bit set), so no group is visited and skipped.
constexpr int kMaxDenseTargets = 12;    // 2^12 amplitudes per group
constexpr int kCtlBlock = 256;
constexpr int kCtlAmps = 4096;          // LDS amplitudes per block: G groups x d rows

struct CtlGeom {
    uint8_t tpos[kMaxDenseTargets];      // local bit b <-> global bit tpos[b]
    uint8_t ins[kMaxDenseTargets + 1];   // targets and control, ascending
};

__device__ __forceinline__ double2 cmac(double2 acc, double2 u, double2 x) {
    acc.x = __fma_rn(u.x, x.x, acc.x);
    acc.x = __fma_rn(-u.y, x.y, acc.x);
    acc.y = __fma_rn(u.x, x.y, acc.y);
    acc.y = __fma_rn(u.y, x.x, acc.y);
    return acc;
}
__device__ __forceinline__ float2 cmac(float2 acc, float2 u, float2 x) {
    acc.x = __fmaf_rn(u.x, x.x, acc.x);
    acc.x = __fmaf_rn(-u.y, x.y, acc.x);
    acc.y = __fmaf_rn(u.x, x.y, acc.y);
    acc.y = __fmaf_rn(u.y, x.x, acc.y);
    return acc;
}

// d = 256 R rows (m = 8 + log2 R); each thread owns rows threadIdx.x + 256 r
// of G = 4096 / d groups staged in LDS.  PT is P transposed (PT[k d + row]):
// for a fixed k the block reads one contiguous row of PT, and every group's
// amplitude k is an LDS broadcast.  PT (d^2 entries, at most 256 MiB) stays
// in L2 / HBM and is read once per G groups.
template <typename A, int R>
__global__ __launch_bounds__(kCtlBlock) void controlled_dense_kernel(A* __restrict__ s, const A* __restrict__ PT,
                                                                     const CtlGeom geo, int m, uint64_t cbit,
                                                                     uint64_t groups) {
    constexpr int d = kCtlBlock * R;
    constexpr int G = kCtlAmps / d;
    __shared__ A col[kCtlAmps];
    uint64_t off[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t l = threadIdx.x + kCtlBlock * r;
        uint64_t o = 0;
        for (int b = 0; b < m; ++b) o |= static_cast<uint64_t>((l >> b) & 1u) << geo.tpos[b];
        off[r] = o;
    }
    auto group_base = [&](uint64_t gi) {
        uint64_t base = gi;
        for (int b = 0; b <= m; ++b) {
            const uint64_t p = geo.ins[b];
            base = ((base >> p) << (p + 1)) | (base & ((uint64_t{1} << p) - 1u));
        }
        return base | cbit;
    };
    const uint64_t supers = (groups + G - 1) / G;
    for (uint64_t sg = blockIdx.x; sg < supers; sg += gridDim.x) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const uint64_t gi = sg * G + g;
            if (gi < groups) {
                const uint64_t base = group_base(gi);
#pragma unroll
                for (int r = 0; r < R; ++r) col[g * d + threadIdx.x + kCtlBlock * r] = s[base | off[r]];
            }
        }
        __syncthreads();
        A acc[R][G];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int g = 0; g < G; ++g) { acc[r][g].x = 0; acc[r][g].y = 0; }
#pragma unroll 2
        for (int k = 0; k < d; ++k) {
            A p[R];
#pragma unroll
            for (int r = 0; r < R; ++r) p[r] = PT[static_cast<size_t>(k) * d + threadIdx.x + kCtlBlock * r];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const A c = col[g * d + k];
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r][g] = cmac(acc[r][g], p[r], c);
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const uint64_t gi = sg * G + g;
            if (gi < groups) {
                const uint64_t base = group_base(gi);
#pragma unroll
                for (int r = 0; r < R; ++r) s[base | off[r]] = acc[r][g];
            }
        }
        __syncthreads();   // col is rewritten by the next super-group
    }
}

template <typename A>
void launch_controlled_dense(qhip_state* s, const void* PT, const CtlGeom& geo, int m, uint64_t cbit) {
    auto& C = ctx();
    const uint64_t groups = uint64_t{1} << (s->n - m - 1);
    const uint64_t supers = (groups + (kCtlAmps >> m) - 1) / (kCtlAmps >> m);
    const int blocks = static_cast<int>(std::min<uint64_t>(supers, static_cast<uint64_t>(C.n_cu) * 4));
    auto* st = static_cast<A*>(s->d);
    auto* pt = static_cast<const A*>(PT);
    switch (m) {
        case 8: hipLaunchKernelGGL((controlled_dense_kernel<A, 1>), dim3(blocks), dim3(kCtlBlock), 0, C.stream, st, pt, geo, m, cbit, groups); break;
        case 9: hipLaunchKernelGGL((controlled_dense_kernel<A, 2>), dim3(blocks), dim3(kCtlBlock), 0, C.stream, st, pt, geo, m, cbit, groups); break;
        case 10: hipLaunchKernelGGL((controlled_dense_kernel<A, 4>), dim3(blocks), dim3(kCtlBlock), 0, C.stream, st, pt, geo, m, cbit, groups); break;
        case 11: hipLaunchKernelGGL((controlled_dense_kernel<A, 8>), dim3(blocks), dim3(kCtlBlock), 0, C.stream, st, pt, geo, m, cbit, groups); break;
        default: hipLaunchKernelGGL((controlled_dense_kernel<A, 16>), dim3(blocks), dim3(kCtlBlock), 0, C.stream, st, pt, geo, m, cbit, groups); break;
    }
    QHIP_CHECK(hipGetLastError());
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
    if (m > kMaxDenseTargets)
        fail(QHIP_CAPABILITY, "qhip_state_controlled_powers: at most " + std::to_string(kMaxDenseTargets) +
                                  " targets (one group of 2^m amplitudes is staged in LDS)");
    const size_t d = size_t{1} << m;
    Mat M(d * d);
    for (size_t i = 0; i < d * d; ++i) M[i] = cplx(matrix[2 * i], matrix[2 * i + 1]);

    // squares[b] = M^(2^b) along one chain; an exponent with several set bits
    // multiplies its squares from the lowest bit up (program.cpp:656-673).
    uint64_t any = 0;
    for (int j = 0; j < n_controls; ++j) any |= exponents[j];
    std::vector<Mat> squares;
    const int n_sq = static_cast<int>(std::bit_width(any));
    if (n_sq > 0) squares.push_back(M);
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

    if (m > kMaxPowerTargets) {
        // Per control, the controlled dense kernel with P transposed in the
        // state's precision; the stream is drained before the host buffer is
        // refilled for the next control.
        auto& C = ctx();
        CtlGeom geo{};
        std::vector<uint32_t> ins(targets, targets + m);
        for (int b = 0; b < m; ++b) geo.tpos[b] = static_cast<uint8_t>(targets[b]);
        const bool c128 = s->precision == QHIP_C128;
        void* dPT = C.misc2.ensure(d * d * s->amp_bytes());
        std::vector<double> h64(c128 ? 2 * d * d : 0);
        std::vector<float> h32(c128 ? 0 : 2 * d * d);
        for (int j = 0; j < n_controls; ++j) {
            const uint64_t e = exponents[j];
            if (e == 0) continue;
            const Mat& P = std::has_single_bit(e) ? squares[static_cast<size_t>(std::countr_zero(e))] : products.at(e);
            for (size_t row = 0; row < d; ++row)
                for (size_t k = 0; k < d; ++k) {
                    const cplx v = P[row * d + k];
                    const size_t t = 2 * (k * d + row);
                    if (c128) { h64[t] = v.real(); h64[t + 1] = v.imag(); }
                    else { h32[t] = static_cast<float>(v.real()); h32[t + 1] = static_cast<float>(v.imag()); }
                }
            std::vector<uint32_t> srt(ins);
            srt.push_back(controls[j]);
            std::sort(srt.begin(), srt.end());
            for (int b = 0; b <= m; ++b) geo.ins[b] = static_cast<uint8_t>(srt[b]);
            QHIP_CHECK(hipMemcpyAsync(dPT, c128 ? static_cast<const void*>(h64.data()) : h32.data(),
                                      d * d * s->amp_bytes(), hipMemcpyHostToDevice, C.stream));
            const uint64_t cbit = uint64_t{1} << controls[j];
            if (c128) launch_controlled_dense<double2>(s, dPT, geo, m, cbit);
            else launch_controlled_dense<float2>(s, dPT, geo, m, cbit);
            QHIP_CHECK(hipStreamSynchronize(C.stream));
        }
        return;
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
