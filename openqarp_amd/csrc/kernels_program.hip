// Structured-program kernels (the reference's Program kernels,
// src/qarpx/include/qarpx/simulator/program.h:17-46):
//  - a basis permutation |x> -> |table[x]> of a qubit subset, applied out of
//    place through a scratch state (apply_permutation, program.cpp:610-646);
//  - controlled matrix powers matrix^e_j on the targets wherever control j
//    is |1>, lowered to controlled dense ops of the ordinary tile pass
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

Mat matmul(const Mat& a, const Mat& b, size_t d) {
    Mat c(d * d, cplx(0.0));
    for (size_t y = 0; y < d; ++y)
        for (size_t t = 0; t < d; ++t) {
            const cplx ay = a[y * d + t];
            for (size_t x = 0; x < d; ++x) c[y * d + x] += ay * b[t * d + x];
        }
    return c;
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
    if (m > kMaxPowerTargets)
        fail(QHIP_CAPABILITY, "qhip_state_controlled_powers: at most " + std::to_string(kMaxPowerTargets) +
                                  " targets (a controlled power runs as one dense op on targets + control)");
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
