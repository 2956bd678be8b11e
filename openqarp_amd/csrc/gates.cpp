// gates.cpp — gate matrices, lowering of commands to device ops, daggers and
// adjoint generators.
//
// Matrices follow docs/contracts/qarp_conventions.md §2-§5 (LSB order, local
// bit b <-> operand b, rotations exp(-i theta/2 P)); the reference obtains the
// same matrices by running apply_command on basis columns
// (QarpSimulator::local_unitary, src/qarpx/src/simulator/qarp_simulator.cpp:92-118).
// Here they are assembled from Pauli/projector algebra instead.
#include "common.h"

#include <cmath>

namespace qhip {

namespace {

constexpr double kPi = 3.14159265358979323846;
const cplx I1{1.0, 0.0}, I0{0.0, 0.0}, IM{0.0, 1.0};

using M2 = std::array<cplx, 4>;   // row-major 2x2

M2 pauli(char p) {
    switch (p) {
        case 'X': return {I0, I1, I1, I0};
        case 'Y': return {I0, -IM, IM, I0};
        case 'Z': return {I1, I0, I0, -I1};
        default:  return {I1, I0, I0, I1};
    }
}

// exp(-i theta/2 P) = cos(theta/2) I - i sin(theta/2) P
M2 rotation(char p, double theta) {
    const double c = std::cos(theta / 2.0), s = std::sin(theta / 2.0);
    M2 P = pauli(p), I = pauli('I'), r;
    for (int i = 0; i < 4; ++i) r[i] = c * I[i] - cplx(0.0, s) * P[i];
    return r;
}

M2 phase_gate(double theta) { return {I1, I0, I0, cplx(std::cos(theta), std::sin(theta))}; }

// U(theta, phi, lambda), conventions §2.5
M2 u_gate(double th, double ph, double la) {
    const double c = std::cos(th / 2.0), s = std::sin(th / 2.0);
    return {cplx(c, 0.0), -std::polar(1.0, la) * s, std::polar(1.0, ph) * s,
            std::polar(1.0, ph + la) * c};
}

bool one_qubit(int gate, const double* p, M2& m) {
    const double r2 = 1.0 / std::sqrt(2.0);
    switch (gate) {
        case QHIP_X: m = pauli('X'); return true;
        case QHIP_Y: m = pauli('Y'); return true;
        case QHIP_Z: m = pauli('Z'); return true;
        case QHIP_ID: m = pauli('I'); return true;
        case QHIP_H: m = {cplx(r2), cplx(r2), cplx(r2), cplx(-r2)}; return true;
        case QHIP_S: m = {I1, I0, I0, IM}; return true;
        case QHIP_SDG: m = {I1, I0, I0, -IM}; return true;
        case QHIP_T: m = {I1, I0, I0, cplx(std::cos(kPi / 4), std::sin(kPi / 4))}; return true;
        case QHIP_TDG: m = {I1, I0, I0, cplx(std::cos(-kPi / 4), std::sin(-kPi / 4))}; return true;
        case QHIP_SX: m = {cplx(.5, .5), cplx(.5, -.5), cplx(.5, -.5), cplx(.5, .5)}; return true;
        case QHIP_SXDG: m = {cplx(.5, -.5), cplx(.5, .5), cplx(.5, .5), cplx(.5, -.5)}; return true;
        case QHIP_RX: m = rotation('X', p[0]); return true;
        case QHIP_RY: m = rotation('Y', p[0]); return true;
        case QHIP_RZ: m = rotation('Z', p[0]); return true;
        case QHIP_P: m = phase_gate(p[0]); return true;
        case QHIP_U: m = u_gate(p[0], p[1], p[2]); return true;
        default: return false;
    }
}

// Target operation of the singly controlled gates (conventions §3.1).
bool controlled_target(int gate, const double* p, M2& t) {
    switch (gate) {
        case QHIP_CX: return one_qubit(QHIP_X, p, t);
        case QHIP_CY: return one_qubit(QHIP_Y, p, t);
        case QHIP_CZ: return one_qubit(QHIP_Z, p, t);
        case QHIP_CH: return one_qubit(QHIP_H, p, t);
        case QHIP_CS: return one_qubit(QHIP_S, p, t);
        case QHIP_CSDG: return one_qubit(QHIP_SDG, p, t);
        case QHIP_CSX: return one_qubit(QHIP_SX, p, t);
        case QHIP_CSXDG: return one_qubit(QHIP_SXDG, p, t);
        case QHIP_CRX: return one_qubit(QHIP_RX, p, t);
        case QHIP_CRY: return one_qubit(QHIP_RY, p, t);
        case QHIP_CRZ: return one_qubit(QHIP_RZ, p, t);
        case QHIP_CP: return one_qubit(QHIP_P, p, t);
        case QHIP_CU: {
            t = u_gate(p[0], p[1], p[2]);
            const cplx g = std::polar(1.0, p[3]);
            for (auto& e : t) e *= g;
            return true;
        }
        default: return false;
    }
}

// |0><0| (x) I + |1><1| (x) T on (control = bit 0, target = bit 1)
std::vector<cplx> controlled(const M2& t) {
    std::vector<cplx> m(16, I0);
    m[0 * 4 + 0] = I1;
    m[2 * 4 + 2] = I1;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) m[(1 + 2 * a) * 4 + (1 + 2 * b)] = t[a * 2 + b];
    return m;
}

// Kronecker product A (x) B with B on the low bit.
std::vector<cplx> kron(const M2& hi, const M2& lo) {
    std::vector<cplx> m(16);
    for (int r1 = 0; r1 < 2; ++r1)
        for (int c1 = 0; c1 < 2; ++c1)
            for (int r0 = 0; r0 < 2; ++r0)
                for (int c0 = 0; c0 < 2; ++c0)
                    m[(r1 * 2 + r0) * 4 + (c1 * 2 + c0)] = hi[r1 * 2 + c1] * lo[r0 * 2 + c0];
    return m;
}

std::vector<cplx> matmul(const std::vector<cplx>& a, const std::vector<cplx>& b, int D) {
    std::vector<cplx> c(static_cast<size_t>(D) * D, I0);
    for (int i = 0; i < D; ++i)
        for (int k = 0; k < D; ++k) {
            const cplx aik = a[i * D + k];
            if (aik == I0) continue;
            for (int j = 0; j < D; ++j) c[i * D + j] += aik * b[k * D + j];
        }
    return c;
}

// Product of a time-ordered 2-qubit sequence: each element is (4x4 matrix).
std::vector<cplx> sequence(std::initializer_list<std::vector<cplx>> steps) {
    std::vector<cplx> acc(16, I0);
    for (int i = 0; i < 4; ++i) acc[i * 4 + i] = I1;
    for (const auto& s : steps) acc = matmul(s, acc, 4);
    return acc;
}

std::vector<cplx> on0(const M2& g) { return kron(pauli('I'), g); }
std::vector<cplx> on1(const M2& g) { return kron(g, pauli('I')); }
std::vector<cplx> cx01() { return controlled(pauli('X')); }
std::vector<cplx> cx10() {   // control bit 1, target bit 0
    std::vector<cplx> m(16, I0);
    m[0] = I1; m[1 * 4 + 1] = I1;
    m[2 * 4 + 3] = I1; m[3 * 4 + 2] = I1;
    return m;
}

}  // namespace

bool named_matrix(int gate, int k, const double* p, std::vector<cplx>& m) {
    const int D = 1 << k;
    if (k == 1) {
        M2 g;
        if (!one_qubit(gate, p, g)) return false;
        m.assign(g.begin(), g.end());
        return true;
    }
    if (gate == QHIP_MCZ && k <= 8) {
        m.assign(static_cast<size_t>(D) * D, I0);
        for (int i = 0; i < D; ++i) m[i * D + i] = I1;
        m[(D - 1) * D + (D - 1)] = -I1;
        return true;
    }
    if (k == 2) {
        M2 t;
        if (controlled_target(gate, p, t)) { m = controlled(t); return true; }
        switch (gate) {
            case QHIP_SWAP:
                m.assign(16, I0);
                m[0] = I1; m[1 * 4 + 2] = I1; m[2 * 4 + 1] = I1; m[3 * 4 + 3] = I1;
                return true;
            case QHIP_RXX: case QHIP_RYY: case QHIP_RZZ: {   // exp(-i theta/2 P(x)P)
                const char c = gate == QHIP_RXX ? 'X' : gate == QHIP_RYY ? 'Y' : 'Z';
                const auto PP = kron(pauli(c), pauli(c));
                const double co = std::cos(p[0] / 2.0), si = std::sin(p[0] / 2.0);
                m.assign(16, I0);
                for (int i = 0; i < 16; ++i) m[i] = cplx(0.0, -si) * PP[i];
                for (int i = 0; i < 4; ++i) m[i * 4 + i] += co;
                return true;
            }
            // iSWAP / iSWAPdg / ECR through the phase-exact decompositions the
            // reference lowers them with (transpiler/decompositions.cpp:391-431).
            case QHIP_ISWAP: {
                const double r2 = 1.0 / std::sqrt(2.0);
                const M2 h = {cplx(r2), cplx(r2), cplx(r2), cplx(-r2)};
                m = sequence({on0({I1, I0, I0, IM}), on1({I1, I0, I0, IM}), on0(h), cx01(),
                              cx10(), on1(h)});
                return true;
            }
            case QHIP_ISWAPDG: {
                const double r2 = 1.0 / std::sqrt(2.0);
                const M2 h = {cplx(r2), cplx(r2), cplx(r2), cplx(-r2)};
                m = sequence({on1(h), cx10(), cx01(), on0(h), on1({I1, I0, I0, -IM}),
                              on0({I1, I0, I0, -IM})});
                return true;
            }
            case QHIP_ECR: {
                const double r2 = 1.0 / std::sqrt(2.0);
                const M2 h = {cplx(r2), cplx(r2), cplx(r2), cplx(-r2)};
                m = sequence({on1(h), cx01(), on1(rotation('Z', kPi / 4)), cx01(), on0(pauli('X')),
                              cx01(), on1(rotation('Z', -kPi / 4)), cx01(), on1(h)});
                return true;
            }
            default: return false;
        }
    }
    if (k == 3 && (gate == QHIP_CCX || gate == QHIP_CSWAP)) {
        m.assign(64, I0);
        for (int i = 0; i < 8; ++i) {
            int j = i;
            if (gate == QHIP_CCX && (i & 3) == 3) j = i ^ 4;                 // flip bit 2
            if (gate == QHIP_CSWAP && (i & 1) && (((i >> 1) & 1) != ((i >> 2) & 1)))
                j = i ^ 6;                                                   // swap bits 1,2
            m[j * 8 + i] = I1;
        }
        return true;
    }
    return false;
}

void classify_diagonal(HostOp& op) {
    if (op.kind != OP_DENSE && op.kind != OP_IP_DENSE) return;
    const int D = 1 << op.k;
    for (int y = 0; y < D; ++y)
        for (int x = 0; x < D; ++x)
            if (x != y && op.m[y * D + x] != I0) return;
    std::vector<cplx> d(D);
    for (int i = 0; i < D; ++i) d[i] = op.m[i * D + i];
    op.m = std::move(d);
    op.kind = op.kind == OP_DENSE ? OP_DIAG : OP_IP_DIAG;
}

static bool is_marker(int g) {
    return g == QHIP_BARRIER || g == QHIP_MEASURE || g == QHIP_RESET || g == QHIP_ID ||
           g == QHIP_BRANCH_BEGIN || g == QHIP_BRANCH_ELSE || g == QHIP_BRANCH_END;
}

static std::vector<cplx> custom_matrix(const Stream& st, int64_t i) {
    const qhip_op& op = st.op(i);
    if (op.unitary_offset < 0 || op.n_qubits < 1 || op.n_qubits > QHIP_MAX_FUSION_QUBITS)
        fail(QHIP_RUNTIME, "QarpSimulator: Custom gate must carry a unitary and 1..8 qubit targets");
    const int D = 1 << op.n_qubits;
    std::vector<cplx> m(static_cast<size_t>(D) * D);
    std::memcpy(m.data(), st.s->unitaries + 2 * op.unitary_offset, sizeof(cplx) * m.size());
    return m;
}

// Local matrix of input op i (a Custom payload or a named gate).
std::vector<cplx> op_matrix(const Stream& st, int64_t i) {
    const qhip_op& op = st.op(i);
    if (op.gate == QHIP_CUSTOM) return custom_matrix(st, i);
    std::vector<cplx> m;
    if (!named_matrix(op.gate, op.n_qubits, op.params, m))
        fail(QHIP_RUNTIME, "QarpSimulator: unsupported gate id " + std::to_string(op.gate) + " on " +
                               std::to_string(op.n_qubits) + " qubit(s)");
    return m;
}

void lower_op(const Stream& st, int64_t i, std::vector<HostOp>& out) {
    const qhip_op& op = st.op(i);
    if (is_marker(op.gate)) return;
    HostOp h;
    if (op.gate == QHIP_GPHASE) {
        h.kind = OP_SCALAR;
        h.phase = std::polar(1.0, op.params[0]);
        out.push_back(std::move(h));
        return;
    }
    if (op.gate == QHIP_MCZ) {
        h.kind = OP_MASKPHASE;
        for (int b = 0; b < op.n_qubits; ++b) h.mask |= uint64_t{1} << st.qubits(i)[b];
        h.phase = cplx(-1.0, 0.0);
        out.push_back(std::move(h));
        return;
    }
    h.kind = OP_DENSE;
    h.k = op.n_qubits;
    if (h.k < 1 || h.k > QHIP_MAX_FUSION_QUBITS)
        fail(QHIP_RUNTIME, "QarpSimulator: gate on " + std::to_string(h.k) + " qubits");
    for (int b = 0; b < h.k; ++b) h.q[b] = st.qubits(i)[b];
    h.m = op_matrix(st, i);
    classify_diagonal(h);
    out.push_back(std::move(h));
}

void lower_item(const Stream& st, const FusedItem& it, int n, std::vector<HostOp>& out,
                bool skip_terminal_markers) {
    (void)n;
    (void)skip_terminal_markers;
    if (it.src >= 0) { lower_op(st, it.src, out); return; }
    HostOp h;
    h.kind = OP_DENSE;
    h.k = it.k;
    h.q = it.q;
    h.m = it.m;
    classify_diagonal(h);
    out.push_back(std::move(h));
}

void lower_dagger(const Stream& st, int64_t i, std::vector<HostOp>& out) {
    const size_t first = out.size();
    lower_op(st, i, out);
    for (size_t j = first; j < out.size(); ++j) {
        HostOp& h = out[j];
        if (h.kind == OP_SCALAR || h.kind == OP_MASKPHASE) { h.phase = std::conj(h.phase); continue; }
        if (h.kind == OP_DIAG) { for (auto& e : h.m) e = std::conj(e); continue; }
        const int D = 1 << h.k;
        std::vector<cplx> a(h.m.size());
        for (int y = 0; y < D; ++y)
            for (int x = 0; x < D; ++x) a[y * D + x] = std::conj(h.m[x * D + y]);
        h.m = std::move(a);
    }
}

// Generators of apply_generator (qarp_simulator.cpp:1352-1418): the gate's
// d/dtheta = c * G * gate, with G a Pauli string, projector or identity.
bool generator(const Stream& st, int64_t i, HostOp& G, cplx& c) {
    const qhip_op& op = st.op(i);
    const uint32_t* q = st.qubits(i);
    const cplx mhalf(0.0, -0.5), plus_i(0.0, 1.0);
    G = HostOp();
    auto single = [&](char p) {
        G.kind = OP_IP_DENSE;
        G.k = 1;
        G.q[0] = q[0];
        const M2 m = pauli(p);
        G.m.assign(m.begin(), m.end());
    };
    switch (op.gate) {
        case QHIP_RX: single('X'); c = mhalf; break;
        case QHIP_RY: single('Y'); c = mhalf; break;
        case QHIP_RZ: single('Z'); c = mhalf; break;
        case QHIP_P:
            G.kind = OP_IP_DENSE; G.k = 1; G.q[0] = q[0];
            G.m = {I0, I0, I0, I1};
            c = plus_i;
            break;
        case QHIP_GPHASE:
            G.kind = OP_IP_DIAG; G.k = 0; G.m = {I1};
            c = plus_i;
            break;
        case QHIP_CRX: case QHIP_CRY: case QHIP_CRZ: {
            const char p = op.gate == QHIP_CRX ? 'X' : op.gate == QHIP_CRY ? 'Y' : 'Z';
            G.kind = OP_IP_DENSE; G.k = 2; G.q[0] = q[0]; G.q[1] = q[1];
            const M2 t = pauli(p);
            G.m.assign(16, I0);   // |1><1|_c (x) P_t
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b) G.m[(1 + 2 * a) * 4 + (1 + 2 * b)] = t[a * 2 + b];
            c = mhalf;
            break;
        }
        case QHIP_CP:
            G.kind = OP_IP_DENSE; G.k = 2; G.q[0] = q[0]; G.q[1] = q[1];
            G.m.assign(16, I0);
            G.m[15] = I1;
            c = plus_i;
            break;
        case QHIP_RXX: case QHIP_RYY: case QHIP_RZZ: {
            const char p = op.gate == QHIP_RXX ? 'X' : op.gate == QHIP_RYY ? 'Y' : 'Z';
            G.kind = OP_IP_DENSE; G.k = 2; G.q[0] = q[0]; G.q[1] = q[1];
            G.m = kron(pauli(p), pauli(p));
            c = mhalf;
            break;
        }
        default:
            return false;
    }
    classify_diagonal(G);
    return true;
}

}  // namespace qhip
