// fusion.cpp — the dispatch-time fusion pass.
//
// Same folding rules as QarpSimulator::fuse_for_dispatch
// (src/qarpx/src/simulator/qarp_simulator.cpp:120-129):
//   k = 0                      -> stream unchanged
//   k = 1 or n < fusion_min    -> consecutive single-qubit gates per wire
//                                 (fuse_single_qubit_gates,
//                                  src/qarpx/src/transpiler/fusion.cpp:196-301)
//   otherwise                  -> greedy wire-front blocks of up to k qubits
//                                 (fuse_for_simulation,
//                                  src/qarpx/src/simulator/fusion.cpp:120-268):
//     a gate joins the most recent open block it can reach without crossing
//     a later block on any of its wires; past two qubits a block widens only
//     through a shared qubit; a lone gate is re-emitted verbatim; a fused
//     1-qubit identity (Frobenius distance < 1e-12) is dropped; GPhase passes
//     through in place; anything unfusible flushes every open block that
//     must precede it (transitively through shared wires), in creation order.
// The output feeds the device scheduler, which packs several fused blocks
// into one HBM pass.
#include "common.h"

#include <algorithm>
#include <cmath>
#include <list>

namespace qhip {

std::vector<cplx> op_matrix(const Stream& st, int64_t i);   // gates.cpp

namespace {

bool is_marker(int g) {
    return g == QHIP_BARRIER || g == QHIP_MEASURE || g == QHIP_RESET ||
           g == QHIP_BRANCH_BEGIN || g == QHIP_BRANCH_ELSE || g == QHIP_BRANCH_END;
}

bool in(const std::vector<uint32_t>& v, uint32_t x) {
    return std::find(v.begin(), v.end(), x) != v.end();
}

// ── single-qubit accumulation ───────────────────────────────────────────────

bool fusible_1q(const qhip_op& op) {
    if (op.n_qubits != 1) return false;
    switch (op.gate) {
        case QHIP_X: case QHIP_Y: case QHIP_Z: case QHIP_H: case QHIP_S: case QHIP_SDG:
        case QHIP_T: case QHIP_TDG: case QHIP_RX: case QHIP_RY: case QHIP_RZ: case QHIP_P:
        case QHIP_U:
            return true;
        default:
            return false;
    }
}

struct Wire1q {
    std::array<cplx, 4> m{cplx(1), cplx(0), cplx(0), cplx(1)};
    int n = 0;
    int64_t first = -1;
};

void flush_wire(Wire1q& w, uint32_t q, std::vector<FusedItem>& out) {
    if (w.n == 1) {
        FusedItem it;
        it.src = w.first;
        out.push_back(std::move(it));
    } else if (w.n > 1) {
        const double tol = 1e-12;
        const bool ident = std::abs(w.m[0] - 1.0) < tol && std::abs(w.m[1]) < tol &&
                           std::abs(w.m[2]) < tol && std::abs(w.m[3] - 1.0) < tol;
        if (!ident) {
            FusedItem it;
            it.k = 1;
            it.q[0] = q;
            it.m.assign(w.m.begin(), w.m.end());
            out.push_back(std::move(it));
        }
    }
    w = Wire1q();
}

std::vector<FusedItem> fuse_1q(const Stream& st) {
    std::vector<FusedItem> out;
    uint32_t max_q = 0;
    for (int64_t i = 0; i < st.size(); ++i)
        for (int b = 0; b < st.op(i).n_qubits; ++b) max_q = std::max(max_q, st.qubits(i)[b]);
    std::vector<Wire1q> wires(max_q + 1);
    auto verbatim = [&](int64_t i) { FusedItem it; it.src = i; out.push_back(std::move(it)); };
    for (int64_t i = 0; i < st.size(); ++i) {
        const qhip_op& op = st.op(i);
        const uint32_t* q = st.qubits(i);
        if (fusible_1q(op)) {
            Wire1q& w = wires[q[0]];
            if (w.n == 0) w.first = i;
            const auto g = op_matrix(st, i);
            const auto a = w.m;   // w.m <- g * w.m
            w.m[0] = g[0] * a[0] + g[1] * a[2];
            w.m[1] = g[0] * a[1] + g[1] * a[3];
            w.m[2] = g[2] * a[0] + g[3] * a[2];
            w.m[3] = g[2] * a[1] + g[3] * a[3];
            ++w.n;
            continue;
        }
        if (op.gate == QHIP_BARRIER) {
            if (op.n_qubits == 0)
                for (uint32_t x = 0; x <= max_q; ++x) flush_wire(wires[x], x, out);
            else
                for (int b = 0; b < op.n_qubits; ++b) flush_wire(wires[q[b]], q[b], out);
            verbatim(i);
            continue;
        }
        if (op.n_qubits == 1 && op.gate != QHIP_MEASURE && op.gate != QHIP_RESET) {
            flush_wire(wires[q[0]], q[0], out);
            verbatim(i);
            continue;
        }
        if (op.n_qubits >= 2) {
            for (int b = 0; b < op.n_qubits; ++b) flush_wire(wires[q[b]], q[b], out);
            verbatim(i);
            continue;
        }
        if ((op.gate == QHIP_MEASURE || op.gate == QHIP_RESET) && op.n_qubits > 0)
            flush_wire(wires[q[0]], q[0], out);
        if (op.gate == QHIP_BRANCH_BEGIN || op.gate == QHIP_BRANCH_ELSE || op.gate == QHIP_BRANCH_END)
            for (uint32_t x = 0; x <= max_q; ++x) flush_wire(wires[x], x, out);
        verbatim(i);
    }
    for (uint32_t x = 0; x <= max_q; ++x) flush_wire(wires[x], x, out);
    return out;
}

// ── k-qubit blocks ──────────────────────────────────────────────────────────

struct Block {
    int64_t id = 0;
    std::vector<uint32_t> qubits;   // local bit b <-> qubits[b]
    std::vector<uint32_t> wires;
    std::vector<cplx> u;            // empty until a second gate folds in
    int n_gates = 0;
    int64_t first = -1;
};

// u <- (I (x) u): appended qubits become the high local bits.
void widen(Block& b, const std::vector<uint32_t>& fresh) {
    const size_t old = size_t{1} << b.qubits.size();
    const size_t copies = size_t{1} << fresh.size();
    const size_t D = old * copies;
    std::vector<cplx> w(D * D, cplx(0.0));
    for (size_t h = 0; h < copies; ++h)
        for (size_t y = 0; y < old; ++y)
            std::copy_n(&b.u[y * old], old, &w[(h * old + y) * D + h * old]);
    b.u.swap(w);
    b.qubits.insert(b.qubits.end(), fresh.begin(), fresh.end());
}

// u <- E(g) u where g acts on block-local positions pos[] (g-local bit j <->
// block bit pos[j]).  Row form: each input row i scatters coeff * u.row(i)
// into the rows reachable through g.
void fold(std::vector<cplx>& u, size_t D, const std::vector<cplx>& g,
          const std::vector<int>& pos) {
    const size_t m = pos.size(), gd = size_t{1} << m;
    std::vector<cplx> r(D * D, cplx(0.0));
    size_t pmask = 0;
    for (int p : pos) pmask |= size_t{1} << p;
    for (size_t i = 0; i < D; ++i) {
        size_t li = 0;
        for (size_t j = 0; j < m; ++j) li |= ((i >> pos[j]) & 1u) << j;
        const size_t rest = i & ~pmask;
        for (size_t lo = 0; lo < gd; ++lo) {
            const cplx c = g[lo * gd + li];
            if (c == cplx(0.0)) continue;
            size_t io = rest;
            for (size_t j = 0; j < m; ++j)
                if ((lo >> j) & 1u) io |= size_t{1} << pos[j];
            cplx* dst = &r[io * D];
            const cplx* src = &u[i * D];
            for (size_t x = 0; x < D; ++x) dst[x] += c * src[x];
        }
    }
    u.swap(r);
}

std::vector<FusedItem> fuse_blocks(const Stream& st, int kmax) {
    std::vector<FusedItem> out;
    uint32_t max_q = 0, max_c = 0;
    bool any_c = false;
    for (int64_t i = 0; i < st.size(); ++i) {
        for (int b = 0; b < st.op(i).n_qubits; ++b) max_q = std::max(max_q, st.qubits(i)[b]);
        if (st.op(i).cbit >= 0) { max_c = std::max(max_c, static_cast<uint32_t>(st.op(i).cbit)); any_c = true; }
    }
    const uint32_t nqw = max_q + 1;
    std::vector<int64_t> front(nqw + (any_c ? max_c + 1 : 0), 0);
    std::list<Block> open;
    int64_t next_id = 1;

    auto verbatim = [&](int64_t i) { FusedItem it; it.src = i; out.push_back(std::move(it)); };
    auto emit = [&](Block& b) {
        if (b.n_gates == 1) { verbatim(b.first); return; }
        if (b.qubits.size() == 1) {
            const double d = std::sqrt(std::norm(b.u[0] - 1.0) + std::norm(b.u[1]) +
                                       std::norm(b.u[2]) + std::norm(b.u[3] - 1.0));
            if (d < 1e-12) return;
        }
        FusedItem it;
        it.k = static_cast<int32_t>(b.qubits.size());
        std::copy(b.qubits.begin(), b.qubits.end(), it.q.begin());
        it.m = std::move(b.u);
        out.push_back(std::move(it));
    };
    auto flush_wires = [&](const std::vector<uint32_t>& ws) {
        std::vector<Block*> blocks;
        for (auto& b : open) blocks.push_back(&b);
        std::vector<char> mark(blocks.size(), 0);
        for (size_t j = 0; j < blocks.size(); ++j)
            for (auto w : ws)
                if (in(blocks[j]->wires, w)) { mark[j] = 1; break; }
        for (size_t j = blocks.size(); j-- > 0;) {
            if (!mark[j]) continue;
            for (size_t l = 0; l < j; ++l) {
                if (mark[l]) continue;
                for (auto w : blocks[l]->wires)
                    if (in(blocks[j]->wires, w)) { mark[l] = 1; break; }
            }
        }
        size_t j = 0;
        for (auto it = open.begin(); it != open.end(); ++j) {
            if (mark[j]) { emit(*it); it = open.erase(it); }
            else ++it;
        }
    };

    for (int64_t i = 0; i < st.size(); ++i) {
        const qhip_op& op = st.op(i);
        const uint32_t* q = st.qubits(i);
        if (op.gate == QHIP_GPHASE) { verbatim(i); continue; }
        if (op.gate == QHIP_BRANCH_BEGIN || op.gate == QHIP_BRANCH_ELSE ||
            op.gate == QHIP_BRANCH_END || (op.gate == QHIP_BARRIER && op.n_qubits == 0)) {
            for (auto& b : open) emit(b);
            open.clear();
            verbatim(i);
            continue;
        }
        std::vector<uint32_t> ws(q, q + op.n_qubits);
        if (op.cbit >= 0) ws.push_back(nqw + static_cast<uint32_t>(op.cbit));

        const bool fusible = op.n_qubits > 0 && op.n_qubits <= kmax && !is_marker(op.gate) &&
                             !(op.gate == QHIP_CUSTOM && op.unitary_offset < 0);
        if (fusible) {
            Block* target = nullptr;
            for (auto it = open.rbegin(); it != open.rend(); ++it) {
                int shared = 0, fresh = 0;
                for (int b = 0; b < op.n_qubits; ++b) (in(it->qubits, q[b]) ? shared : fresh)++;
                if (fresh > 0) {
                    const int width = static_cast<int>(it->qubits.size()) + fresh;
                    if (width > kmax || (shared == 0 && width > 2)) continue;
                }
                bool reachable = true;
                for (auto w : ws)
                    if (front[w] > it->id) { reachable = false; break; }
                if (reachable) { target = &*it; break; }
            }
            if (!target) {
                Block b;
                b.id = next_id++;
                b.qubits.assign(q, q + op.n_qubits);
                b.wires = ws;
                b.n_gates = 1;
                b.first = i;
                for (auto w : ws) front[w] = b.id;
                open.push_back(std::move(b));
                continue;
            }
            Block& b = *target;
            if (b.n_gates == 1) b.u = op_matrix(st, b.first);
            std::vector<uint32_t> fresh;
            for (int j = 0; j < op.n_qubits; ++j)
                if (!in(b.qubits, q[j])) fresh.push_back(q[j]);
            if (!fresh.empty()) widen(b, fresh);
            std::vector<int> pos(op.n_qubits);
            for (int j = 0; j < op.n_qubits; ++j)
                pos[j] = static_cast<int>(std::find(b.qubits.begin(), b.qubits.end(), q[j]) - b.qubits.begin());
            fold(b.u, size_t{1} << b.qubits.size(), op_matrix(st, i), pos);
            ++b.n_gates;
            for (auto w : ws) {
                if (!in(b.wires, w)) b.wires.push_back(w);
                front[w] = b.id;
            }
            continue;
        }
        flush_wires(ws);
        verbatim(i);
        const int64_t id = next_id++;
        for (auto w : ws) front[w] = id;
    }
    for (auto& b : open) emit(b);
    return out;
}

}  // namespace

std::vector<FusedItem> fuse_for_dispatch(const Stream& st, int n, int fusion_max, int fusion_min) {
    if (fusion_max < 0 || fusion_max > QHIP_MAX_FUSION_QUBITS)
        fail(QHIP_INVALID, "QarpSimulator::set_fusion_max_qubits: " + std::to_string(fusion_max) +
                               " exceeds the maximum block width 8");
    if (fusion_max == 0 || st.size() == 0) {
        std::vector<FusedItem> out(static_cast<size_t>(st.size()));
        for (int64_t i = 0; i < st.size(); ++i) out[i].src = i;
        return out;
    }
    if (fusion_max == 1 || std::max(n, 0) < fusion_min) return fuse_1q(st);
    return fuse_blocks(st, fusion_max);
}

}  // namespace qhip
