// derive.cpp — derivation helpers of the structured-program planner: the
// basis permutation a command stream applies to a qubit subset, exact
// including its phase, or none (permutation_table, src/qarpx/src/simulator/
// program.cpp:71-402, 714-724), and the FNV-1a mixing of the planner's
// command digests (program.cpp:739-829).
//
// Host code over small registers, as in the reference: a classical stream
// (X, CX, CCX, SWAP, CSWAP, the H·MCZ·H multi-controlled X) is followed bit
// by bit over every basis state; otherwise the qubits only ever used as
// controls, phase partners or moved classically among themselves ("fixed")
// are enumerated and the rest is simulated column by column on 2^|rest|-wide
// blocks.  The stream arrives remapped onto local qubits [0, k) and free of
// symbolic, conditioned and non-unitary commands (the Python wrapper rejects
// those first, as to_local does, program.cpp:404-436).
#include "common.h"

#include <algorithm>
#include <bit>
#include <cmath>
#include <optional>
#include <string>
#include <vector>

namespace qhip {
namespace {

constexpr int kMaxTableQubits = 26;        // program.cpp:35
constexpr int kMaxDenseBlockQubits = 8;    // widest dense command (fusion cap)
constexpr double kZero = 1e-14;            // "this matrix entry is zero"
constexpr double kBasisTolerance = 1e-10;  // a column is a basis vector times 1

uint64_t deposit(uint64_t x, const std::vector<uint32_t>& qs) {
    uint64_t i = 0;
    for (size_t b = 0; b < qs.size(); ++b) i |= ((x >> b) & 1u) << qs[b];
    return i;
}

struct ClassicalStep {
    uint64_t controls = 0;
    uint32_t target = 0, other = 0;
    bool swap = false;
};

// The stream as classical steps, or nothing when a command is not a basis
// permutation with trivial phase gate by gate (program.cpp:79-127).
std::optional<std::vector<ClassicalStep>> classical_steps(const Stream& st) {
    std::vector<ClassicalStep> steps;
    for (int64_t i = 0; i < st.size(); ++i) {
        const qhip_op& op = st.op(i);
        const uint32_t* q = st.qubits(i);
        switch (op.gate) {
            case QHIP_BARRIER: case QHIP_ID: break;
            case QHIP_X: steps.push_back({0, q[0], 0, false}); break;
            case QHIP_CX: steps.push_back({uint64_t{1} << q[0], q[1], 0, false}); break;
            case QHIP_CCX: steps.push_back({(uint64_t{1} << q[0]) | (uint64_t{1} << q[1]), q[2], 0, false}); break;
            case QHIP_SWAP: steps.push_back({0, q[0], q[1], true}); break;
            case QHIP_CSWAP: steps.push_back({uint64_t{1} << q[0], q[1], q[2], true}); break;
            case QHIP_H: {
                // H(t) · MCZ(controls..., t) · H(t): a multi-controlled X.
                if (i + 2 >= st.size()) return std::nullopt;
                const qhip_op& mcz = st.op(i + 1);
                const qhip_op& h2 = st.op(i + 2);
                const uint32_t t = q[0];
                if (mcz.gate != QHIP_MCZ || h2.gate != QHIP_H || st.qubits(i + 2)[0] != t || mcz.n_qubits == 0 ||
                    st.qubits(i + 1)[mcz.n_qubits - 1] != t)
                    return std::nullopt;
                uint64_t ctrl = 0;
                for (int b = 0; b + 1 < mcz.n_qubits; ++b) ctrl |= uint64_t{1} << st.qubits(i + 1)[b];
                steps.push_back({ctrl, t, 0, false});
                i += 2;
                break;
            }
            default: return std::nullopt;
        }
    }
    return steps;
}

std::vector<uint64_t> classical_table(const std::vector<ClassicalStep>& steps, int k) {
    const uint64_t n = uint64_t{1} << k;
    std::vector<uint64_t> table(n);
    for (uint64_t x = 0; x < n; ++x) {
        uint64_t v = x;
        for (const auto& s : steps) {
            if ((v & s.controls) != s.controls) continue;
            if (!s.swap) v ^= uint64_t{1} << s.target;
            else if (((v >> s.target) & 1) != ((v >> s.other) & 1))
                v ^= (uint64_t{1} << s.target) | (uint64_t{1} << s.other);
        }
        table[x] = v;
    }
    return table;
}

bool is_classical_gate(int g) {
    return g == QHIP_X || g == QHIP_CX || g == QHIP_CCX || g == QHIP_SWAP || g == QHIP_CSWAP;
}

uint64_t apply_classical(int gate, const std::vector<uint32_t>& q, uint64_t v) {
    auto bit = [&](uint32_t a) { return (v >> a) & 1; };
    auto swap_bits = [&](uint32_t a, uint32_t b) {
        if (bit(a) != bit(b)) v ^= (uint64_t{1} << a) | (uint64_t{1} << b);
    };
    switch (gate) {
        case QHIP_X: v ^= uint64_t{1} << q[0]; break;
        case QHIP_CX: if (bit(q[0])) v ^= uint64_t{1} << q[1]; break;
        case QHIP_CCX: if (bit(q[0]) && bit(q[1])) v ^= uint64_t{1} << q[2]; break;
        case QHIP_SWAP: swap_bits(q[0], q[1]); break;
        case QHIP_CSWAP: if (bit(q[0])) swap_bits(q[1], q[2]); break;
        default: break;
    }
    return v;
}

// A command on a small register: MCZ as its sign flip, anything else as its
// matrix over its operands (local bit b <-> qubits[b]).
struct SmallOp {
    bool mcz = false;
    std::vector<uint32_t> qubits;
    std::vector<cplx> m;   // row-major 2^k x 2^k
};

void apply_small(const SmallOp& op, std::vector<cplx>& v) {
    const uint64_t len = v.size();
    if (op.mcz) {
        uint64_t mask = 0;
        for (uint32_t q : op.qubits) mask |= uint64_t{1} << q;
        for (uint64_t i = 0; i < len; ++i)
            if ((i & mask) == mask) v[i] = -v[i];
        return;
    }
    const int k = static_cast<int>(op.qubits.size());
    const uint64_t D = uint64_t{1} << k;
    uint64_t mask = 0;
    for (uint32_t q : op.qubits) mask |= uint64_t{1} << q;
    std::vector<uint64_t> off(D);
    for (uint64_t l = 0; l < D; ++l) off[l] = deposit(l, op.qubits);
    std::vector<cplx> in(D);
    for (uint64_t base = 0; base < len; ++base) {
        if (base & mask) continue;
        for (uint64_t l = 0; l < D; ++l) in[l] = v[base | off[l]];
        for (uint64_t r = 0; r < D; ++r) {
            cplx s = 0.0;
            for (uint64_t c = 0; c < D; ++c) s += op.m[r * D + c] * in[c];
            v[base | off[r]] = s;
        }
    }
}

// Matrix of stream op i over its operands (named gates and Custom).
bool op_matrix(const Stream& st, int64_t i, std::vector<cplx>& m) {
    const qhip_op& op = st.op(i);
    if (op.gate == QHIP_CUSTOM) {
        if (op.unitary_offset < 0) return false;
        const uint64_t D = uint64_t{1} << op.n_qubits;
        const double* u = st.s->unitaries + 2 * op.unitary_offset;
        m.resize(D * D);
        for (uint64_t e = 0; e < D * D; ++e) m[e] = cplx(u[2 * e], u[2 * e + 1]);
        return true;
    }
    return named_matrix(op.gate, op.n_qubits, op.params, m);
}

struct GateView {
    int64_t idx = 0;
    int gate = 0;
    std::vector<uint32_t> qubits;
    std::vector<cplx> u;
    uint64_t changes = 0;   // bit b: couples states differing in qubits[b]
    bool classical = false;
};

// program.cpp:193-402.
std::optional<std::vector<uint64_t>> restricted_table(const Stream& st, int k, uint32_t max_rest,
                                                      uint32_t max_work_log2) {
    std::vector<GateView> gates;
    cplx global_phase(1.0, 0.0);
    for (int64_t i = 0; i < st.size(); ++i) {
        const qhip_op& op = st.op(i);
        if (op.gate == QHIP_BARRIER || op.gate == QHIP_ID) continue;
        if (op.gate == QHIP_GPHASE) {
            global_phase *= std::polar(1.0, op.params[0]);
            continue;
        }
        GateView g;
        g.idx = i;
        g.gate = op.gate;
        g.qubits.assign(st.qubits(i), st.qubits(i) + op.n_qubits);
        g.classical = is_classical_gate(op.gate);
        const int q = op.n_qubits;
        if (op.gate != QHIP_MCZ) {
            if (q == 0 || q > kMaxDenseBlockQubits) return std::nullopt;
            if (!op_matrix(st, i, g.u)) return std::nullopt;
            const uint64_t D = uint64_t{1} << q;
            for (int b = 0; b < q; ++b) {
                bool couples = false;
                for (uint64_t r = 0; !couples && r < D; ++r)
                    for (uint64_t c = 0; c < D; ++c)
                        if ((((r ^ c) >> b) & 1) && std::abs(g.u[r * D + c]) > kZero) {
                            couples = true;
                            break;
                        }
                if (couples) g.changes |= uint64_t{1} << b;
            }
        }
        gates.push_back(std::move(g));
    }
    // Rest: changed by a non-classical gate, or by a classical gate that also
    // reads a rest qubit (closure).
    std::vector<uint8_t> is_rest(static_cast<size_t>(k), 0);
    for (const auto& g : gates)
        if (!g.classical)
            for (size_t b = 0; b < g.qubits.size(); ++b)
                if ((g.changes >> b) & 1) is_rest[g.qubits[b]] = 1;
    for (bool grew = true; grew;) {
        grew = false;
        for (const auto& g : gates) {
            if (!g.classical) continue;
            bool reads_rest = false;
            for (uint32_t q : g.qubits) reads_rest |= is_rest[q] != 0;
            if (!reads_rest) continue;
            for (size_t b = 0; b < g.qubits.size(); ++b)
                if (((g.changes >> b) & 1) && !is_rest[g.qubits[b]]) {
                    is_rest[g.qubits[b]] = 1;
                    grew = true;
                }
        }
    }
    std::vector<uint32_t> fixed, rest, rest_index(static_cast<size_t>(k), UINT32_MAX);
    for (uint32_t q = 0; q < static_cast<uint32_t>(k); ++q) {
        if (is_rest[q]) {
            rest_index[q] = static_cast<uint32_t>(rest.size());
            rest.push_back(q);
        } else {
            fixed.push_back(q);
        }
    }
    if (rest.size() > max_rest) return std::nullopt;
    const uint64_t R = uint64_t{1} << rest.size();

    // Commands touching no fixed qubit, remapped once; a run of them longer
    // than one dense pass on the rest costs is compressed into that pass.
    struct Op {
        bool dependent = false;
        size_t gate = 0;
        std::vector<SmallOp> commands;
    };
    auto small_of = [&](const GateView& g) {
        SmallOp s;
        s.mcz = g.gate == QHIP_MCZ;
        for (uint32_t q : g.qubits) s.qubits.push_back(rest_index[q]);
        s.m = g.u;
        return s;
    };
    std::vector<Op> ops;
    for (size_t i = 0; i < gates.size(); ++i) {
        bool dependent = false;
        for (uint32_t q : gates[i].qubits) dependent |= !is_rest[q];
        if (dependent) {
            ops.push_back(Op{true, i, {}});
            continue;
        }
        if (ops.empty() || ops.back().dependent) ops.push_back(Op{false, 0, {}});
        ops.back().commands.push_back(small_of(gates[i]));
    }
    std::vector<cplx> column(R);
    double units = 0.0;   // per column: amplitude sweeps of length R
    for (auto& op : ops) {
        if (op.dependent) {
            units += 1.0;
            continue;
        }
        if (rest.size() <= static_cast<size_t>(kMaxDenseBlockQubits) &&
            static_cast<double>(op.commands.size()) > static_cast<double>(R)) {
            SmallOp dense;
            for (uint32_t b = 0; b < rest.size(); ++b) dense.qubits.push_back(b);
            dense.m.assign(R * R, 0.0);
            for (uint64_t c = 0; c < R; ++c) {
                std::fill(column.begin(), column.end(), cplx(0.0, 0.0));
                column[c] = 1.0;
                for (const auto& cmd : op.commands) apply_small(cmd, column);
                for (uint64_t r = 0; r < R; ++r) dense.m[r * R + c] = column[r];
            }
            op.commands.assign(1, std::move(dense));
            units += static_cast<double>(R);
        } else {
            units += static_cast<double>(op.commands.size());
        }
    }
    // Work: 2^|fixed| assignments x R columns x `units` sweeps of R amplitudes.
    const double work = std::ldexp(units * static_cast<double>(R) * static_cast<double>(R),
                                   static_cast<int>(fixed.size()));
    if (work > std::ldexp(static_cast<double>(std::max<int64_t>(st.size(), 1)), static_cast<int>(max_work_log2)))
        return std::nullopt;

    const uint64_t n_fixed = uint64_t{1} << fixed.size();
    std::vector<uint64_t> table(uint64_t{1} << k);
    std::vector<uint8_t> hit(R);
    std::vector<SmallOp> reduced;
    std::vector<cplx> block(R * R);
    for (uint64_t a = 0; a < n_fixed; ++a) {
        const uint64_t start = deposit(a, fixed);
        uint64_t cur = start;
        cplx phase = global_phase;
        reduced.clear();
        for (const auto& op : ops) {
            if (!op.dependent) {
                reduced.insert(reduced.end(), op.commands.begin(), op.commands.end());
                continue;
            }
            const GateView& g = gates[op.gate];
            const auto& qs = g.qubits;
            bool touches_rest = false;
            for (uint32_t q : qs) touches_rest |= is_rest[q] != 0;
            if (g.classical && !touches_rest) {
                cur = apply_classical(g.gate, qs, cur);
                continue;
            }
            std::vector<uint32_t> live;
            std::vector<size_t> live_pos;
            uint64_t base = 0;
            bool active = true;
            for (size_t b = 0; b < qs.size(); ++b) {
                if (!is_rest[qs[b]]) {
                    const uint64_t bit = (cur >> qs[b]) & 1;
                    base |= bit << b;
                    active &= bit != 0;
                } else {
                    live.push_back(rest_index[qs[b]]);
                    live_pos.push_back(b);
                }
            }
            if (g.gate == QHIP_MCZ) {
                // -1 exactly where every qubit is 1: the fixed ones must read 1.
                if (!active) continue;
                if (live.empty()) {
                    phase = -phase;
                } else {
                    SmallOp z;
                    z.mcz = true;   // Z on one qubit is the one-qubit sign flip
                    z.qubits = live;
                    reduced.push_back(std::move(z));
                }
                continue;
            }
            const uint64_t D = uint64_t{1} << qs.size();
            const uint64_t d = uint64_t{1} << live_pos.size();
            std::vector<cplx> sub(d * d);
            for (uint64_t r = 0; r < d; ++r)
                for (uint64_t c = 0; c < d; ++c) {
                    uint64_t rr = base, cc = base;
                    for (size_t j = 0; j < live_pos.size(); ++j) {
                        rr |= ((r >> j) & 1) << live_pos[j];
                        cc |= ((c >> j) & 1) << live_pos[j];
                    }
                    sub[r * d + c] = g.u[rr * D + cc];
                }
            if (live.empty()) {
                phase *= sub[0];
                continue;
            }
            SmallOp custom;
            custom.qubits = live;
            custom.m = std::move(sub);
            reduced.push_back(std::move(custom));
        }
        // All R columns at once: entry (row, x) at row + R x of a 2|rest|-qubit
        // vector, so a command on rest qubit j (< |rest|) acts on every column.
        std::fill(block.begin(), block.end(), cplx(0.0, 0.0));
        for (uint64_t x = 0; x < R; ++x) block[x + R * x] = 1.0;
        for (const auto& c : reduced) apply_small(c, block);
        std::fill(hit.begin(), hit.end(), 0);
        for (uint64_t x = 0; x < R; ++x) {
            const cplx* col = block.data() + R * x;
            uint64_t y = 0;
            double best = -1.0;
            for (uint64_t j = 0; j < R; ++j) {
                const double m = std::norm(col[j]);
                if (m > best) {
                    best = m;
                    y = j;
                }
            }
            if (std::abs(col[y] * phase - cplx(1.0, 0.0)) > kBasisTolerance) return std::nullopt;
            for (uint64_t j = 0; j < R; ++j)
                if (j != y && std::abs(col[j]) > kBasisTolerance) return std::nullopt;
            if (hit[y]) return std::nullopt;
            hit[y] = 1;
            table[start | deposit(x, rest)] = cur | deposit(y, rest);
        }
    }
    return table;
}

}  // namespace

bool permutation_table(const qhip_stream* stream, int k, uint32_t max_rest, uint32_t max_work_log2,
                       uint64_t* out) {
    if (k < 1 || k > kMaxTableQubits) return false;
    const Stream st{stream};
    for (int64_t i = 0; i < st.size(); ++i)
        for (int b = 0; b < st.op(i).n_qubits; ++b)
            if (st.qubits(i)[b] >= static_cast<uint32_t>(k))
                fail(QHIP_INVALID, "qhip_permutation_table: command qubit outside the k local qubits");
    std::optional<std::vector<uint64_t>> t;
    if (auto steps = classical_steps(st)) t = classical_table(*steps, k);
    else t = restricted_table(st, k, max_rest, max_work_log2);
    if (!t) return false;
    std::copy(t->begin(), t->end(), out);
    return true;
}

uint64_t fnv1a_words(const uint64_t* words, int64_t n, uint64_t h) {
    for (int64_t i = 0; i < n; ++i)
        for (int b = 0; b < 8; ++b) {
            h ^= (words[i] >> (8 * b)) & 0xff;
            h *= 0x100000001b3ULL;
        }
    return h;
}

}  // namespace qhip
