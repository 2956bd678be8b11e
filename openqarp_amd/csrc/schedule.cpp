// schedule.cpp — packs an op sequence into tile passes.
//
// A pass streams the whole register through the chip once: every workgroup
// loads one tile of 2^T amplitudes (the tile qubits are the `low` contiguous
// qubits, for coalesced 16-B loads, plus up to T-low other qubits), applies a
// run of ops in LDS, and writes the tile back.  A dense op can join a pass
// only if all its operands are tile qubits; diagonal / mask-phase / scalar
// ops know each amplitude's full index and join any pass.  An op may run
// once every earlier op sharing a qubit with it has run (ops on disjoint
// qubits commute, which is what lets them overtake); each pass takes the
// tile qubits that let the most such ops run (pick_pass below), and runs
// them in program order.
#include "common.h"

#include <algorithm>
#include <bit>
#include <map>

namespace qhip {

static const cplx I0{0.0, 0.0}, I1{1.0, 0.0};

int default_tile_qubits(int precision, bool two_states) {
    // 32 KiB of LDS per state (2^11 complex128 or 2^12 complex64
    // amplitudes): several workgroups share a CU's 160 KiB, so one tile's
    // ops overlap other tiles' loads and stores.
    (void)two_states;
    return precision == QHIP_C128 ? 11 : 12;
}

static double op_flops(const HostOp& op) {
    switch (op.kind) {
        case OP_DENSE: return 8.0 * static_cast<double>(1 << op.k);
        case OP_IP_DENSE: return 8.0 * static_cast<double>(1 << op.k) + 8.0;
        case OP_DIAG: case OP_MASKPHASE: case OP_SCALAR: return 6.0;
        case OP_IP_DIAG: return 14.0;
        case OP_PHASE: return 36.0;
        case OP_IP_PHASE: return 24.0;
        default: return 0.0;
    }
}

// Local positions of a dense op's group-number bits (the tile positions that
// are not operands).  The tile kernel keeps LDS slot l at
// l ^ (((l >> 4) ^ (l >> 8)) & 15), so position p < 12 lands on slot bit
// p % 4.  Group bits 0-3 index lanes 0-15 of a wavefront: they take one free
// position per residue 0..3 where the operands leave one, so those 16 lanes
// touch 16 distinct 16-B slots modulo 16 (no bank conflict); the other
// positions follow in ascending order.
static void group_positions(DevOp& d, int T) {
    bool operand[64] = {};
    for (int b = 0; b < d.k; ++b) operand[d.pos[b]] = true;
    std::vector<int> free;
    for (int p = 0; p < T; ++p)
        if (!operand[p]) free.push_back(p);
    std::vector<int> order;
    std::vector<bool> used(free.size(), false);
    for (int r = 0; r < 4; ++r)
        for (size_t i = 0; i < free.size(); ++i)
            if (!used[i] && free[i] < 12 && free[i] % 4 == r) {
                order.push_back(free[i]);
                used[i] = true;
                break;
            }
    for (size_t i = 0; i < free.size(); ++i)
        if (!used[i]) order.push_back(free[i]);
    if (order.size() > sizeof(d.gpos)) fail(QHIP_RUNTIME, "schedule: tile too wide for the group map");
    d.gmask = 0;
    for (size_t b = 0; b < order.size(); ++b) {
        d.gpos[b] = static_cast<int8_t>(order[b]);
        if (b >= 8) d.gmask |= uint64_t{1} << order[b];
    }
    // Positions of bits >= 8 ascend, so the kernel can step through their
    // deposits with (x - mask) & mask.
}

// MFMA fragments of a dense k-qubit op (k = 3..5) for v_mfma_f64_16x16x4.
// k = 4, 5 use three real products (see the k >= 4 branch); k = 3 the
// realified matrix below.
// Real form: M[(b,q)][(a,p)] over output amplitude b / part q and input
// amplitude a / part p (part 0 = re, 1 = im): re U_ba, -im U_ba, im U_ba,
// re U_ba for (q,p) = (0,0), (0,1), (1,0), (1,1).  The instruction holds
// A[i][c] in lane (i = lane % 16, c = lane / 16) and returns D[4r + lane/16]
// [lane % 16] in register r.  Rows: M-tile mt, row i = 4r + g is output
// amplitude 8mt + 2g + (r >> 1), part r & 1, so lane group g receives two
// whole amplitudes.  K-steps: step s, column c is input amplitude
// 4(s >> 1) + c, part s & 1, so lane group c feeds whole amplitudes c,
// c + 4, ....  Stored as complex pairs (steps 2t, 2t+1) at
// [(mt * D/4 + t) * 64 + lane].
void mfma_fragments(const std::vector<cplx>& U, int k, std::vector<cplx>& out, bool f32) {
    const int D = 1 << k;
    if (k >= 4) {   // three real products (Gauss), layout of the k = 4, 5 kernel
        const int NT = D / 16, HP = D / 8;
        auto u = [&](int b, int a) { return U[static_cast<size_t>(b) * D + a]; };
        for (int mt = 0; mt < NT; ++mt)
            for (int t = 0; t < HP; ++t)
                for (int p = 0; p < 3; ++p)
                    for (int lane = 0; lane < 64; ++lane) {
                        const int b = 16 * mt + lane % 16, c = lane / 16;
                        double v[2];
                        for (int h = 0; h < 2; ++h) {
                            const cplx e = u(b, 4 * (2 * t + h) + c);
                            v[h] = p == 0 ? e.real() : p == 1 ? e.imag() : e.real() + e.imag();
                        }
                        out.emplace_back(v[0], v[1]);
                    }
        return;
    }
    const int NT = D / 8, S = D / 2;
    auto M = [&](int b, int q, int a, int p) -> double {
        const cplx u = U[static_cast<size_t>(b) * D + a];
        if (q == 0) return p == 0 ? u.real() : -u.imag();
        return p == 0 ? u.imag() : u.real();
    };
    auto frag = [&](int mt, int s, int lane) {
        const int i = lane % 16, c = lane / 16;
        const int g = f32 ? i / 4 : i % 4, r = f32 ? i % 4 : i / 4;   // result row i = 4g + r or 4r + g
        return M(8 * mt + 2 * g + (r >> 1), r & 1, 4 * (s >> 1) + c, s & 1);
    };
    for (int mt = 0; mt < NT; ++mt)
        for (int t = 0; t < S / 2; ++t)
            for (int lane = 0; lane < 64; ++lane)
                out.emplace_back(frag(mt, 2 * t, lane), frag(mt, 2 * t + 1, lane));
}

// Inverse of mfma_fragments (schedule inspection).
std::vector<cplx> mfma_matrix(const cplx* f, int k, bool f32) {
    const int D = 1 << k;
    std::vector<cplx> U(static_cast<size_t>(D) * D);
    if (k >= 4) {
        const int NT = D / 16, HP = D / 8;
        for (int mt = 0; mt < NT; ++mt)
            for (int t = 0; t < HP; ++t)
                for (int lane = 0; lane < 64; ++lane) {
                    const int b = 16 * mt + lane % 16, c = lane / 16;
                    const cplx re = f[((mt * HP + t) * 3 + 0) * 64 + lane];
                    const cplx im = f[((mt * HP + t) * 3 + 1) * 64 + lane];
                    U[static_cast<size_t>(b) * D + 4 * (2 * t) + c] = cplx(re.real(), im.real());
                    U[static_cast<size_t>(b) * D + 4 * (2 * t + 1) + c] = cplx(re.imag(), im.imag());
                }
        return U;
    }
    const int NT = D / 8, S = D / 2;
    for (int mt = 0; mt < NT; ++mt)
        for (int t = 0; t < S / 2; ++t)
            for (int lane = 0; lane < 64; ++lane) {
                const cplx v = f[(mt * (S / 2) + t) * 64 + lane];
                const int i = lane % 16, c = lane / 16, g = f32 ? i / 4 : i % 4, r = f32 ? i % 4 : i / 4;
                const int b = 8 * mt + 2 * g + (r >> 1), q = r & 1, a = 4 * t + c;   // step 2t: part 0
                auto& u = U[static_cast<size_t>(b) * D + a];
                if (q == 0) u.real(v.real());
                else u.imag(v.real());
            }
    return U;
}

// The records of a merged phase op (additive = false: factors exp(i w),
// constant exp(i c), row-pair table of products) or of an inner-product
// phase op (additive: weights w + i wi, constant HostOp::phase, row-pair
// table of signed sums).  Terms by where their operands sit: an operand at
// tile position p < 8 (thread bit p) reads bit 56 + p of the thread's key,
// one outside the tile its global bit (n <= 55), a missing one bit 55
// (always clear); an operand at position 8 + m is row bit m.  Each term
// takes two records, its weight then {operand code, 0}.  cnt: the none-row
// term count, 1 when a row-pair table follows, the term counts of rows 0..5.
static void emit_phase_terms(Plan& plan, const HostOp& op, const int* local_of, const int* pos, int T,
                             bool additive, uint64_t cnt[8]) {
    std::vector<cplx> none_row, rows[6], pairs;
    auto where = [&](int q, int& row) {
        row = -1;
        if (q < 0) return 55;
        const int l = local_of[q];
        if (l < 0) return pos[q];
        if (l >= 8) {
            row = l - 8;
            return 0;
        }
        return 56 + l;
    };
    auto put = [&](std::vector<cplx>& v, const PhaseTerm& t, int code) {
        v.push_back(additive ? cplx(t.w, t.wi) : std::polar(1.0, t.w));
        v.emplace_back(static_cast<double>(code), 0.0);
    };
    for (const PhaseTerm& t : op.terms) {
        int ra, rb;
        const int ka = where(t.a, ra), kb = where(t.b, rb);
        if (ra < 0 && rb < 0) put(none_row, t, ka | kb << 8);
        else if (ra >= 0 && rb >= 0) put(pairs, t, ra | rb << 8);
        else if (ra >= 0) put(rows[ra], t, kb);
        else put(rows[rb], t, ka);
    }
    plan.coefs.push_back(additive ? op.phase : std::polar(1.0, op.angle0));
    plan.coefs.insert(plan.coefs.end(), none_row.begin(), none_row.end());
    for (const auto& r : rows) plan.coefs.insert(plan.coefs.end(), r.begin(), r.end());
    // Terms with both operands on row bits depend on the row code alone:
    // one table over the 2^(T-8) row codes.
    const int n_rows = T > 8 ? T - 8 : 0;
    if (!pairs.empty())
        for (uint32_t rr = 0; rr < (1u << n_rows); ++rr) {
            cplx f = additive ? cplx(0.0, 0.0) : cplx(1.0, 0.0);
            for (size_t t = 0; t < pairs.size(); t += 2) {
                const int code = static_cast<int>(pairs[t + 1].real());
                const bool odd = ((rr >> (code & 0xff)) ^ (rr >> (code >> 8))) & 1u;
                if (additive) f += odd ? -pairs[t] : pairs[t];
                else f *= odd ? std::conj(pairs[t]) : pairs[t];
            }
            plan.coefs.push_back(f);
        }
    cnt[0] = none_row.size() / 2;
    cnt[1] = pairs.empty() ? 0 : 1;
    for (int m = 0; m < 6; ++m) cnt[2 + m] = rows[m].size() / 2;
}

// local_of[q]: tile position of qubit q (-1 outside the tile); pos[q]: the
// global bit that holds qubit q (the identity unless passes relabel).
static void emit_op(Plan& plan, const HostOp& op, const int* local_of, const int* pos, int n, int T,
                    const ScheduleOptions& opt) {
    DevOp d{};
    d.kind = op.kind;
    d.k = op.k;
    d.coef = static_cast<int64_t>(plan.coefs.size());
    d.lmask = d.gmask = 0;
    switch (op.kind) {
        case OP_DENSE: case OP_IP_DENSE: {
            int sorted[8];
            for (int b = 0; b < op.k; ++b) sorted[b] = d.pos[b] = static_cast<int8_t>(local_of[op.q[b]]);
            std::sort(sorted, sorted + op.k);
            for (int b = 0; b < op.k; ++b) d.spos[b] = static_cast<int8_t>(sorted[b]);
            group_positions(d, T);
            if (op.kind == OP_DENSE && op.k >= 3 && op.k <= 5 && opt.mfma != 0 && T - op.k >= 4) {
                d.kind = OP_MFMA;
                mfma_fragments(op.m, op.k, plan.coefs, opt.precision != QHIP_C128);
                ++plan.mfma_ops;
            } else {
                plan.coefs.insert(plan.coefs.end(), op.m.begin(), op.m.end());
            }
            ++plan.dense_ops;
            break;
        }
        case OP_DIAG: case OP_IP_DIAG: {
            int lb[8], gb[8], kl = 0, kg = 0;   // operand indices inside / outside the tile
            for (int b = 0; b < op.k; ++b) (local_of[op.q[b]] >= 0 ? lb[kl++] : gb[kg++]) = b;
            if (T > 0 && kl <= 2 && kg <= 6) {
                // Grouped form: no per-amplitude gather of the entries.
                d.kind = op.kind == OP_DIAG ? OP_DIAGG : OP_IP_DIAGG;
                d.k = kl;
                for (int i = 0; i < kl; ++i) d.pos[i] = static_cast<int8_t>(local_of[op.q[lb[i]]]);
                if (kl > 0) group_positions(d, T);
                for (int j = 0; j < kg; ++j) d.spos[j] = static_cast<int8_t>(pos[op.q[gb[j]]]);
                d.spos[7] = static_cast<int8_t>(kg);
                for (uint32_t gi = 0; gi < (1u << kg); ++gi)
                    for (uint32_t li = 0; li < (1u << kl); ++li) {
                        uint32_t idx = 0;
                        for (int i = 0; i < kl; ++i) idx |= ((li >> i) & 1u) << lb[i];
                        for (int j = 0; j < kg; ++j) idx |= ((gi >> j) & 1u) << gb[j];
                        plan.coefs.push_back(op.m[idx]);
                    }
            } else {
                for (int b = 0; b < op.k; ++b)
                    d.pos[b] = static_cast<int8_t>(local_of[op.q[b]] >= 0 ? local_of[op.q[b]]
                                                                           : 32 + pos[op.q[b]]);
                plan.coefs.insert(plan.coefs.end(), op.m.begin(), op.m.end());
            }
            ++plan.diag_ops;
            break;
        }
        case OP_MASKPHASE:
            for (int q = 0; q < n; ++q)
                if ((op.mask >> q) & 1) {
                    if (local_of[q] >= 0) d.lmask |= uint64_t{1} << local_of[q];
                    else d.gmask |= uint64_t{1} << pos[q];
                }
            plan.coefs.push_back(op.phase);
            ++plan.diag_ops;
            break;
        case OP_SCALAR:
            plan.coefs.push_back(op.phase);
            ++plan.diag_ops;
            break;
        case OP_PHASE: {
            uint64_t cnt[8];
            emit_phase_terms(plan, op, local_of, pos, T, false, cnt);
            d.lmask = cnt[0] | cnt[1] << 16 | cnt[2] << 32 | cnt[3] << 48;
            d.gmask = cnt[4] | cnt[5] << 16 | cnt[6] << 32 | cnt[7] << 48;
            d.k = 0;
            ++plan.diag_ops;
            break;
        }
        case OP_IP_PHASE: {
            uint64_t cnt[8];
            emit_phase_terms(plan, op, local_of, pos, T, true, cnt);
            d.gmask = cnt[0] | cnt[1] << 16 | cnt[2] << 32 | cnt[3] << 48;
            const uint64_t g0 = cnt[4] | cnt[5] << 16 | cnt[6] << 32 | cnt[7] << 48;
            std::memcpy(d.gpos, &g0, sizeof(g0));
            d.k = 0;
            ++plan.diag_ops;
            break;
        }
    }
    if (op.kind == OP_IP_DENSE || op.kind == OP_IP_DIAG || op.kind == OP_IP_PHASE)
        d.lmask = static_cast<uint64_t>(op.slot);
    plan.ops.push_back(d);
    plan.flops += op_flops(op) * static_cast<double>(uint64_t{1} << n);
}

// Merged phase ops.  A diagonal op on one or two qubits, a mask phase on
// one or two, or a scalar, all with entries of modulus one, is
// exp(i theta) with theta a polynomial in the z_q = 1 - 2 j_q of its
// qubits: the Walsh coefficients of its entries' angles.  A run of at least
// kMinPhaseRun such ops, consecutive in a pass's op order, becomes one
// OP_PHASE (one LDS round trip and a few complex products per amplitude
// instead of one round trip per op); its terms on equal qubit sets are
// summed.
constexpr int kMinPhaseRun = 4;
constexpr size_t kMaxPhaseTerms = 4096;

using PhaseAcc = std::map<std::pair<int, int>, double>;

static bool unit_modulus(cplx v) { return std::abs(std::abs(v) - 1.0) <= 1e-14; }

static bool phase_walsh(const HostOp& op, double& c, PhaseAcc* acc) {
    auto add = [&](int a, int b, double w) {
        if (acc && w != 0.0) (*acc)[{a, b}] += w;
    };
    switch (op.kind) {
        case OP_SCALAR:
            if (!unit_modulus(op.phase)) return false;
            c += std::arg(op.phase);
            return true;
        case OP_MASKPHASE: {
            const int nq = std::popcount(op.mask);
            if (nq < 1 || nq > 2 || !unit_modulus(op.phase)) return false;
            const double f = std::arg(op.phase);   // angle where every operand bit is set
            const int a = std::countr_zero(op.mask);
            if (nq == 1) {
                c += f / 2;
                add(a, -1, -f / 2);
                return true;
            }
            const int b = 63 - std::countl_zero(op.mask);
            c += f / 4;
            add(a, -1, -f / 4);
            add(b, -1, -f / 4);
            add(a, b, f / 4);
            return true;
        }
        case OP_DIAG: {
            if (op.k < 1 || op.k > 2) return false;
            for (const cplx& e : op.m)
                if (!unit_modulus(e)) return false;
            const int q0 = static_cast<int>(op.q[0]);
            if (op.k == 1) {
                const double f0 = std::arg(op.m[0]), f1 = std::arg(op.m[1]);
                c += (f0 + f1) / 2;
                add(q0, -1, (f0 - f1) / 2);
                return true;
            }
            const int q1 = static_cast<int>(op.q[1]);
            // entry index = bit of q0 | bit of q1 << 1
            const double f0 = std::arg(op.m[0]), f1 = std::arg(op.m[1]), f2 = std::arg(op.m[2]),
                         f3 = std::arg(op.m[3]);
            c += (f0 + f1 + f2 + f3) / 4;
            add(q0, -1, (f0 - f1 + f2 - f3) / 4);
            add(q1, -1, (f0 + f1 - f2 - f3) / 4);
            add(std::min(q0, q1), std::max(q0, q1), (f0 - f1 - f2 + f3) / 4);
            return true;
        }
        default:
            return false;
    }
}

// The pass's ops (chosen, in order) with each run of mergeable phase ops
// replaced by one OP_PHASE; false (out untouched) when no run qualifies.
static bool merge_phase_runs(const std::vector<HostOp>& ops, const std::vector<size_t>& chosen,
                             std::vector<HostOp>& out) {
    std::vector<char> ok(chosen.size());
    double scratch = 0.0;
    bool any = false;
    for (size_t i = 0, run = 0; i < chosen.size(); ++i) {
        ok[i] = phase_walsh(ops[chosen[i]], scratch, nullptr);
        run = ok[i] ? run + 1 : 0;
        any = any || run >= static_cast<size_t>(kMinPhaseRun);
    }
    if (!any) return false;
    out.clear();
    size_t i = 0;
    while (i < chosen.size()) {
        size_t j = i;
        while (j < chosen.size() && ok[j]) ++j;
        if (j - i < static_cast<size_t>(kMinPhaseRun)) {
            for (size_t e = i; e < std::max(j, i + 1); ++e) out.push_back(ops[chosen[e]]);
            i = std::max(j, i + 1);
            continue;
        }
        HostOp m;
        PhaseAcc acc;
        auto flush = [&] {
            m.kind = OP_PHASE;
            for (const auto& [key, w] : acc) m.terms.push_back(PhaseTerm{key.first, key.second, w});
            out.push_back(std::move(m));
            m = HostOp();
            acc.clear();
        };
        for (size_t e = i; e < j; ++e) {
            if (acc.size() + 3 > kMaxPhaseTerms) flush();
            const HostOp& op = ops[chosen[e]];
            phase_walsh(op, m.angle0, &acc);
            m.mask |= op.qmask();
        }
        flush();
        i = j;
    }
    return true;
}

// Adjoint runs.  In the reverse sweep every parametric gate contributes
// its generator's inner product, then its un-apply.  Where a run of such
// entries is all diagonal (the un-applies are mergeable phase ops, the
// generators real diagonals on at most two qubits: Z, ZZ, projectors,
// |1><1| (x) Z, GPhase's identity), the matrices commute, so every inner
// product of the run may be taken before any of its un-applies.  Per
// parameter p the run's generators then add up to one diagonal D_p = sum of
// alpha c G over the gates depending on p (alpha the dependency coefficient,
// c the generator's prefactor), whose Walsh coefficients are complex: one
// OP_IP_PHASE per parameter, its slot adding 2 Re <phi|D_p|psi> to p, then
// one OP_PHASE for the un-applies.  Done where it pays: four or more inner
// products per parameter (a QAOA cost layer: 51 RZZ gates, one parameter).
using PhaseAccC = std::map<std::pair<int, int>, cplx>;
constexpr size_t kMaxAdjointRun = 1024;   // entries per merged run (terms stay below 4096)

static bool real_diag_ip(const HostOp& op) {
    if (op.kind != OP_IP_DIAG || op.k > 2) return false;
    for (const cplx& e : op.m)
        if (e.imag() != 0.0) return false;
    return true;
}

// Adds f times the Walsh coefficients of a real diagonal generator.
static void ip_walsh(const HostOp& op, cplx f, cplx& k0, PhaseAccC& acc) {
    auto add = [&](int a, int b, cplx w) {
        if (w != cplx(0.0, 0.0)) acc[{a, b}] += w;
    };
    const std::vector<cplx>& e = op.m;
    if (op.k == 0) {
        k0 += f * e[0].real();
        return;
    }
    const int q0 = static_cast<int>(op.q[0]);
    if (op.k == 1) {
        k0 += f * ((e[0].real() + e[1].real()) / 2);
        add(q0, -1, f * ((e[0].real() - e[1].real()) / 2));
        return;
    }
    const int q1 = static_cast<int>(op.q[1]);
    const double e0 = e[0].real(), e1 = e[1].real(), e2 = e[2].real(), e3 = e[3].real();
    k0 += f * ((e0 + e1 + e2 + e3) / 4);
    add(q0, -1, f * ((e0 - e1 + e2 - e3) / 4));
    add(q1, -1, f * ((e0 + e1 - e2 - e3) / 4));
    add(std::min(q0, q1), std::max(q0, q1), f * ((e0 - e1 - e2 + e3) / 4));
}

void merge_adjoint_runs(const Stream& st, int n, std::vector<GradSlot>& slots, std::vector<HostOp>& back) {
    if (n > 55) return;   // the thread key of the phase ops holds global bits below 55
    std::vector<HostOp> out;
    out.reserve(back.size());
    double scratch = 0.0;
    size_t i = 0;
    while (i < back.size()) {
        size_t j = i;
        size_t n_ip = 0;
        while (j < back.size() && j - i < kMaxAdjointRun) {
            if (real_diag_ip(back[j])) ++n_ip;
            else if (!phase_walsh(back[j], scratch, nullptr)) break;
            ++j;
        }
        std::vector<int32_t> params;
        for (size_t e = i; e < j && n_ip >= 4; ++e) {
            if (back[e].kind != OP_IP_DIAG) continue;
            const qhip_op& g = st.op(slots[back[e].slot].op);
            for (int d = 0; d < g.n_deps; ++d) {
                const int32_t p = st.s->dep_index[g.dep_offset + d];
                if (p >= 0 && std::find(params.begin(), params.end(), p) == params.end()) params.push_back(p);
            }
        }
        if (n_ip < 4 || params.empty() || params.size() * 4 > n_ip) {
            const size_t stop = std::max(j, i + 1);
            for (size_t e = i; e < stop; ++e) out.push_back(std::move(back[e]));
            i = stop;
            continue;
        }
        uint64_t mask = 0;
        for (size_t e = i; e < j; ++e) mask |= back[e].qmask();
        for (const int32_t p : params) {
            HostOp ip;
            ip.kind = OP_IP_PHASE;
            ip.mask = mask;
            ip.phase = cplx(0.0, 0.0);
            PhaseAccC acc;
            for (size_t e = i; e < j; ++e) {
                if (back[e].kind != OP_IP_DIAG) continue;
                const GradSlot& gs = slots[back[e].slot];
                const qhip_op& g = st.op(gs.op);
                for (int d = 0; d < g.n_deps; ++d)
                    if (st.s->dep_index[g.dep_offset + d] == p)
                        ip_walsh(back[e], gs.c * st.s->dep_alpha[g.dep_offset + d], ip.phase, acc);
            }
            for (const auto& [key, w] : acc) ip.terms.push_back(PhaseTerm{key.first, key.second, w.real(), w.imag()});
            ip.slot = static_cast<int32_t>(slots.size());
            slots.push_back(GradSlot{cplx(1.0, 0.0), -1, p});
            out.push_back(std::move(ip));
        }
        HostOp ph;   // the un-applies
        ph.kind = OP_PHASE;
        PhaseAcc pacc;
        for (size_t e = i; e < j; ++e) {
            if (back[e].kind == OP_IP_DIAG) {
                slots[back[e].slot].param = kMergedSlot;
                continue;
            }
            phase_walsh(back[e], ph.angle0, &pacc);
            ph.mask |= back[e].qmask();
        }
        for (const auto& [key, w] : pacc) ph.terms.push_back(PhaseTerm{key.first, key.second, w});
        out.push_back(std::move(ph));
        i = j;
    }
    back = std::move(out);   // every entry was moved into out
}

// Register blocks.  Dense ops on at most two qubits are gathered into blocks
// whose qubit union has at most three qubits; the kernel loads a block's
// 2^3-amplitude groups into registers once, applies every member op there
// and stores once, so LDS traffic and addressing are paid per block, not per
// op.  An op joins the latest block that touches its qubits (it must follow
// that block, and no later unit touches them, so moving it up commutes it
// past disjoint ops only), or else a block after the latest unit touching
// its qubits (a merged phase op touches many qubits and is no block; the
// ops after it still share blocks); an op touching no unit may join any
// block.
struct Unit {
    std::vector<size_t> ops;
    uint64_t qset = 0;   // tile qubits (a block's bits)
    uint64_t qall = 0;   // every qubit its ops touch (ordering)
    bool block_ok = false;
};

static bool is_ip_host(const HostOp& op) {
    return op.kind == OP_IP_DENSE || op.kind == OP_IP_DIAG || op.kind == OP_IP_PHASE;
}

// Members: dense ops on at most two qubits, and diagonal ops with at most
// two operands on tile qubits (tile mask tq) and up to four outside the
// tile, whose bits select a sub-table once per tile; in two-state passes
// also the inner-product forms of both (at most kMaxBlockIp per block),
// which a block evaluates on its registers.
static std::vector<Unit> form_units(const std::vector<HostOp>& ops, const std::vector<size_t>& chosen, bool enable,
                                    bool two, uint64_t tq) {
    std::vector<Unit> units;
    std::vector<int> n_ip;
    for (size_t idx : chosen) {
        const HostOp& op = ops[idx];
        const uint64_t Q = op.qmask(), Ql = Q & tq;
        const bool ip = is_ip_host(op);
        // Mask phases (CZ, MCZ) join a block as diagonal members.
        const bool diag = op.kind == OP_DIAG || (two && op.kind == OP_IP_DIAG) || op.kind == OP_MASKPHASE;
        const bool kind_ok = (op.kind == OP_DENSE && op.k <= 2) || (two && op.kind == OP_IP_DENSE && op.k <= 2) ||
                             (diag && std::popcount(Ql) <= 2 && std::popcount(Q & ~tq) <= 4);
        const bool mergeable = enable && kind_ok && (op.k >= 1 || op.kind == OP_MASKPHASE);
        int last = -1;
        for (int u = static_cast<int>(units.size()) - 1; u >= 0; --u)
            if (op.kind == OP_SCALAR || (units[u].qall & Q)) { last = u; break; }
        int target = -1;
        auto fits = [&](int u) {
            return units[u].block_ok && std::popcount(units[u].qset | Ql) <= 3 && (!ip || n_ip[u] < kMaxBlockIp);
        };
        if (mergeable) {
            // The latest unit touching the op's qubits first (it keeps the
            // op with them), else the latest block after it (no unit after
            // `last` touches the op's qubits, so the op commutes there).
            if (last >= 0 && fits(last)) {
                target = last;
            } else {
                for (int u = static_cast<int>(units.size()) - 1; u > last; --u)
                    if (fits(u)) { target = u; break; }
            }
        }
        if (target >= 0) {
            units[target].ops.push_back(idx);
            units[target].qset |= Ql;
            units[target].qall |= Q;
            n_ip[target] += ip;
        } else {
            Unit u;
            u.ops.push_back(idx);
            u.qset = Ql;
            u.qall = Q;
            u.block_ok = mergeable;
            units.push_back(std::move(u));
            n_ip.push_back(ip);
        }
    }
    return units;
}

// Emits a block of >= 2 member ops: the OP_BLOCK header, then one OP_SUB per
// member with its matrix in block-bit order.
static void emit_block(Plan& plan, const std::vector<HostOp>& ops, const Unit& u, const int* local_of, const int* pos,
                       int n, int T) {
    int bp[3];
    int m = 0;
    for (int q = 0; q < n; ++q)
        if ((u.qset >> q) & 1) bp[m++] = local_of[q];
    // Pad to three block bits with the lowest free tile positions.
    for (int p = 0; m < 3 && p < T; ++p) {
        bool used = false;
        for (int i = 0; i < m; ++i) used |= bp[i] == p;
        if (!used) bp[m++] = p;
    }
    std::sort(bp, bp + 3);
    DevOp h{};
    h.kind = OP_BLOCK;
    h.k = 3;
    for (int b = 0; b < 3; ++b) h.pos[b] = h.spos[b] = static_cast<int8_t>(bp[b]);
    group_positions(h, T);
    h.lmask = u.ops.size();
    h.coef = static_cast<int64_t>(plan.coefs.size());
    const size_t hdr_index = plan.ops.size();   // spos[7] receives the inner-product member count
    plan.ops.push_back(h);
    auto bit_of = [&](uint32_t q) {
        for (int b = 0; b < 3; ++b)
            if (bp[b] == local_of[q]) return b;
        fail(QHIP_RUNTIME, "schedule: block operand outside the block");
    };
    int n_ip = 0;
    auto sw = [](int x) { return ((x & 1) << 1) | ((x >> 1) & 1); };
    for (size_t idx : u.ops) {
        HostOp as_diag;   // a mask phase as the diagonal over its operands
        if (ops[idx].kind == OP_MASKPHASE) {
            const HostOp& mp = ops[idx];
            as_diag.kind = OP_DIAG;
            for (int q = 0; q < 64; ++q)
                if ((mp.mask >> q) & 1) as_diag.q[as_diag.k++] = static_cast<uint32_t>(q);
            as_diag.m.assign(size_t{1} << as_diag.k, I1);
            as_diag.m.back() = mp.phase;
        }
        const HostOp& op = ops[idx].kind == OP_MASKPHASE ? as_diag : ops[idx];
        const bool diag = op.kind == OP_DIAG || op.kind == OP_IP_DIAG;
        DevOp d{};
        d.kind = is_ip_host(op) ? OP_IPSUB : OP_SUB;
        d.coef = static_cast<int64_t>(plan.coefs.size());
        if (diag) {
            // Local operands index the entries within a block group (lower
            // block bit first), global ones pick a sub-table per tile.
            int lb[2], gb[4], kl = 0, kg = 0;
            for (int b = 0; b < op.k; ++b) (local_of[op.q[b]] >= 0 ? lb[kl++] : gb[kg++]) = b;
            if (kl == 2 && bit_of(op.q[lb[0]]) > bit_of(op.q[lb[1]])) std::swap(lb[0], lb[1]);
            int bits[2] = {0, 0};
            for (int i = 0; i < kl; ++i) bits[i] = bit_of(op.q[lb[i]]);
            d.k = kl == 0 ? 12 : kl == 1 ? 6 + bits[0] : 9 + (bits[0] == 0 ? (bits[1] == 1 ? 0 : 1) : 2);
            // Sign patterns with every operand in the block: Z as a generator,
            // and -1 where both bits are set (CZ).
            if (kg == 0 && kl == 1 && d.kind == OP_IPSUB && op.m[0] == I1 && op.m[1] == -I1) d.k = kZGen + bits[0];
            if (kg == 0 && kl == 2 && d.kind == OP_SUB && op.m[0] == I1 && op.m[1] == I1 && op.m[2] == I1 &&
                op.m[3] == -I1)
                d.k = kNegPair + (d.k - 9);
            for (int i = 0; i < kl; ++i) d.pos[i] = static_cast<int8_t>(local_of[op.q[lb[i]]]);
            for (int j = 0; j < kg; ++j) d.spos[j] = static_cast<int8_t>(pos[op.q[gb[j]]]);
            d.spos[7] = static_cast<int8_t>(kg);
            for (uint32_t gi = 0; gi < (1u << kg); ++gi)
                for (uint32_t li = 0; li < (1u << kl); ++li) {
                    uint32_t e = 0;
                    for (int i = 0; i < kl; ++i) e |= ((li >> i) & 1u) << lb[i];
                    for (int j = 0; j < kg; ++j) e |= ((gi >> j) & 1u) << gb[j];
                    plan.coefs.push_back(op.m[e]);
                }
        } else {
            const std::vector<cplx>& M = op.m;
            if (op.k == 1) {
                d.k = bit_of(op.q[0]);
                const cplx i1(0.0, 1.0);
                if (d.kind == OP_SUB && M[0].imag() == 0.0 && M[1].imag() == 0.0 && M[2].imag() == 0.0 &&
                    M[3].imag() == 0.0)
                    d.k += kReal2;   // real rotation (Ry): real-times-complex products
                else if (d.kind == OP_SUB && M[0].imag() == 0.0 && M[3].imag() == 0.0 && M[1].real() == 0.0 &&
                         M[2].real() == 0.0)
                    d.k += kImag2;   // Rx, Y: real diagonal, imaginary off-diagonal
                else if (d.kind == OP_IPSUB && M[0] == I0 && M[1] == I1 && M[2] == I1 && M[3] == I0)
                    d.k += kXGen;
                else if (d.kind == OP_IPSUB && M[0] == I0 && M[1] == -i1 && M[2] == i1 && M[3] == I0)
                    d.k += kYGen;
                plan.coefs.insert(plan.coefs.end(), M.begin(), M.end());
            } else {
                const int b0 = bit_of(op.q[0]), b1 = bit_of(op.q[1]);
                const int lo = std::min(b0, b1), hi = std::max(b0, b1);
                d.k = lo == 0 ? (hi == 1 ? 3 : 4) : 5;
                if (b0 < b1) {
                    plan.coefs.insert(plan.coefs.end(), M.begin(), M.end());
                } else {   // swap the two operand bits of the 4x4 matrix
                    for (int y = 0; y < 4; ++y)
                        for (int x = 0; x < 4; ++x) plan.coefs.push_back(M[sw(y) * 4 + sw(x)]);
                }
            }
            for (int b = 0; b < op.k; ++b) d.pos[b] = static_cast<int8_t>(local_of[op.q[b]]);
        }
        if (d.kind == OP_IPSUB) {
            d.lmask = static_cast<uint64_t>(op.slot);
            ++n_ip;
        }
        plan.ops.push_back(d);
        ++(diag ? plan.diag_ops : plan.dense_ops);
        plan.flops += op_flops(op) * static_cast<double>(uint64_t{1} << n);
    }
    plan.ops[hdr_index].spos[7] = static_cast<int8_t>(n_ip);
}

// Dense ops wider than the tile kernel handles (k > 5) run as passes of
// their own, straight from HBM.
constexpr int kMaxTileK = 5;

// Pass selection.  Two ops conflict when they share a qubit; an op may run
// in a pass once every earlier conflicting op has run in an earlier pass or
// earlier in the same one (ops on disjoint qubits commute).  A pass is built
// best-fit: starting from the forced low qubits (plus a seed op's qubits),
// every ready op whose qubits are already tile qubits is taken, then the
// ready op that adds the fewest new tile qubits, until none fits in T.  The
// pass applying the most ops over a few seeds (none, and each of the first
// ready dense ops) wins.  Candidates come from a window of the first
// kWindow remaining ops, which keeps the cost linear in the circuit length.
namespace {
constexpr size_t kWindow = 512;
constexpr int kSeeds = 16;

struct PassPick {
    std::vector<size_t> chosen;   // positions in `remaining`, ascending
    uint64_t Q = 0;
};

struct SegmentDag {
    const std::vector<HostOp>& ops;
    std::vector<std::vector<int>> preds;   // per segment position: earlier conflicting positions
    std::vector<char> done;
    SegmentDag(const std::vector<HostOp>& o, const std::vector<size_t>& seg) : ops(o), preds(seg.size()),
                                                                               done(seg.size(), 0) {
        int last[64];
        for (int q = 0; q < 64; ++q) last[q] = -1;
        for (size_t i = 0; i < seg.size(); ++i) {
            uint64_t qm = ops[seg[i]].qmask();
            while (qm) {
                const int q = std::countr_zero(qm);
                qm &= qm - 1;
                if (last[q] >= 0 && (preds[i].empty() || preds[i].back() != last[q])) preds[i].push_back(last[q]);
                last[q] = static_cast<int>(i);
            }
        }
    }
};

PassPick pick_pass(const SegmentDag& dag, const std::vector<size_t>& seg, const std::vector<int>& remaining,
                   uint64_t Q0, int T, int seed, const ScheduleOptions& opt) {
    const size_t W = std::min(remaining.size(), kWindow);
    std::vector<char> in_pass(seg.size(), 0);
    PassPick pk;
    pk.Q = Q0;
    double flops = 0.0;
    int n_ip = 0, n_dense = 0;
    auto ready = [&](int i) {
        for (int p : dag.preds[i])
            if (!dag.done[p] && !in_pass[p]) return false;
        return true;
    };
    auto take = [&](size_t r) {
        const HostOp& op = dag.ops[seg[remaining[r]]];
        in_pass[remaining[r]] = 1;
        pk.chosen.push_back(r);
        pk.Q |= op.needs_local() ? op.qmask() : 0;
        flops += op_flops(op);
        n_ip += is_ip_host(op);
        n_dense += op.needs_local();
    };
    auto admissible = [&](const HostOp& op) {
        if (is_ip_host(op) && n_ip >= kMaxIpPerPass) return false;
        if (opt.max_ops > 0 && n_dense >= opt.max_ops && op.needs_local()) return false;
        if (opt.max_flops_per_amp > 0 && !pk.chosen.empty() && flops + op_flops(op) > opt.max_flops_per_amp)
            return false;
        return true;
    };
    if (seed >= 0) {
        pk.Q |= dag.ops[seg[remaining[seed]]].qmask();
        take(static_cast<size_t>(seed));
    }
    for (;;) {
        bool progress = false;
        size_t best = W;
        int best_need = 65;
        for (size_t r = 0; r < W; ++r) {
            const int i = remaining[r];
            if (in_pass[i] || !ready(i)) continue;
            const HostOp& op = dag.ops[seg[i]];
            if (!admissible(op)) continue;
            const int need = op.needs_local() ? std::popcount(op.qmask() & ~pk.Q) : 0;
            if (need == 0) {
                take(r);
                progress = true;
            } else if (std::popcount(pk.Q) + need <= T && need < best_need) {
                best = r;
                best_need = need;
            }
        }
        if (progress) continue;
        if (best == W) break;
        take(best);
    }
    std::sort(pk.chosen.begin(), pk.chosen.end());
    return pk;
}
}  // namespace

// Qubit relabeling.  A tile keeps its low qubits at the lowest global bits,
// so every thread's loads and stores are contiguous runs.  Which qubits sit
// there is free: a pass may store its tile through a permutation of tile
// positions (PassDesc::sperm) that moves the qubits the next pass wants to
// the low bits, so that pass need not spend tile positions on the forced
// ones.  Layouts are transpositions of low slot i with the home bit of the
// qubit that occupies it; a relabel is possible when every bit it changes is
// a bit of the current tile, and among the possible ones (moving the first
// m qubits the next pass uses, for each m) the one whose low qubits let the
// following pass apply the most ops is taken.  The identity layout is restored by the last pass
// of the segment (or by one pass of its own), so a plan starts and ends on
// the caller's qubit order.
namespace {
struct Layout {
    int low = 0;
    int at[64];    // at[b] = qubit held by global bit b
    int pos[64];   // pos[q] = global bit holding qubit q
    explicit Layout(int l) : low(l) {
        for (int q = 0; q < 64; ++q) at[q] = pos[q] = q;
    }
    // Low slot i holds a[i] (a[i] == i or a[i] >= low, all distinct).
    static Layout of(const int* a, int low) {
        Layout L(low);
        for (int i = 0; i < low; ++i)
            if (a[i] != i) {
                L.at[i] = a[i];
                L.at[a[i]] = i;
                L.pos[a[i]] = i;
                L.pos[i] = a[i];
            }
        return L;
    }
    uint64_t low_qubits() const {
        uint64_t m = 0;
        for (int i = 0; i < low; ++i) m |= uint64_t{1} << at[i];
        return m;
    }
    bool identity() const {
        for (int i = 0; i < low; ++i)
            if (at[i] != i) return false;
        return true;
    }
    // Global bits whose qubit differs between this layout and `to`.
    uint64_t changed(const Layout& to) const {
        uint64_t m = 0;
        for (int i = 0; i < low; ++i) {
            if (at[i] != to.at[i]) m |= uint64_t{1} << i;
            for (int x : {at[i], to.at[i]})
                if (at[x] != to.at[x]) m |= uint64_t{1} << x;
        }
        return m;
    }
};

// The low-slot assignment that holds as many of `want` (qubits in the order
// the next pass uses them, first m of them) as the slots take, keeping
// qubits that already sit in a slot.  Returns false when it is not a valid
// assignment.
bool assign_low(const Layout& cur, const std::vector<int>& want, size_t m, int* a) {
    const int low = cur.low;
    std::vector<int> rest;
    for (int i = 0; i < low; ++i) a[i] = -1;
    for (size_t j = 0; j < m; ++j) {
        const int x = want[j];
        if (x < low) a[x] = x;
        else rest.push_back(x);
    }
    for (int i = 0; i < low; ++i)
        if (a[i] < 0) {
            auto it = std::find(rest.begin(), rest.end(), cur.at[i]);
            if (it != rest.end()) {
                a[i] = *it;
                rest.erase(it);
            }
        }
    for (int i = 0; i < low; ++i)
        if (a[i] < 0 && !rest.empty()) {
            a[i] = rest.front();
            rest.erase(rest.begin());
        }
    // Free slots keep a displaced occupant where they can (no move), else
    // take their own qubit back.
    for (int i = 0; i < low; ++i)
        if (a[i] < 0) {
            const int x = cur.at[i];
            a[i] = (x >= low && std::find(a, a + low, x) == a + low) ? x : i;
        }
    for (int i = 0; i < low; ++i)
        for (int j = 0; j < i; ++j)
            if (a[i] == a[j]) return false;
    return true;
}
}  // namespace

static void schedule_segment(Plan& plan, const std::vector<HostOp>& ops, const std::vector<size_t>& seg, int n,
                             int T, int low, const ScheduleOptions& opt) {
    const uint64_t all_mask = n >= 64 ? ~uint64_t{0} : ((uint64_t{1} << n) - 1);
    SegmentDag dag(ops, seg);
    std::vector<int> remaining(seg.size());
    for (size_t i = 0; i < seg.size(); ++i) remaining[i] = static_cast<int>(i);
    Layout cur(low);
    const bool relabel = opt.relabel && low > 0;

    // One pass over tile bits `tile` (global bits) applying `chosen`, whose
    // store leaves layout `next`.
    auto emit_pass = [&](uint64_t tile, const std::vector<size_t>& chosen, const Layout& next) {
        PassDesc P{};
        P.n = n;
        P.T = T;
        P.tmask = tile;
        int local_of[64];
        for (int q = 0; q < 64; ++q) local_of[q] = -1;
        int l = 0;
        for (int b = 0; b < n; ++b)
            if ((tile >> b) & 1) {
                P.qpos[l] = b;
                local_of[cur.at[b]] = l;
                if (l >= 8) P.hmask |= uint64_t{1} << b;
                ++l;
            }
        for (int i = 0; i < 16; ++i) P.sperm[i] = static_cast<int8_t>(i);
        for (int i = 0; i < l; ++i) {
            const int src = local_of[next.at[P.qpos[i]]];
            if (src < 0) fail(QHIP_RUNTIME, "schedule: relabel moves a qubit from outside the tile");
            P.sperm[i] = static_cast<int8_t>(src);
        }
        P.op_begin = static_cast<int32_t>(plan.ops.size());
        P.kmax = 0;
        uint64_t Q = 0;   // tile qubits
        for (int b = 0; b < n; ++b)
            if ((tile >> b) & 1) Q |= uint64_t{1} << cur.at[b];
        const bool blocks = opt.register_blocks && T >= 3;
        // Single-state passes merge runs of phase ops (the thread key of
        // OP_PHASE holds global bits below 55).
        std::vector<HostOp> merged;
        std::vector<size_t> merged_sel;
        const bool use_merged = !opt.two_states && n <= 55 && merge_phase_runs(ops, chosen, merged);
        if (use_merged)
            for (size_t i = 0; i < merged.size(); ++i) merged_sel.push_back(i);
        const std::vector<HostOp>& pops = use_merged ? merged : ops;
        const std::vector<size_t>& psel = use_merged ? merged_sel : chosen;
        for (const Unit& u : form_units(pops, psel, blocks, opt.two_states, Q)) {
            if (u.ops.size() >= 2) {
                emit_block(plan, pops, u, local_of, cur.pos, n, T);   // runs in every kernel variant
            } else {
                emit_op(plan, pops[u.ops[0]], local_of, cur.pos, n, T, opt);
                // kmax sizes the VALU dense path; matrix-core ops run in every variant.
                if (pops[u.ops[0]].needs_local() && plan.ops.back().kind != OP_MFMA)
                    P.kmax = std::max(P.kmax, pops[u.ops[0]].k);
            }
        }
        P.op_end = static_cast<int32_t>(plan.ops.size());
        for (int o = P.op_begin; o < P.op_end; ++o)
            if (plan.ops[o].kind == OP_PHASE || plan.ops[o].kind == OP_IP_PHASE) P.kmax |= kPhaseKernel;
        P.frag = -1;
        const bool widest = T == max_tile_qubits(opt.precision, false);
        if (!opt.two_states && (opt.stream > 0 || (opt.stream < 0 && widest))) {
            P.frag = 0;
            for (int o = P.op_begin; o < P.op_end; ++o)
                if (plan.ops[o].kind == OP_MFMA)
                    P.frag += staged_fragment_elems(plan.ops[o].k, opt.precision != QHIP_C128);
            plan.stream_passes += streaming_pass(P, opt.precision) ? 1 : 0;
        }
        plan.passes.push_back(P);
        cur = next;
    };
    // Tile bits of qubits Q plus `extra` bits, filled up to T with the
    // lowest free bits; false when Q and extra need more than T.
    auto tile_bits = [&](uint64_t Q, uint64_t extra, uint64_t& tile) {
        tile = extra;
        for (int q = 0; q < n; ++q)
            if ((Q >> q) & 1) tile |= uint64_t{1} << cur.pos[q];
        if (std::popcount(tile) > T) return false;
        for (int b = 0; b < n && std::popcount(tile) < T; ++b) tile |= uint64_t{1} << b;
        tile &= all_mask;
        return true;
    };

    // The pass applying the most ops from tile qubits Q0 over the seeds.
    auto best_pick = [&](uint64_t Q0) {
        PassPick best = pick_pass(dag, seg, remaining, Q0, T, -1, opt);
        int seeds = 0;
        const size_t W = std::min(remaining.size(), kWindow);
        for (size_t r = 0; r < W && seeds < kSeeds; ++r) {
            const int i = remaining[r];
            const HostOp& op = ops[seg[i]];
            if (!op.needs_local()) continue;
            bool ready = true;
            for (int p : dag.preds[i]) ready = ready && dag.done[p];
            if (!ready) continue;
            ++seeds;
            if (std::popcount(Q0 | op.qmask()) > T) continue;
            PassPick pk = pick_pass(dag, seg, remaining, Q0, T, static_cast<int>(r), opt);
            if (pk.chosen.size() > best.chosen.size()) best = std::move(pk);
        }
        return best;
    };

    while (!remaining.empty()) {
        PassPick best = best_pick(cur.low_qubits());
        if (best.chosen.empty()) {
            // The first remaining op (always ready) needs more than the free
            // tile qubits: it displaces forced low qubits (the tile is then
            // less coalesced, never wrong), so every pass makes progress.
            const uint64_t qm = ops[seg[remaining.front()]].qmask();
            uint64_t Q0 = qm;
            for (int i = 0; i < low && std::popcount(Q0) < T; ++i) Q0 |= uint64_t{1} << cur.at[i];
            best = pick_pass(dag, seg, remaining, Q0, T, 0, opt);
        }
        std::vector<size_t> chosen;
        chosen.reserve(best.chosen.size());
        std::vector<int> rest;
        rest.reserve(remaining.size() - best.chosen.size());
        size_t c = 0;
        for (size_t r = 0; r < remaining.size(); ++r) {
            if (c < best.chosen.size() && best.chosen[c] == r) {
                chosen.push_back(seg[remaining[r]]);
                dag.done[remaining[r]] = 1;
                ++c;
            } else {
                rest.push_back(remaining[r]);
            }
        }
        if (chosen.empty()) fail(QHIP_RUNTIME, "schedule: no op could join a pass");
        remaining.swap(rest);

        uint64_t tile = 0;
        if (!tile_bits(best.Q, 0, tile)) fail(QHIP_RUNTIME, "schedule: pass wider than the tile");
        Layout next = cur;
        if (relabel && !remaining.empty()) {
            // Look ahead: the qubits of the pass that would follow, in the
            // order its ops use them, go to the low slots as far as the
            // current tile can carry the move.
            const PassPick nx = best_pick(0);
            std::vector<int> want;
            for (size_t r : nx.chosen) {
                uint64_t qm = ops[seg[remaining[r]]].qmask();
                while (qm) {
                    const int q = std::countr_zero(qm);
                    qm &= qm - 1;
                    if (std::find(want.begin(), want.end(), q) == want.end()) want.push_back(q);
                }
            }
            // Each feasible assignment is scored by the ops the following
            // pass would apply from it; ties keep the larger move.
            size_t best_score = best_pick(cur.low_qubits()).chosen.size();
            bool moved = false;
            for (size_t m = std::min<size_t>(low, want.size()); m > 0; --m) {
                int a[64];
                if (!assign_low(cur, want, m, a)) continue;
                const Layout cand = Layout::of(a, low);
                uint64_t t2;
                if (!tile_bits(best.Q, cur.changed(cand), t2)) continue;
                const size_t score = best_pick(cand.low_qubits()).chosen.size();
                if (score > best_score || (score == best_score && !moved)) {
                    best_score = score;
                    tile = t2;
                    next = cand;
                    moved = true;
                }
            }
        } else if (relabel && !cur.identity()) {
            // Last pass of the segment: restore the identity if it fits.
            const Layout id(low);
            uint64_t t2;
            if (tile_bits(best.Q, cur.changed(id), t2)) {
                tile = t2;
                next = id;
            }
        }
        emit_pass(tile, chosen, next);
    }
    if (!cur.identity()) {
        // A pass of its own that only moves the qubits home.
        const Layout id(low);
        uint64_t tile;
        if (!tile_bits(0, cur.changed(id), tile)) fail(QHIP_RUNTIME, "schedule: relabel restore wider than the tile");
        emit_pass(tile, {}, id);
    }
}

Plan schedule(const std::vector<HostOp>& ops, int n, int precision, const ScheduleOptions& opt_in) {
    Plan plan;
    ScheduleOptions opt = opt_in;
    opt.precision = precision;
    int T = opt.tile_qubits > 0 ? opt.tile_qubits : default_tile_qubits(precision, opt.two_states);
    if (T > max_tile_qubits(precision, opt.two_states))
        fail(QHIP_INVALID, "tile_qubits " + std::to_string(T) + " exceeds the LDS limit of " +
                               std::to_string(max_tile_qubits(precision, opt.two_states)) + " for this precision" +
                               (opt.two_states ? " in two-state passes" : ""));
    if (T == max_tile_qubits(precision, false) && (opt.mfma == 0 || opt.stream == 0))
        fail(QHIP_INVALID, "tile_qubits " + std::to_string(T) +
                               " runs only on the streaming tile kernel (stream_tiles and use_mfma not 0)");
    T = std::max(T, std::min(kMaxTileK, n));
    T = std::min(T, n);
    int kdense = 0;
    for (const auto& op : ops)
        if (op.needs_local() && op.k <= kMaxTileK) kdense = std::max(kdense, op.k);
    int low = opt.low_qubits > 0 ? opt.low_qubits : 4;
    low = std::max(0, std::min(low, T - kdense));
    std::vector<size_t> seg;
    for (size_t i = 0; i < ops.size(); ++i) {
        const HostOp& op = ops[i];
        if (op.needs_local() && op.k > kMaxTileK) {
            schedule_segment(plan, ops, seg, n, T, low, opt);
            seg.clear();
            if (op.kind != OP_DENSE) fail(QHIP_RUNTIME, "inner-product op wider than 5 qubits");
            PassDesc P{};
            P.n = n;
            P.T = 0;   // wide pass
            P.kmax = op.k;
            P.frag = -1;
            for (int i = 0; i < 16; ++i) P.sperm[i] = static_cast<int8_t>(i);
            int local_of[64];
            for (int q = 0; q < 64; ++q) local_of[q] = q;   // positions are global qubits
            P.op_begin = static_cast<int32_t>(plan.ops.size());
            emit_op(plan, op, local_of, local_of, n, 0, opt);   // no tile: the wide kernel reads spos
            P.op_end = static_cast<int32_t>(plan.ops.size());
            plan.passes.push_back(P);
            continue;
        }
        seg.push_back(i);
    }
    schedule_segment(plan, ops, seg, n, T, low, opt);
    return plan;
}

}  // namespace qhip
