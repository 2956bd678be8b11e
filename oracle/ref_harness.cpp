// ref_harness.cpp — C entry points over the OpenQARP reference simulator.
//
// TEST INFRASTRUCTURE ONLY.  Compiled (by oracle/Makefile) together with the
// reference's own simulator sources where they lie in the OpenQARP tree, into
// oracle/_ref/libqarp_ref.so.  Used to write golden vectors, to pin the C
// restatement (qarp_oracle.c), and as the timed CPU baseline.  Never linked
// by the product.
//
// Each function converts the plain-C stream types of include/qarp_hip.h into
// qarpx::Command / PauliObservable and calls the reference method named in
// its comment (src/qarpx/include/qarpx/simulator/qarp_simulator.h).
#include "../include/qarp_hip.h"

#include "qarpx/core/command.h"
#include "qarpx/core/errors.h"
#include "qarpx/device/noise_model.h"
#include "qarpx/simulator/fusion.h"
#include "qarpx/simulator/pauli_expectation.h"
#include "qarpx/simulator/program.h"
#include "qarpx/simulator/qarp_simulator.h"
#include "qarpx/transpiler/fusion.h"

#include <Eigen/Dense>

#include <algorithm>
#include <complex>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <unordered_map>
#include <vector>

using namespace qarpx;
using cd = std::complex<double>;

namespace {

thread_local std::string g_err;

int catch_all(const std::exception& e, int code) {
    g_err = e.what();
    return code;
}

std::string sym_name(int k) { return "p" + std::to_string(k); }

// Build a qarpx::Command for op i.  When `symbolic` and the op has deps, the
// first parameter becomes sum_k alpha_k * p_k + offset with the offset chosen
// so the value at `values` equals params[0].
Command to_command(const qhip_stream* s, int64_t i, bool symbolic, const double* values) {
    const qhip_op& op = s->ops[i];
    Command c;
    c.gate = static_cast<GateType>(op.gate);
    for (int b = 0; b < op.n_qubits; ++b) c.qubits.push_back(s->qubits[op.qubit_offset + b]);
    for (int p = 0; p < op.n_params; ++p) c.params.push_back(Param(op.params[p]));
    if (op.cbit >= 0) c.cbits.push_back(static_cast<uint32_t>(op.cbit));
    if (op.gate == QHIP_CUSTOM && op.unitary_offset >= 0) {
        const int D = 1 << op.n_qubits;
        auto m = std::make_shared<Eigen::MatrixXcd>(D, D);
        const double* u = s->unitaries + 2 * op.unitary_offset;
        for (int y = 0; y < D; ++y)
            for (int x = 0; x < D; ++x) (*m)(y, x) = cd(u[2 * (y * D + x)], u[2 * (y * D + x) + 1]);
        c.unitary = m;
    }
    if (symbolic && op.n_deps > 0 && op.n_params > 0) {
        double offset = op.params[0];
        for (int d = 0; d < op.n_deps; ++d)
            offset -= s->dep_alpha[op.dep_offset + d] * values[s->dep_index[op.dep_offset + d]];
        Param p = Param::linear(s->dep_alpha[op.dep_offset], sym_name(s->dep_index[op.dep_offset]), offset);
        for (int d = 1; d < op.n_deps; ++d)
            p = p + Param::linear(s->dep_alpha[op.dep_offset + d], sym_name(s->dep_index[op.dep_offset + d]), 0.0);
        c.params[0] = p;
    }
    return c;
}

std::vector<Command> to_commands(const qhip_stream* s, bool symbolic = false, const double* values = nullptr) {
    std::vector<Command> out;
    out.reserve(static_cast<size_t>(s->n_ops));
    for (int64_t i = 0; i < s->n_ops; ++i) out.push_back(to_command(s, i, symbolic, values));
    return out;
}

PauliObservable to_observable(const qhip_observable* o) {
    PauliObservable obs;
    for (int64_t t = 0; t < o->n_terms; ++t) {
        const auto& term = o->terms[t];
        std::vector<std::pair<uint32_t, char>> letters;
        for (int l = 0; l < term.n_letters; ++l)
            letters.emplace_back(o->qubits[term.offset + l], o->letters[term.offset + l]);
        obs.emplace_back(std::move(letters), cd(term.coeff_re, term.coeff_im));
    }
    return obs;
}

std::optional<std::vector<cd>> to_state(const double* init, int n) {
    if (!init) return std::nullopt;
    const uint64_t dim = uint64_t{1} << n;
    std::vector<cd> v(dim);
    std::memcpy(v.data(), init, sizeof(cd) * dim);
    return v;
}

QarpSimulator make_sim(int fusion_max, int fusion_min) {
    QarpSimulator sim;
    sim.set_fusion_max_qubits(static_cast<std::size_t>(fusion_max));
    sim.set_fusion_min_qubits(static_cast<std::size_t>(fusion_min));
    return sim;
}

}  // namespace

extern "C" {

const char* ref_last_error(void) { return g_err.c_str(); }

// QarpSimulator::statevector
int ref_statevector(const qhip_stream* s, int n, int fusion_max, int fusion_min,
                    const double* init, double* out) {
    try {
        auto sim = make_sim(fusion_max, fusion_min);
        auto sv = sim.statevector(to_commands(s), n, to_state(init, n));
        std::memcpy(out, sv.data(), sizeof(cd) * sv.size());
        return 0;
    } catch (const std::invalid_argument& e) { return catch_all(e, QHIP_INVALID);
    } catch (const capability_error& e) { return catch_all(e, QHIP_CAPABILITY);
    } catch (const std::exception& e) { return catch_all(e, QHIP_RUNTIME); }
}

// QarpSimulator::expectation
int ref_expectation(const qhip_stream* s, int n, int fusion_max, int fusion_min,
                    const qhip_observable* obs, const double* init, double* out) {
    try {
        auto sim = make_sim(fusion_max, fusion_min);
        *out = sim.expectation(to_commands(s), n, to_observable(obs), to_state(init, n));
        return 0;
    } catch (const std::invalid_argument& e) { return catch_all(e, QHIP_INVALID);
    } catch (const std::exception& e) { return catch_all(e, QHIP_RUNTIME); }
}

// QarpSimulator::statevector followed by pauli_transition on the result,
// reduced to a fingerprint so full-size states (16 GiB at 30 qubits) need no
// second host copy: the amplitudes at idx[0..n_idx), the FP64 sum of |a|^2 in
// index order, <Z_q> for every qubit (stats[1..n]), and <obs> (stats[n+1],
// skipped when obs is NULL).  stats[0] is the norm.
int ref_statevector_fingerprint(const qhip_stream* s, int n, int fusion_max, int fusion_min,
                                const uint64_t* idx, int64_t n_idx, double* amps,
                                const qhip_observable* obs, double* stats) {
    try {
        auto sim = make_sim(fusion_max, fusion_min);
        const auto sv = sim.statevector(to_commands(s), n);
        for (int64_t j = 0; j < n_idx; ++j) {
            amps[2 * j] = sv[idx[j]].real();
            amps[2 * j + 1] = sv[idx[j]].imag();
        }
        const uint64_t dim = uint64_t{1} << n;
        std::vector<double> z(static_cast<size_t>(n) + 1, 0.0);
        for (uint64_t i = 0; i < dim; ++i) {
            const double p = std::norm(sv[i]);
            z[0] += p;
            for (int q = 0; q < n; ++q) z[q + 1] += ((i >> q) & 1) ? -p : p;
        }
        for (int q = 0; q <= n; ++q) stats[q] = z[q];
        if (obs) stats[n + 1] = pauli_transition(sv.data(), sv.data(), n, to_observable(obs)).real();
        return 0;
    } catch (const std::invalid_argument& e) { return catch_all(e, QHIP_INVALID);
    } catch (const std::exception& e) { return catch_all(e, QHIP_RUNTIME); }
}

// pauli_transition
int ref_transition(const double* bra, const double* ket, int n, const qhip_observable* obs,
                   double out[2]) {
    try {
        const cd v = pauli_transition(reinterpret_cast<const cd*>(bra), reinterpret_cast<const cd*>(ket),
                                      n, to_observable(obs));
        out[0] = v.real();
        out[1] = v.imag();
        return 0;
    } catch (const std::invalid_argument& e) { return catch_all(e, QHIP_INVALID);
    } catch (const std::exception& e) { return catch_all(e, QHIP_RUNTIME); }
}

// The fused stream of fuse_for_dispatch: fuse_for_simulation for k >= 2 and
// n >= fusion_min, fuse_single_qubit_gates otherwise.  Writes per fused op:
// gate id, operand count, operands (8 max), params (4), and for Custom the
// matrix (k <= 5) into mats at op_index * 1024 complex entries.
int ref_fuse(const qhip_stream* s, int n, int fusion_max, int fusion_min, int64_t cap_ops,
             int64_t* n_out, int32_t* gates, int32_t* nqs, uint32_t* qubits, double* params,
             double* mats) {
    try {
        auto sim = make_sim(fusion_max, fusion_min);
        const auto cmds = to_commands(s);
        std::vector<Command> fused;
        if (fusion_max == 0) fused = cmds;
        else if (fusion_max == 1 || n < fusion_min) fused = fuse_single_qubit_gates(cmds);
        else fused = fuse_for_simulation(cmds, static_cast<std::size_t>(fusion_max),
                                         [&sim](const Command& c) { return sim.local_unitary(c); });
        *n_out = static_cast<int64_t>(fused.size());
        if (static_cast<int64_t>(fused.size()) > cap_ops) return 0;
        for (size_t i = 0; i < fused.size(); ++i) {
            const auto& c = fused[i];
            gates[i] = static_cast<int32_t>(c.gate);
            nqs[i] = static_cast<int32_t>(c.qubits.size());
            for (size_t b = 0; b < c.qubits.size() && b < 8; ++b) qubits[i * 8 + b] = c.qubits[b];
            for (size_t p = 0; p < 4; ++p)
                params[i * 4 + p] = p < c.params.size() && c.params[p].is_concrete() ? c.params[p].value() : 0.0;
            if (c.gate == GateType::Custom && c.unitary) {
                const auto& U = *c.unitary;
                const int D = static_cast<int>(U.rows());
                if (D <= 32)
                    for (int y = 0; y < D; ++y)
                        for (int x = 0; x < D; ++x) {
                            mats[(i * 1024 + y * D + x) * 2] = U(y, x).real();
                            mats[(i * 1024 + y * D + x) * 2 + 1] = U(y, x).imag();
                        }
            }
        }
        return 0;
    } catch (const std::exception& e) { return catch_all(e, QHIP_RUNTIME); }
}

// QarpSimulator::run with a recorded terminal Measure on every qubit, so the
// per-shot cbit_history (draw order) gives each shot's outcome index.
int ref_sample_state(const double* sv, int n, int n_shots, uint32_t seed, uint64_t* outcomes) {
    try {
        QarpSimulator sim;
        std::vector<Command> cmds;
        for (int q = 0; q < n; ++q) {
            Command m(GateType::Measure, static_cast<uint32_t>(q));
            m.cbits.push_back(static_cast<uint32_t>(q));
            cmds.push_back(m);
        }
        auto res = sim.run(cmds, n, n_shots, seed, to_state(sv, n));
        if (n == 0) {
            for (int i = 0; i < n_shots; ++i) outcomes[i] = 0;
            return 0;
        }
        for (int i = 0; i < n_shots; ++i) {
            uint64_t v = 0;
            for (int q = 0; q < n; ++q) if (res.cbit_history[i][q]) v |= uint64_t{1} << q;
            outcomes[i] = v;
        }
        return 0;
    } catch (const std::invalid_argument& e) { return catch_all(e, QHIP_INVALID);
    } catch (const std::exception& e) { return catch_all(e, QHIP_RUNTIME); }
}

// QarpSimulator::run counts (no recorded measures): the fast path itself.
int ref_run_counts(const qhip_stream* s, int n, int n_shots, uint32_t seed, int fusion_max,
                   int fusion_min, int64_t cap, int64_t* n_keys, uint64_t* keys, int32_t* counts) {
    try {
        auto sim = make_sim(fusion_max, fusion_min);
        auto res = sim.run(to_commands(s), n, n_shots, seed);
        *n_keys = static_cast<int64_t>(res.counts.size());
        int64_t i = 0;
        for (const auto& [k, v] : res.counts) {
            if (i >= cap) break;
            keys[i] = k;
            counts[i] = v;
            ++i;
        }
        return 0;
    } catch (const std::invalid_argument& e) { return catch_all(e, QHIP_INVALID);
    } catch (const std::exception& e) { return catch_all(e, QHIP_RUNTIME); }
}

// QarpSimulator::run on any circuit (the trajectory path when measurements are
// not terminal): counts and the per-shot classical registers (shots x n_cbits
// bytes, capacity hist_cap).
// Conditions ride beside the stream: cond_count[i] (bit, value) pairs per op,
// flattened in op order.
int ref_run_trajectory(const qhip_stream* s, const int32_t* cond_count, const uint32_t* cond_bits,
                       const uint8_t* cond_values, int n, int n_shots, uint32_t seed, int64_t cap, int64_t* n_keys,
                       uint64_t* keys, int32_t* counts, int32_t* n_cbits, uint8_t* hist, int64_t hist_cap) {
    try {
        auto sim = make_sim(3, 12);
        auto cmds = to_commands(s);
        int64_t k0 = 0;
        for (size_t i = 0; i < cmds.size(); ++i)
            for (int j = 0; j < cond_count[i]; ++j, ++k0) {
                cmds[i].condition_bits.push_back(cond_bits[k0]);
                cmds[i].condition_values.push_back(cond_values[k0] != 0);
            }
        auto res = sim.run(cmds, n, n_shots, seed);
        *n_keys = static_cast<int64_t>(res.counts.size());
        std::vector<std::pair<uint64_t, int>> kv(res.counts.begin(), res.counts.end());
        std::sort(kv.begin(), kv.end());
        for (int64_t i = 0; i < static_cast<int64_t>(kv.size()) && i < cap; ++i) {
            keys[i] = kv[i].first;
            counts[i] = kv[i].second;
        }
        *n_cbits = res.n_cbits;
        int64_t k = 0;
        for (const auto& reg : res.cbit_history)
            for (bool bit : reg)
                if (k < hist_cap) hist[k++] = bit ? 1 : 0;
        return 0;
    } catch (const std::invalid_argument& e) { return catch_all(e, QHIP_INVALID);
    } catch (const std::exception& e) { return catch_all(e, QHIP_RUNTIME); }
}

// One builder call of a noise model, as the tests describe it.
struct ref_noise_entry {
    int32_t        kind;        // 0 make_depolarizing(p0), 1 make_pauli(p0, p1, p2),
                                // 2 make_bit_flip(p0), 3 make_amplitude_damping(p0)
    int32_t        n_gates;
    double         p[3];
    const int32_t* gates;
};

// QarpSimulator(NoiseModel)::run with the model composed (operator+) from the
// reference's own builders, and `enabled`; otherwise as ref_run_trajectory.
int ref_run_noisy(const qhip_stream* s, const int32_t* cond_count, const uint32_t* cond_bits,
                  const uint8_t* cond_values, const ref_noise_entry* entries, int n_entries, int enabled, int n,
                  int n_shots, uint32_t seed, int64_t cap, int64_t* n_keys, uint64_t* keys, int32_t* counts,
                  int32_t* n_cbits, uint8_t* hist, int64_t hist_cap) {
    try {
        NoiseModel model;
        for (int e = 0; e < n_entries; ++e) {
            std::vector<GateType> gates;
            for (int g = 0; g < entries[e].n_gates; ++g) gates.push_back(static_cast<GateType>(entries[e].gates[g]));
            const double* p = entries[e].p;
            switch (entries[e].kind) {
                case 0: model += make_depolarizing(p[0], gates); break;
                case 1: model += make_pauli(p[0], p[1], p[2], gates); break;
                case 2: model += make_bit_flip(p[0], gates); break;
                default: model += make_amplitude_damping(p[0], gates);
            }
        }
        model.enabled = enabled != 0;
        QarpSimulator sim(model);
        auto cmds = to_commands(s);
        int64_t k0 = 0;
        for (size_t i = 0; i < cmds.size(); ++i)
            for (int j = 0; j < cond_count[i]; ++j, ++k0) {
                cmds[i].condition_bits.push_back(cond_bits[k0]);
                cmds[i].condition_values.push_back(cond_values[k0] != 0);
            }
        auto res = sim.run(cmds, n, n_shots, seed);
        *n_keys = static_cast<int64_t>(res.counts.size());
        std::vector<std::pair<uint64_t, int>> kv(res.counts.begin(), res.counts.end());
        std::sort(kv.begin(), kv.end());
        for (int64_t i = 0; i < static_cast<int64_t>(kv.size()) && i < cap; ++i) {
            keys[i] = kv[i].first;
            counts[i] = kv[i].second;
        }
        *n_cbits = res.n_cbits;
        int64_t k = 0;
        for (const auto& reg : res.cbit_history)
            for (bool bit : reg)
                if (k < hist_cap) hist[k++] = bit ? 1 : 0;
        return 0;
    } catch (const std::invalid_argument& e) { return catch_all(e, QHIP_INVALID);
    } catch (const std::exception& e) { return catch_all(e, QHIP_RUNTIME); }
}

// std::uniform_int_distribution<uint64_t>(lo, hi) over std::mt19937(seed), n draws.
void ref_uniform_int(uint32_t seed, uint64_t lo, uint64_t hi, int64_t n, uint64_t* out) {
    std::mt19937 g(seed);
    std::uniform_int_distribution<uint64_t> d(lo, hi);
    for (int64_t i = 0; i < n; ++i) out[i] = d(g);
}

// The toolchain's std::mt19937 words and generate_canonical<double,53> values.
void ref_mt19937(uint32_t seed, int64_t n, uint32_t* words) {
    std::mt19937 g(seed);
    for (int64_t i = 0; i < n; ++i) words[i] = static_cast<uint32_t>(g());
}
void ref_uniforms(uint32_t seed, int64_t n, double* u) {
    std::mt19937 g(seed);
    for (int64_t i = 0; i < n; ++i) u[i] = std::generate_canonical<double, 53>(g);
}
// std::discrete_distribution over given weights, draws in order.
void ref_discrete(const double* w, int64_t len, uint32_t seed, int64_t n, uint64_t* out) {
    std::mt19937 g(seed);
    std::discrete_distribution<uint64_t> d(w, w + len);
    for (int64_t i = 0; i < n; ++i) out[i] = d(g);
}

// QarpSimulator::run_gradient with symbols p<k> built from the dep pools.
int ref_run_gradient(const qhip_stream* s, int n, const qhip_observable* obs, int n_params,
                     const double* values, const double* init, double* grad) {
    try {
        QarpSimulator sim;
        std::unordered_map<std::string, double> params;
        std::vector<std::string> order;
        for (int k = 0; k < n_params; ++k) { params[sym_name(k)] = values[k]; order.push_back(sym_name(k)); }
        auto g = sim.run_gradient(to_commands(s, true, values), n, to_observable(obs), params, order,
                                  to_state(init, n));
        std::memcpy(grad, g.data(), sizeof(double) * g.size());
        return 0;
    } catch (const std::invalid_argument& e) { return catch_all(e, QHIP_INVALID);
    } catch (const capability_error& e) { return catch_all(e, QHIP_CAPABILITY);
    } catch (const std::exception& e) { return catch_all(e, QHIP_RUNTIME); }
}

// QarpSimulator::run_gradient_phi
int ref_run_gradient_phi(const qhip_stream* s, int n, const double* phi, int n_params,
                         const double* values, const double* init, double* grad) {
    try {
        QarpSimulator sim;
        std::unordered_map<std::string, double> params;
        std::vector<std::string> order;
        for (int k = 0; k < n_params; ++k) { params[sym_name(k)] = values[k]; order.push_back(sym_name(k)); }
        const uint64_t dim = uint64_t{1} << n;
        std::vector<cd> ph(dim);
        std::memcpy(ph.data(), phi, sizeof(cd) * dim);
        auto g = sim.run_gradient_phi(to_commands(s, true, values), n, ph, params, order, to_state(init, n));
        std::memcpy(grad, g.data(), sizeof(double) * g.size());
        return 0;
    } catch (const std::invalid_argument& e) { return catch_all(e, QHIP_INVALID);
    } catch (const capability_error& e) { return catch_all(e, QHIP_CAPABILITY);
    } catch (const std::exception& e) { return catch_all(e, QHIP_RUNTIME); }
}

// One kernel of a structured program, as the tests describe it.
struct ref_kernel {
    int32_t            kind;        // 0 gates, 1 permutation, 2 dense, 3 controlled powers
    int32_t            k;           // permutation / dense qubits, or targets
    int32_t            n_controls;
    int32_t            unused;
    const uint32_t*    qubits;      // permutation / dense qubits, or targets
    const uint64_t*    table;       // permutation
    const double*      matrix;      // dense / controlled powers, row-major (re, im)
    const uint32_t*    controls;
    const uint64_t*    exponents;
    const qhip_stream* gates;
};

// QarpSimulator::program_statevector over a Program built with the
// reference's own add_* methods (permutations fold as they do there).
int ref_program_statevector(const ref_kernel* ks, int nk, int n, int fusion_max, int fusion_min,
                            const double* init, double* out) {
    try {
        auto sim = make_sim(fusion_max, fusion_min);
        Program prog;
        auto matrix = [](const double* m, int k) {
            const int D = 1 << k;
            Eigen::MatrixXcd M(D, D);
            for (int y = 0; y < D; ++y)
                for (int x = 0; x < D; ++x) M(y, x) = cd(m[2 * (y * D + x)], m[2 * (y * D + x) + 1]);
            return M;
        };
        for (int i = 0; i < nk; ++i) {
            const ref_kernel& K = ks[i];
            std::vector<uint32_t> qs(K.qubits, K.qubits + (K.kind == 0 ? 0 : K.k));
            switch (K.kind) {
                case 0: prog.add_gates(to_commands(K.gates)); break;
                case 1: prog.add_permutation(qs, std::vector<uint64_t>(K.table, K.table + (uint64_t{1} << K.k)));
                        break;
                case 2: prog.add_dense(qs, matrix(K.matrix, K.k)); break;
                default:
                    prog.add_controlled_powers(std::vector<uint32_t>(K.controls, K.controls + K.n_controls),
                                               std::vector<uint64_t>(K.exponents, K.exponents + K.n_controls),
                                               qs, matrix(K.matrix, K.k));
            }
        }
        auto sv = sim.program_statevector(prog, n, to_state(init, n));
        std::memcpy(out, sv.data(), sizeof(cd) * sv.size());
        return 0;
    } catch (const std::invalid_argument& e) { return catch_all(e, QHIP_INVALID);
    } catch (const std::exception& e) { return catch_all(e, QHIP_RUNTIME); }
}

// permutation_table (program.cpp:714-724): *found = 0 for nullopt, else the
// 2^k table.
int ref_permutation_table(const qhip_stream* s, const uint32_t* qubits, int k, uint32_t max_rest,
                          uint32_t max_work_log2, int32_t* found, uint64_t* out) {
    try {
        auto t = permutation_table(to_commands(s), std::vector<uint32_t>(qubits, qubits + k), max_rest,
                                   max_work_log2);
        *found = t ? 1 : 0;
        if (t) std::copy(t->begin(), t->end(), out);
        return 0;
    } catch (const std::invalid_argument& e) { return catch_all(e, QHIP_INVALID);
    } catch (const std::exception& e) { return catch_all(e, QHIP_RUNTIME); }
}

// local_unitary_of (program.cpp:726-738), row-major (re, im).
int ref_local_unitary_of(const qhip_stream* s, const uint32_t* qubits, int k, double* out) {
    try {
        const Eigen::MatrixXcd U = local_unitary_of(to_commands(s), std::vector<uint32_t>(qubits, qubits + k));
        const Eigen::Index D = U.rows();
        for (Eigen::Index r = 0; r < D; ++r)
            for (Eigen::Index c = 0; c < D; ++c) {
                out[2 * (r * D + c)] = U(r, c).real();
                out[2 * (r * D + c) + 1] = U(r, c).imag();
            }
        return 0;
    } catch (const std::invalid_argument& e) { return catch_all(e, QHIP_INVALID);
    } catch (const std::exception& e) { return catch_all(e, QHIP_RUNTIME); }
}

// commands_digest / local_commands_digest (program.cpp:782-803).
uint64_t ref_commands_digest(const qhip_stream* s) { return commands_digest(to_commands(s)); }
uint64_t ref_local_commands_digest(const qhip_stream* s, const uint32_t* qubits, int k) {
    return local_commands_digest(to_commands(s), std::vector<uint32_t>(qubits, qubits + k));
}

}  // extern "C"
