/*
 * qarp_oracle.h — CPU restatement of the OpenQARP statevector hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP backend:
 * only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark may load it.  The product (libqarp_hip.so and the openqarp_amd
 * package) never links, imports or calls it.
 *
 * Every function restates one reference function (cited at its definition
 * in qarp_oracle.c) over the plain-C stream types of include/qarp_hip.h.
 * Statevectors are complex128, (re, im) interleaved, LSB-indexed.
 */
#ifndef QARP_ORACLE_H
#define QARP_ORACLE_H

#include "../include/qarp_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failure (static buffer). */
const char* orc_last_error(void);

/* 2^k x 2^k row-major matrix of one concrete command over its own operands. */
int orc_gate_matrix(const qhip_stream* s, int64_t op_index, double* out, int64_t capacity);

/* The fusion pass of QarpSimulator::fuse_for_dispatch.  Result handle. */
typedef struct orc_fused orc_fused;
int orc_fuse(const qhip_stream* in, int n_qubits, int fusion_max, int fusion_min,
             orc_fused** out);
int64_t orc_fused_count(const orc_fused* f);
/* src >= 0: op re-emitted verbatim (index into the input stream); src < 0: a
 * Custom block with k operands and a 2^k x 2^k matrix. */
int orc_fused_get(const orc_fused* f, int64_t i, int* src, int* k, uint32_t* qubits,
                  double* matrix, int64_t capacity);
void orc_fused_free(orc_fused* f);

/* QarpSimulator::statevector.  init may be NULL (|0...0>); out has 2^n entries. */
int orc_statevector(const qhip_stream* s, int n, int fusion_max, int fusion_min,
                    const double* init, double* out);
/* Apply the stream to `state` in place without fusion (apply_command loop). */
int orc_apply_unfused(const qhip_stream* s, int n, double* state);

/* pauli_transition: <bra|H|ket>. */
int orc_transition(const double* bra, const double* ket, int n, const qhip_observable* obs,
                   double out[2]);
/* apply_observable: out = H in. */
int orc_apply_observable(const double* in, int n, const qhip_observable* obs, double* out);

/* std::mt19937 words and the generate_canonical<double,53> uniforms. */
void orc_mt19937(uint32_t seed, int64_t n, uint32_t* words);
void orc_uniforms(uint32_t seed, int64_t n, double* u);
/* sample_terminal draws: outcomes in draw order. */
int orc_sample(const double* sv, int n, int64_t n_shots, uint32_t seed, uint64_t* outcomes);
int orc_sample_uniforms(const double* sv, int n, const double* u, int64_t n_draws,
                        uint64_t* outcomes);
/* The libstdc++ CDF itself (2^n doubles). */
int orc_cdf(const double* sv, int n, double* cdf);

/* run_gradient / run_gradient_phi (no fusion, as the reference). */
int orc_run_gradient(const qhip_stream* s, int n, const qhip_observable* obs, int n_params,
                     const double* init, double* grad);
int orc_run_gradient_phi(const qhip_stream* s, int n, const double* phi, int n_params,
                         const double* init, double* grad);

/* apply_permutation (program.cpp:610-646): |x> -> |table[x]> on `qubits`
 * (local bit b <-> qubits[b]), out of place. */
int orc_permute(const double* in, int n, const uint32_t* qubits, int k, const uint64_t* table,
                double* out);
/* apply_controlled_powers (program.cpp:648-712), in place. */
int orc_controlled_powers(double* state, int n, const uint32_t* controls, const uint64_t* exponents,
                          int n_controls, const uint32_t* targets, int m, const double* matrix);

#ifdef __cplusplus
}
#endif
#endif
