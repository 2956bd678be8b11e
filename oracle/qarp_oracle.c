/*
 * qarp_oracle.c — CPU restatement of the OpenQARP statevector hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see qarp_oracle.h): the checker of the HIP
 * backend, never part of it.
 *
 * References are paths inside the OpenQARP source tree:
 *   qarp_simulator.cpp  = src/qarpx/src/simulator/qarp_simulator.cpp
 *   sim fusion.cpp      = src/qarpx/src/simulator/fusion.cpp
 *   tr fusion.cpp       = src/qarpx/src/transpiler/fusion.cpp
 *   dense_kernel.cpp    = src/qarpx/src/simulator/dense_kernel.cpp
 *   pauli_expectation.cpp = src/qarpx/src/simulator/pauli_expectation.cpp
 *   conventions         = docs/contracts/qarp_conventions.md
 * Sampling follows libstdc++ (GCC 11) std::mt19937, generate_canonical and
 * discrete_distribution, which sample_terminal (qarp_simulator.cpp:751-783)
 * calls; mt19937 is the published MT19937 generator (Matsumoto & Nishimura,
 * ACM TOMACS 8(1), 1998) that the C++ standard specifies.
 *
 * Built with gcc -O2 and no -march, so no fused multiply-add is formed, as in
 * the reference's x86-64 release build.
 */
#include "qarp_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char g_err[512];
static int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}
const char* orc_last_error(void) { return g_err; }

typedef struct { double re, im; } cplx;
static inline cplx cmul(cplx a, cplx b) {
    cplx r = { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
    return r;
}
static inline cplx cadd(cplx a, cplx b) { cplx r = { a.re + b.re, a.im + b.im }; return r; }
static inline cplx mk(double re, double im) { cplx r = { re, im }; return r; }

/* ── gate matrices (conventions §2-§5; csim kernels called by apply_command,
 *    qarp_simulator.cpp:356-596) ──────────────────────────────────────────── */

static const uint32_t* op_qubits(const qhip_stream* s, const qhip_op* op) {
    return s->qubits + op->qubit_offset;
}

static int is_marker(int g) {
    return g == QHIP_BARRIER || g == QHIP_MEASURE || g == QHIP_RESET ||
           g == QHIP_BRANCH_BEGIN || g == QHIP_BRANCH_ELSE || g == QHIP_BRANCH_END;
}

/* Set 2x2 row-major m = [a b; c d]. */
static void set2(cplx* m, cplx a, cplx b, cplx c, cplx d) { m[0] = a; m[1] = b; m[2] = c; m[3] = d; }

/* 1-qubit matrices (conventions §2.1-§2.6). */
static int mat1(int gate, const double* p, cplx* m) {
    const double r2 = 1.0 / sqrt(2.0);
    const double pi = 3.14159265358979323846;
    double c, s;
    switch (gate) {
        case QHIP_X:   set2(m, mk(0,0), mk(1,0), mk(1,0), mk(0,0)); return 0;
        case QHIP_Y:   set2(m, mk(0,0), mk(0,-1), mk(0,1), mk(0,0)); return 0;
        case QHIP_Z:   set2(m, mk(1,0), mk(0,0), mk(0,0), mk(-1,0)); return 0;
        case QHIP_H:   set2(m, mk(r2,0), mk(r2,0), mk(r2,0), mk(-r2,0)); return 0;
        case QHIP_S:   set2(m, mk(1,0), mk(0,0), mk(0,0), mk(0,1)); return 0;
        case QHIP_SDG: set2(m, mk(1,0), mk(0,0), mk(0,0), mk(0,-1)); return 0;
        case QHIP_T:   set2(m, mk(1,0), mk(0,0), mk(0,0), mk(cos(pi/4), sin(pi/4))); return 0;
        case QHIP_TDG: set2(m, mk(1,0), mk(0,0), mk(0,0), mk(cos(-pi/4), sin(-pi/4))); return 0;
        case QHIP_SX:  set2(m, mk(.5,.5), mk(.5,-.5), mk(.5,-.5), mk(.5,.5)); return 0;
        case QHIP_SXDG:set2(m, mk(.5,-.5), mk(.5,.5), mk(.5,.5), mk(.5,-.5)); return 0;
        case QHIP_ID:  set2(m, mk(1,0), mk(0,0), mk(0,0), mk(1,0)); return 0;
        case QHIP_RX: c = cos(p[0]/2); s = sin(p[0]/2);
            set2(m, mk(c,0), mk(0,-s), mk(0,-s), mk(c,0)); return 0;
        case QHIP_RY: c = cos(p[0]/2); s = sin(p[0]/2);
            set2(m, mk(c,0), mk(-s,0), mk(s,0), mk(c,0)); return 0;
        case QHIP_RZ: c = cos(p[0]/2); s = sin(p[0]/2);
            set2(m, mk(c,-s), mk(0,0), mk(0,0), mk(c,s)); return 0;
        case QHIP_P:
            set2(m, mk(1,0), mk(0,0), mk(0,0), mk(cos(p[0]), sin(p[0]))); return 0;
        case QHIP_U: {
            const double th = p[0], ph = p[1], la = p[2];
            c = cos(th/2); s = sin(th/2);
            set2(m, mk(c,0), mk(-cos(la)*s, -sin(la)*s),
                    mk(cos(ph)*s, sin(ph)*s), mk(cos(ph+la)*c, sin(ph+la)*c));
            return 0;
        }
        default: return -1;
    }
}

/* Target matrix of a controlled 1q gate (conventions §3.1). */
static int controlled_target(int gate, const double* p, cplx* m) {
    switch (gate) {
        case QHIP_CX:    return mat1(QHIP_X, p, m);
        case QHIP_CY:    return mat1(QHIP_Y, p, m);
        case QHIP_CZ:    return mat1(QHIP_Z, p, m);
        case QHIP_CH:    return mat1(QHIP_H, p, m);
        case QHIP_CS:    return mat1(QHIP_S, p, m);
        case QHIP_CSDG:  return mat1(QHIP_SDG, p, m);
        case QHIP_CSX:   return mat1(QHIP_SX, p, m);
        case QHIP_CSXDG: return mat1(QHIP_SXDG, p, m);
        case QHIP_CRX:   return mat1(QHIP_RX, p, m);
        case QHIP_CRY:   return mat1(QHIP_RY, p, m);
        case QHIP_CRZ:   return mat1(QHIP_RZ, p, m);
        case QHIP_CP:    return mat1(QHIP_P, p, m);
        case QHIP_CU: {
            if (mat1(QHIP_U, p, m)) return -1;
            const cplx ph = mk(cos(p[3]), sin(p[3]));
            for (int i = 0; i < 4; ++i) m[i] = cmul(m[i], ph);
            return 0;
        }
        default: return -1;
    }
}

/* Multiply a k-qubit matrix (row-major) by an l-qubit gate g acting on the
 * given local positions: out = E(g) * acc.  Used to build decompositions. */
static void embed_left(int k, cplx* acc, int l, const int* pos, const cplx* g) {
    const int D = 1 << k, gd = 1 << l;
    cplx* out = calloc((size_t)D * D, sizeof(cplx));
    for (int i = 0; i < D; ++i) {
        int li = 0, rest = i;
        for (int b = 0; b < l; ++b) { li |= ((i >> pos[b]) & 1) << b; rest &= ~(1 << pos[b]); }
        for (int lo = 0; lo < gd; ++lo) {
            const cplx c = g[lo * gd + li];
            if (c.re == 0.0 && c.im == 0.0) continue;
            int io = rest;
            for (int b = 0; b < l; ++b) if ((lo >> b) & 1) io |= 1 << pos[b];
            for (int x = 0; x < D; ++x) out[io * D + x] = cadd(out[io * D + x], cmul(c, acc[i * D + x]));
        }
    }
    memcpy(acc, out, sizeof(cplx) * D * D);
    free(out);
}

static void identity(int k, cplx* m) {
    const int D = 1 << k;
    memset(m, 0, sizeof(cplx) * D * D);
    for (int i = 0; i < D; ++i) m[i * D + i] = mk(1, 0);
}

/* 2-qubit gates built from two-qubit primitives (conventions §3). */
static void cx_on(cplx* acc, int c, int t) {
    cplx g[16]; identity(2, g);
    /* local (b0 = control, b1 = target): swap |c=1,t=0> (1) <-> |c=1,t=1> (3) */
    g[1*4+1] = mk(0,0); g[3*4+3] = mk(0,0); g[1*4+3] = mk(1,0); g[3*4+1] = mk(1,0);
    int pos[2] = { c, t };
    embed_left(2, acc, 2, pos, g);
}
static void one_on(cplx* acc, int gate, double angle, int q) {
    cplx g[4]; double p[4] = { angle, 0, 0, 0 };
    mat1(gate, p, g);
    int pos[1] = { q };
    embed_left(2, acc, 1, pos, g);
}

/* Fill m (2^k x 2^k) for a concrete named command over its own operands.
 * ECR / iSWAP / iSWAPdg follow the decompositions apply_command lowers them
 * through (qarp_simulator.cpp:404-414; transpiler/decompositions.cpp:391-431). */
static int named_matrix(int gate, int k, const double* p, cplx* m) {
    const int D = 1 << k;
    if (k == 1) return mat1(gate, p, m);
    if (gate == QHIP_MCZ) {
        identity(k, m);
        m[(D - 1) * D + (D - 1)] = mk(-1, 0);
        return 0;
    }
    if (k == 2) {
        cplx t[4];
        if (controlled_target(gate, p, t) == 0) {
            identity(2, m);
            /* control = local bit 0, target = local bit 1: indices 1 (t=0), 3 (t=1) */
            m[1*4+1] = t[0]; m[1*4+3] = t[1]; m[3*4+1] = t[2]; m[3*4+3] = t[3];
            return 0;
        }
        const double c = cos(p[0] / 2), s = sin(p[0] / 2);
        switch (gate) {
            case QHIP_SWAP:
                identity(2, m);
                m[1*4+1] = mk(0,0); m[2*4+2] = mk(0,0); m[1*4+2] = mk(1,0); m[2*4+1] = mk(1,0);
                return 0;
            case QHIP_RZZ:  /* exp(-i θ/2 Z⊗Z), conventions §3.3 */
                memset(m, 0, sizeof(cplx) * 16);
                m[0] = mk(c, -s); m[5] = mk(c, s); m[10] = mk(c, s); m[15] = mk(c, -s);
                return 0;
            case QHIP_RXX:  /* cI - i s X⊗X */
                memset(m, 0, sizeof(cplx) * 16);
                for (int i = 0; i < 4; ++i) m[i*4+i] = mk(c, 0);
                m[0*4+3] = mk(0,-s); m[3*4+0] = mk(0,-s); m[1*4+2] = mk(0,-s); m[2*4+1] = mk(0,-s);
                return 0;
            case QHIP_RYY:  /* cI - i s Y⊗Y; Y⊗Y|00> = -|11>, Y⊗Y|01> = |10> */
                memset(m, 0, sizeof(cplx) * 16);
                for (int i = 0; i < 4; ++i) m[i*4+i] = mk(c, 0);
                m[0*4+3] = mk(0,s); m[3*4+0] = mk(0,s); m[1*4+2] = mk(0,-s); m[2*4+1] = mk(0,-s);
                return 0;
            case QHIP_ISWAP:     /* S(0) S(1) H(0) CX(0,1) CX(1,0) H(1) */
                identity(2, m);
                one_on(m, QHIP_S, 0, 0); one_on(m, QHIP_S, 0, 1); one_on(m, QHIP_H, 0, 0);
                cx_on(m, 0, 1); cx_on(m, 1, 0); one_on(m, QHIP_H, 0, 1);
                return 0;
            case QHIP_ISWAPDG:   /* H(1) CX(1,0) CX(0,1) H(0) Sdg(1) Sdg(0) */
                identity(2, m);
                one_on(m, QHIP_H, 0, 1); cx_on(m, 1, 0); cx_on(m, 0, 1); one_on(m, QHIP_H, 0, 0);
                one_on(m, QHIP_SDG, 0, 1); one_on(m, QHIP_SDG, 0, 0);
                return 0;
            case QHIP_ECR: {     /* H(1) CX Rz(π/4)(1) CX X(0) CX Rz(-π/4)(1) CX H(1) */
                const double pi4 = 3.14159265358979323846 / 4.0;
                identity(2, m);
                one_on(m, QHIP_H, 0, 1); cx_on(m, 0, 1); one_on(m, QHIP_RZ, pi4, 1);
                cx_on(m, 0, 1); one_on(m, QHIP_X, 0, 0); cx_on(m, 0, 1);
                one_on(m, QHIP_RZ, -pi4, 1); cx_on(m, 0, 1); one_on(m, QHIP_H, 0, 1);
                return 0;
            }
            default: return -1;
        }
    }
    if (k == 3) {
        identity(3, m);
        if (gate == QHIP_CCX) {          /* controls bits 0,1; target bit 2 */
            m[3*8+3] = mk(0,0); m[7*8+7] = mk(0,0); m[3*8+7] = mk(1,0); m[7*8+3] = mk(1,0);
            return 0;
        }
        if (gate == QHIP_CSWAP) {        /* control bit 0; swap bits 1 and 2 */
            m[3*8+3] = mk(0,0); m[5*8+5] = mk(0,0); m[3*8+5] = mk(1,0); m[5*8+3] = mk(1,0);
            return 0;
        }
    }
    return -1;
}

static int op_matrix(const qhip_stream* s, const qhip_op* op, cplx* m) {
    const int k = op->n_qubits;
    if (op->gate == QHIP_CUSTOM) {
        if (op->unitary_offset < 0) return fail(QHIP_RUNTIME, "Custom gate must carry a unitary");
        memcpy(m, s->unitaries + 2 * op->unitary_offset, sizeof(cplx) * (size_t)(1 << k) * (1 << k));
        return 0;
    }
    if (named_matrix(op->gate, k, op->params, m))
        return fail(QHIP_RUNTIME, "unsupported gate %d on %d qubits", op->gate, k);
    return 0;
}

int orc_gate_matrix(const qhip_stream* s, int64_t i, double* out, int64_t capacity) {
    const qhip_op* op = &s->ops[i];
    if (op->n_qubits < 1 || op->n_qubits > 8) return fail(QHIP_INVALID, "bad width");
    const int64_t D = (int64_t)1 << op->n_qubits;
    if (capacity < D * D) return fail(QHIP_INVALID, "capacity");
    return op_matrix(s, op, (cplx*)out);
}

/* ── apply_dense_k (dense_kernel.cpp:23-78) ─────────────────────────────── */

static int cmp_u32(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
    return (x > y) - (x < y);
}

static void apply_dense(const uint32_t* qubits, int K, const cplx* U, double* state, uint64_t dim) {
    const int D = 1 << K;
    uint32_t sorted[64];
    memcpy(sorted, qubits, sizeof(uint32_t) * K);
    qsort(sorted, K, sizeof(uint32_t), cmp_u32);
    uint64_t* masks = malloc(sizeof(uint64_t) * D);
    for (int l = 0; l < D; ++l) {
        uint64_t m = 0;
        for (int b = 0; b < K; ++b) if ((l >> b) & 1) m |= (uint64_t)1 << qubits[b];
        masks[l] = m;
    }
    double* mr = malloc(sizeof(double) * D * D);
    double* mi = malloc(sizeof(double) * D * D);
    for (int i = 0; i < D * D; ++i) { mr[i] = U[i].re; mi[i] = U[i].im; }
    double* inr = malloc(sizeof(double) * D);
    double* ini = malloc(sizeof(double) * D);
    const uint64_t outer = dim >> K;
    for (uint64_t o = 0; o < outer; ++o) {
        uint64_t base = o;
        for (int b = 0; b < K; ++b) {
            const uint64_t p = sorted[b];
            const uint64_t low = base & (((uint64_t)1 << p) - 1);
            base = ((base >> p) << (p + 1)) | low;
        }
        for (int l = 0; l < D; ++l) {
            inr[l] = state[2 * (base | masks[l])];
            ini[l] = state[2 * (base | masks[l]) + 1];
        }
        for (int y = 0; y < D; ++y) {
            const double* rr = mr + y * D;
            const double* ii = mi + y * D;
            double sr = 0.0, si = 0.0;
            for (int x = 0; x < D; ++x) {
                sr += rr[x] * inr[x] - ii[x] * ini[x];
                si += rr[x] * ini[x] + ii[x] * inr[x];
            }
            state[2 * (base | masks[y])] = sr;
            state[2 * (base | masks[y]) + 1] = si;
        }
    }
    free(masks); free(mr); free(mi); free(inr); free(ini);
}

/* apply_command (qarp_simulator.cpp:356-596) for a concrete op. */
static int apply_op(const qhip_stream* s, const qhip_op* op, double* state, int n) {
    const uint64_t dim = (uint64_t)1 << n;
    const uint32_t* q = op_qubits(s, op);
    const int k = op->n_qubits;
    for (int b = 0; b < k; ++b)
        if (q[b] >= (uint32_t)n) return fail(QHIP_INVALID, "qubit %u out of range for %d qubits", q[b], n);
    switch (op->gate) {
        case QHIP_BARRIER: case QHIP_MEASURE: case QHIP_RESET: case QHIP_ID:
            return 0;
        case QHIP_GPHASE: {   /* :551-556 */
            const double c = cos(op->params[0]), sn = sin(op->params[0]);
            for (uint64_t i = 0; i < dim; ++i) {
                const double re = state[2*i], im = state[2*i+1];
                state[2*i] = re * c - im * sn;
                state[2*i+1] = re * sn + im * c;
            }
            return 0;
        }
        case QHIP_MCZ: {      /* phase -1 where every operand bit is 1 */
            uint64_t mask = 0;
            for (int b = 0; b < k; ++b) mask |= (uint64_t)1 << q[b];
            for (uint64_t i = 0; i < dim; ++i)
                if ((i & mask) == mask) { state[2*i] = -state[2*i]; state[2*i+1] = -state[2*i+1]; }
            return 0;
        }
        default: break;
    }
    if (k < 1 || k > 8) return fail(QHIP_RUNTIME, "gate %d on %d qubits", op->gate, k);
    const int D = 1 << k;
    cplx* m = malloc(sizeof(cplx) * D * D);
    int rc = op_matrix(s, op, m);
    if (rc == 0) apply_dense(q, k, m, state, dim);
    free(m);
    return rc;
}

/* ── fuse_single_qubit_gates (tr fusion.cpp:196-301) ───────────────────── */

typedef struct orc_fop {
    int       src;        /* input op index when re-emitted verbatim, else -1 */
    int       k;
    uint32_t  q[8];
    cplx*     m;          /* Custom matrix (owned) */
} orc_fop;

struct orc_fused {
    orc_fop* ops;
    int64_t  n, cap;
};

static void push_fop(orc_fused* f, orc_fop op) {
    if (f->n == f->cap) {
        f->cap = f->cap ? 2 * f->cap : 64;
        f->ops = realloc(f->ops, sizeof(orc_fop) * f->cap);
    }
    f->ops[f->n++] = op;
}
static void push_src(orc_fused* f, int src) {
    orc_fop op; memset(&op, 0, sizeof op); op.src = src; push_fop(f, op);
}

static int is_fusible_1q(const qhip_op* op) {
    if (op->n_qubits != 1) return 0;
    switch (op->gate) {
        case QHIP_X: case QHIP_Y: case QHIP_Z: case QHIP_H:
        case QHIP_S: case QHIP_SDG: case QHIP_T: case QHIP_TDG:
        case QHIP_RX: case QHIP_RY: case QHIP_RZ: case QHIP_P: case QHIP_U:
            return 1;   /* parameters are concrete in a stream handed to the simulator */
        default: return 0;
    }
}

typedef struct { cplx m[4]; int n_gates; int first; } acc1q;

static void mat2_lmul(cplx* A, const cplx* B) {    /* A = B * A */
    cplx c00 = cadd(cmul(B[0], A[0]), cmul(B[1], A[2]));
    cplx c01 = cadd(cmul(B[0], A[1]), cmul(B[1], A[3]));
    cplx c10 = cadd(cmul(B[2], A[0]), cmul(B[3], A[2]));
    cplx c11 = cadd(cmul(B[2], A[1]), cmul(B[3], A[3]));
    A[0] = c00; A[1] = c01; A[2] = c10; A[3] = c11;
}

static double cabs_(cplx a) { return hypot(a.re, a.im); }

static void flush1q(acc1q* a, uint32_t q, orc_fused* out) {
    if (a->n_gates == 0) return;
    if (a->n_gates == 1) {
        push_src(out, a->first);
    } else {
        const double tol = 1e-12;
        const int ident = cabs_(mk(a->m[0].re - 1.0, a->m[0].im)) < tol && cabs_(a->m[1]) < tol &&
                          cabs_(a->m[2]) < tol && cabs_(mk(a->m[3].re - 1.0, a->m[3].im)) < tol;
        if (!ident) {
            orc_fop op; memset(&op, 0, sizeof op);
            op.src = -1; op.k = 1; op.q[0] = q;
            op.m = malloc(sizeof(cplx) * 4);
            memcpy(op.m, a->m, sizeof(cplx) * 4);
            push_fop(out, op);
        }
    }
    identity(1, a->m);
    a->n_gates = 0;
}

static void fuse_1q(const qhip_stream* s, orc_fused* out) {
    uint32_t max_q = 0;
    for (int64_t i = 0; i < s->n_ops; ++i)
        for (int b = 0; b < s->ops[i].n_qubits; ++b)
            if (op_qubits(s, &s->ops[i])[b] > max_q) max_q = op_qubits(s, &s->ops[i])[b];
    acc1q* accs = calloc(max_q + 1, sizeof(acc1q));
    for (uint32_t q = 0; q <= max_q; ++q) identity(1, accs[q].m);
    for (int64_t i = 0; i < s->n_ops; ++i) {
        const qhip_op* op = &s->ops[i];
        const uint32_t* q = op_qubits(s, op);
        if (is_fusible_1q(op)) {
            acc1q* a = &accs[q[0]];
            if (a->n_gates == 0) a->first = (int)i;
            cplx g[4];
            mat1(op->gate, op->params, g);
            mat2_lmul(a->m, g);
            a->n_gates++;
            continue;
        }
        if (op->gate == QHIP_BARRIER) {
            if (op->n_qubits == 0) { for (uint32_t w = 0; w <= max_q; ++w) flush1q(&accs[w], w, out); }
            else for (int b = 0; b < op->n_qubits; ++b) flush1q(&accs[q[b]], q[b], out);
            push_src(out, (int)i);
            continue;
        }
        if (op->n_qubits == 1 && op->gate != QHIP_MEASURE && op->gate != QHIP_RESET) {
            flush1q(&accs[q[0]], q[0], out);
            push_src(out, (int)i);
            continue;
        }
        if (op->n_qubits >= 2) {
            for (int b = 0; b < op->n_qubits; ++b) flush1q(&accs[q[b]], q[b], out);
            push_src(out, (int)i);
            continue;
        }
        if ((op->gate == QHIP_MEASURE || op->gate == QHIP_RESET) && op->n_qubits > 0)
            flush1q(&accs[q[0]], q[0], out);
        if (op->gate == QHIP_BRANCH_BEGIN || op->gate == QHIP_BRANCH_ELSE || op->gate == QHIP_BRANCH_END)
            for (uint32_t w = 0; w <= max_q; ++w) flush1q(&accs[w], w, out);
        push_src(out, (int)i);
    }
    for (uint32_t w = 0; w <= max_q; ++w) flush1q(&accs[w], w, out);
    free(accs);
}

/* ── fuse_for_simulation (sim fusion.cpp:120-268) ──────────────────────── */

typedef struct block {
    int64_t  id;
    int      nq;           /* local bit b <-> qubits[b] */
    uint32_t qubits[8];
    int      nw;
    uint32_t wires[72];
    cplx*    u;            /* NULL until a second gate folds in */
    int      n_gates;
    int      first;
} block;

static int has_u32(const uint32_t* v, int n, uint32_t x) {
    for (int i = 0; i < n; ++i) if (v[i] == x) return 1;
    return 0;
}

static int fusible(const qhip_op* op, int k) {
    if (op->n_qubits == 0 || op->n_qubits > k) return 0;
    if (is_marker(op->gate) || op->gate == QHIP_GPHASE) return 0;
    if (op->gate == QHIP_CUSTOM && op->unitary_offset < 0) return 0;
    return 1;   /* concrete stream */
}

/* can_join (sim fusion.cpp:57-65) */
static int can_join(const block* b, const uint32_t* gq, int gn, int kmax) {
    int shared = 0, fresh = 0;
    for (int i = 0; i < gn; ++i) (has_u32(b->qubits, b->nq, gq[i]) ? shared++ : fresh++);
    if (fresh == 0) return 1;
    const int width = b->nq + fresh;
    if (width > kmax) return 0;
    return shared > 0 || width <= 2;
}

/* extend (sim fusion.cpp:69-77): U <- I ⊗ U, new qubits as high local bits */
static void extend(block* b, const uint32_t* fresh, int nf) {
    const int old = 1 << b->nq, copies = 1 << nf, D = old * copies;
    cplx* out = calloc((size_t)D * D, sizeof(cplx));
    for (int h = 0; h < copies; ++h)
        for (int y = 0; y < old; ++y)
            for (int x = 0; x < old; ++x)
                out[(h * old + y) * D + h * old + x] = b->u[y * old + x];
    free(b->u);
    b->u = out;
    for (int i = 0; i < nf; ++i) b->qubits[b->nq++] = fresh[i];
}

/* left_apply (sim fusion.cpp:81-105): U <- E(g) U, row form */
static void left_apply(cplx** Up, int nq, const cplx* g, const int* pos, int m) {
    const int dim = 1 << nq, gdim = 1 << m;
    cplx* U = *Up;
    cplx* out = calloc((size_t)dim * dim, sizeof(cplx));
    for (int i = 0; i < dim; ++i) {
        int li = 0, rest = i;
        for (int b = 0; b < m; ++b) { li |= ((i >> pos[b]) & 1) << b; rest &= ~(1 << pos[b]); }
        for (int lo = 0; lo < gdim; ++lo) {
            const cplx c = g[lo * gdim + li];
            if (c.re == 0.0 && c.im == 0.0) continue;
            int io = rest;
            for (int b = 0; b < m; ++b) if ((lo >> b) & 1) io |= 1 << pos[b];
            for (int x = 0; x < dim; ++x) out[io * dim + x] = cadd(out[io * dim + x], cmul(c, U[i * dim + x]));
        }
    }
    free(U);
    *Up = out;
}

static void emit_block(const qhip_stream* s, block* b, orc_fused* out) {
    (void)s;
    if (b->n_gates == 1) { push_src(out, b->first); return; }
    if (b->nq == 1) {
        double n2 = 0;
        const cplx d[4] = { mk(b->u[0].re - 1, b->u[0].im), b->u[1], b->u[2], mk(b->u[3].re - 1, b->u[3].im) };
        for (int i = 0; i < 4; ++i) n2 += d[i].re * d[i].re + d[i].im * d[i].im;
        if (sqrt(n2) < 1e-12) { free(b->u); b->u = NULL; return; }
    }
    orc_fop op; memset(&op, 0, sizeof op);
    op.src = -1; op.k = b->nq;
    memcpy(op.q, b->qubits, sizeof(uint32_t) * b->nq);
    op.m = b->u; b->u = NULL;
    push_fop(out, op);
}

static int fuse_blocks(const qhip_stream* s, int kmax, orc_fused* out) {
    uint32_t max_q = 0, max_c = 0; int any_c = 0;
    for (int64_t i = 0; i < s->n_ops; ++i) {
        const qhip_op* op = &s->ops[i];
        for (int b = 0; b < op->n_qubits; ++b) if (op_qubits(s, op)[b] > max_q) max_q = op_qubits(s, op)[b];
        if (op->cbit >= 0) { if ((uint32_t)op->cbit > max_c) max_c = (uint32_t)op->cbit; any_c = 1; }
    }
    const uint32_t nqw = max_q + 1;
    const size_t n_wires = nqw + (any_c ? max_c + 1 : 0);
    int64_t* front = calloc(n_wires, sizeof(int64_t));
    block* open = NULL; int n_open = 0, cap_open = 0;
    int64_t next_id = 1;

    for (int64_t i = 0; i < s->n_ops; ++i) {
        const qhip_op* op = &s->ops[i];
        const uint32_t* q = op_qubits(s, op);
        if (op->gate == QHIP_GPHASE) { push_src(out, (int)i); continue; }
        if (op->gate == QHIP_BRANCH_BEGIN || op->gate == QHIP_BRANCH_ELSE || op->gate == QHIP_BRANCH_END ||
            (op->gate == QHIP_BARRIER && op->n_qubits == 0)) {
            for (int j = 0; j < n_open; ++j) emit_block(s, &open[j], out);
            n_open = 0;
            push_src(out, (int)i);
            continue;
        }
        uint32_t ws[72]; int nws = 0;
        for (int b = 0; b < op->n_qubits; ++b) ws[nws++] = q[b];
        if (op->cbit >= 0) ws[nws++] = nqw + (uint32_t)op->cbit;

        if (fusible(op, kmax)) {
            block* target = NULL;
            for (int j = n_open - 1; j >= 0; --j) {
                if (!can_join(&open[j], q, op->n_qubits, kmax)) continue;
                int ok = 1;
                for (int w = 0; w < nws; ++w) if (front[ws[w]] > open[j].id) { ok = 0; break; }
                if (ok) { target = &open[j]; break; }
            }
            if (!target) {
                if (n_open == cap_open) { cap_open = cap_open ? 2 * cap_open : 16; open = realloc(open, sizeof(block) * cap_open); }
                block* b = &open[n_open++];
                memset(b, 0, sizeof *b);
                b->id = next_id++;
                b->nq = op->n_qubits;
                memcpy(b->qubits, q, sizeof(uint32_t) * op->n_qubits);
                b->nw = nws; memcpy(b->wires, ws, sizeof(uint32_t) * nws);
                b->n_gates = 1; b->first = (int)i;
                for (int w = 0; w < nws; ++w) front[ws[w]] = b->id;
                continue;
            }
            block* b = target;
            if (b->n_gates == 1) {
                const int D = 1 << b->nq;
                b->u = malloc(sizeof(cplx) * D * D);
                int rc = op_matrix(s, &s->ops[b->first], b->u);
                if (rc) return rc;
            }
            uint32_t fresh[8]; int nf = 0;
            for (int bq = 0; bq < op->n_qubits; ++bq) if (!has_u32(b->qubits, b->nq, q[bq])) fresh[nf++] = q[bq];
            if (nf) extend(b, fresh, nf);
            int pos[8];
            for (int bq = 0; bq < op->n_qubits; ++bq)
                for (int l = 0; l < b->nq; ++l) if (b->qubits[l] == q[bq]) pos[bq] = l;
            const int gd = 1 << op->n_qubits;
            cplx* g = malloc(sizeof(cplx) * gd * gd);
            int rc = op_matrix(s, op, g);
            if (rc) { free(g); return rc; }
            left_apply(&b->u, b->nq, g, pos, op->n_qubits);
            free(g);
            b->n_gates++;
            for (int w = 0; w < nws; ++w) {
                if (!has_u32(b->wires, b->nw, ws[w])) b->wires[b->nw++] = ws[w];
                front[ws[w]] = b->id;
            }
            continue;
        }
        /* flush_wires (sim fusion.cpp:176-192) */
        if (n_open) {
            int* mark = calloc(n_open, sizeof(int));
            for (int j = 0; j < n_open; ++j)
                for (int w = 0; w < nws; ++w) if (has_u32(open[j].wires, open[j].nw, ws[w])) { mark[j] = 1; break; }
            for (int j = n_open - 1; j >= 0; --j)
                if (mark[j])
                    for (int l = 0; l < j; ++l) {
                        if (mark[l]) continue;
                        for (int w = 0; w < open[l].nw; ++w)
                            if (has_u32(open[j].wires, open[j].nw, open[l].wires[w])) { mark[l] = 1; break; }
                    }
            int keep = 0;
            for (int j = 0; j < n_open; ++j) {
                if (mark[j]) emit_block(s, &open[j], out);
                else open[keep++] = open[j];
            }
            n_open = keep;
            free(mark);
        }
        push_src(out, (int)i);
        const int64_t id = next_id++;
        for (int w = 0; w < nws; ++w) front[ws[w]] = id;
    }
    for (int j = 0; j < n_open; ++j) emit_block(s, &open[j], out);
    free(open);
    free(front);
    return 0;
}

int orc_fuse(const qhip_stream* in, int n_qubits, int fusion_max, int fusion_min, orc_fused** outp) {
    orc_fused* out = calloc(1, sizeof(orc_fused));
    int rc = 0;
    /* fuse_for_dispatch (qarp_simulator.cpp:120-129) */
    if (fusion_max == 0 || in->n_ops == 0) {
        for (int64_t i = 0; i < in->n_ops; ++i) push_src(out, (int)i);
    } else if (fusion_max == 1 || (n_qubits < 0 ? 0 : n_qubits) < fusion_min) {
        fuse_1q(in, out);
    } else {
        rc = fuse_blocks(in, fusion_max, out);
    }
    *outp = out;
    return rc;
}

int64_t orc_fused_count(const orc_fused* f) { return f->n; }

int orc_fused_get(const orc_fused* f, int64_t i, int* src, int* k, uint32_t* qubits,
                  double* matrix, int64_t capacity) {
    const orc_fop* op = &f->ops[i];
    *src = op->src;
    *k = op->k;
    if (op->src < 0) {
        memcpy(qubits, op->q, sizeof(uint32_t) * op->k);
        const int64_t D = (int64_t)1 << op->k;
        if (matrix && capacity >= D * D) memcpy(matrix, op->m, sizeof(cplx) * D * D);
    }
    return 0;
}

void orc_fused_free(orc_fused* f) {
    if (!f) return;
    for (int64_t i = 0; i < f->n; ++i) free(f->ops[i].m);
    free(f->ops);
    free(f);
}

static int apply_fused(const qhip_stream* s, const orc_fused* f, double* state, int n) {
    for (int64_t i = 0; i < f->n; ++i) {
        const orc_fop* op = &f->ops[i];
        int rc;
        if (op->src >= 0) rc = apply_op(s, &s->ops[op->src], state, n);
        else { apply_dense(op->q, op->k, op->m, state, (uint64_t)1 << n); rc = 0; }
        if (rc) return rc;
    }
    return 0;
}

static int validate_stream(const qhip_stream* s, int n) {
    if (n < 0 || n > 40) return fail(QHIP_INVALID, "n_qubits %d outside the oracle's range", n);
    for (int64_t i = 0; i < s->n_ops; ++i) {
        const qhip_op* op = &s->ops[i];
        if (op->gate == QHIP_RESET || (op->gate == QHIP_MEASURE && op->cbit >= 0) ||
            (op->flags & QHIP_OP_CONDITIONAL) || op->gate == QHIP_BRANCH_BEGIN)
            return fail(QHIP_RUNTIME, "QarpSimulator::statevector: circuit contains mid-circuit "
                        "measurement, Reset, or classical condition");
        for (int b = 0; b < op->n_qubits; ++b)
            if (op_qubits(s, op)[b] >= (uint32_t)n)
                return fail(QHIP_INVALID, "qubit %u out of range for %d qubits", op_qubits(s, op)[b], n);
    }
    return 0;
}

/* QarpSimulator::statevector (qarp_simulator.cpp:901-943) */
int orc_statevector(const qhip_stream* s, int n, int fusion_max, int fusion_min,
                    const double* init, double* out) {
    int rc = validate_stream(s, n);
    if (rc) return rc;
    const uint64_t dim = (uint64_t)1 << n;
    if (init) memcpy(out, init, sizeof(double) * 2 * dim);
    else { memset(out, 0, sizeof(double) * 2 * dim); out[0] = 1.0; }
    orc_fused* f = NULL;
    rc = orc_fuse(s, n, fusion_max, fusion_min, &f);
    if (rc == 0) rc = apply_fused(s, f, out, n);
    orc_fused_free(f);
    return rc;
}

int orc_apply_unfused(const qhip_stream* s, int n, double* state) {
    for (int64_t i = 0; i < s->n_ops; ++i) {
        int rc = apply_op(s, &s->ops[i], state, n);
        if (rc) return rc;
    }
    return 0;
}

/* ── pauli_transition (pauli_expectation.cpp:34-167) ───────────────────── */

typedef struct {
    uint64_t x;
    int      nz, cap;
    uint64_t* z;
    double*  cr;
    double*  ci;
} group;

static inline double parity_sign(uint64_t v) { return (__builtin_popcountll(v) & 1) ? -1.0 : 1.0; }

/* build_groups (:34-71): terms keyed by X/Y flip mask in first-appearance
 * order, i^{n_Y} folded into the coefficient. */
static int build_groups(const qhip_observable* obs, int n, group** gout, int* ngout) {
    group* g = NULL; int ng = 0, cap = 0;
    for (int64_t t = 0; t < obs->n_terms; ++t) {
        const qhip_pauli_term* term = &obs->terms[t];
        uint64_t x = 0, z = 0, seen = 0; int ny = 0;
        for (int l = 0; l < term->n_letters; ++l) {
            const uint32_t q = obs->qubits[term->offset + l];
            const char c = obs->letters[term->offset + l];
            if (q >= (uint32_t)n) return fail(QHIP_INVALID, "pauli_transition: observable qubit %u out of range for %d qubits", q, n);
            const uint64_t bit = (uint64_t)1 << q;
            if (seen & bit) return fail(QHIP_INVALID, "pauli_transition: qubit %u repeated within one Pauli string", q);
            seen |= bit;
            switch (c) {
                case 'X': x |= bit; break;
                case 'Z': z |= bit; break;
                case 'Y': x |= bit; z |= bit; ++ny; break;
                case 'I': break;
                default: return fail(QHIP_INVALID, "pauli_transition: invalid Pauli '%c'", c);
            }
        }
        static const double ipr[4] = { 1, 0, -1, 0 }, ipi[4] = { 0, 1, 0, -1 };
        const cplx c = cmul(mk(term->coeff_re, term->coeff_im), mk(ipr[ny & 3], ipi[ny & 3]));
        int gi = -1;
        for (int j = 0; j < ng; ++j) if (g[j].x == x) { gi = j; break; }
        if (gi < 0) {
            if (ng == cap) { cap = cap ? 2 * cap : 16; g = realloc(g, sizeof(group) * cap); }
            gi = ng++;
            memset(&g[gi], 0, sizeof(group));
            g[gi].x = x;
        }
        group* G = &g[gi];
        if (G->nz == G->cap) {
            G->cap = G->cap ? 2 * G->cap : 4;
            G->z = realloc(G->z, sizeof(uint64_t) * G->cap);
            G->cr = realloc(G->cr, sizeof(double) * G->cap);
            G->ci = realloc(G->ci, sizeof(double) * G->cap);
        }
        G->z[G->nz] = z; G->cr[G->nz] = c.re; G->ci[G->nz] = c.im; G->nz++;
    }
    *gout = g; *ngout = ng;
    return 0;
}

static void free_groups(group* g, int ng) {
    for (int j = 0; j < ng; ++j) { free(g[j].z); free(g[j].cr); free(g[j].ci); }
    free(g);
}

int orc_transition(const double* bra, const double* ket, int n, const qhip_observable* obs, double out[2]) {
    if (n < 0 || n > 62) return fail(QHIP_INVALID, "pauli_transition: n_qubits must be in [0, 62], got %d", n);
    group* g = NULL; int ng = 0;
    int rc = build_groups(obs, n, &g, &ng);
    if (rc) return rc;
    out[0] = out[1] = 0.0;
    if (ng == 0) { free(g); return 0; }
    const uint64_t dim = (uint64_t)1 << n;
    const cplx* B = (const cplx*)bra;
    const cplx* K = (const cplx*)ket;
    if (dim < 64) {   /* transition_small (:89-106) */
        cplx total = mk(0, 0);
        for (uint64_t i = 0; i < dim; ++i) {
            cplx acc = mk(0, 0);
            for (int j = 0; j < ng; ++j) {
                double wr = 0, wi = 0;
                for (int t = 0; t < g[j].nz; ++t) {
                    const double sg = parity_sign(i & g[j].z[t]);
                    wr += sg * g[j].cr[t]; wi += sg * g[j].ci[t];
                }
                const cplx b = B[i ^ g[j].x];
                acc = cadd(acc, cmul(mk(b.re, -b.im), mk(wr, wi)));
            }
            total = cadd(total, cmul(acc, K[i]));
        }
        out[0] = total.re; out[1] = total.im;
        free_groups(g, ng);
        return 0;
    }
    /* transition_blocked (:110-155), one thread */
    double re = 0.0, im = 0.0;
    for (uint64_t base = 0; base < dim; base += 64) {
        double acc_r[64] = {0}, acc_i[64] = {0};
        for (int j = 0; j < ng; ++j) {
            double wr[64] = {0}, wi[64] = {0};
            for (int t = 0; t < g[j].nz; ++t) {
                const double sg = parity_sign(base & g[j].z[t]);
                const double cr = sg * g[j].cr[t], ci = sg * g[j].ci[t];
                const uint64_t zl = g[j].z[t] & 63;
                for (int lo = 0; lo < 64; ++lo) {
                    const double row = parity_sign(zl & (uint64_t)lo);
                    wr[lo] += cr * row;
                    wi[lo] += ci * row;
                }
            }
            const uint64_t bra_base = base ^ (g[j].x & ~(uint64_t)63);
            const uint64_t xl = g[j].x & 63;
            for (int lo = 0; lo < 64; ++lo) {
                const cplx b = B[bra_base | ((uint64_t)lo ^ xl)];
                const double pr = b.re, pi = -b.im;
                acc_r[lo] += pr * wr[lo] - pi * wi[lo];
                acc_i[lo] += pr * wi[lo] + pi * wr[lo];
            }
        }
        for (int lo = 0; lo < 64; ++lo) {
            const cplx v = cmul(mk(acc_r[lo], acc_i[lo]), K[base + lo]);
            re += v.re; im += v.im;
        }
    }
    out[0] = re; out[1] = im;
    free_groups(g, ng);
    return 0;
}

/* apply_observable (qarp_simulator.cpp:1459-1479): out = sum_t c_t P_t in */
int orc_apply_observable(const double* in, int n, const qhip_observable* obs, double* out) {
    const uint64_t dim = (uint64_t)1 << n;
    memset(out, 0, sizeof(double) * 2 * dim);
    const cplx* I = (const cplx*)in;
    cplx* O = (cplx*)out;
    for (int64_t t = 0; t < obs->n_terms; ++t) {
        const qhip_pauli_term* term = &obs->terms[t];
        uint64_t x = 0, z = 0; int ny = 0;
        for (int l = 0; l < term->n_letters; ++l) {
            const uint32_t q = obs->qubits[term->offset + l];
            const char c = obs->letters[term->offset + l];
            if (q >= (uint32_t)n) return fail(QHIP_INVALID, "run_gradient: observable qubit %u out of range", q);
            const uint64_t bit = (uint64_t)1 << q;
            if (c == 'X') x |= bit;
            else if (c == 'Z') z |= bit;
            else if (c == 'Y') { x |= bit; z |= bit; ++ny; }
            else if (c != 'I') return fail(QHIP_RUNTIME, "run_gradient: invalid Pauli '%c'", c);
        }
        static const double ipr[4] = { 1, 0, -1, 0 }, ipi[4] = { 0, 1, 0, -1 };
        const cplx c = cmul(mk(term->coeff_re, term->coeff_im), mk(ipr[ny & 3], ipi[ny & 3]));
        /* (P psi)[j] = i^{nY} (-1)^{popcount((j^x) & z)} psi[j ^ x] */
        for (uint64_t j = 0; j < dim; ++j) {
            const uint64_t src = j ^ x;
            const double sg = parity_sign(src & z);
            const cplx v = cmul(c, mk(sg * I[src].re, sg * I[src].im));
            O[j] = cadd(O[j], v);
        }
    }
    return 0;
}

/* ── sampling: std::mt19937, generate_canonical<double,53>,
 *    discrete_distribution (libstdc++ bits/random.tcc) ────────────────── */

typedef struct { uint32_t mt[624]; int idx; } mt19937;

static void mt_seed(mt19937* g, uint32_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}

static uint32_t mt_next(mt19937* g) {
    if (g->idx >= 624) {
        for (int i = 0; i < 624; ++i) {
            const uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
            g->mt[i] = g->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        g->idx = 0;
    }
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

void orc_mt19937(uint32_t seed, int64_t n, uint32_t* words) {
    mt19937 g; mt_seed(&g, seed);
    for (int64_t i = 0; i < n; ++i) words[i] = mt_next(&g);
}

/* generate_canonical<double, 53>: two 32-bit words, sum = w0 + w1 2^32 in
 * double, divided by 2^64, clamped below 1. */
static double canonical(mt19937* g) {
    const double r = 4294967296.0;
    double sum = 0.0, tmp = 1.0;
    sum += (double)mt_next(g) * tmp; tmp *= r;
    sum += (double)mt_next(g) * tmp; tmp *= r;
    double ret = sum / tmp;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret;
}

void orc_uniforms(uint32_t seed, int64_t n, double* u) {
    mt19937 g; mt_seed(&g, seed);
    for (int64_t i = 0; i < n; ++i) u[i] = canonical(&g);
}

/* discrete_distribution::param_type::_M_initialize: sum = accumulate(p, 0.0),
 * p_i / sum, partial_sum, last entry forced to 1.0; sizes < 2 clear. */
int orc_cdf(const double* sv, int n, double* cdf) {
    const uint64_t dim = (uint64_t)1 << n;
    double sum = 0.0;
    for (uint64_t i = 0; i < dim; ++i) {
        const double p = sv[2*i] * sv[2*i] + sv[2*i+1] * sv[2*i+1];   /* std::norm */
        cdf[i] = p;
        sum = sum + p;
    }
    double run = 0.0;
    for (uint64_t i = 0; i < dim; ++i) {
        const double q = cdf[i] / sum;
        run = (i == 0) ? q : run + q;
        cdf[i] = run;
    }
    cdf[dim - 1] = 1.0;
    return 0;
}

static uint64_t lower_bound(const double* cdf, uint64_t len, double u) {
    uint64_t lo = 0, count = len;
    while (count > 0) {
        const uint64_t step = count / 2, mid = lo + step;
        if (cdf[mid] < u) { lo = mid + 1; count -= step + 1; }
        else count = step;
    }
    return lo;
}

int orc_sample_uniforms(const double* sv, int n, const double* u, int64_t n_draws, uint64_t* outcomes) {
    const uint64_t dim = (uint64_t)1 << n;
    if (dim < 2) { for (int64_t i = 0; i < n_draws; ++i) outcomes[i] = 0; return 0; }
    double* cdf = malloc(sizeof(double) * dim);
    if (!cdf) return fail(QHIP_NOMEM, "cdf");
    orc_cdf(sv, n, cdf);
    for (int64_t i = 0; i < n_draws; ++i) outcomes[i] = lower_bound(cdf, dim, u[i]);
    free(cdf);
    return 0;
}

/* sample_terminal (qarp_simulator.cpp:751-783) */
int orc_sample(const double* sv, int n, int64_t n_shots, uint32_t seed, uint64_t* outcomes) {
    const uint64_t dim = (uint64_t)1 << n;
    if (dim < 2) { for (int64_t i = 0; i < n_shots; ++i) outcomes[i] = 0; return 0; }
    double* cdf = malloc(sizeof(double) * dim);
    if (!cdf) return fail(QHIP_NOMEM, "cdf");
    orc_cdf(sv, n, cdf);
    mt19937 g; mt_seed(&g, seed);
    for (int64_t i = 0; i < n_shots; ++i) outcomes[i] = lower_bound(cdf, dim, canonical(&g));
    free(cdf);
    return 0;
}

/* ── adjoint gradient (qarp_simulator.cpp:1352-1605) ───────────────────── */

/* apply_generator (:1352-1418): returns 1 with prefactor c, or 0 when the
 * gate has no generator rule. */
static int apply_generator(const qhip_stream* s, const qhip_op* op, double* st, int n, cplx* c) {
    const uint64_t dim = (uint64_t)1 << n;
    const uint32_t* q = op_qubits(s, op);
    cplx* S = (cplx*)st;
    const cplx mhalf = mk(0, -0.5), plus_i = mk(0, 1);
    qhip_op tmp; memset(&tmp, 0, sizeof tmp);
    qhip_stream ts = *s;
    uint32_t tq[2];
    ts.qubits = tq;
    tmp.n_qubits = 1; tmp.qubit_offset = 0; tmp.unitary_offset = -1; tmp.cbit = -1;
#define PAULI(G, Q) do { tmp.gate = (G); tq[0] = (Q); apply_op(&ts, &tmp, st, n); } while (0)
    switch (op->gate) {
        case QHIP_RX: PAULI(QHIP_X, q[0]); *c = mhalf; return 1;
        case QHIP_RY: PAULI(QHIP_Y, q[0]); *c = mhalf; return 1;
        case QHIP_RZ: PAULI(QHIP_Z, q[0]); *c = mhalf; return 1;
        case QHIP_P: {
            const uint64_t bit = (uint64_t)1 << q[0];
            for (uint64_t i = 0; i < dim; ++i) if (!(i & bit)) S[i] = mk(0, 0);
            *c = plus_i; return 1;
        }
        case QHIP_GPHASE: *c = plus_i; return 1;
        case QHIP_CRX: case QHIP_CRY: case QHIP_CRZ: {
            const uint64_t bit = (uint64_t)1 << q[0];
            for (uint64_t i = 0; i < dim; ++i) if (!(i & bit)) S[i] = mk(0, 0);
            PAULI(op->gate == QHIP_CRX ? QHIP_X : op->gate == QHIP_CRY ? QHIP_Y : QHIP_Z, q[1]);
            *c = mhalf; return 1;
        }
        case QHIP_CP: {
            const uint64_t both = ((uint64_t)1 << q[0]) | ((uint64_t)1 << q[1]);
            for (uint64_t i = 0; i < dim; ++i) if ((i & both) != both) S[i] = mk(0, 0);
            *c = plus_i; return 1;
        }
        case QHIP_RXX: case QHIP_RYY: case QHIP_RZZ: {
            const int g = op->gate == QHIP_RXX ? QHIP_X : op->gate == QHIP_RYY ? QHIP_Y : QHIP_Z;
            PAULI(g, q[0]); PAULI(g, q[1]);
            *c = mhalf; return 1;
        }
        default: return 0;
    }
#undef PAULI
}

/* Command::dagger (core/command.cpp:19-59) applied in place. */
static int apply_dagger(const qhip_stream* s, const qhip_op* op, double* st, int n) {
    qhip_op d = *op;
    switch (op->gate) {
        case QHIP_S: d.gate = QHIP_SDG; break;   case QHIP_SDG: d.gate = QHIP_S; break;
        case QHIP_T: d.gate = QHIP_TDG; break;   case QHIP_TDG: d.gate = QHIP_T; break;
        case QHIP_SX: d.gate = QHIP_SXDG; break; case QHIP_SXDG: d.gate = QHIP_SX; break;
        case QHIP_ISWAP: d.gate = QHIP_ISWAPDG; break; case QHIP_ISWAPDG: d.gate = QHIP_ISWAP; break;
        case QHIP_CS: d.gate = QHIP_CSDG; break; case QHIP_CSDG: d.gate = QHIP_CS; break;
        case QHIP_CSX: d.gate = QHIP_CSXDG; break; case QHIP_CSXDG: d.gate = QHIP_CSX; break;
        case QHIP_U:
            d.params[0] = -op->params[0]; d.params[1] = -op->params[2]; d.params[2] = -op->params[1];
            break;
        case QHIP_CU:
            d.params[0] = -op->params[0]; d.params[1] = -op->params[2]; d.params[2] = -op->params[1];
            d.params[3] = -op->params[3];
            break;
        case QHIP_RX: case QHIP_RY: case QHIP_RZ: case QHIP_P: case QHIP_CRX: case QHIP_CRY:
        case QHIP_CRZ: case QHIP_CP: case QHIP_RZZ: case QHIP_RXX: case QHIP_RYY: case QHIP_GPHASE:
            for (int i = 0; i < op->n_params; ++i) d.params[i] = -op->params[i];
            break;
        case QHIP_CUSTOM: {
            const int D = 1 << op->n_qubits;
            const cplx* U = (const cplx*)(s->unitaries + 2 * op->unitary_offset);
            cplx* A = malloc(sizeof(cplx) * D * D);
            for (int y = 0; y < D; ++y)
                for (int x = 0; x < D; ++x) A[y * D + x] = mk(U[x * D + y].re, -U[x * D + y].im);
            apply_dense(op_qubits(s, op), op->n_qubits, A, st, (uint64_t)1 << n);
            free(A);
            return 0;
        }
        default: break;   /* self-adjoint */
    }
    return apply_op(s, &d, st, n);
}

static int gradient_core(const qhip_stream* s, int n, int n_params, const double* init,
                         const qhip_observable* obs, const double* phi_in, double* grad) {
    for (int64_t i = 0; i < s->n_ops; ++i) {
        const qhip_op* op = &s->ops[i];
        if (op->gate == QHIP_RESET || (op->gate == QHIP_MEASURE && op->cbit >= 0) ||
            (op->flags & QHIP_OP_CONDITIONAL) || op->gate == QHIP_BRANCH_BEGIN)
            return fail(QHIP_RUNTIME, "QarpSimulator::run_gradient: trajectory commands (Measure / "
                        "Reset / classical condition) are not differentiable.");
    }
    const uint64_t dim = (uint64_t)1 << n;
    double* psi = malloc(sizeof(double) * 2 * dim);
    double* phi = malloc(sizeof(double) * 2 * dim);
    double* scratch = malloc(sizeof(double) * 2 * dim);
    if (!psi || !phi || !scratch) { free(psi); free(phi); free(scratch); return fail(QHIP_NOMEM, "gradient buffers"); }
    if (init) memcpy(psi, init, sizeof(double) * 2 * dim);
    else { memset(psi, 0, sizeof(double) * 2 * dim); psi[0] = 1.0; }
    int rc = orc_apply_unfused(s, n, psi);
    if (!rc) {
        if (obs) rc = orc_apply_observable(psi, n, obs, phi);
        else memcpy(phi, phi_in, sizeof(double) * 2 * dim);
    }
    for (int k = 0; k < n_params; ++k) grad[k] = 0.0;
    for (int64_t idx = s->n_ops - 1; idx >= 0 && !rc; --idx) {
        const qhip_op* op = &s->ops[idx];
        if (op->n_deps > 0) {
            memcpy(scratch, psi, sizeof(double) * 2 * dim);
            cplx c;
            if (!apply_generator(s, op, scratch, n, &c)) {
                rc = fail(QHIP_CAPABILITY, "run_gradient: gate %d carries a symbolic parameter but has no "
                          "adjoint generator rule", op->gate);
                break;
            }
            double ipr = 0, ipi = 0;
            for (uint64_t i = 0; i < dim; ++i) {
                const double ar = phi[2*i], ai = -phi[2*i+1];
                const double br = scratch[2*i], bi = scratch[2*i+1];
                ipr += ar * br - ai * bi;
                ipi += ar * bi + ai * br;
            }
            const double base = 2.0 * (c.re * ipr - c.im * ipi);
            for (int d = 0; d < op->n_deps; ++d) {
                const int p = s->dep_index[op->dep_offset + d];
                if (p >= 0 && p < n_params) grad[p] += s->dep_alpha[op->dep_offset + d] * base;
            }
        }
        rc = apply_dagger(s, op, psi, n);
        if (!rc) rc = apply_dagger(s, op, phi, n);
    }
    free(psi); free(phi); free(scratch);
    return rc;
}

int orc_run_gradient(const qhip_stream* s, int n, const qhip_observable* obs, int n_params,
                     const double* init, double* grad) {
    return gradient_core(s, n, n_params, init, obs, NULL, grad);
}

int orc_run_gradient_phi(const qhip_stream* s, int n, const double* phi, int n_params,
                         const double* init, double* grad) {
    return gradient_core(s, n, n_params, init, NULL, phi, grad);
}

/* ── structured-program kernels (program.cpp:610-712) ──────────────────── */

/* apply_permutation (program.cpp:610-646), scatter form: out[(i & keep) |
 * deposit(table[extract(i)])] = in[i].  The caller validated the table. */
int orc_permute(const double* in, int n, const uint32_t* qubits, int k, const uint64_t* table,
                double* out) {
    if (k < 1 || k > 26) return fail(QHIP_INVALID, "orc_permute: 1..26 qubits required");
    const uint64_t dim = (uint64_t)1 << n;
    uint64_t keep = ~(uint64_t)0;
    for (int b = 0; b < k; ++b) keep &= ~((uint64_t)1 << qubits[b]);
    for (uint64_t i = 0; i < dim; ++i) {
        uint64_t x = 0;
        for (int b = 0; b < k; ++b) x |= ((i >> qubits[b]) & 1u) << b;
        const uint64_t y = table[x];
        uint64_t j = i & keep;
        for (int b = 0; b < k; ++b) j |= ((y >> b) & 1u) << qubits[b];
        out[2 * j] = in[2 * i];
        out[2 * j + 1] = in[2 * i + 1];
    }
    return 0;
}

static void mat_mul(const cplx* a, const cplx* b, cplx* c, uint64_t d) {
    for (uint64_t y = 0; y < d; ++y)
        for (uint64_t x = 0; x < d; ++x) {
            cplx s = {0, 0};
            for (uint64_t t = 0; t < d; ++t) {
                const cplx p = cmul(a[y * d + t], b[t * d + x]);
                s.re += p.re;
                s.im += p.im;
            }
            c[y * d + x] = s;
        }
}

/* matrix^e by the squaring chain of apply_controlled_powers: squares of the
 * matrix, multiplied from the lowest set bit of e up. */
static void mat_power(const cplx* m, uint64_t e, cplx* out, uint64_t d) {
    cplx* sq = malloc(sizeof(cplx) * d * d);
    cplx* tmp = malloc(sizeof(cplx) * d * d);
    memcpy(sq, m, sizeof(cplx) * d * d);
    int first = 1;
    for (int b = 0; b < 64 && (e >> b); ++b) {
        if (b > 0) {
            mat_mul(sq, sq, tmp, d);
            memcpy(sq, tmp, sizeof(cplx) * d * d);
        }
        if (!((e >> b) & 1)) continue;
        if (first) {
            memcpy(out, sq, sizeof(cplx) * d * d);
            first = 0;
        } else {
            mat_mul(out, sq, tmp, d);
            memcpy(out, tmp, sizeof(cplx) * d * d);
        }
    }
    free(sq);
    free(tmp);
}

/* apply_controlled_powers (program.cpp:648-712): for each control j in order
 * with e_j != 0, P = matrix^e_j applied to the target sub-vector of every
 * basis state whose control bit is set. */
int orc_controlled_powers(double* state, int n, const uint32_t* controls, const uint64_t* exponents,
                          int n_controls, const uint32_t* targets, int m, const double* matrix) {
    if (m < 1 || m > 26) return fail(QHIP_INVALID, "orc_controlled_powers: 1..26 targets");
    const uint64_t dim = (uint64_t)1 << n, d = (uint64_t)1 << m;
    uint64_t tmask = 0;
    for (int b = 0; b < m; ++b) tmask |= (uint64_t)1 << targets[b];
    cplx* M = malloc(sizeof(cplx) * d * d);
    cplx* P = malloc(sizeof(cplx) * d * d);
    cplx* col = malloc(sizeof(cplx) * d);
    uint64_t* off = malloc(sizeof(uint64_t) * d);
    for (uint64_t i = 0; i < d * d; ++i) M[i] = (cplx){matrix[2 * i], matrix[2 * i + 1]};
    for (uint64_t l = 0; l < d; ++l) {
        off[l] = 0;
        for (int b = 0; b < m; ++b) off[l] |= ((l >> b) & 1u) << targets[b];
    }
    for (int j = 0; j < n_controls; ++j) {
        if (exponents[j] == 0) continue;
        mat_power(M, exponents[j], P, d);
        const uint64_t cbit = (uint64_t)1 << controls[j];
        for (uint64_t base = 0; base < dim; ++base) {
            if ((base & tmask) || !(base & cbit)) continue;
            for (uint64_t l = 0; l < d; ++l) col[l] = (cplx){state[2 * (base | off[l])], state[2 * (base | off[l]) + 1]};
            for (uint64_t y = 0; y < d; ++y) {
                cplx s = {0, 0};
                for (uint64_t x = 0; x < d; ++x) {
                    const cplx p = cmul(P[y * d + x], col[x]);
                    s.re += p.re;
                    s.im += p.im;
                }
                state[2 * (base | off[y])] = s.re;
                state[2 * (base | off[y]) + 1] = s.im;
            }
        }
    }
    free(M); free(P); free(col); free(off);
    return 0;
}
