/*
 * qi_oracle.c — CPU FP64 restatement of quant-iron's state-vector hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see qi_oracle.h): the engine never links or calls
 * this file; tests, the smoke check and the benchmark's CPU baseline do.
 *
 * Arithmetic follows num-complex 0.4.6 (the reference's Complex type):
 *   mul (a,b)*(c,d) = (a*c - b*d, a*d + b*c); f64*Complex scales both parts;
 *   exp(z) = from_polar(exp(re), im) = (e*cos im, e*sin im);
 *   cosh(z) = (cosh re * cos im, sinh re * sin im);
 *   sinh(z) = (sinh re * cos im, cosh re * sin im).
 * Every expression keeps the reference's operand order, so with
 * -ffp-contract=off the per-gate results are the reference's bit for bit
 * (up to the sign of zero). Transcendentals come from the host libm (glibc),
 * as Rust's f64 methods do on Linux.
 *
 * Pinning, beyond this restatement (tests/test_oracle_*.py):
 *   - reference known-answer tests: operator_tests.rs, pauli_string_tests.rs,
 *     time_evolution_tests.rs, gate_tests.rs, circuit_tests.rs,
 *     measurement_tests.rs, state_tests.rs (values transcribed as fixtures);
 *   - dense-matrix brute force (numpy Kronecker products) for every gate kind,
 *     Pauli strings, exp(P) (scipy expm) and SumOp expectations, n <= 8;
 *   - closed forms: QFT = R*F*R, exp(-i t P) = cos t I - i sin t P,
 *     invariants H^2 = I, S S^dag = I, T^2 = S, SWAP = 3 CNOTs;
 *   - measurement: brute-force histogram and the sequential-cumsum rule.
 * Parity unpinned: rayon's reduction order (bit level only) and rand's draws
 * (replaced by injected uniforms).
 */
#include "qi_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cx;

static inline cx cx_make(double re, double im) { cx r = {re, im}; return r; }
static inline cx cx_mul(cx a, cx b) { return cx_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cx cx_add(cx a, cx b) { return cx_make(a.re + b.re, a.im + b.im); }
static inline cx cx_sub(cx a, cx b) { return cx_make(a.re - b.re, a.im - b.im); }
static inline cx cx_neg(cx a) { return cx_make(-a.re, -a.im); }
static inline cx cx_scale(double f, cx a) { return cx_make(f * a.re, f * a.im); }  /* f64 * Complex */
static inline cx cx_scale_r(cx a, double f) { return cx_make(a.re * f, a.im * f); } /* Complex * f64 */
static inline cx cx_conj(cx a) { return cx_make(a.re, -a.im); }
static inline double cx_norm_sqr(cx a) { return a.re * a.re + a.im * a.im; }
static inline cx cx_exp(cx z) { double e = exp(z.re); return cx_make(e * cos(z.im), e * sin(z.im)); }
static inline cx cx_cosh(cx z) { return cx_make(cosh(z.re) * cos(z.im), sinh(z.re) * sin(z.im)); }
static inline cx cx_sinh(cx z) { return cx_make(sinh(z.re) * cos(z.im), cosh(z.re) * sin(z.im)); }

#define AMP(p, i) (((cx *)(p))[i])

/* ---------------------------------------------------------------- errors -- */
/* Display strings follow the #[error] attributes of errors.rs:4-97. */
static int set_err(qi_error *err, int code, uint64_t a, uint64_t b, double re, double im) {
    if (!err) return code;
    memset(err, 0, sizeof(*err));
    err->code = code; err->a = a; err->b = b; err->re = re; err->im = im;
    switch (code) {
    case QI_ERR_INVALID_NUMBER_OF_MEASUREMENTS:
        snprintf(err->msg, sizeof err->msg, "Invalid number of measurements: %llu", (unsigned long long)a); break;
    case QI_ERR_OVERLAPPING_CONTROL_AND_TARGET:
        snprintf(err->msg, sizeof err->msg, "Control qubit index %llu overlaps with target qubit index %llu",
                 (unsigned long long)a, (unsigned long long)b); break;
    case QI_ERR_INVALID_NUMBER_OF_QUBITS:
        snprintf(err->msg, sizeof err->msg, "Invalid number of qubits: %llu", (unsigned long long)a); break;
    case QI_ERR_INVALID_QUBIT_INDEX:
        snprintf(err->msg, sizeof err->msg, "Invalid qubit index: %llu for %llu qubits",
                 (unsigned long long)a, (unsigned long long)b); break;
    case QI_ERR_STATE_NOT_NORMALISED: snprintf(err->msg, sizeof err->msg, "State vector is not normalised"); break;
    case QI_ERR_NON_UNITARY_MATRIX: snprintf(err->msg, sizeof err->msg, "Non-unitary matrix"); break;
    case QI_ERR_UNKNOWN: snprintf(err->msg, sizeof err->msg, "An unknown error occurred"); break;
    case QI_ERR_ZERO_NORM: snprintf(err->msg, sizeof err->msg, "The state cannot be normalised because it has zero norm."); break;
    case QI_ERR_INVALID_PAULI_COEFFICIENT:
        snprintf(err->msg, sizeof err->msg, "Invalid Pauli String coefficient: %g+%gi", re, im); break;
    case QI_ERR_INVALID_ARGUMENT: snprintf(err->msg, sizeof err->msg, "Invalid argument"); break;
    default: break;
    }
    return code;
}

/* ------------------------------------------------------------ validation -- */
/* validate_qubits, operator.rs:214-273. */
static int validate_qubits(uint32_t n, const uint64_t *t, uint32_t nt, const uint64_t *c, uint32_t nc,
                           uint32_t expected, qi_error *err) {
    if (nt != expected) return set_err(err, QI_ERR_INVALID_NUMBER_OF_QUBITS, nt, 0, 0, 0);
    for (uint32_t i = 0; i < nt; i++)
        if (t[i] >= n) return set_err(err, QI_ERR_INVALID_QUBIT_INDEX, t[i], n, 0, 0);
    for (uint32_t j = 0; j < nc; j++) {
        if (c[j] >= n) return set_err(err, QI_ERR_INVALID_QUBIT_INDEX, c[j], n, 0, 0);
        for (uint32_t i = 0; i < nt; i++)
            if (c[j] == t[i]) return set_err(err, QI_ERR_OVERLAPPING_CONTROL_AND_TARGET, c[j], t[i], 0, 0);
    }
    if (expected > 1)
        for (uint32_t i = 0; i < nt; i++)
            for (uint32_t j = i + 1; j < nt; j++)
                if (t[i] == t[j]) return set_err(err, QI_ERR_INVALID_QUBIT_INDEX, t[i], n, 0, 0);
    return QI_OK;
}

/* Unitary2::new, operator.rs:2092-2118. */
static int check_unitary2(const double *m, qi_error *err) {
    double tol = DBL_EPSILON * 2.0;
    cx a = cx_make(m[0], m[1]), b = cx_make(m[2], m[3]), c = cx_make(m[4], m[5]), d = cx_make(m[6], m[7]);
    if (fabs((cx_norm_sqr(a) + cx_norm_sqr(b)) - 1.0) > tol) return set_err(err, QI_ERR_NON_UNITARY_MATRIX, 0, 0, 0, 0);
    if (fabs((cx_norm_sqr(c) + cx_norm_sqr(d)) - 1.0) > tol) return set_err(err, QI_ERR_NON_UNITARY_MATRIX, 0, 0, 0, 0);
    if (cx_norm_sqr(cx_add(cx_mul(a, cx_conj(c)), cx_mul(b, cx_conj(d)))) > tol * tol)
        return set_err(err, QI_ERR_NON_UNITARY_MATRIX, 0, 0, 0, 0);
    return QI_OK;
}

/* Lowest mask bit at or above n: the qubit index an op on it would reject. */
static int check_mask(uint32_t n, uint64_t mask, qi_error *err) {
    uint64_t bad = n >= 64 ? 0 : (mask >> n);
    if (bad) {
        unsigned q = n + (unsigned)__builtin_ctzll(bad);
        return set_err(err, QI_ERR_INVALID_QUBIT_INDEX, q, n, 0, 0);
    }
    return QI_OK;
}

static int validate_measure(uint32_t n, const uint64_t *q, uint32_t m, qi_error *err) {
    /* State::measure, state.rs:525-560: empty = all; count and index checks. */
    if (m > n) return set_err(err, QI_ERR_INVALID_NUMBER_OF_QUBITS, n, 0, 0, 0);
    for (uint32_t i = 0; i < m; i++)
        if (q[i] >= n) return set_err(err, QI_ERR_INVALID_QUBIT_INDEX, q[i], n, 0, 0);
    return QI_OK;
}

int or_validate_gate(uint32_t n, const qi_gate *g, qi_error *err) {
    if (!g) return set_err(err, QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0);
    uint32_t nt = g->ntargets, nc = g->ncontrols;
    if (nc > QI_MAX_CONTROLS) return set_err(err, QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0);
    int st;
    switch (g->kind) {
    case QI_GATE_H: case QI_GATE_X: case QI_GATE_Y: case QI_GATE_Z: case QI_GATE_I:
    case QI_GATE_S: case QI_GATE_T: case QI_GATE_SDAG: case QI_GATE_TDAG: case QI_GATE_P:
    case QI_GATE_RX: case QI_GATE_RY: case QI_GATE_RZ:
        return validate_qubits(n, g->targets, nt, g->controls, nc, 1, err);
    case QI_GATE_U2:
        if ((st = check_unitary2(g->matrix, err))) return st;
        return validate_qubits(n, g->targets, nt, g->controls, nc, 1, err);
    case QI_GATE_CNOT: /* operator.rs:667-685 */
        if ((st = validate_qubits(n, g->targets, nt, g->controls, nc, 1, err))) return st;
        if (nc != 1) return set_err(err, QI_ERR_INVALID_NUMBER_OF_QUBITS, nc, 0, 0, 0);
        return QI_OK;
    case QI_GATE_TOFFOLI: /* operator.rs:1055-1075 */
        if ((st = validate_qubits(n, g->targets, nt, g->controls, nc, 1, err))) return st;
        if (nc != 2) return set_err(err, QI_ERR_INVALID_NUMBER_OF_QUBITS, nc, 0, 0, 0);
        if (g->controls[0] == g->controls[1]) return set_err(err, QI_ERR_INVALID_NUMBER_OF_QUBITS, nc, 0, 0, 0);
        return QI_OK;
    case QI_GATE_SWAP:
        return validate_qubits(n, g->targets, nt, g->controls, nc, 2, err);
    case QI_GATE_MATCHGATE: /* operator.rs:893-907 */
        if ((st = validate_qubits(n, g->targets, nt, g->controls, nc, 1, err))) return st;
        if (g->targets[0] == (uint64_t)n - 1) return set_err(err, QI_ERR_INVALID_QUBIT_INDEX, g->targets[0], n, 0, 0);
        return QI_OK;
    case QI_GATE_MEASURE:
        if (g->basis == QI_BASIS_CUSTOM && (st = check_unitary2(g->matrix, err))) return st;
        if (g->basis > QI_BASIS_CUSTOM) return set_err(err, QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0);
        return validate_measure(n, g->controls, nc, err);
    case QI_GATE_PAULI_STRING:
        return check_mask(n, g->x_mask | g->z_mask, err);
    case QI_GATE_PAULI_EVOLUTION:
        if (g->params[1] != 0.0)
            return set_err(err, QI_ERR_INVALID_PAULI_COEFFICIENT, 0, 0, g->params[0], g->params[1]);
        return check_mask(n, g->x_mask | g->z_mask, err);
    default:
        return set_err(err, QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0);
    }
}

/* ----------------------------------------------------------------- gates -- */
/* check_controls, operator.rs:195-199. */
static inline int controls_ok(uint64_t i, const uint64_t *c, uint32_t nc) {
    for (uint32_t k = 0; k < nc; k++)
        if (((i >> c[k]) & 1) == 0) return 0;
    return 1;
}

/* Pairwise 1-qubit gates in the sequential form of operator.rs (e.g. Hadamard
 * :396-412): every i with target bit 0 and controls met updates (i, i|2^t)
 * from the old pair. In place is equivalent to the reference's clone because
 * each pair is read before it is written and pairs are disjoint. */
typedef enum { PK_H, PK_X, PK_Y, PK_RX, PK_RY, PK_U2 } pair_kind;

static void apply_pair(double *psi, uint32_t n, uint64_t t, const uint64_t *c, uint32_t nc,
                       pair_kind k, const double *prm, const double *mat) {
    uint64_t dim = 1ULL << n, tb = 1ULL << t;
    const double s2 = 1.0 / sqrt(2.0);                  /* operator.rs:316 */
    const cx i_c = cx_make(0.0, 1.0), mi_c = cx_neg(i_c); /* operator.rs:489 */
    double ch = 0, sh = 0;
    if (k == PK_RX || k == PK_RY) { double h = prm[0] / 2.0; ch = cos(h); sh = sin(h); }
    cx m00 = {0}, m01 = {0}, m10 = {0}, m11 = {0};
    if (k == PK_U2) {
        m00 = cx_make(mat[0], mat[1]); m01 = cx_make(mat[2], mat[3]);
        m10 = cx_make(mat[4], mat[5]); m11 = cx_make(mat[6], mat[7]);
    }
    for (uint64_t i = 0; i < dim; i++) {
        if (i & tb) continue;
        if (!controls_ok(i, c, nc)) continue;
        uint64_t j = i | tb;
        cx a = AMP(psi, i), b = AMP(psi, j), na, nb;
        switch (k) {
        case PK_H: na = cx_scale(s2, cx_add(a, b)); nb = cx_scale(s2, cx_sub(a, b)); break;
        case PK_X: na = b; nb = a; break;
        case PK_Y: na = cx_mul(mi_c, b); nb = cx_mul(i_c, a); break;             /* operator.rs:539-540 */
        case PK_RX: /* operator.rs:1756-1757 */
            na = cx_sub(cx_scale(ch, a), cx_mul(cx_scale_r(i_c, sh), b));
            nb = cx_add(cx_mul(cx_scale_r(mi_c, sh), a), cx_scale(ch, b));
            break;
        case PK_RY: /* operator.rs:1897-1898 */
            na = cx_sub(cx_scale(ch, a), cx_scale(sh, b));
            nb = cx_add(cx_scale(sh, a), cx_scale(ch, b));
            break;
        default: /* U2, operator.rs:2253-2254 */
            na = cx_add(cx_mul(m00, a), cx_mul(m01, b));
            nb = cx_add(cx_mul(m10, a), cx_mul(m11, b));
            break;
        }
        AMP(psi, i) = na; AMP(psi, j) = nb;
    }
}

/* Diagonal gates: Z negates, S/T/Sdag/Tdag/P multiply target-1 amplitudes by
 * a phase (operator.rs:1191-1207 and siblings), RZ multiplies both halves
 * (operator.rs:2013-2028). */
static void apply_diag(double *psi, uint32_t n, uint64_t t, const uint64_t *c, uint32_t nc, int kind,
                       double angle) {
    uint64_t dim = 1ULL << n, tb = 1ULL << t;
    const double r2 = 1.0 / sqrt(2.0);
    cx ph1 = {0}, ph0 = {0};
    switch (kind) {
    case QI_GATE_S: ph1 = cx_make(0.0, 1.0); break;
    case QI_GATE_SDAG: ph1 = cx_make(0.0, -1.0); break;
    case QI_GATE_T: ph1 = cx_make(r2, r2); break;        /* operator.rs:1288-1289 */
    case QI_GATE_TDAG: ph1 = cx_make(r2, -r2); break;
    case QI_GATE_P: ph1 = cx_make(cos(angle), sin(angle)); break;
    case QI_GATE_RZ: { double h = angle / 2.0; ph0 = cx_make(cos(h), -sin(h)); ph1 = cx_make(cos(h), sin(h)); } break;
    default: break;
    }
    for (uint64_t i = 0; i < dim; i++) {
        int one = (i & tb) != 0;
        if (kind != QI_GATE_RZ && !one) continue;
        if (!controls_ok(i, c, nc)) continue;
        cx a = AMP(psi, i);
        if (kind == QI_GATE_Z) AMP(psi, i) = cx_neg(a);
        else AMP(psi, i) = cx_mul(a, one ? ph1 : ph0);
    }
}

/* SWAP, operator.rs:800-815: i < j, bits differ, controls on i. */
static void apply_swap(double *psi, uint32_t n, uint64_t a, uint64_t b, const uint64_t *c, uint32_t nc) {
    uint64_t dim = 1ULL << n;
    for (uint64_t i = 0; i < dim; i++) {
        uint64_t ba = (i >> a) & 1, bb = (i >> b) & 1;
        if (ba == bb) continue;
        uint64_t j = i ^ (1ULL << a) ^ (1ULL << b);
        if (i < j && controls_ok(i, c, nc)) { cx t = AMP(psi, i); AMP(psi, i) = AMP(psi, j); AMP(psi, j) = t; }
    }
}

/* Matchgate, operator.rs:980-1006 (the CPU form, e^{i phi1} on both
 * off-diagonal terms; controls tested on i01 for the rotation and on i11 for
 * the phase). */
static void apply_matchgate(double *psi, uint32_t n, uint64_t q1, const uint64_t *c, uint32_t nc,
                            const double *prm) {
    uint64_t q2 = q1 + 1;
    double ch = cos(prm[0] / 2.0), sh = sin(prm[0] / 2.0);
    cx e1 = cx_exp(cx_make(0.0, prm[1])), e2 = cx_exp(cx_make(0.0, prm[2]));
    uint64_t cnt = 1ULL << (n - 2);
    for (uint64_t i = 0; i < cnt; i++) {
        uint64_t k = ((i >> q1) << (q1 + 1)) | (i & ((1ULL << q1) - 1));
        uint64_t l = ((k >> (q2 - 1)) << q2) | (k & ((1ULL << (q2 - 1)) - 1));
        uint64_t i01 = l | (1ULL << q1), i10 = l | (1ULL << q2), i11 = i01 | (1ULL << q2);
        if (controls_ok(i01, c, nc)) {
            cx a01 = AMP(psi, i01), a10 = AMP(psi, i10);
            AMP(psi, i01) = cx_sub(cx_scale(ch, a01), cx_mul(cx_scale_r(e1, sh), a10));
            AMP(psi, i10) = cx_add(cx_scale(sh, a01), cx_mul(cx_scale_r(e1, ch), a10));
        }
        if (controls_ok(i11, c, nc)) AMP(psi, i11) = cx_mul(AMP(psi, i11), e2);
    }
}

/* Operator::apply without validation. */
static void apply_operator(double *psi, uint32_t n, const qi_gate *g) {
    const uint64_t *c = g->controls;
    uint32_t nc = g->ncontrols;
    uint64_t t = g->targets[0];
    switch (g->kind) {
    case QI_GATE_H: apply_pair(psi, n, t, c, nc, PK_H, 0, 0); break;
    case QI_GATE_X: case QI_GATE_CNOT: case QI_GATE_TOFFOLI: apply_pair(psi, n, t, c, nc, PK_X, 0, 0); break;
    case QI_GATE_Y: apply_pair(psi, n, t, c, nc, PK_Y, 0, 0); break;
    case QI_GATE_RX: apply_pair(psi, n, t, c, nc, PK_RX, g->params, 0); break;
    case QI_GATE_RY: apply_pair(psi, n, t, c, nc, PK_RY, g->params, 0); break;
    case QI_GATE_U2: apply_pair(psi, n, t, c, nc, PK_U2, 0, g->matrix); break;
    case QI_GATE_Z: case QI_GATE_S: case QI_GATE_T: case QI_GATE_SDAG: case QI_GATE_TDAG:
    case QI_GATE_P: case QI_GATE_RZ: apply_diag(psi, n, t, c, nc, g->kind, g->params[0]); break;
    case QI_GATE_I: break;
    case QI_GATE_SWAP: apply_swap(psi, n, g->targets[0], g->targets[1], c, nc); break;
    case QI_GATE_MATCHGATE: apply_matchgate(psi, n, t, c, nc, g->params); break;
    default: break;
    }
}

/* ------------------------------------------------------------ Pauli ops -- */
/* Pauli ops of a string applied one qubit at a time (pauli_string.rs:172-184),
 * ascending qubit order; single-qubit Paulis on distinct qubits commute
 * exactly, so the HashMap order of the reference does not matter. */
static void apply_pauli_ops(double *psi, uint32_t n, uint64_t x, uint64_t z) {
    for (uint32_t q = 0; q < n; q++) {
        int hx = (x >> q) & 1, hz = (z >> q) & 1;
        if (!hx && !hz) continue;
        qi_gate g; memset(&g, 0, sizeof g);
        g.kind = hx && hz ? QI_GATE_Y : (hx ? QI_GATE_X : QI_GATE_Z);
        g.ntargets = 1; g.targets[0] = q;
        apply_operator(psi, n, &g);
    }
}

static void scale_state(double *psi, uint32_t n, cx f) { /* State * Complex, state.rs:2687-2705 */
    uint64_t dim = 1ULL << n;
    for (uint64_t i = 0; i < dim; i++) AMP(psi, i) = cx_mul(AMP(psi, i), f);
}

int or_pauli_apply(const double *in, double *out, uint32_t n, const qi_pauli_term *t, qi_error *err) {
    int st = check_mask(n, t->x_mask | t->z_mask, err);
    if (st) return st;
    uint64_t dim = 1ULL << n;
    if (out != in) memcpy(out, in, dim * sizeof(cx));
    apply_pauli_ops(out, n, t->x_mask, t->z_mask);
    scale_state(out, n, cx_make(t->c_re, t->c_im));
    return QI_OK;
}

int or_normalise(double *psi, uint32_t n, qi_error *err) { /* state.rs:924-945 */
    uint64_t dim = 1ULL << n;
    double s = 0.0;
    for (uint64_t i = 0; i < dim; i++) s += cx_norm_sqr(AMP(psi, i));
    double norm = sqrt(s);
    if (norm == 0.0) return set_err(err, QI_ERR_ZERO_NORM, 0, 0, 0, 0);
    if (norm == 1.0) return QI_OK;
    for (uint64_t i = 0; i < dim; i++) { cx a = AMP(psi, i); AMP(psi, i) = cx_make(a.re / norm, a.im / norm); }
    return QI_OK;
}

int or_pauli_apply_normalised(const double *in, double *out, uint32_t n, const qi_pauli_term *t, qi_error *err) {
    int st = check_mask(n, t->x_mask | t->z_mask, err);
    if (st) return st;
    uint64_t dim = 1ULL << n;
    if (out != in) memcpy(out, in, dim * sizeof(cx));
    apply_pauli_ops(out, n, t->x_mask, t->z_mask);
    return or_normalise(out, n, err);
}

/* apply_exp_factor, pauli_string.rs:237-262. */
int or_pauli_exp(double *psi, uint32_t n, const qi_pauli_term *t, double f_re, double f_im, qi_error *err) {
    cx alpha = cx_mul(cx_make(t->c_re, t->c_im), cx_make(f_re, f_im));
    uint64_t dim = 1ULL << n;
    if ((t->x_mask | t->z_mask) == 0) { scale_state(psi, n, cx_exp(alpha)); return QI_OK; }
    int st = check_mask(n, t->x_mask | t->z_mask, err);
    if (st) return st;
    cx *p = (cx *)malloc(dim * sizeof(cx));
    if (!p) return set_err(err, QI_ERR_OUT_OF_MEMORY, 0, 0, 0, 0);
    memcpy(p, psi, dim * sizeof(cx));
    apply_pauli_ops((double *)p, n, t->x_mask, t->z_mask);
    cx ch = cx_cosh(alpha), sh = cx_sinh(alpha);
    for (uint64_t i = 0; i < dim; i++) AMP(psi, i) = cx_add(cx_mul(AMP(psi, i), ch), cx_mul(p[i], sh));
    free(p);
    return QI_OK;
}

int or_pauli_exp_neg_i_dt(double *psi, uint32_t n, const qi_pauli_term *t, double dt, qi_error *err) {
    if (t->c_im != 0.0) return set_err(err, QI_ERR_INVALID_PAULI_COEFFICIENT, 0, 0, t->c_re, t->c_im);
    return or_pauli_exp(psi, n, t, 0.0, -dt, err);
}

/* SumOp::apply, pauli_string.rs:453-466: sum of term images, folded in term order. */
int or_sumop_apply(const double *in, double *out, uint32_t n, const qi_pauli_term *terms, size_t nt, qi_error *err) {
    uint64_t dim = 1ULL << n;
    if (nt == 0) { memcpy(out, in, dim * sizeof(cx)); scale_state(out, n, cx_make(0.0, 0.0)); return QI_OK; }
    for (size_t k = 0; k < nt; k++) {
        int st = check_mask(n, terms[k].x_mask | terms[k].z_mask, err);
        if (st) return st;
    }
    cx *tmp = (cx *)malloc(dim * sizeof(cx));
    if (!tmp) return set_err(err, QI_ERR_OUT_OF_MEMORY, 0, 0, 0, 0);
    or_pauli_apply(in, out, n, &terms[0], err);
    for (size_t k = 1; k < nt; k++) {
        or_pauli_apply(in, (double *)tmp, n, &terms[k], err);
        for (uint64_t i = 0; i < dim; i++) AMP(out, i) = cx_add(AMP(out, i), tmp[i]);
    }
    free(tmp);
    return QI_OK;
}

void or_inner_product(const double *a, const double *b, uint32_t n, double *re, double *im) {
    uint64_t dim = 1ULL << n;
    cx s = cx_make(0.0, 0.0);
    for (uint64_t i = 0; i < dim; i++) s = cx_add(s, cx_mul(cx_conj(AMP(a, i)), AMP(b, i)));
    *re = s.re; *im = s.im;
}

double or_norm_sqr(const double *psi, uint32_t n) {
    uint64_t dim = 1ULL << n;
    double s = 0.0;
    for (uint64_t i = 0; i < dim; i++) s += cx_norm_sqr(AMP(psi, i));
    return s;
}

/* SumOp::expectation_value, pauli_string.rs:485-507. */
int or_sumop_expectation(const double *psi, uint32_t n, const qi_pauli_term *terms, size_t nt,
                         double *re, double *im, qi_error *err) {
    *re = 0.0; *im = 0.0;
    if (nt == 0) return QI_OK;
    uint64_t dim = 1ULL << n;
    for (size_t k = 0; k < nt; k++) {
        int st = check_mask(n, terms[k].x_mask | terms[k].z_mask, err);
        if (st) return st;
    }
    cx *phi = (cx *)malloc(dim * sizeof(cx));
    if (!phi) return set_err(err, QI_ERR_OUT_OF_MEMORY, 0, 0, 0, 0);
    cx tot = cx_make(0.0, 0.0);
    for (size_t k = 0; k < nt; k++) {
        double r, i;
        or_pauli_apply(psi, (double *)phi, n, &terms[k], err);
        or_inner_product(psi, (double *)phi, n, &r, &i);
        tot = cx_add(tot, cx_make(r, i));
    }
    free(phi);
    *re = tot.re; *im = tot.im;
    return QI_OK;
}

/* first_order_trotter_step / second_order_trotter_step / trotter_evolve_state,
 * time_evolution.rs:45-166. */
int or_trotter_evolve(double *psi, uint32_t n, const qi_pauli_term *terms, size_t nt, double dt,
                      uint64_t steps, int order, qi_error *err) {
    if (nt == 0) return set_err(err, QI_ERR_INVALID_NUMBER_OF_QUBITS, 0, 0, 0, 0);
    for (uint64_t s = 0; s < steps; s++) {
        if (order == 2) {
            double h = -dt / 2.0;
            for (size_t k = 0; k < nt; k++) { int st = or_pauli_exp(psi, n, &terms[k], 0.0, h, err); if (st) return st; }
            for (size_t k = nt; k-- > 0;) { int st = or_pauli_exp(psi, n, &terms[k], 0.0, h, err); if (st) return st; }
        } else {
            for (size_t k = 0; k < nt; k++) { int st = or_pauli_exp(psi, n, &terms[k], 0.0, -dt, err); if (st) return st; }
        }
    }
    return QI_OK;
}

/* ---------------------------------------------------------- constructors -- */
int or_state_init(double *psi, uint32_t n, int kind, uint64_t basis, qi_error *err) {
    uint64_t dim = 1ULL << n;
    if (kind == QI_INIT_BASIS && basis >= dim) return set_err(err, QI_ERR_INVALID_QUBIT_INDEX, basis, n, 0, 0);
    if (n == 0) return set_err(err, QI_ERR_INVALID_NUMBER_OF_QUBITS, 0, 0, 0, 0);
    memset(psi, 0, dim * sizeof(cx));
    switch (kind) {
    case QI_INIT_ZERO: AMP(psi, 0) = cx_make(1.0, 0.0); break;
    case QI_INIT_BASIS: AMP(psi, basis) = cx_make(1.0, 0.0); break;
    case QI_INIT_PLUS: { double a = 1.0 / sqrt((double)dim); for (uint64_t i = 0; i < dim; i++) AMP(psi, i) = cx_make(a, 0.0); } break;
    case QI_INIT_MINUS: {
        double a = 1.0 / sqrt((double)dim);
        for (uint64_t i = 0; i < dim; i++) AMP(psi, i) = (__builtin_popcountll(i) % 2 == 0) ? cx_make(a, 0.0) : cx_neg(cx_make(a, 0.0));
    } break;
    case QI_INIT_GHZ: {
        double a = 0.70710678118654752440; /* FRAC_1_SQRT_2 */
        AMP(psi, 0) = cx_make(a, 0.0); AMP(psi, dim - 1) = cx_make(a, 0.0);
    } break;
    default: return set_err(err, QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0);
    }
    return QI_OK;
}

int or_state_check_norm(const double *psi, uint32_t n, qi_error *err) {
    uint64_t dim = 1ULL << n;
    double s = or_norm_sqr(psi, n);
    double tol = DBL_EPSILON * (double)dim;
    if (fabs(s - 1.0) > tol) return set_err(err, QI_ERR_STATE_NOT_NORMALISED, 0, 0, 0, 0);
    return QI_OK;
}

int or_fidelity(const double *a, const double *b, uint32_t n, double *out, qi_error *err) {
    uint64_t dim = 1ULL << n;
    cx *na = (cx *)malloc(dim * sizeof(cx)), *nb = (cx *)malloc(dim * sizeof(cx));
    if (!na || !nb) { free(na); free(nb); return set_err(err, QI_ERR_OUT_OF_MEMORY, 0, 0, 0, 0); }
    memcpy(na, a, dim * sizeof(cx)); memcpy(nb, b, dim * sizeof(cx));
    int st = or_normalise((double *)na, n, err);
    if (!st) st = or_normalise((double *)nb, n, err);
    if (!st) { double r, i; or_inner_product((double *)na, (double *)nb, n, &r, &i); *out = r * r + i * i; }
    free(na); free(nb);
    return st;
}

/* fs_fidelity / fs_dist (state.rs:470-498) on states of na and nb qubits, in
 * the reference's order: normalise self, normalise other (ZeroNorm each,
 * state.rs:924-935), then inner_product's checks (state.rs:890-897:
 * InvalidNumberOfQubits(0) for an empty state, InvalidNumberOfQubits(na) when
 * the lengths differ). `dist` != 0 returns acos(|<a|b>|) (Complex::norm is
 * hypot), else |<a|b>|^2 (Complex::norm_sqr). */
int or_fs_metric(const double *a, uint32_t na, const double *b, uint32_t nb, int dist, double *out, qi_error *err) {
    uint64_t da = 1ULL << na, db = 1ULL << nb;
    cx *ca = (cx *)malloc(da * sizeof(cx)), *cb = (cx *)malloc(db * sizeof(cx));
    if (!ca || !cb) { free(ca); free(cb); return set_err(err, QI_ERR_OUT_OF_MEMORY, 0, 0, 0, 0); }
    memcpy(ca, a, da * sizeof(cx)); memcpy(cb, b, db * sizeof(cx));
    int st = or_normalise((double *)ca, na, err);
    if (!st) st = or_normalise((double *)cb, nb, err);
    if (!st && (na == 0 || nb == 0)) st = set_err(err, QI_ERR_INVALID_NUMBER_OF_QUBITS, 0, 0, 0, 0);
    if (!st && da != db) st = set_err(err, QI_ERR_INVALID_NUMBER_OF_QUBITS, na, 0, 0, 0);
    if (!st) {
        double r, i;
        or_inner_product((double *)ca, (double *)cb, na, &r, &i);
        *out = dist ? acos(hypot(r, i)) : r * r + i * i;
    }
    free(ca); free(cb);
    return st;
}

/* impl Add / Sub for State (state.rs:2773-2800, 2836-2862): elementwise a + b
 * or a - b; impl Mul<Complex<f64>> (state.rs:2687-2705): amplitude * rhs. */
void or_state_add(const double *a, const double *b, double *out, uint32_t n, int sub) {
    uint64_t dim = 1ULL << n;
    for (uint64_t i = 0; i < dim; i++) AMP(out, i) = sub ? cx_sub(AMP(a, i), AMP(b, i)) : cx_add(AMP(a, i), AMP(b, i));
}

void or_state_scale(double *psi, uint32_t n, double re, double im) { scale_state(psi, n, cx_make(re, im)); }

/* ----------------------------------------------------------- measurement -- */
static uint64_t outcome_of(uint64_t idx, const uint64_t *q, uint32_t m) {
    uint64_t o = 0;
    for (uint32_t b = 0; b < m; b++)
        if ((idx >> q[b]) & 1) o |= 1ULL << b;
    return o;
}

int or_measure_probs(const double *psi, uint32_t n, const uint64_t *qubits, uint32_t m, double *probs, qi_error *err) {
    uint64_t all[64];
    if (m == 0) { for (uint32_t i = 0; i < n; i++) all[i] = i; qubits = all; m = n; }
    int st = validate_measure(n, qubits, m, err);
    if (st) return st;
    uint64_t dim = 1ULL << n, nb = 1ULL << m;
    memset(probs, 0, nb * sizeof(double));
    for (uint64_t i = 0; i < dim; i++) probs[outcome_of(i, qubits, m)] += cx_norm_sqr(AMP(psi, i));
    return QI_OK;
}

uint64_t or_sample_outcome(const double *probs, uint64_t nb, double u) {
    double total = 0.0;
    for (uint64_t i = 0; i < nb; i++) total += probs[i];
    double cum = 0.0;
    uint64_t pick = 0;
    int found = 0;
    for (uint64_t i = 0; i < nb; i++) {
        cum += probs[i] / total;
        if (u < cum) { pick = i; found = 1; break; }
    }
    if (!found && u >= cum && nb > 0) pick = nb - 1;
    return pick;
}

/* The cumulative probabilities C_i of that rule (sequential sums of
 * p_i / total), the bin boundaries the uniform draw is compared with. */
void or_sample_cdf(const double *probs, uint64_t nb, double *cdf) {
    double total = 0.0;
    for (uint64_t i = 0; i < nb; i++) total += probs[i];
    double cum = 0.0;
    for (uint64_t i = 0; i < nb; i++) { cum += probs[i] / total; cdf[i] = cum; }
}

/* Collapse and renormalise, state.rs:622-667 (including State::new's check). */
int or_measure_collapse(const double *psi, double *out, uint32_t n, const uint64_t *qubits, uint32_t m,
                        uint64_t outcome, qi_error *err) {
    uint64_t all[64];
    if (m == 0) { for (uint32_t i = 0; i < n; i++) all[i] = i; qubits = all; m = n; }
    uint64_t dim = 1ULL << n;
    double ns = 0.0;
    for (uint64_t i = 0; i < dim; i++) {
        if (outcome_of(i, qubits, m) == outcome) { AMP(out, i) = AMP(psi, i); ns += cx_norm_sqr(AMP(psi, i)); }
        else AMP(out, i) = cx_make(0.0, 0.0);
    }
    if (ns > DBL_EPSILON) {
        double f = sqrt(ns);
        for (uint64_t i = 0; i < dim; i++) { cx a = AMP(out, i); AMP(out, i) = cx_make(a.re / f, a.im / f); }
    }
    return or_state_check_norm(out, n, err);
}

static void multi_1q(double *psi, uint32_t n, int kind, const uint64_t *q, uint32_t m, const double *mat) {
    for (uint32_t k = 0; k < m; k++) {
        qi_gate g; memset(&g, 0, sizeof g);
        g.kind = kind; g.ntargets = 1; g.targets[0] = q[k];
        if (mat) memcpy(g.matrix, mat, sizeof g.matrix);
        apply_operator(psi, n, &g);
    }
}

/* calculate_adjoint of a 2x2 matrix (conjugate transpose). */
static void adjoint2(const double *u, double *ud) {
    ud[0] = u[0]; ud[1] = -u[1]; ud[2] = u[4]; ud[3] = -u[5];
    ud[4] = u[2]; ud[5] = -u[3]; ud[6] = u[6]; ud[7] = -u[7];
}

/* State::measure, state.rs:525-730, with the uniform injected. */
int or_measure(const double *psi, uint32_t n, int basis, const double *custom_u, const uint64_t *qubits,
               uint32_t m, double uniform, double *out_psi, uint64_t *outcome, qi_error *err) {
    uint64_t all[64];
    if (m == 0) { for (uint32_t i = 0; i < n; i++) all[i] = i; qubits = all; m = n; }
    int st = validate_measure(n, qubits, m, err);
    if (st) return st;
    if (basis == QI_BASIS_CUSTOM && (st = check_unitary2(custom_u, err))) return st;
    uint64_t dim = 1ULL << n, nb = 1ULL << m;
    cx *w = (cx *)malloc(dim * sizeof(cx));
    double *probs = (double *)malloc(nb * sizeof(double));
    if (!w || !probs) { free(w); free(probs); return set_err(err, QI_ERR_OUT_OF_MEMORY, 0, 0, 0, 0); }
    memcpy(w, psi, dim * sizeof(cx));
    if (basis == QI_BASIS_X) multi_1q((double *)w, n, QI_GATE_H, qubits, m, 0);
    else if (basis == QI_BASIS_Y) { multi_1q((double *)w, n, QI_GATE_SDAG, qubits, m, 0); multi_1q((double *)w, n, QI_GATE_H, qubits, m, 0); }
    else if (basis == QI_BASIS_CUSTOM) multi_1q((double *)w, n, QI_GATE_U2, qubits, m, custom_u);
    for (uint64_t i = 0; i < nb; i++) probs[i] = 0.0;
    for (uint64_t i = 0; i < dim; i++) probs[outcome_of(i, qubits, m)] += cx_norm_sqr(w[i]);
    double total = 0.0;
    for (uint64_t i = 0; i < nb; i++) total += probs[i];
    if (total < DBL_EPSILON) { free(w); free(probs); return set_err(err, QI_ERR_UNKNOWN, 0, 0, 0, 0); }
    uint64_t o = or_sample_outcome(probs, nb, uniform);
    free(probs);
    st = or_measure_collapse((double *)w, out_psi, n, qubits, m, o, err);
    free(w);
    if (st) return st;
    if (basis == QI_BASIS_X) multi_1q(out_psi, n, QI_GATE_H, qubits, m, 0);
    else if (basis == QI_BASIS_Y) { multi_1q(out_psi, n, QI_GATE_H, qubits, m, 0); multi_1q(out_psi, n, QI_GATE_S, qubits, m, 0); }
    else if (basis == QI_BASIS_CUSTOM) { double ud[8]; adjoint2(custom_u, ud); multi_1q(out_psi, n, QI_GATE_U2, qubits, m, ud); }
    *outcome = o;
    return QI_OK;
}

int or_measure_n(const double *psi, uint32_t n, int basis, const double *custom_u, const uint64_t *qubits,
                 uint32_t m, const double *uniforms, size_t shots, uint64_t *outcomes, qi_error *err) {
    if (shots == 0) return set_err(err, QI_ERR_INVALID_NUMBER_OF_MEASUREMENTS, 0, 0, 0, 0);
    uint64_t dim = 1ULL << n;
    double *tmp = (double *)malloc(dim * sizeof(cx));
    if (!tmp) return set_err(err, QI_ERR_OUT_OF_MEMORY, 0, 0, 0, 0);
    for (size_t s = 0; s < shots; s++) {
        int st = or_measure(psi, n, basis, custom_u, qubits, m, uniforms[s], tmp, &outcomes[s], err);
        if (st) { free(tmp); return st; }
    }
    free(tmp);
    return QI_OK;
}

/* ------------------------------------------------------------- circuits -- */
int or_apply_gate(double *psi, uint32_t n, const qi_gate *g, qi_error *err) {
    int st = or_validate_gate(n, g, err);
    if (st) return st;
    switch (g->kind) {
    case QI_GATE_MEASURE: {
        uint64_t dim = 1ULL << n, o;
        double *tmp = (double *)malloc(dim * sizeof(cx));
        if (!tmp) return set_err(err, QI_ERR_OUT_OF_MEMORY, 0, 0, 0, 0);
        st = or_measure(psi, n, (int)g->basis, g->matrix, g->controls, g->ncontrols, g->params[0], tmp, &o, err);
        if (!st) memcpy(psi, tmp, dim * sizeof(cx));
        free(tmp);
        return st;
    }
    case QI_GATE_PAULI_STRING: {
        qi_pauli_term t = {g->x_mask, g->z_mask, g->params[0], g->params[1]};
        return or_pauli_apply_normalised(psi, psi, n, &t, err);
    }
    case QI_GATE_PAULI_EVOLUTION: {
        qi_pauli_term t = {g->x_mask, g->z_mask, g->params[0], g->params[1]};
        return or_pauli_exp_neg_i_dt(psi, n, &t, g->params[2], err);
    }
    default:
        apply_operator(psi, n, g);
        return QI_OK;
    }
}

int or_apply_gates(double *psi, uint32_t n_state, uint32_t n_circuit, const qi_gate *gates, size_t ng, qi_error *err) {
    if (n_state != n_circuit) return set_err(err, QI_ERR_INVALID_NUMBER_OF_QUBITS, n_state, 0, 0, 0);
    for (size_t k = 0; k < ng; k++) {
        int st = or_apply_gate(psi, n_state, &gates[k], err);
        if (st) return st;
    }
    return QI_OK;
}

/* ------------------------------------------- reference-structured baseline -- */
/* The rayon path of operator.rs, restated for timing: per gate the state is
 * cloned, pair gates build an (index, value) update list in parallel over all
 * 2^n indices (filter_map + collect) and scatter it serially, diagonal gates
 * run an in-place pass over the clone. */
typedef struct { uint64_t idx; cx val; } upd;

typedef struct {
    const cx *src; cx *dst; const qi_gate *g; uint32_t n;
    uint64_t lo, hi; upd *list; uint64_t count;
    cx ph0, ph1, m[4]; int kind;
} rf_job;

static void *rf_worker(void *arg) {
    rf_job *J = (rf_job *)arg;
    const qi_gate *g = J->g;
    uint64_t tb = 1ULL << g->targets[0];
    uint64_t cnt = 0;
    for (uint64_t i = J->lo; i < J->hi; i++) {
        switch (J->kind) {
        case 0: { /* pair gate given by m[] (general 2x2) */
            if ((i & tb) || !controls_ok(i, g->controls, g->ncontrols)) break;
            uint64_t j = i | tb;
            cx a = J->src[i], b = J->src[j];
            J->list[cnt].idx = i; J->list[cnt].val = cx_add(cx_mul(J->m[0], a), cx_mul(J->m[1], b)); cnt++;
            J->list[cnt].idx = j; J->list[cnt].val = cx_add(cx_mul(J->m[2], a), cx_mul(J->m[3], b)); cnt++;
        } break;
        case 1: { /* diagonal, in place on the clone */
            if (!controls_ok(i, g->controls, g->ncontrols)) break;
            J->dst[i] = cx_mul(J->src[i], (i & tb) ? J->ph1 : J->ph0);
        } break;
        case 2: { /* swap */
            uint64_t a = g->targets[0], b = g->targets[1];
            if (((i >> a) & 1) == ((i >> b) & 1)) break;
            uint64_t j = i ^ (1ULL << a) ^ (1ULL << b);
            if (!(i < j && controls_ok(i, g->controls, g->ncontrols))) break;
            J->list[cnt].idx = i; J->list[cnt].val = J->src[j]; cnt++;
            J->list[cnt].idx = j; J->list[cnt].val = J->src[i]; cnt++;
        } break;
        }
    }
    J->count = cnt;
    return 0;
}

int rf_apply_gates(double **psi_io, uint32_t n, const qi_gate *gates, size_t ng, int nthreads, qi_error *err) {
    uint64_t dim = 1ULL << n;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    rf_job jobs[256];
    const double s2 = 1.0 / sqrt(2.0), r2 = s2;
    for (size_t k = 0; k < ng; k++) {
        const qi_gate *g = &gates[k];
        int st = or_validate_gate(n, g, err);
        if (st) return st;
        cx *src = (cx *)*psi_io;
        cx *dst = (cx *)malloc(dim * sizeof(cx)); /* clone per gate */
        if (!dst) return set_err(err, QI_ERR_OUT_OF_MEMORY, 0, 0, 0, 0);
        memcpy(dst, src, dim * sizeof(cx));
        int kind = -1;
        cx m[4] = {{0}}, ph0 = {1, 0}, ph1 = {1, 0};
        double h = g->params[0] / 2.0;
        switch (g->kind) {
        case QI_GATE_H: kind = 0; m[0] = cx_make(s2, 0); m[1] = cx_make(s2, 0); m[2] = cx_make(s2, 0); m[3] = cx_make(-s2, 0); break;
        case QI_GATE_X: case QI_GATE_CNOT: case QI_GATE_TOFFOLI: kind = 0; m[1] = cx_make(1, 0); m[2] = cx_make(1, 0); break;
        case QI_GATE_Y: kind = 0; m[1] = cx_make(0, -1); m[2] = cx_make(0, 1); break;
        case QI_GATE_RX: kind = 0; m[0] = cx_make(cos(h), 0); m[1] = cx_make(0, -sin(h)); m[2] = cx_make(0, -sin(h)); m[3] = cx_make(cos(h), 0); break;
        case QI_GATE_RY: kind = 0; m[0] = cx_make(cos(h), 0); m[1] = cx_make(-sin(h), 0); m[2] = cx_make(sin(h), 0); m[3] = cx_make(cos(h), 0); break;
        case QI_GATE_U2: kind = 0; for (int q = 0; q < 4; q++) m[q] = cx_make(g->matrix[2 * q], g->matrix[2 * q + 1]); break;
        case QI_GATE_Z: kind = 1; ph1 = cx_make(-1, 0); break;
        case QI_GATE_S: kind = 1; ph1 = cx_make(0, 1); break;
        case QI_GATE_SDAG: kind = 1; ph1 = cx_make(0, -1); break;
        case QI_GATE_T: kind = 1; ph1 = cx_make(r2, r2); break;
        case QI_GATE_TDAG: kind = 1; ph1 = cx_make(r2, -r2); break;
        case QI_GATE_P: kind = 1; ph1 = cx_make(cos(g->params[0]), sin(g->params[0])); break;
        case QI_GATE_RZ: kind = 1; ph0 = cx_make(cos(h), -sin(h)); ph1 = cx_make(cos(h), sin(h)); break;
        case QI_GATE_SWAP: kind = 2; break;
        case QI_GATE_I: break;
        default: { /* other kinds: sequential oracle on the clone */
            st = or_apply_gate((double *)dst, n, g, err);
            if (st) { free(dst); return st; }
        } break;
        }
        if (kind >= 0) {
            uint64_t chunk = (dim + nthreads - 1) / nthreads;
            int oom = 0;
            for (int w = 0; w < nthreads; w++) {
                rf_job *J = &jobs[w];
                J->src = src; J->dst = dst; J->g = g; J->n = n; J->kind = kind;
                J->lo = (uint64_t)w * chunk; if (J->lo > dim) J->lo = dim;
                J->hi = J->lo + chunk > dim ? dim : J->lo + chunk;
                J->list = NULL; J->count = 0;
                /* each worker's share of the collected update list (at most two
                 * entries per index in its range; pages are touched only as filled) */
                if (kind != 1 && J->hi > J->lo && !(J->list = (upd *)malloc(2 * (J->hi - J->lo) * sizeof(upd)))) oom = 1;
                J->ph0 = ph0; J->ph1 = ph1; memcpy(J->m, m, sizeof m);
            }
            if (oom) {
                for (int w = 0; w < nthreads; w++) free(jobs[w].list);
                free(dst);
                return set_err(err, QI_ERR_OUT_OF_MEMORY, 0, 0, 0, 0);
            }
            for (int w = 0; w < nthreads; w++) pthread_create(&th[w], 0, rf_worker, &jobs[w]);
            for (int w = 0; w < nthreads; w++) pthread_join(th[w], 0);
            for (int w = 0; w < nthreads; w++) { /* serial scatter */
                for (uint64_t u = 0; u < jobs[w].count; u++) dst[jobs[w].list[u].idx] = jobs[w].list[u].val;
                free(jobs[w].list);
            }
        }
        free(src);
        *psi_io = (double *)dst;
    }
    return QI_OK;
}
