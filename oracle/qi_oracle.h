/*
 * qi_oracle.h — CPU FP64 restatement of quant-iron's state-vector hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the checker the engine is compared
 * against; only tests, the smoke check and the benchmark's CPU-baseline leg
 * may load it. The engine never calls it and has no CPU fallback.
 *
 * Every function restates a reference function (file:line given at each
 * declaration) with the same per-element arithmetic, in sequential index order,
 * on host arrays of interleaved (re, im) doubles. Where the reference uses a
 * rayon reduction (inner_product, histograms, norms) this oracle sums in
 * ascending index order; rayon's own order is unspecified, so results agree to
 * rounding, not to bits, for those reductions only.
 *
 * Pinning (what checks each function beyond restating it): see the header of
 * qi_oracle.c and the CPU tests under tests/.
 */
#ifndef QI_ORACLE_H
#define QI_ORACLE_H

#include "../include/qiron.h"

#ifdef __cplusplus
extern "C" {
#endif

/* validate_qubits (operator.rs:214-273) plus the per-operator checks. */
int or_validate_gate(uint32_t n, const qi_gate *g, qi_error *err);
/* Operator::apply / Gate::apply for one gate, in place. */
int or_apply_gate(double *psi, uint32_t n, const qi_gate *g, qi_error *err);
/* Circuit::execute (circuit.rs:160-172). */
int or_apply_gates(double *psi, uint32_t n_state, uint32_t n_circuit, const qi_gate *gates,
                   size_t ngates, qi_error *err);

/* State constructors (state.rs:99-300). */
int or_state_init(double *psi, uint32_t n, int kind, uint64_t basis, qi_error *err);
/* State::new's normalisation check (state.rs:99-127). */
int or_state_check_norm(const double *psi, uint32_t n, qi_error *err);

/* Reductions (state.rs:470-498, 890-945). */
void or_inner_product(const double *a, const double *b, uint32_t n, double *re, double *im);
double or_norm_sqr(const double *psi, uint32_t n);
int or_normalise(double *psi, uint32_t n, qi_error *err);
int or_fidelity(const double *a, const double *b, uint32_t n, double *out, qi_error *err);
/* fs_fidelity (dist = 0) / fs_dist (dist = 1) on states of na and nb qubits,
 * with the reference's error order (state.rs:470-498, 890-945). */
int or_fs_metric(const double *a, uint32_t na, const double *b, uint32_t nb, int dist, double *out,
                 qi_error *err);
/* State Add / Sub / Mul<Complex> (state.rs:2687-2862). */
void or_state_add(const double *a, const double *b, double *out, uint32_t n, int sub);
void or_state_scale(double *psi, uint32_t n, double re, double im);

/* PauliString / SumOp (pauli_string.rs:139-287, 453-507). */
int or_pauli_apply(const double *in, double *out, uint32_t n, const qi_pauli_term *t,
                   qi_error *err);
int or_pauli_apply_normalised(const double *in, double *out, uint32_t n, const qi_pauli_term *t,
                              qi_error *err);
int or_pauli_exp(double *psi, uint32_t n, const qi_pauli_term *t, double f_re, double f_im,
                 qi_error *err);
int or_pauli_exp_neg_i_dt(double *psi, uint32_t n, const qi_pauli_term *t, double dt,
                          qi_error *err);
int or_sumop_apply(const double *in, double *out, uint32_t n, const qi_pauli_term *terms,
                   size_t nterms, qi_error *err);
int or_sumop_expectation(const double *psi, uint32_t n, const qi_pauli_term *terms,
                         size_t nterms, double *re, double *im, qi_error *err);
/* trotter_evolve_state (time_evolution.rs:140-166). */
int or_trotter_evolve(double *psi, uint32_t n, const qi_pauli_term *terms, size_t nterms,
                      double dt, uint64_t steps, int order, qi_error *err);

/* Measurement (state.rs:525-784), uniforms injected. */
int or_measure_probs(const double *psi, uint32_t n, const uint64_t *qubits, uint32_t m,
                     double *probs, qi_error *err);
int or_measure(const double *psi, uint32_t n, int basis, const double *custom_u,
               const uint64_t *qubits, uint32_t m, double uniform, double *out_psi,
               uint64_t *outcome, qi_error *err);
int or_measure_n(const double *psi, uint32_t n, int basis, const double *custom_u,
                 const uint64_t *qubits, uint32_t m, const double *uniforms, size_t shots,
                 uint64_t *outcomes, qi_error *err);
int or_measure_collapse(const double *psi, double *out, uint32_t n, const uint64_t *qubits,
                        uint32_t m, uint64_t outcome, qi_error *err);

/* Sequential cumulative-probability rule of State::measure (state.rs:590-619):
 * index of the first bin with u < C_i, else the last. */
uint64_t or_sample_outcome(const double *probs, uint64_t nbins, double u);
/* Its cumulative probabilities C_i, the boundaries u is compared with. */
void or_sample_cdf(const double *probs, uint64_t nbins, double *cdf);

/* Reference-structured CPU path used as the timing baseline: clone per gate,
 * update list + serial scatter for pair gates, in-place pass on a clone for
 * diagonal gates, `nthreads` worker threads (the rayon path, operator.rs). */
int rf_apply_gates(double **psi_io, uint32_t n, const qi_gate *gates, size_t ngates,
                   int nthreads, qi_error *err);

#ifdef __cplusplus
}
#endif
#endif
