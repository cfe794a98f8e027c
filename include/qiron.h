/*
 * qiron.h — C ABI of the MI355X (gfx950) state-vector engine.
 *
 * This is the drop-in boundary for quant-iron's state-vector hot path. It
 * replaces the OpenCL seam `execute_on_gpu` (src/components/operator.rs:24-148)
 * and the CPU bodies behind it, and exports one entry point per public
 * operation of that path:
 *   - the `State` gate methods (src/components/state.rs:1019-2345) and the
 *     `Operator::apply` impls behind them (src/components/operator.rs:279-2285),
 *   - `Circuit::execute` / `Gate::apply` (src/circuit.rs:160-172,
 *     src/components/gate.rs:99-122),
 *   - `PauliString` / `SumOp` apply, exp and expectation
 *     (src/components/pauli_string.rs:139-287, 453-507),
 *   - `State::measure` / `measure_n` (src/components/state.rs:525-784).
 *
 * Conventions (all follow the reference):
 *   - A state is 2^n complex amplitudes, interleaved (re, im); qubit k is bit k
 *     of the amplitude index (little-endian).
 *   - Errors mirror `enum Error` (src/errors.rs:4-97): every call returns a
 *     `qi_status` and, when `err` is non-NULL, fills the variant's payload.
 *   - The reference returns a new `State` from every operation; the engine is
 *     in place on a device-resident handle. `qi_state_clone` is the explicit copy.
 *   - All calls are synchronous with respect to the host unless stated.
 *   - Calls on one context are serialised by a per-context mutex (the reference
 *     serialises all GPU work through one global mutex, operator.rs:32).
 *
 * Plain C: no C++ or torch types cross this boundary.
 */
#ifndef QIRON_H
#define QIRON_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QI_ABI_VERSION 1
#define QI_MAX_TARGETS 4
#define QI_MAX_CONTROLS 62
#define QI_MAX_QUBITS 40

/* ---- status codes: one per `enum Error` variant (src/errors.rs:4-97) ---- */
typedef enum qi_status {
    QI_OK = 0,
    QI_ERR_INVALID_NUMBER_OF_MEASUREMENTS = 1, /* InvalidNumberOfMeasurements(a)      */
    QI_ERR_OVERLAPPING_CONTROL_AND_TARGET = 2, /* OverlappingControlAndTargetQubits(a,b) */
    QI_ERR_INVALID_NUMBER_OF_QUBITS = 3,       /* InvalidNumberOfQubits(a)            */
    QI_ERR_INVALID_QUBIT_INDEX = 4,            /* InvalidQubitIndex(a, b)             */
    QI_ERR_STATE_NOT_NORMALISED = 5,           /* StateVectorNotNormalised            */
    QI_ERR_NON_UNITARY_MATRIX = 6,             /* NonUnitaryMatrix                    */
    QI_ERR_INVALID_NUMBER_OF_INPUTS = 7,       /* InvalidNumberOfInputs(a, b)         */
    QI_ERR_MISMATCHED_PARAMETERS = 8,          /* MismatchedNumberOfParameters{a,b}   */
    QI_ERR_UNKNOWN = 9,                        /* UnknownError                        */
    QI_ERR_DEVICE = 10,                        /* replaces OpenCLError(String)        */
    QI_ERR_LOCK = 11,                          /* GpuContextLockError                 */
    QI_ERR_CIRCUIT_MACRO = 12,                 /* CircuitMacroError(String)           */
    QI_ERR_INVALID_INPUT_VALUE = 13,           /* InvalidInputValue(a)                */
    QI_ERR_ZERO_NORM = 14,                     /* ZeroNorm                            */
    QI_ERR_INVALID_PAULI_COEFFICIENT = 15,     /* InvalidPauliStringCoefficient(re,im)*/
    /* engine-only codes (no reference variant) */
    QI_ERR_OUT_OF_MEMORY = 16,                 /* device allocation failed            */
    QI_ERR_INVALID_ARGUMENT = 17,              /* NULL pointer, bad enum, size limit   */
    QI_ERR_PRECISION_MISMATCH = 18             /* two states of different precision   */
} qi_status;

typedef struct qi_error {
    int32_t code;     /* qi_status                                        */
    uint64_t a, b;    /* usize payloads, in the order of the Rust variant */
    double re, im;    /* Complex payload (InvalidPauliStringCoefficient)  */
    char msg[256];    /* Display text of the variant (errors.rs #[error]) */
} qi_error;

typedef struct qi_ctx_s *qi_ctx_t;
typedef struct qi_state_s *qi_state_t;

typedef enum qi_prec { QI_C64 = 0, QI_C128 = 1 } qi_prec;

typedef enum qi_init {
    QI_INIT_ZERO = 0,  /* State::new_zero  (state.rs:165-178) */
    QI_INIT_BASIS = 1, /* State::new_basis_n (state.rs:194-210) */
    QI_INIT_PLUS = 2,  /* State::new_plus  (state.rs:225-237) */
    QI_INIT_MINUS = 3, /* State::new_minus (state.rs:240-275) */
    QI_INIT_GHZ = 4    /* State::new_ghz   (state.rs:277-300) */
} qi_init;

/* Gate kinds: one per `Operator` impl (operator.rs) plus the non-operator
 * `Gate` variants (gate.rs:13-52). */
typedef enum qi_gate_kind {
    QI_GATE_H = 0,         /* Hadamard      operator.rs:279-439  */
    QI_GATE_X = 1,         /* Pauli::X      operator.rs:443-637  */
    QI_GATE_Y = 2,         /* Pauli::Y      operator.rs:443-637  */
    QI_GATE_Z = 3,         /* Pauli::Z      operator.rs:443-637  */
    QI_GATE_I = 4,         /* Identity      operator.rs:1096-1138 */
    QI_GATE_S = 5,         /* PhaseS        operator.rs:1144-1235 */
    QI_GATE_T = 6,         /* PhaseT        operator.rs:1237-1334 */
    QI_GATE_SDAG = 7,      /* PhaseSdag     operator.rs:1336-1427 */
    QI_GATE_TDAG = 8,      /* PhaseTdag     operator.rs:1429-1526 */
    QI_GATE_P = 9,         /* PhaseShift(params[0]) operator.rs:1528-1643 */
    QI_GATE_RX = 10,       /* RotateX(params[0])  operator.rs:1645-1786 */
    QI_GATE_RY = 11,       /* RotateY(params[0])  operator.rs:1788-1927 */
    QI_GATE_RZ = 12,       /* RotateZ(params[0])  operator.rs:1929-2056 */
    QI_GATE_U2 = 13,       /* Unitary2(matrix)    operator.rs:2058-2284 */
    QI_GATE_CNOT = 14,     /* CNOT          operator.rs:643-700   */
    QI_GATE_SWAP = 15,     /* SWAP          operator.rs:706-835   */
    QI_GATE_MATCHGATE = 16,/* Matchgate(theta, phi1, phi2) operator.rs:852-1025 */
    QI_GATE_TOFFOLI = 17,  /* Toffoli       operator.rs:1031-1090 */
    QI_GATE_MEASURE = 18,  /* Gate::Measurement (gate.rs:103-105)   */
    QI_GATE_PAULI_STRING = 19,    /* Gate::PauliString -> apply_normalised (gate.rs:115-117) */
    QI_GATE_PAULI_EVOLUTION = 20, /* Gate::PauliTimeEvolution -> apply_exp_neg_i_dt (gate.rs:118-120) */
    QI_GATE_KIND_COUNT = 21
} qi_gate_kind;

typedef enum qi_basis {
    QI_BASIS_COMPUTATIONAL = 0, /* MeasurementBasis::Computational (measurement.rs:76-86) */
    QI_BASIS_X = 1,
    QI_BASIS_Y = 2,
    QI_BASIS_CUSTOM = 3         /* MeasurementBasis::Custom(matrix) */
} qi_basis;

/* One gate = one `Gate` value. `targets`/`controls` carry the index slices of
 * `Gate::Operator(op, targets, controls)`; `ntargets` may exceed what the
 * operator expects (validation then reports InvalidNumberOfQubits(ntargets),
 * as validate_qubits does, operator.rs:214-273), but at most QI_MAX_TARGETS
 * entries are read.
 *   QI_GATE_MEASURE:  controls[0..ncontrols) are the measured qubits (empty =
 *                     all), `basis`, custom basis in `matrix`, and the injected
 *                     uniform draw in params[0] (replaces rand::rng(), state.rs:601);
 *                     NaN draws from the context's stream each time the gate runs.
 *   QI_GATE_PAULI_STRING / QI_GATE_PAULI_EVOLUTION: x_mask/z_mask (Y = X&Z),
 *                     coefficient (params[0], params[1]); evolution time params[2].
 *   QI_GATE_U2:       matrix = [[m00, m01], [m10, m11]] row-major, each (re, im). */
typedef struct qi_gate {
    int32_t kind;
    uint32_t ntargets;
    uint32_t ncontrols;
    uint32_t basis;
    uint64_t targets[QI_MAX_TARGETS];
    uint64_t controls[QI_MAX_CONTROLS];
    double params[3];
    double matrix[8];
    uint64_t x_mask, z_mask;
} qi_gate;

/* PauliString as masks: qubit q carries X if x&~z, Z if z&~x, Y if x&z; the
 * product runs over distinct qubits, so the HashMap iteration order of the
 * reference (pauli_string.rs:147) cannot change the result. */
typedef struct qi_pauli_term {
    uint64_t x_mask, z_mask;
    double c_re, c_im;
} qi_pauli_term;

/* Executor options for qi_apply_gates. */
typedef enum qi_exec_mode {
    QI_EXEC_PER_GATE = 0,  /* one streaming kernel per gate                  */
    QI_EXEC_TILED = 1,     /* tile passes: the gates on a 2^k-amplitude qubit set per HBM round trip,
                              the tile held in registers (transposed through LDS) */
    QI_EXEC_FUSED_MFMA = 2,/* dense 2^k x 2^k blocks applied as complex GEMM on MFMA */
    QI_EXEC_AUTO = 3       /* tiled unless the state is too small for tiles   */
} qi_exec_mode;

typedef struct qi_exec_opts {
    int32_t mode;        /* qi_exec_mode                                    */
    int32_t tile_qubits; /* tile width k for QI_EXEC_TILED: 10..13 (c64), 10..12 (c128); 0 = 12 (c64) / 11 (c128) */
    int32_t block_qubits;/* dense block width k (<= 6) for QI_EXEC_FUSED_MFMA */
    int32_t use_graph;   /* capture the launch sequence in a hipGraph         */
    int32_t tile_threads_log2; /* QI_EXEC_TILED threads per tile: 6 (one wave; tile_qubits <= 12 for complex64,
                                  <= 11 for complex128), 8 or 9; 0 = default (6 where allowed, else 8; 9 for
                                  complex64 13 / complex128 12 tile qubits) */
} qi_exec_opts;

/* ---------------- context ---------------- */
/* Replaces GpuContext::new + the global Lazy<Mutex<..>> (gpu_context.rs:94-168). */
qi_status qi_ctx_create(int device, qi_ctx_t *out, qi_error *err);
qi_status qi_ctx_destroy(qi_ctx_t ctx);
qi_status qi_ctx_synchronize(qi_ctx_t ctx, qi_error *err);
/* Seeds the context's SplitMix64 stream of uniform draws, used by
 * measurements whose draw is not injected (a NaN uniform, NULL uniforms).
 * Replaces the OS-seeded rand::rng() of State::measure (state.rs:601-602);
 * an unseeded context starts from an OS-seeded state. */
qi_status qi_ctx_seed(qi_ctx_t ctx, uint64_t seed, qi_error *err);
/* Device name and HBM size, for reports. */
qi_status qi_ctx_info(qi_ctx_t ctx, char *name, size_t name_len, uint64_t *hbm_bytes,
                      int *cu_count, qi_error *err);

/* ---------------- states ---------------- */
/* Allocates 2^n amplitudes (8 or 16 bytes each) on the context's device. */
qi_status qi_state_create(qi_ctx_t ctx, uint32_t n, qi_prec prec, qi_state_t *out, qi_error *err);
qi_status qi_state_destroy(qi_state_t s);
/* `State::clone` (derive(Clone), state.rs:74). */
qi_status qi_state_clone(qi_state_t src, qi_state_t *out, qi_error *err);
/* dst <- src (same n and precision). */
qi_status qi_state_copy(qi_state_t dst, qi_state_t src, qi_error *err);
uint32_t qi_state_num_qubits(qi_state_t s);
int32_t qi_state_precision(qi_state_t s);
/* Device address of the amplitude array (interop; owned by the handle). */
void *qi_state_device_ptr(qi_state_t s);
/* Constructors: new_zero / new_basis_n / new_plus / new_minus / new_ghz
 * (state.rs:165-300); `basis` is used by QI_INIT_BASIS only. */
qi_status qi_state_init(qi_state_t s, qi_init kind, uint64_t basis, qi_error *err);
/* Host -> device from 2^(n+1) doubles. check_norm != 0 applies `State::new`'s
 * normalisation check (|sum|a|^2 - 1| <= f64::EPSILON * len, state.rs:99-127). */
qi_status qi_state_upload(qi_state_t s, const double *amps, int check_norm, qi_error *err);
/* Device -> host into 2^(n+1) doubles (c64 amplitudes are widened). */
qi_status qi_state_download(qi_state_t s, double *amps, qi_error *err);

/* ---------------- gates ---------------- */
/* `Operator::apply` for one gate, in place (operator.rs:151-190 and the impls
 * it dispatches to; also `Gate::apply`, gate.rs:99-122, for the non-operator
 * kinds). Validation order and payloads follow validate_qubits
 * (operator.rs:214-273) and the per-gate checks (CNOT :674-679, Toffoli
 * :1064-1072, Matchgate :903-905, Unitary2::new :2092-2118). */
qi_status qi_apply_gate(qi_state_t s, const qi_gate *g, qi_error *err);
/* `Circuit::execute` (circuit.rs:160-172): all gates in list order, in place,
 * no per-gate host synchronisation. `n_circuit` is the circuit's qubit count
 * (InvalidNumberOfQubits(n_state) when it differs). opts may be NULL
 * (= QI_EXEC_AUTO). The list is validated before any amplitude changes. */
qi_status qi_apply_gates(qi_state_t s, uint32_t n_circuit, const qi_gate *gates, size_t ngates,
                         const qi_exec_opts *opts, qi_error *err);

/* ---------------- Pauli algebra ---------------- */
/* out <- PauliString::apply(in) (pauli_string.rs:139-151). in == out allowed. */
qi_status qi_pauli_apply(qi_state_t in, qi_state_t out, const qi_pauli_term *t, qi_error *err);
/* out <- PauliString::apply_normalised(in) (pauli_string.rs:165-168). */
qi_status qi_pauli_apply_normalised(qi_state_t in, qi_state_t out, const qi_pauli_term *t,
                                    qi_error *err);
/* s <- PauliString::apply_exp_factor(s, factor) (pauli_string.rs:237-262);
 * factor = 1 gives apply_exp (:199-223). */
qi_status qi_pauli_exp(qi_state_t s, const qi_pauli_term *t, double factor_re, double factor_im,
                       qi_error *err);
/* s <- PauliString::apply_exp_neg_i_dt(s, dt) (pauli_string.rs:281-287). */
qi_status qi_pauli_exp_neg_i_dt(qi_state_t s, const qi_pauli_term *t, double dt, qi_error *err);
/* out <- SumOp::apply(in) (pauli_string.rs:453-466). in != out. */
qi_status qi_sumop_apply(qi_state_t in, qi_state_t out, const qi_pauli_term *terms, size_t nterms,
                         qi_error *err);
/* SumOp::expectation_value (pauli_string.rs:485-507): FP64 compensated sums. */
qi_status qi_sumop_expectation(qi_state_t s, const qi_pauli_term *terms, size_t nterms,
                               double *re, double *im, qi_error *err);
/* qi_sumop_expectation with no allocation of any kind: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid it). The term
 * table, block partials and group results live in a fixed static device
 * array (one per device, 76 KiB, held until the process ends) instead
 * of the context's growable buffers. At most 512 terms (more: InvalidArgument).
 * Same kernels, same summation order and the same result bits as
 * qi_sumop_expectation. Calls on one context are serialised by its lock;
 * calls that overlap on one device from different contexts race on the
 * array. */
qi_status qi_sumop_expectation_noalloc(qi_state_t s, const qi_pauli_term *terms, size_t nterms,
                                       double *re, double *im, qi_error *err);
/* trotter_evolve_state (time_evolution.rs:140-166); order 1 = First, 2 = Second. */
qi_status qi_trotter_evolve(qi_state_t s, const qi_pauli_term *terms, size_t nterms, double dt,
                            uint64_t steps, int order, qi_error *err);

/* ---------------- reductions and arithmetic (FP64 accumulation) ---------------- */
/* State::inner_product (state.rs:890-917): sum conj(a_i) b_i. */
qi_status qi_inner_product(qi_state_t a, qi_state_t b, double *re, double *im, qi_error *err);
/* sum |a_i|^2 */
qi_status qi_norm_sqr(qi_state_t s, double *out, qi_error *err);
/* State::normalise (state.rs:924-945), in place. */
qi_status qi_normalise(qi_state_t s, qi_error *err);
/* State::fs_fidelity (state.rs:492-498): |<a/|a| | b/|b|>|^2. Errors in the
 * reference's order: ZeroNorm for a, then for b (normalise, state.rs:924-935),
 * then InvalidNumberOfQubits(n_a) when the qubit counts differ
 * (inner_product, state.rs:890-897). */
qi_status qi_fidelity(qi_state_t a, qi_state_t b, double *out, qi_error *err);
/* State::fs_dist (state.rs:470-478): acos(|<a/|a| | b/|b|>|), not clamped
 * (as in the reference); errors as qi_fidelity. */
qi_status qi_fs_dist(qi_state_t a, qi_state_t b, double *out, qi_error *err);
/* State * Complex (state.rs:2687-2705), in place. */
qi_status qi_state_scale(qi_state_t s, double re, double im, qi_error *err);
/* y <- y + alpha * x (a scaled State Add, state.rs:2773-2800). */
qi_status qi_state_axpy(qi_state_t y, double alpha_re, double alpha_im, qi_state_t x,
                        qi_error *err);
/* out <- a + b (impl Add for State, state.rs:2773-2800) and out <- a - b
 * (impl Sub, state.rs:2836-2862), elementwise with no norm check; out may be
 * a or b. A qubit-count mismatch (a panic in the reference) returns
 * InvalidNumberOfQubits(n_a). */
qi_status qi_state_add(qi_state_t a, qi_state_t b, qi_state_t out, qi_error *err);
qi_status qi_state_sub(qi_state_t a, qi_state_t b, qi_state_t out, qi_error *err);

/* ---------------- measurement ---------------- */
/* Outcome probabilities over `qubits` (bit b of the outcome = qubit qubits[b]),
 * the histogram of State::measure (state.rs:559-588), FP64, 2^m entries. */
qi_status qi_measure_probs(qi_state_t s, const uint64_t *qubits, uint32_t m, double *probs,
                           qi_error *err);
/* State::measure (state.rs:525-730) with the uniform draw injected:
 * outcome = first i with u < C_i (sequential cumsum of normalised
 * probabilities), else the last bin. The collapsed state is written to
 * `out_state` (may equal s); `*outcome` holds the outcome bits. A NaN
 * `uniform` is drawn from the context's stream (qi_ctx_seed). */
qi_status qi_measure(qi_state_t s, qi_basis basis, const double *custom_u, const uint64_t *qubits,
                     uint32_t m, double uniform, qi_state_t out_state, uint64_t *outcome,
                     qi_error *err);
/* State::measure_n (state.rs:750-784): `shots` independent measurements of the
 * same state with injected uniforms (NULL: `shots` draws of the context's
 * stream); outcomes only. Shot k's MeasurementResult::new_state is
 * qi_measure_result_state of outcomes[k]. */
qi_status qi_measure_n(qi_state_t s, qi_basis basis, const double *custom_u,
                       const uint64_t *qubits, uint32_t m, const double *uniforms, size_t shots,
                       uint64_t *outcomes, qi_error *err);
/* Collapse onto a given computational outcome and renormalise (state.rs:622-654). */
qi_status qi_measure_collapse(qi_state_t src, qi_state_t dst, const uint64_t *qubits, uint32_t m,
                              uint64_t outcome, qi_error *err);
/* MeasurementResult::new_state (measurement.rs:15-72) of a measurement of
 * `src` in `basis` that gave `outcome`: the basis change, the collapse onto
 * the outcome with renormalisation and State::new's check, and the basis
 * change back (state.rs:622-728). dst may be src. */
qi_status qi_measure_result_state(qi_state_t src, qi_basis basis, const double *custom_u,
                                  const uint64_t *qubits, uint32_t m, uint64_t outcome,
                                  qi_state_t dst, qi_error *err);

/* ---------------- parametric circuits ---------------- */
/* A circuit whose RX / RY / RZ / P gates may read their angle from an angle
 * table: Gate::Parametric (gate.rs:107-114) over shared Parameter<N> values
 * (parameter.rs:12-68), for loops (VQE) that re-run one circuit with new
 * parameters. The circuit is validated and planned once (tiled or per gate,
 * as qi_apply_gates; opts may be NULL); angle_slots[g] >= 0 makes gate g read
 * table entry angle_slots[g] (NULL: no parametric gates). The table starts
 * from the gates' own angles. qi_param_circuit_set_angles replaces the whole
 * table (MismatchedNumberOfParameters{expected, given} on a length
 * mismatch); qi_param_circuit_run applies the circuit to a state of the same
 * qubit count and precision in place (Circuit::execute, circuit.rs:160-172),
 * replaying one captured launch sequence per state. Operator gates only. */
typedef struct qi_param_circuit_s *qi_param_circuit_t;
qi_status qi_param_circuit_create(qi_ctx_t ctx, uint32_t n, qi_prec prec, const qi_gate *gates, size_t ngates,
                                  const int32_t *angle_slots, const qi_exec_opts *opts, qi_param_circuit_t *out,
                                  qi_error *err);
qi_status qi_param_circuit_destroy(qi_param_circuit_t pc);
qi_status qi_param_circuit_num_slots(qi_param_circuit_t pc, uint64_t *nslots, qi_error *err);
qi_status qi_param_circuit_set_angles(qi_param_circuit_t pc, const double *angles, size_t nangles, qi_error *err);
qi_status qi_param_circuit_run(qi_param_circuit_t pc, qi_state_t s, qi_error *err);

/* ---------------- OpenQASM 3.0 input ---------------- */
/* Reads an OpenQASM 3.0 program into gate records: the programs written by
 * Circuit::to_qasm (src/circuit.rs:244-278, src/compiler/qasm.rs,
 * src/compiler/ir.rs) and the same gate set in standard syntax (stdgates.inc
 * gates, ctrl / negctrl / inv modifiers, register broadcasting, measure in the
 * computational, X (xmeasure) and Y (ymeasure) bases). The reference exports
 * QASM but has no reader; this is its input direction, for Circuit::execute
 * (qi_apply_gates) and qi_batch_add. Qubits are numbered across the qubit
 * registers in declaration order. Measurements become QI_GATE_MEASURE records
 * whose uniform is drawn when they run (NaN). `input float`/`input angle`
 * declarations are circuit parameters, numbered in declaration order: a gate
 * whose whole angle is parameter k gets angle_slots[g] = k and params[0] =
 * param_values[k] (0 when param_values is NULL or shorter); every other gate
 * gets -1. With gates == NULL only *ngates, *nqubits and *nparams are
 * returned (the size query). Errors: QI_ERR_INVALID_INPUT_VALUE with the line
 * number in err->a and the reason in err->msg. */
qi_status qi_qasm_parse(const char *src, size_t len, const double *param_values, size_t nvalues,
                        qi_gate *gates, int32_t *angle_slots, size_t capacity, size_t *ngates,
                        uint32_t *nqubits, uint32_t *nparams, qi_error *err);

/* ---------------- batched circuits ---------------- */
/* A batch of parametrised circuits of uneven qubit counts, each run from
 * |0..0> (State::new_zero, state.rs:165-178) through Circuit::execute
 * (circuit.rs:160-172) and reduced to SumOp::expectation_value
 * (pauli_string.rs:485-507) of its own Hamiltonian. The whole batch is one
 * hipGraph launch. No reference counterpart as a single call: it is the
 * reference's per-circuit path applied to many circuits at once. */
typedef struct qi_batch_s *qi_batch_t;
qi_status qi_batch_create(qi_ctx_t ctx, qi_prec prec, qi_batch_t *out, qi_error *err);
qi_status qi_batch_destroy(qi_batch_t b);
/* Adds one circuit (operator gates only, validated as qi_apply_gate does).
 * angle_slots[g] >= 0 makes gate g (RX/RY/RZ/P) take its angle from slot
 * angle_slots[g] of the batch's angle table instead of params[0]; NULL = none.
 * Each term's X/Y support must fit one LDS tile (at most 9 qubits outside the
 * 4 lowest for complex64, 8 outside the 3 lowest for complex128). */
qi_status qi_batch_add(qi_batch_t b, uint32_t n, const qi_gate *gates, size_t ngates,
                       const int32_t *angle_slots, const qi_pauli_term *terms, size_t nterms,
                       qi_error *err);
/* Replaces the angle table (one upload; the captured graph is reused). */
qi_status qi_batch_set_angles(qi_batch_t b, const double *angles, size_t nangles, qi_error *err);
/* Replaces every circuit's term coefficients: (re, im) pairs in the order the
 * terms were added, circuit by circuit (no replanning). */
qi_status qi_batch_set_term_coeffs(qi_batch_t b, const double *coeffs, size_t nterms_total,
                                   qi_error *err);
/* Runs the batch; energies[2c], energies[2c+1] = Re, Im of circuit c's energy. */
qi_status qi_batch_run(qi_batch_t b, double *energies, int use_graph, qi_error *err);
qi_status qi_batch_info(qi_batch_t b, uint64_t *ncircuits, uint64_t *nsmall, uint64_t *nlevels,
                        double *last_ms, qi_error *err);
/* Schedule of the batch: circuits with n <= ks qubits run LDS-resident (one
 * workgroup each); larger ones run tiled passes of 2^k amplitudes.
 * 10 <= k <= ks <= 13 (complex64) / 12 (complex128). Default k = ks = 13 / 12. */
qi_status qi_batch_set_schedule(qi_batch_t b, uint32_t k, uint32_t ks, qi_error *err);
qi_status qi_batch_get_schedule(qi_batch_t b, uint32_t *k, uint32_t *ks, qi_error *err);
/* Executor of the circuits larger than ks qubits. QI_BATCH_WAVE (default):
 * circuits of 12+ (complex64) / 11+ (complex128) qubits run the register-
 * resident one-wave passes of the tiled executor (a wave holds a 2^12 / 2^11
 * amplitude tile; pass p of every circuit in one launch), the others LDS
 * tiles of width k. QI_BATCH_LDS: every one of them runs LDS tiles of width k.
 * No reference counterpart. */
typedef enum qi_batch_exec { QI_BATCH_LDS = 0, QI_BATCH_WAVE = 1 } qi_batch_exec;
qi_status qi_batch_set_executor(qi_batch_t b, qi_batch_exec exec, qi_error *err);
qi_status qi_batch_get_executor(qi_batch_t b, qi_batch_exec *exec, qi_error *err);
/* Autotuner: plans, captures and times every schedule on this batch (both
 * executors, every k and ks; best of `trials` replays after a warm-up, HIP-event
 * device time), keeps the fastest and leaves the batch planned with it (the
 * executor it picked: qi_batch_get_executor). No reference counterpart. */
qi_status qi_batch_autotune(qi_batch_t b, uint32_t trials, uint32_t *k, uint32_t *ks, double *best_ms,
                            qi_error *err);

/* ---------------- energy surrogate (MLP) ---------------- */
/* An MLP n_in -> H -> H -> H -> 1 with GELU, BF16 weights on the matrix
 * cores with FP32 accumulation and FP32 Adam masters. It screens ansatz
 * parameters; exact simulation (qi_batch_*) stays the judge of every value.
 * No counterpart in the reference. Parameter layout for get/set/gradients:
 * W1 [k_in x H] (k_in = n_in rounded up to 16; rows >= n_in are zero), b1 [H],
 * W2 [H x H], b2, W3 [H x H], b3, w4 [H], b4 [1]; row-major. */
typedef struct qi_mlp_s *qi_mlp_t;
qi_status qi_mlp_create(qi_ctx_t ctx, uint32_t n_in, uint32_t hidden, uint64_t seed, qi_mlp_t *out,
                        qi_error *err);
qi_status qi_mlp_destroy(qi_mlp_t m);
qi_status qi_mlp_num_params(qi_mlp_t m, uint64_t *n, qi_error *err);
qi_status qi_mlp_get_params(qi_mlp_t m, float *out, qi_error *err);
qi_status qi_mlp_set_params(qi_mlp_t m, const float *in, qi_error *err);
/* Gradient of the mean squared error over rows samples (rows % 16 == 0). */
qi_status qi_mlp_gradients(qi_mlp_t m, const float *x, const float *t, size_t rows, float *grads,
                           float *loss, qi_error *err);
/* Uploads a training set once; qi_mlp_train_epoch runs Adam over shuffled
 * mini-batches of `batch` rows (batch % 16 == 0) with no host round trip
 * inside the epoch, and returns the epoch's mean squared error. */
qi_status qi_mlp_set_data(qi_mlp_t m, const float *x, const float *t, size_t rows, qi_error *err);
qi_status qi_mlp_train_epoch(qi_mlp_t m, uint32_t batch, float lr, uint64_t shuffle_seed,
                             float *mean_loss, qi_error *err);
qi_status qi_mlp_predict(qi_mlp_t m, const float *x, size_t rows, float *y, qi_error *err);
/* Predictions for `count` ansatz candidates generated on the device (n_in must
 * be 2 nang + 3). Candidate c's angles are SplitMix64 draws
 * draw0 + c*nang + j + 1 of the stream `seed` (theta_j = 2 pi u), its
 * couplings the draws draw0 + count*nang + 3c + i + 1 (J_i = 0.5 + u); the
 * inputs are (cos theta_j, sin theta_j, J_i - 1). The caller regenerates the
 * parameters of the candidates it keeps from their indices. */
qi_status qi_mlp_predict_ansatz(qi_mlp_t m, uint64_t seed, uint64_t draw0, uint64_t count, uint32_t nang,
                                float *y, qi_error *err);

/* Predictions for `count` candidates of a local optimisation round: angles
 * theta_j = theta0[j] + sigma z_j with z_2p = rho cos(2 pi u2), z_2p+1 =
 * rho sin(2 pi u2), rho = sqrt(-2 ln(1 - u1)), u1, u2 the SplitMix64 draws
 * d + 1 and d + 2 of the stream `seed`, d = draw0 + 2 (c P + p),
 * P = ceil(nang / 2) (one Box-Muller pair per two angles); couplings J[0..2]
 * fixed for the round. */
qi_status qi_mlp_predict_ansatz_local(qi_mlp_t m, uint64_t seed, uint64_t draw0, uint64_t count, uint32_t nang,
                                      const double *theta0, double sigma, const double *J, float *y,
                                      qi_error *err);

/* ---------------- timing hooks (for benchmarks) ---------------- */
/* Device time of the last qi_apply_gates call's launch sequence, from HIP
 * events recorded on the context's stream (milliseconds). */
qi_status qi_last_exec_ms(qi_ctx_t ctx, double *ms, qi_error *err);
/* Number of launch steps (kernels or host-orchestrated gates) the last
 * qi_apply_gates call planned. */
qi_status qi_last_exec_steps(qi_ctx_t ctx, uint64_t *steps, qi_error *err);

#ifdef __cplusplus
}
#endif
#endif /* QIRON_H */
