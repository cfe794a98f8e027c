// api.hip — the C ABI (include/qiron.h): validation, error payloads, state
// handles and the host-side orchestration of the kernels. Every entry point
// locks its context's mutex and runs on the context's single in-order stream.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <random>

#include "qi_internal.h"
#include "exec_plan.h"

namespace qi {


void fill_err(qi_error *err, const Status &st) {
    if (!err) return;
    memset(err, 0, sizeof(*err));
    err->code = st.code; err->a = st.a; err->b = st.b; err->re = st.re; err->im = st.im;
    snprintf(err->msg, sizeof err->msg, "%s", st.msg.c_str());
}

Status ensure(DevBuf &b, size_t bytes) {
    if (b.bytes >= bytes && b.p) return Status{};
    if (b.p) {
        QI_HIP_TRY(hipDeviceSynchronize());  // earlier launches may still read it
        QI_HIP_TRY(hipFree(b.p));
        b.p = nullptr; b.bytes = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) { b.p = nullptr; return oom_err(e); }
    b.bytes = want;
    return Status{};
}

// ------------------------------------------------------------- validation --
// validate_qubits, operator.rs:214-273.
static Status validate_qubits(uint32_t n, const uint64_t *t, uint32_t nt, const uint64_t *c, uint32_t nc, uint32_t expected) {
    if (nt != expected) return make_err(QI_ERR_INVALID_NUMBER_OF_QUBITS, nt);
    for (uint32_t i = 0; i < nt; i++)
        if (t[i] >= n) return make_err(QI_ERR_INVALID_QUBIT_INDEX, t[i], n);
    for (uint32_t j = 0; j < nc; j++) {
        if (c[j] >= n) return make_err(QI_ERR_INVALID_QUBIT_INDEX, c[j], n);
        for (uint32_t i = 0; i < nt; i++)
            if (c[j] == t[i]) return make_err(QI_ERR_OVERLAPPING_CONTROL_AND_TARGET, c[j], t[i]);
    }
    if (expected > 1)
        for (uint32_t i = 0; i < nt; i++)
            for (uint32_t j = i + 1; j < nt; j++)
                if (t[i] == t[j]) return make_err(QI_ERR_INVALID_QUBIT_INDEX, t[i], n);
    return Status{};
}

// Unitary2::new, operator.rs:2092-2118 (tolerance 2 * f64::EPSILON).
static Status check_unitary(const double *m) {
    const double tol = DBL_EPSILON * 2.0;
    double ar = m[0], ai = m[1], br = m[2], bi = m[3], cr = m[4], ci = m[5], dr = m[6], di = m[7];
    double n0 = (ar * ar + ai * ai) + (br * br + bi * bi);
    double n1 = (cr * cr + ci * ci) + (dr * dr + di * di);
    if (std::fabs(n0 - 1.0) > tol || std::fabs(n1 - 1.0) > tol) return make_err(QI_ERR_NON_UNITARY_MATRIX);
    // a * conj(c) + b * conj(d)
    double xr = (ar * cr - ai * -ci) + (br * dr - bi * -di);
    double xi = (ar * -ci + ai * cr) + (br * -di + bi * dr);
    if (xr * xr + xi * xi > tol * tol) return make_err(QI_ERR_NON_UNITARY_MATRIX);
    return Status{};
}

static Status check_mask(uint32_t n, uint64_t mask) {
    uint64_t bad = n >= 64 ? 0 : mask >> n;
    if (bad) return make_err(QI_ERR_INVALID_QUBIT_INDEX, n + (uint64_t)__builtin_ctzll(bad), n);
    return Status{};
}

// State::measure argument checks, state.rs:525-557.
static Status check_measure(uint32_t n, const uint64_t *q, uint32_t m) {
    if (m > n) return make_err(QI_ERR_INVALID_NUMBER_OF_QUBITS, n);
    for (uint32_t i = 0; i < m; i++)
        if (q[i] >= n) return make_err(QI_ERR_INVALID_QUBIT_INDEX, q[i], n);
    return Status{};
}

Status validate_gate(uint32_t n, const qi_gate &g) {
    uint32_t nt = g.ntargets, nc = g.ncontrols;
    if (nc > QI_MAX_CONTROLS) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "too many controls");
    Status s;
    switch (g.kind) {
    case QI_GATE_H: case QI_GATE_X: case QI_GATE_Y: case QI_GATE_Z: case QI_GATE_I: case QI_GATE_S:
    case QI_GATE_T: case QI_GATE_SDAG: case QI_GATE_TDAG: case QI_GATE_P: case QI_GATE_RX: case QI_GATE_RY:
    case QI_GATE_RZ:
        return validate_qubits(n, g.targets, nt, g.controls, nc, 1);
    case QI_GATE_U2:
        if (!(s = check_unitary(g.matrix)).ok()) return s;
        return validate_qubits(n, g.targets, nt, g.controls, nc, 1);
    case QI_GATE_CNOT:  // operator.rs:674-679
        if (!(s = validate_qubits(n, g.targets, nt, g.controls, nc, 1)).ok()) return s;
        if (nc != 1) return make_err(QI_ERR_INVALID_NUMBER_OF_QUBITS, nc);
        return Status{};
    case QI_GATE_TOFFOLI:  // operator.rs:1062-1072
        if (!(s = validate_qubits(n, g.targets, nt, g.controls, nc, 1)).ok()) return s;
        if (nc != 2 || g.controls[0] == g.controls[1]) return make_err(QI_ERR_INVALID_NUMBER_OF_QUBITS, nc);
        return Status{};
    case QI_GATE_SWAP:
        return validate_qubits(n, g.targets, nt, g.controls, nc, 2);
    case QI_GATE_MATCHGATE:  // operator.rs:900-905
        if (!(s = validate_qubits(n, g.targets, nt, g.controls, nc, 1)).ok()) return s;
        if (g.targets[0] == uint64_t(n) - 1) return make_err(QI_ERR_INVALID_QUBIT_INDEX, g.targets[0], n);
        return Status{};
    case QI_GATE_MEASURE:
        if (g.basis > QI_BASIS_CUSTOM) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "measurement basis");
        if (g.basis == QI_BASIS_CUSTOM && !(s = check_unitary(g.matrix)).ok()) return s;
        // a NaN uniform is drawn from the context's stream when the gate runs
        if (!std::isnan(g.params[0]) && !(g.params[0] >= 0.0 && g.params[0] < 1.0))
            return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "uniform draw outside [0,1)");
        return check_measure(n, g.controls, nc);
    case QI_GATE_PAULI_STRING:
        return check_mask(n, g.x_mask | g.z_mask);
    case QI_GATE_PAULI_EVOLUTION:  // pauli_string.rs:281-287
        if (g.params[1] != 0.0) return make_err(QI_ERR_INVALID_PAULI_COEFFICIENT, 0, 0, g.params[0], g.params[1]);
        return check_mask(n, g.x_mask | g.z_mask);
    default:
        return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "unknown gate kind");
    }
}

static bool is_operator(int kind) { return kind >= QI_GATE_H && kind <= QI_GATE_TOFFOLI; }

// ------------------------------------------------------------ complex math --
struct cxd { double re, im; };
static cxd cmul(cxd a, cxd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
// num-complex exp / cosh / sinh (from_polar form)
static cxd cexp_(cxd z) { double e = std::exp(z.re); return {e * std::cos(z.im), e * std::sin(z.im)}; }
static cxd ccosh_(cxd z) { return {std::cosh(z.re) * std::cos(z.im), std::sinh(z.re) * std::sin(z.im)}; }
static cxd csinh_(cxd z) { return {std::sinh(z.re) * std::cos(z.im), std::cosh(z.re) * std::sin(z.im)}; }

// ---------------------------------------------------------------- helpers --
static Status norm_sqr(State &s, double *out) {
    double r[2];
    Status e = reduce_state(s, nullptr, RK_NORM, r, s.ctx->stream);
    *out = r[0];
    return e;
}

// State::normalise, state.rs:924-945 (norm in FP64 on the device).
static Status normalise(State &s) {
    double ns;
    Status e = norm_sqr(s, &ns);
    if (!e.ok()) return e;
    double norm = std::sqrt(ns);
    if (norm == 0.0) return make_err(QI_ERR_ZERO_NORM);
    if (norm == 1.0) return Status{};
    return launch_div_real(s, norm, s.ctx->stream);
}

// State::new's check (state.rs:117-122) on a device state: |sum |a|^2 - 1|
// <= tol. complex128 uses the reference's tolerance, f64::EPSILON * len.
// complex64 amplitudes that were just renormalised carry one FP32 rounding
// of the quotient and one of the divisor, so sum |a|^2 (summed in FP64) is
// within about 2 f32::EPSILON of 1 at any size; the bound 16 f32::EPSILON
// does not grow with the state, so a collapse onto an outcome of zero
// probability (an all-zero state) is rejected at every size.
static Status check_normalised(State &s) {
    double ns;
    Status e = norm_sqr(s, &ns);
    if (!e.ok()) return e;
    double tol = s.prec == QI_C128 ? DBL_EPSILON * double(uint64_t(1) << s.n) : 16.0 * (double)FLT_EPSILON;
    if (std::fabs(ns - 1.0) > tol) return make_err(QI_ERR_STATE_NOT_NORMALISED);
    return Status{};
}

static Status alloc_state(Ctx &C, uint32_t n, int prec, State &out) {
    out.ctx = &C; out.n = n; out.prec = prec; out.d = nullptr;
    hipError_t e = hipMalloc(&out.d, out.bytes());
    if (e != hipSuccess) { out.d = nullptr; return oom_err(e); }
    return Status{};
}

struct TmpState {
    State s;
    ~TmpState() { if (s.d) (void)hipFree(s.d); }
};

static Status apply_1q_list(State &s, int kind, const uint64_t *q, uint32_t m, const double *mat) {
    for (uint32_t k = 0; k < m; k++) {
        qi_gate g{};
        g.kind = kind; g.ntargets = 1; g.targets[0] = q[k];
        if (mat) memcpy(g.matrix, mat, sizeof g.matrix);
        Lowered L;
        Status e = lower_gate(s.n, g, L);
        if (!e.ok()) return e;
        if (!(e = launch_gate(s, L, s.ctx->stream)).ok()) return e;
    }
    return Status{};
}

// The basis change before and after a computational measurement
// (state.rs:670-728).
static Status basis_in(State &w, int basis, const double *u, const uint64_t *q, uint32_t m) {
    if (basis == QI_BASIS_X) return apply_1q_list(w, QI_GATE_H, q, m, nullptr);
    if (basis == QI_BASIS_Y) {
        Status e = apply_1q_list(w, QI_GATE_SDAG, q, m, nullptr);
        if (!e.ok()) return e;
        return apply_1q_list(w, QI_GATE_H, q, m, nullptr);
    }
    if (basis == QI_BASIS_CUSTOM) return apply_1q_list(w, QI_GATE_U2, q, m, u);
    return Status{};
}
static Status basis_out(State &w, int basis, const double *u, const uint64_t *q, uint32_t m) {
    if (basis == QI_BASIS_X) return apply_1q_list(w, QI_GATE_H, q, m, nullptr);
    if (basis == QI_BASIS_Y) {
        Status e = apply_1q_list(w, QI_GATE_H, q, m, nullptr);
        if (!e.ok()) return e;
        return apply_1q_list(w, QI_GATE_S, q, m, nullptr);
    }
    if (basis == QI_BASIS_CUSTOM) {  // calculate_adjoint: conjugate transpose
        double ud[8] = {u[0], -u[1], u[4], -u[5], u[2], -u[3], u[6], -u[7]};
        Status e = check_unitary(ud);
        if (!e.ok()) return e;
        return apply_1q_list(w, QI_GATE_U2, q, m, ud);
    }
    return Status{};
}

// State::measure with an injected uniform; `dst` may be `s`.
static Status measure_one(State &s, int basis, const double *u, const uint64_t *qubits, uint32_t m, double uniform,
                          State &dst, uint64_t *outcome) {
    std::vector<uint64_t> all;
    if (m == 0) { for (uint32_t i = 0; i < s.n; i++) all.push_back(i); qubits = all.data(); m = s.n; }
    Status e = check_measure(s.n, qubits, m);
    if (!e.ok()) return e;
    if (basis == QI_BASIS_CUSTOM && !(e = check_unitary(u)).ok()) return e;
    if (std::isnan(uniform)) uniform = splitmix_uniform(s.ctx->rng);  // rand::rng(), state.rs:601-602
    if (!(uniform >= 0.0 && uniform < 1.0)) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "uniform draw outside [0,1)");
    hipStream_t st = s.ctx->stream;
    if (dst.d != s.d && !(e = launch_copy(dst, s, st)).ok()) return e;
    if (!(e = basis_in(dst, basis, u, qubits, m)).ok()) return e;
    if (!(e = measure_sample_collapse(dst, dst, qubits, m, &uniform, 1, outcome, true, st)).ok()) return e;
    if (!(e = check_normalised(dst)).ok()) return e;
    return basis_out(dst, basis, u, qubits, m);
}

// MeasurementResult::new_state of a given outcome (state.rs:622-728): the
// basis change, the collapse onto `outcome` with renormalisation and
// State::new's check, and the basis change back. `dst` may be `s`.
static Status measured_state(State &s, int basis, const double *u, const uint64_t *qubits, uint32_t m,
                             uint64_t outcome, State &dst) {
    hipStream_t st = s.ctx->stream;
    Status e;
    if (dst.d != s.d && !(e = launch_copy(dst, s, st)).ok()) return e;
    if (!(e = basis_in(dst, basis, u, qubits, m)).ok()) return e;
    if (!(e = collapse_outcome(dst, dst, qubits, m, outcome, st)).ok()) return e;
    if (!(e = check_normalised(dst)).ok()) return e;
    return basis_out(dst, basis, u, qubits, m);
}

// apply_exp_factor (pauli_string.rs:237-262) on the device.
static Status pauli_exp(State &s, const qi_pauli_term &t, cxd factor) {
    cxd alpha = cmul(cxd{t.c_re, t.c_im}, factor);
    if ((t.x_mask | t.z_mask) == 0) {
        cxd e = cexp_(alpha);
        return launch_scale(s, e.re, e.im, s.ctx->stream);
    }
    Status e = check_mask(s.n, t.x_mask | t.z_mask);
    if (!e.ok()) return e;
    cxd ch = ccosh_(alpha), sh = csinh_(alpha);
    return launch_pauli_exp(s, t, ch.re, ch.im, sh.re, sh.im, s.ctx->stream);
}

static Status pauli_apply_normalised(State &in, State &out, const qi_pauli_term &t) {
    Status e = check_mask(in.n, t.x_mask | t.z_mask);
    if (!e.ok()) return e;
    qi_pauli_term u = t;
    u.c_re = 1.0; u.c_im = 0.0;  // apply_operators drops the coefficient
    if (!(e = launch_pauli_apply(in, out, u, in.ctx->stream)).ok()) return e;
    return normalise(out);
}

// Gate::apply for one validated gate.
static Status apply_validated(State &s, const qi_gate &g) {
    switch (g.kind) {
    case QI_GATE_MEASURE: {
        uint64_t o;
        return measure_one(s, (int)g.basis, g.matrix, g.controls, g.ncontrols, g.params[0], s, &o);
    }
    case QI_GATE_PAULI_STRING: {
        qi_pauli_term t{g.x_mask, g.z_mask, g.params[0], g.params[1]};
        return pauli_apply_normalised(s, s, t);
    }
    case QI_GATE_PAULI_EVOLUTION: {
        qi_pauli_term t{g.x_mask, g.z_mask, g.params[0], g.params[1]};
        return pauli_exp(s, t, cxd{0.0, -g.params[2]});
    }
    default: {
        Lowered L;
        Status e = lower_gate(s.n, g, L);
        if (!e.ok()) return e;
        return launch_gate(s, L, s.ctx->stream);
    }
    }
}

static Status device_check(hipStream_t st) {
    QI_HIP_TRY(hipGetLastError());
    QI_HIP_TRY(hipStreamSynchronize(st));
    return Status{};
}

// Runs a launch sequence, optionally captured into a hipGraph that is cached
// for an identical call (same gates, state, options).
template <typename F>
static Status run_maybe_graph(Ctx &C, State &s, bool use_graph, const void *key, size_t key_len, F &&launch) {
    hipStream_t st = C.stream;
    QI_HIP_TRY(hipEventRecord(C.ev0, st));
    if (!use_graph) {
        Status e = launch();
        if (!e.ok()) return e;
    } else {
        bool hit = C.graph_exec && C.graph_state == (uint64_t)(uintptr_t)s.d && C.graph_key.size() == key_len &&
                   memcmp(C.graph_key.data(), key, key_len) == 0;
        if (!hit) {
            if (C.graph_exec) { (void)hipGraphExecDestroy(C.graph_exec); C.graph_exec = nullptr; }
            hipGraph_t graph = nullptr;
            QI_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            Status e = launch();
            hipError_t ce = hipStreamEndCapture(st, &graph);
            if (!e.ok()) { if (graph) (void)hipGraphDestroy(graph); return e; }
            if (ce != hipSuccess) return make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(ce));
            hipError_t ie = hipGraphInstantiate(&C.graph_exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) { C.graph_exec = nullptr; return make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(ie)); }
            C.graph_key.assign((const unsigned char *)key, (const unsigned char *)key + key_len);
            C.graph_state = (uint64_t)(uintptr_t)s.d;
        }
        QI_HIP_TRY(hipGraphLaunch(C.graph_exec, st));
    }
    QI_HIP_TRY(hipEventRecord(C.ev1, st));
    Status e = device_check(st);
    if (!e.ok()) return e;
    float ms = 0.f;
    QI_HIP_TRY(hipEventElapsedTime(&ms, C.ev0, C.ev1));
    C.last_exec_ms = ms;
    return Status{};
}

// Plans the launch sequence of a gate list: runs of operator gates go through
// the chosen executor, other Gate variants (measurement, Pauli strings) become
// host-orchestrated steps between them.
// `slots` (may be NULL): the angle slot of each gate, -1 for none (parametric circuits).
static Status build_plan(State &s, const qi_gate *gates, size_t ng, int mode, const qi_exec_opts &o, ExecPlan &plan,
                         const int32_t *slots = nullptr, bool allow_oop = false) {
    size_t k = 0;
    while (k < ng) {
        if (!is_operator(gates[k].kind)) { plan.steps.push_back(ExecStep::host(k)); k++; continue; }
        size_t j = k;
        std::vector<Lowered> seg;
        while (j < ng && is_operator(gates[j].kind)) {
            Lowered L;
            Status e = lower_gate(s.n, gates[j], L);
            if (!e.ok()) return e;
            if (slots) L.slot = slots[j];
            if (L.kind != LK_NOP) seg.push_back(L);
            j++;
        }
        Status e;
        if (mode == QI_EXEC_TILED) e = plan_tiled(s, seg, o.tile_qubits, o.tile_threads_log2, plan, allow_oop);
        else if (mode == QI_EXEC_FUSED_MFMA) e = plan_fused(s, seg, o.block_qubits, plan);
        else for (const Lowered &L : seg) plan.steps.push_back(ExecStep::gate(L));
        if (!e.ok()) return e;
        k = j;
    }
    return Status{};
}

// `blob`: the device copy of the plan's op tables (reg_ops, then matrices).
// `alt`: the second state buffer of out-of-place plans (Ctx::alt; may be null
// for plans without moving passes): a moving pass switches the current buffer,
// the steps between run on it, and every plan ends in the state's own buffer.
static Status launch_plan(State &s, const ExecPlan &plan, const qi_gate *gates, const void *blob, void *alt = nullptr) {
    Ctx &C = *s.ctx;
    hipStream_t st = C.stream;
    const RegOp *dops = (const RegOp *)blob;
    const unsigned char *dmats = (const unsigned char *)blob + plan.mats_offset();
    void *cur = s.d;
    for (const ExecStep &step : plan.steps) {
        Status e;
        State on = s;
        on.d = cur;
        switch (step.type) {
        case ExecStep::GATE: e = launch_gate(on, step.L, st); break;
        case ExecStep::REG: {
            void *dst = cur;
            if (step.reg.oop) {
                if (!alt) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "out-of-place plan without a second buffer");
                dst = cur == s.d ? alt : s.d;
            }
            e = launch_reg_pass(on, step.reg, dops, st, cur, dst);
            cur = dst;
        } break;
        case ExecStep::FUSED:
            if (cur != s.d) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "dense block on the second buffer");
            e = launch_fused_block(s, step.fused, dmats + step.fused.mat_off, st);
            break;
        default:
            if (cur != s.d) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "host step on the second buffer");
            e = apply_validated(s, gates[step.host_gate]);
            break;
        }
        if (!e.ok()) return e;
    }
    if (cur != s.d) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "plan ends in the second buffer");
    return Status{};
}

// `internal`: a gate list the library built itself (Trotter steps), whose
// scaled basis changes are deliberately not unitary; it skips validation.
static Status apply_gates(State &s, uint32_t n_circuit, const qi_gate *gates, size_t ng, const qi_exec_opts *opts,
                          bool internal = false) {
    if (n_circuit != s.n) return make_err(QI_ERR_INVALID_NUMBER_OF_QUBITS, s.n);  // circuit.rs:161-163
    if (ng > 0 && !gates) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "gates is NULL");
    for (size_t k = 0; k < ng && !internal; k++) {
        Status e = validate_gate(s.n, gates[k]);
        if (!e.ok()) return e;
    }
    qi_exec_opts o{};
    o.mode = QI_EXEC_AUTO;
    if (opts) o = *opts;
    if (o.mode < QI_EXEC_PER_GATE || o.mode > QI_EXEC_AUTO) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "exec mode");
    int mode = o.mode;
    if (mode == QI_EXEC_AUTO) mode = s.n >= 12 ? QI_EXEC_TILED : QI_EXEC_PER_GATE;
    bool all_ops = true;
    for (size_t k = 0; k < ng; k++) all_ops = all_ops && is_operator(gates[k].kind);
    Ctx &C = *s.ctx;
    const bool use_graph = o.use_graph && all_ops;
    // A second state buffer lets a tiled plan with relabelled SWAPs end in one
    // permuting pass instead of a chain of in-place ones: taken for one-wave
    // tiles when the list has uncontrolled SWAPs and the device keeps 16 GiB
    // free beside it (the buffer stays with the context).
    bool allow_oop = false;
    if (mode == QI_EXEC_TILED && (o.tile_threads_log2 <= 0 || o.tile_threads_log2 == kLaneBits) && s.n >= 16) {
        bool swaps = false;
        for (size_t k = 0; k < ng && !swaps; k++) swaps = gates[k].kind == QI_GATE_SWAP && gates[k].ncontrols == 0;
        if (swaps && C.alt.bytes >= s.bytes()) {
            allow_oop = true;
        } else if (swaps) {
            size_t fr = 0, tot = 0;
            if (hipMemGetInfo(&fr, &tot) == hipSuccess && fr > s.bytes() + (size_t(16) << 30))
                allow_oop = ensure(C.alt, s.bytes()).ok();
        }
    }
    // plan key: options, gates, state shape and the buffer choice; the same gate
    // list again reuses the last plan and its op tables in the plan buffer (no
    // planning, no upload)
    std::vector<unsigned char> key(sizeof(qi_exec_opts) + ng * sizeof(qi_gate) + 9);
    memcpy(key.data(), &o, sizeof o);
    if (ng) memcpy(key.data() + sizeof o, gates, ng * sizeof(qi_gate));
    const uint64_t np = (uint64_t)s.n | ((uint64_t)s.prec << 32);
    memcpy(key.data() + sizeof o + ng * sizeof(qi_gate), &np, 8);
    key[sizeof o + ng * sizeof(qi_gate) + 8] = allow_oop ? 1 : 0;
    std::shared_ptr<ExecPlan> plan;
    if (C.cached_plan && C.plan_key == key) {
        plan = std::static_pointer_cast<ExecPlan>(C.cached_plan);
    } else {
        plan = std::make_shared<ExecPlan>();
        Status e = build_plan(s, gates, ng, mode, o, *plan, nullptr, allow_oop);
        if (!e.ok()) return e;
        C.cached_plan.reset();
        C.plan_key.clear();
        if (plan->blob_bytes() > 0) {
            if (!(e = ensure(C.plan, plan->blob_bytes())).ok()) return e;
            QI_HIP_TRY(hipStreamSynchronize(C.stream));  // the previous plan may still be in use
            // a cached graph reads the op table about to be overwritten
            if (C.graph_exec) { (void)hipGraphExecDestroy(C.graph_exec); C.graph_exec = nullptr; C.graph_key.clear(); }
            std::vector<unsigned char> blob(plan->blob_bytes(), 0);
            if (!plan->reg_ops.empty()) memcpy(blob.data(), plan->reg_ops.data(), plan->reg_ops.size() * sizeof(RegOp));
            if (!plan->mats.empty()) memcpy(blob.data() + plan->mats_offset(), plan->mats.data(), plan->mats.size());
            QI_HIP_TRY(hipMemcpy(C.plan.p, blob.data(), blob.size(), hipMemcpyHostToDevice));
        }
        C.cached_plan = plan;
        C.plan_key = key;
    }
    C.last_plan_steps = plan->steps.size();
    // graph key: the plan key and the op-table and second-state buffers the graph reads
    const uint64_t pp[2] = {(uint64_t)(uintptr_t)C.plan.p, (uint64_t)(uintptr_t)(allow_oop ? C.alt.p : nullptr)};
    key.insert(key.end(), (const unsigned char *)pp, (const unsigned char *)pp + sizeof pp);
    void *alt = allow_oop ? C.alt.p : nullptr;
    auto launch = [&]() -> Status { return launch_plan(s, *plan, gates, C.plan.p, alt); };
    return run_maybe_graph(C, s, use_graph, key.data(), key.size(), launch);
}

static Status check_pair(const State &a, const State &b) {
    if (a.n != b.n) return make_err(QI_ERR_INVALID_NUMBER_OF_QUBITS, a.n);
    if (a.prec != b.prec) return make_err(QI_ERR_PRECISION_MISMATCH);
    if (a.ctx != b.ctx) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "states on different contexts");
    return Status{};
}

}  // namespace qi

using namespace qi;

#define QI_RET(st)               \
    do {                         \
        Status s_ = (st);        \
        fill_err(err, s_);       \
        return (qi_status)s_.code; \
    } while (0)

#define QI_HIP_TRY_C(expr)                                                                  \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) QI_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e_))); \
    } while (0)

#define QI_NULLCHK(p)                                                                             \
    do {                                                                                          \
        if (!(p)) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, #p " is NULL"));          \
    } while (0)

#define QI_LOCK(ctxp) std::lock_guard<std::mutex> lk_((ctxp)->mu)

// A circuit with parametric gates (Gate::Parametric, gate.rs:107-114, whose
// angles come from shared Parameter<N> values, parameter.rs:12-68), planned
// once: new angles rewrite the coefficients of the parametric ops in the op
// tables (one upload), and each run replays the captured launch sequence.
struct qi_param_circuit_s {
    Ctx *ctx = nullptr;
    uint32_t n = 0;
    int prec = QI_C64;
    std::vector<qi_gate> gates;
    std::vector<int32_t> slots;
    std::vector<double> angles;  // the angle table: entry k is slot k's angle
    ExecPlan plan;
    DevBuf blob;                 // device copy of plan.reg_ops
    bool dirty = true;           // plan.reg_ops changed since the last upload
    bool graph_ok = true;        // false when parametric gates run as streaming steps (angles in launch arguments)
    hipGraphExec_t graph = nullptr;
    void *graph_state = nullptr;
};

namespace qi {

// New coefficients of every parametric op of the plan for the angle table.
static void patch_param_plan(qi_param_circuit_s &pc) {
    for (RegOp &op : pc.plan.reg_ops)
        if (op.param > 0) param_op_coeffs(op, pc.angles[(size_t)op.param - 1]);
    static const int gate_of[LK_NOP + 1] = {-1, -1, -1, QI_GATE_RX, QI_GATE_RY, -1, -1, QI_GATE_P, QI_GATE_RZ, -1, -1, -1};
    for (ExecStep &st : pc.plan.steps) {
        if (st.type != ExecStep::GATE || st.L.slot < 0) continue;
        qi_gate g{};
        g.kind = gate_of[st.L.kind];
        g.ntargets = 1;
        g.params[0] = pc.angles[(size_t)st.L.slot];
        Lowered M;
        (void)lower_gate(pc.n, g, M);
        memcpy(st.L.c, M.c, sizeof M.c);
    }
    pc.dirty = true;
}

}  // namespace qi

extern "C" {

qi_status qi_ctx_create(int device, qi_ctx_t *out, qi_error *err) {
    QI_NULLCHK(out);
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) QI_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, e != hipSuccess ? hipGetErrorString(e) : "no HIP device"));
    if (device < 0 || device >= ndev) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "device index"));
    std::unique_ptr<qi_ctx_s> c(new qi_ctx_s());
    Ctx &C = c->c;
    C.device = device;
    if ((e = hipSetDevice(device)) != hipSuccess) QI_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    if ((e = hipStreamCreateWithFlags(&C.stream, hipStreamNonBlocking)) != hipSuccess)
        QI_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    if ((e = hipEventCreate(&C.ev0)) != hipSuccess || (e = hipEventCreate(&C.ev1)) != hipSuccess)
        QI_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    if ((e = hipHostMalloc((void **)&C.host_pinned, 64 * sizeof(double))) != hipSuccess)
        QI_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    C.host_pinned_words = 64;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) C.cu_count = prop.multiProcessorCount;
    // the no-allocation scratch array's address on this device (the device is current)
    {
        Status ns = noalloc_scratch(device, &C.noalloc);
        if (!ns.ok()) QI_RET(ns);
    }
    // an unseeded context draws from an OS-seeded stream, as rand::rng() does
    C.rng = (uint64_t(std::random_device{}()) << 32) ^ uint64_t(std::random_device{}());
    *out = c.release();
    QI_RET(Status{});
}

qi_status qi_ctx_destroy(qi_ctx_t ctx) {
    if (!ctx) return QI_OK;
    Ctx &C = ctx->c;
    (void)hipSetDevice(C.device);
    (void)hipStreamSynchronize(C.stream);
    if (C.graph_exec) (void)hipGraphExecDestroy(C.graph_exec);
    for (DevBuf *b : {&C.red, &C.red_out, &C.ops, &C.probs, &C.aux, &C.plan, &C.trot, &C.alt})
        if (b->p) (void)hipFree(b->p);
    if (C.host_pinned) (void)hipHostFree(C.host_pinned);
    if (C.ev0) (void)hipEventDestroy(C.ev0);
    if (C.ev1) (void)hipEventDestroy(C.ev1);
    if (C.stream) (void)hipStreamDestroy(C.stream);
    delete ctx;
    return QI_OK;
}

qi_status qi_ctx_seed(qi_ctx_t ctx, uint64_t seed, qi_error *err) {
    QI_NULLCHK(ctx);
    QI_LOCK(&ctx->c);
    ctx->c.rng = seed;
    QI_RET(Status{});
}

qi_status qi_ctx_synchronize(qi_ctx_t ctx, qi_error *err) {
    QI_NULLCHK(ctx);
    QI_LOCK(&ctx->c);
    QI_RET(device_check(ctx->c.stream));
}

qi_status qi_ctx_info(qi_ctx_t ctx, char *name, size_t name_len, uint64_t *hbm_bytes, int *cu_count, qi_error *err) {
    QI_NULLCHK(ctx);
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, ctx->c.device);
    if (e != hipSuccess) QI_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    if (name && name_len) snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    if (cu_count) *cu_count = prop.multiProcessorCount;
    QI_RET(Status{});
}

qi_status qi_state_create(qi_ctx_t ctx, uint32_t n, qi_prec prec, qi_state_t *out, qi_error *err) {
    QI_NULLCHK(ctx);
    QI_NULLCHK(out);
    if (n == 0) QI_RET(make_err(QI_ERR_INVALID_NUMBER_OF_QUBITS, 0));  // state.rs:166-168
    if (n > QI_MAX_QUBITS) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "more qubits than the engine supports"));
    if (prec != QI_C64 && prec != QI_C128) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "precision"));
    QI_LOCK(&ctx->c);
    std::unique_ptr<qi_state_s> s(new qi_state_s());
    Status e = alloc_state(ctx->c, n, prec, s->s);
    if (!e.ok()) QI_RET(e);
    *out = s.release();
    QI_RET(Status{});
}

qi_status qi_state_destroy(qi_state_t s) {
    if (!s) return QI_OK;
    {
        QI_LOCK(s->s.ctx);
        (void)hipStreamSynchronize(s->s.ctx->stream);
        if (s->s.d) (void)hipFree(s->s.d);
    }
    delete s;
    return QI_OK;
}

qi_status qi_state_clone(qi_state_t src, qi_state_t *out, qi_error *err) {
    QI_NULLCHK(src);
    QI_NULLCHK(out);
    QI_LOCK(src->s.ctx);
    std::unique_ptr<qi_state_s> s(new qi_state_s());
    Status e = alloc_state(*src->s.ctx, src->s.n, src->s.prec, s->s);
    if (!e.ok()) QI_RET(e);
    if (!(e = launch_copy(s->s, src->s, src->s.ctx->stream)).ok()) { (void)hipFree(s->s.d); QI_RET(e); }
    if (!(e = device_check(src->s.ctx->stream)).ok()) { (void)hipFree(s->s.d); QI_RET(e); }
    *out = s.release();
    QI_RET(Status{});
}

qi_status qi_state_copy(qi_state_t dst, qi_state_t src, qi_error *err) {
    QI_NULLCHK(dst);
    QI_NULLCHK(src);
    Status e = check_pair(dst->s, src->s);
    if (!e.ok()) QI_RET(e);
    QI_LOCK(src->s.ctx);
    if (!(e = launch_copy(dst->s, src->s, src->s.ctx->stream)).ok()) QI_RET(e);
    QI_RET(device_check(src->s.ctx->stream));
}

uint32_t qi_state_num_qubits(qi_state_t s) { return s ? s->s.n : 0; }
int32_t qi_state_precision(qi_state_t s) { return s ? s->s.prec : -1; }
void *qi_state_device_ptr(qi_state_t s) { return s ? s->s.d : nullptr; }

qi_status qi_state_init(qi_state_t s, qi_init kind, uint64_t basis, qi_error *err) {
    QI_NULLCHK(s);
    State &S = s->s;
    if (kind < QI_INIT_ZERO || kind > QI_INIT_GHZ) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "init kind"));
    if (kind == QI_INIT_BASIS && basis >= (uint64_t(1) << S.n)) QI_RET(make_err(QI_ERR_INVALID_QUBIT_INDEX, basis, S.n));  // state.rs:196-198
    QI_LOCK(S.ctx);
    Status e = launch_init(S, kind, basis, S.ctx->stream);
    if (!e.ok()) QI_RET(e);
    QI_RET(device_check(S.ctx->stream));
}

qi_status qi_state_upload(qi_state_t s, const double *amps, int check_norm, qi_error *err) {
    QI_NULLCHK(s);
    QI_NULLCHK(amps);
    State &S = s->s;
    uint64_t dim = uint64_t(1) << S.n;
    if (check_norm) {  // State::new, state.rs:117-122
        double ns = 0.0;
        for (uint64_t i = 0; i < 2 * dim; i++) ns += amps[i] * amps[i];
        if (std::fabs(ns - 1.0) > DBL_EPSILON * double(dim)) QI_RET(make_err(QI_ERR_STATE_NOT_NORMALISED));
    }
    QI_LOCK(S.ctx);
    hipStream_t st = S.ctx->stream;
    if (S.prec == QI_C128) {
        hipError_t e = hipMemcpyAsync(S.d, amps, S.bytes(), hipMemcpyHostToDevice, st);
        if (e != hipSuccess) QI_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    } else {
        std::vector<float> f(2 * dim);
        for (uint64_t i = 0; i < 2 * dim; i++) f[i] = (float)amps[i];
        hipError_t e = hipMemcpyAsync(S.d, f.data(), S.bytes(), hipMemcpyHostToDevice, st);
        if (e != hipSuccess) QI_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
        QI_RET(device_check(st));
    }
    QI_RET(device_check(st));
}

qi_status qi_state_download(qi_state_t s, double *amps, qi_error *err) {
    QI_NULLCHK(s);
    QI_NULLCHK(amps);
    State &S = s->s;
    uint64_t dim = uint64_t(1) << S.n;
    QI_LOCK(S.ctx);
    hipStream_t st = S.ctx->stream;
    if (S.prec == QI_C128) {
        hipError_t e = hipMemcpyAsync(amps, S.d, S.bytes(), hipMemcpyDeviceToHost, st);
        if (e != hipSuccess) QI_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
        QI_RET(device_check(st));
    }
    std::vector<float> f(2 * dim);
    hipError_t e = hipMemcpyAsync(f.data(), S.d, S.bytes(), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) QI_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    Status ds = device_check(st);
    if (!ds.ok()) QI_RET(ds);
    for (uint64_t i = 0; i < 2 * dim; i++) amps[i] = f[i];
    QI_RET(Status{});
}

qi_status qi_apply_gate(qi_state_t s, const qi_gate *g, qi_error *err) {
    QI_NULLCHK(s);
    QI_NULLCHK(g);
    State &S = s->s;
    Status e = validate_gate(S.n, *g);
    if (!e.ok()) QI_RET(e);
    QI_LOCK(S.ctx);
    if (!(e = apply_validated(S, *g)).ok()) QI_RET(e);
    QI_RET(device_check(S.ctx->stream));
}

qi_status qi_apply_gates(qi_state_t s, uint32_t n_circuit, const qi_gate *gates, size_t ngates, const qi_exec_opts *opts,
                         qi_error *err) {
    QI_NULLCHK(s);
    QI_LOCK(s->s.ctx);
    QI_RET(apply_gates(s->s, n_circuit, gates, ngates, opts));
}

qi_status qi_pauli_apply(qi_state_t in, qi_state_t out, const qi_pauli_term *t, qi_error *err) {
    QI_NULLCHK(in); QI_NULLCHK(out); QI_NULLCHK(t);
    Status e = check_pair(in->s, out->s);
    if (!e.ok()) QI_RET(e);
    if (!(e = check_mask(in->s.n, t->x_mask | t->z_mask)).ok()) QI_RET(e);
    QI_LOCK(in->s.ctx);
    if (!(e = launch_pauli_apply(in->s, out->s, *t, in->s.ctx->stream)).ok()) QI_RET(e);
    QI_RET(device_check(in->s.ctx->stream));
}

qi_status qi_pauli_apply_normalised(qi_state_t in, qi_state_t out, const qi_pauli_term *t, qi_error *err) {
    QI_NULLCHK(in); QI_NULLCHK(out); QI_NULLCHK(t);
    Status e = check_pair(in->s, out->s);
    if (!e.ok()) QI_RET(e);
    QI_LOCK(in->s.ctx);
    if (!(e = pauli_apply_normalised(in->s, out->s, *t)).ok()) QI_RET(e);
    QI_RET(device_check(in->s.ctx->stream));
}

qi_status qi_pauli_exp(qi_state_t s, const qi_pauli_term *t, double f_re, double f_im, qi_error *err) {
    QI_NULLCHK(s); QI_NULLCHK(t);
    QI_LOCK(s->s.ctx);
    Status e = pauli_exp(s->s, *t, cxd{f_re, f_im});
    if (!e.ok()) QI_RET(e);
    QI_RET(device_check(s->s.ctx->stream));
}

qi_status qi_pauli_exp_neg_i_dt(qi_state_t s, const qi_pauli_term *t, double dt, qi_error *err) {
    QI_NULLCHK(s); QI_NULLCHK(t);
    if (t->c_im != 0.0) QI_RET(make_err(QI_ERR_INVALID_PAULI_COEFFICIENT, 0, 0, t->c_re, t->c_im));
    QI_LOCK(s->s.ctx);
    Status e = pauli_exp(s->s, *t, cxd{0.0, -dt});
    if (!e.ok()) QI_RET(e);
    QI_RET(device_check(s->s.ctx->stream));
}

qi_status qi_sumop_apply(qi_state_t in, qi_state_t out, const qi_pauli_term *terms, size_t nterms, qi_error *err) {
    QI_NULLCHK(in); QI_NULLCHK(out);
    if (nterms && !terms) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "terms is NULL"));
    if (in == out) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "in and out must differ"));
    Status e = check_pair(in->s, out->s);
    if (!e.ok()) QI_RET(e);
    for (size_t k = 0; k < nterms; k++)
        if (!(e = check_mask(in->s.n, terms[k].x_mask | terms[k].z_mask)).ok()) QI_RET(e);
    QI_LOCK(in->s.ctx);
    hipStream_t st = in->s.ctx->stream;
    if (nterms == 0) {  // state.clone() * 0.0 (pauli_string.rs:454-456)
        if (!(e = launch_copy(out->s, in->s, st)).ok()) QI_RET(e);
        if (!(e = launch_scale(out->s, 0.0, 0.0, st)).ok()) QI_RET(e);
    } else if (!(e = launch_sumop_apply(in->s, out->s, terms, nterms, st)).ok()) {
        QI_RET(e);
    }
    QI_RET(device_check(st));
}

qi_status qi_sumop_expectation(qi_state_t s, const qi_pauli_term *terms, size_t nterms, double *re, double *im,
                               qi_error *err) {
    QI_NULLCHK(s); QI_NULLCHK(re); QI_NULLCHK(im);
    if (nterms && !terms) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "terms is NULL"));
    *re = 0.0; *im = 0.0;
    if (nterms == 0) QI_RET(Status{});  // pauli_string.rs:486-488
    Status e;
    for (size_t k = 0; k < nterms; k++)
        if (!(e = check_mask(s->s.n, terms[k].x_mask | terms[k].z_mask)).ok()) QI_RET(e);
    QI_LOCK(s->s.ctx);
    QI_RET(sumop_expectation(s->s, terms, nterms, re, im, s->s.ctx->stream));
}

// exp(-i theta P) for a Pauli string P as gates, so the tiled executor fuses
// a whole Trotter step into a few HBM passes. With X or Y factors, P is
// conjugated to X on one pivot qubit p (the highest X/Y factor):
// V = S^dag on each Y factor (S^dag Y S = X), CNOT(p -> q) on each other X/Y
// factor q (X_p X_q -> X_p) and CZ(p, q) on each Z factor q (X_p Z_q -> X_p),
// so exp(-i theta P) = V^dag RX(2 theta)_p V. A Z string takes a CNOT ladder
// gathering its parity on its last qubit, RZ(2 theta) there and the ladder
// undone. Every gate but the one rotation is a permutation or a diagonal of
// units (exact in floating point), so a step adds no norm drift beyond the
// rotations' rounding. The same unitary as the reference's cos/sin form
// (pauli_string.rs:237-262, real coefficient), up to rounding.
static void pauli_exp_gates(const qi_pauli_term &t, double theta, std::vector<qi_gate> &out) {
    const uint64_t sup = t.x_mask | t.z_mask;
    int qs[64], m = 0;
    for (uint64_t b = sup; b; b &= b - 1) qs[m++] = __builtin_ctzll(b);
    auto one = [&](int kind, int q, double angle) {
        qi_gate g{};
        g.kind = kind; g.ntargets = 1; g.targets[0] = (uint64_t)q; g.params[0] = angle;
        out.push_back(g);
    };
    auto ctl = [&](int kind, int c, int q) {
        qi_gate g{};
        g.kind = kind; g.ntargets = 1; g.targets[0] = (uint64_t)q; g.ncontrols = 1; g.controls[0] = (uint64_t)c;
        out.push_back(g);
    };
    if (t.x_mask == 0) {
        for (int k = 0; k + 1 < m; k++) ctl(QI_GATE_CNOT, qs[k], qs[m - 1]);
        one(QI_GATE_RZ, qs[m - 1], 2.0 * theta);
        for (int k = m - 2; k >= 0; k--) ctl(QI_GATE_CNOT, qs[k], qs[m - 1]);
        return;
    }
    const int p = 63 - __builtin_clzll(t.x_mask);
    auto y_to_x = [&](bool undo) {  // S^dag on each Y factor, S to undo
        for (int k = 0; k < m; k++)
            if ((t.x_mask >> qs[k]) & (t.z_mask >> qs[k]) & 1) one(undo ? QI_GATE_S : QI_GATE_SDAG, qs[k], 0.0);
    };
    auto gather = [&]() {  // self-inverse, mutually commuting (one control p, distinct targets)
        for (int k = 0; k < m; k++) {
            const int q = qs[k];
            if (q == p) continue;
            ctl((t.x_mask >> q) & 1 ? QI_GATE_CNOT : QI_GATE_Z, p, q);  // CNOT, or CZ on a Z factor
        }
    };
    y_to_x(false);
    gather();
    one(QI_GATE_RX, p, 2.0 * theta);
    gather();
    y_to_x(true);
}

// A term X_pX_q, Y_pY_q or Z_pZ_q: 1, 2 or 3 (0: anything else).
static int bond_kind(const qi_pauli_term &t) {
    if (__builtin_popcountll(t.x_mask | t.z_mask) != 2) return 0;
    if (t.z_mask == 0) return 1;
    if (t.x_mask == t.z_mask) return 2;
    return t.x_mask == 0 ? 3 : 0;
}

// exp(-i (a X_pX_q + b Y_pY_q + c Z_pZ_q)) for a run of such terms on one
// qubit pair (they commute, so their product is exact in any order): with
// C = CNOT(p -> q), C X_pX_q C = X_p, C Y_pY_q C = -X_p Z_q and C Z_pZ_q C = Z_q,
// and CZ X_p CZ = X_p Z_q, the run is C RZ_q(2c) RX_p(2a) CZ RX_p(-2b) CZ C:
// 7 gates for the 13 of three separate terms, every rotation uncontrolled (the
// tile executor's norm-preserving shear forms in complex64).
static void bond_exp_gates(int p, int q, double a, double b, double c, std::vector<qi_gate> &out) {
    auto push = [&](int kind, int t, int ctl, double angle) {
        qi_gate g{};
        g.kind = kind; g.ntargets = 1; g.targets[0] = (uint64_t)t; g.params[0] = angle;
        if (ctl >= 0) { g.ncontrols = 1; g.controls[0] = (uint64_t)ctl; }
        out.push_back(g);
    };
    push(QI_GATE_CNOT, q, p, 0.0);
    if (c != 0.0) push(QI_GATE_RZ, q, -1, 2.0 * c);
    if (a != 0.0) push(QI_GATE_RX, p, -1, 2.0 * a);
    if (b != 0.0) {
        push(QI_GATE_Z, q, p, 0.0);  // CZ
        push(QI_GATE_RX, p, -1, -2.0 * b);
        push(QI_GATE_Z, q, p, 0.0);
    }
    push(QI_GATE_CNOT, q, p, 0.0);
}

// One Trotter step's exponentials as gates, terms in list order or reversed
// (angle f times each coefficient): runs of consecutive XX / YY / ZZ terms on
// one qubit pair become one bond_exp_gates, every other term pauli_exp_gates.
static void trotter_step_gates(const qi_pauli_term *terms, size_t nterms, double f, bool reverse,
                               std::vector<qi_gate> &out) {
    const auto order = [&](size_t i) { return reverse ? nterms - 1 - i : i; };
    for (size_t i = 0; i < nterms;) {
        const qi_pauli_term &t = terms[order(i)];
        const uint64_t sup = t.x_mask | t.z_mask;
        size_t j = i;
        double abc[4] = {0, 0, 0, 0};
        while (j < nterms) {
            const qi_pauli_term &u = terms[order(j)];
            const int k = bond_kind(u);
            if (!k || (u.x_mask | u.z_mask) != sup) break;
            abc[k] += f * u.c_re;
            j++;
        }
        if (j > i + 1) {
            const int p = 63 - __builtin_clzll(sup), q = __builtin_ctzll(sup);
            bond_exp_gates(p, q, abc[1], abc[2], abc[3], out);
            i = j;
        } else {
            pauli_exp_gates(t, f * t.c_re, out);
            i++;
        }
    }
}

// States from this size on run Trotter steps through the tiled executor.
constexpr uint32_t kFusedTrotterQubits = 16;

// Two Pauli strings commute iff they anticommute on an even number of qubits.
static bool pauli_commute(uint64_t x1, uint64_t z1, uint64_t x2, uint64_t z2) {
    return (__builtin_popcountll((x1 & z2) ^ (z1 & x2)) & 1) == 0;
}

// One Trotter step on the streaming kernels as a sequence of pieces: a
// non-diagonal term exp(a P), or a run of diagonal terms applied in one pass.
// A diagonal term joins the last diagonal run when it commutes with every
// term after that run (moving it back past them changes nothing), and
// consecutive exponentials of the same string add their factors; both are
// exact rewrites of the reference's term-by-term product (time_evolution.rs:
// 45-115), so results differ only by rounding.
struct TrotterPiece {
    bool diag = false;
    std::vector<std::pair<qi_pauli_term, cxd>> terms;  // (term, factor applied to its coefficient)
    DiagPass pass{};
    size_t table_off = 0;  // first table entry of this piece in the step's table buffer
};

static bool diag_fits(const qi_pauli_term &t) { return t.x_mask == 0 && __builtin_popcountll(t.z_mask) <= kDiagTabBits; }

static std::vector<TrotterPiece> trotter_pieces(const qi_pauli_term *terms, size_t nt, double dt, int order) {
    std::vector<std::pair<qi_pauli_term, cxd>> seq;
    if (order == 1) {
        for (size_t k = 0; k < nt; k++) seq.push_back({terms[k], cxd{0.0, -dt}});
    } else {
        for (size_t k = 0; k < nt; k++) seq.push_back({terms[k], cxd{0.0, -dt / 2.0}});
        for (size_t k = nt; k-- > 0;) seq.push_back({terms[k], cxd{0.0, -dt / 2.0}});
    }
    std::vector<TrotterPiece> out;
    for (const auto &tf : seq) {
        const qi_pauli_term &t = tf.first;
        if (diag_fits(t)) {
            // walk back over non-diagonal pieces the term commutes with
            size_t j = out.size();
            while (j > 0 && !out[j - 1].diag &&
                   pauli_commute(0, t.z_mask, out[j - 1].terms[0].first.x_mask, out[j - 1].terms[0].first.z_mask))
                j--;
            if (j > 0 && out[j - 1].diag) { out[j - 1].terms.push_back(tf); continue; }
        } else if (!out.empty() && !out.back().diag && out.back().terms[0].first.x_mask == t.x_mask &&
                   out.back().terms[0].first.z_mask == t.z_mask && out.back().terms[0].first.c_re == t.c_re &&
                   out.back().terms[0].first.c_im == t.c_im) {
            // exp(a P) exp(b P) = exp((a + b) P)
            out.back().terms[0].second.re += tf.second.re;
            out.back().terms[0].second.im += tf.second.im;
            continue;
        }
        TrotterPiece p;
        p.diag = diag_fits(t);
        p.terms.push_back(tf);
        out.push_back(p);
    }
    return out;
}

// Splits a diagonal run into passes of at most kDiagTabMax tables over at most
// kDiagTabBits qubits each, and fills the tables: entry u of a table is the
// product over its terms of cosh(a) + s sinh(a), s = (-1)^{popcount(u's bits & z)}.
static void diag_tables(const std::vector<std::pair<qi_pauli_term, cxd>> &run, std::vector<TrotterPiece> &passes,
                        std::vector<double> &tab) {
    struct Table { uint64_t mask; std::vector<size_t> members; };
    std::vector<Table> cur;
    auto flush = [&]() {
        if (cur.empty()) return;
        TrotterPiece p;
        p.diag = true;
        p.table_off = tab.size() / 2;
        int entries = 0;
        for (size_t b = 0; b < cur.size(); b++) {
            DiagTab &d = p.pass.tab[b];
            d.bits = 0;
            d.off = entries;
            for (uint64_t m = cur[b].mask; m; m &= m - 1) d.pos[d.bits++] = (uint8_t)__builtin_ctzll(m);
            for (uint32_t u = 0; u < (1u << d.bits); u++) {
                uint64_t pattern = 0;
                for (int k = 0; k < d.bits; k++) if ((u >> k) & 1) pattern |= uint64_t(1) << d.pos[k];
                cxd prod{1.0, 0.0};
                for (size_t i : cur[b].members) {
                    const qi_pauli_term &t = run[i].first;
                    const cxd a = cmul(cxd{t.c_re, t.c_im}, run[i].second);
                    const double s = (__builtin_popcountll(pattern & t.z_mask) & 1) ? -1.0 : 1.0;
                    const cxd ch = ccosh_(a), sh = csinh_(a);
                    prod = cmul(prod, cxd{ch.re + s * sh.re, ch.im + s * sh.im});
                }
                tab.push_back(prod.re);
                tab.push_back(prod.im);
            }
            entries += 1 << d.bits;
        }
        p.pass.ntab = (int)cur.size();
        p.pass.entries = entries;
        passes.push_back(p);
        cur.clear();
    };
    for (size_t i = 0; i < run.size(); i++) {
        const uint64_t z = run[i].first.z_mask;
        size_t b = 0;
        while (b < cur.size() && __builtin_popcountll(cur[b].mask | z) > kDiagTabBits) b++;
        if (b == cur.size()) {
            if (cur.size() == (size_t)kDiagTabMax) { flush(); b = 0; }
            cur.push_back(Table{0, {}});
        }
        cur[b].mask |= z;
        cur[b].members.push_back(i);
    }
    flush();
}

qi_status qi_sumop_expectation_noalloc(qi_state_t s, const qi_pauli_term *terms, size_t nterms, double *re, double *im,
                                       qi_error *err) {
    QI_NULLCHK(s); QI_NULLCHK(re); QI_NULLCHK(im);
    if (nterms && !terms) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "terms is NULL"));
    Status e;
    for (size_t k = 0; k < nterms; k++)
        if (!(e = check_mask(s->s.n, terms[k].x_mask | terms[k].z_mask)).ok()) QI_RET(e);
    QI_LOCK(s->s.ctx);
    QI_RET(sumop_expectation_noalloc(s->s, terms, nterms, re, im, s->s.ctx->stream));
}

qi_status qi_trotter_evolve(qi_state_t s, const qi_pauli_term *terms, size_t nterms, double dt, uint64_t steps, int order,
                            qi_error *err) {
    QI_NULLCHK(s);
    if (nterms == 0) QI_RET(make_err(QI_ERR_INVALID_NUMBER_OF_QUBITS, 0));  // time_evolution.rs:147-149
    QI_NULLCHK(terms);
    if (order != 1 && order != 2) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "trotter order"));
    Status e;
    for (size_t k = 0; k < nterms; k++)
        if (!(e = check_mask(s->s.n, terms[k].x_mask | terms[k].z_mask)).ok()) QI_RET(e);
    QI_LOCK(s->s.ctx);
    State &S = s->s;
    // Hermitian terms on a large state: one step as one gate list, planned
    // once and replayed (cached plan and graph) for every step
    bool fused = S.n >= kFusedTrotterQubits && steps > 0;
    for (size_t k = 0; k < nterms; k++)
        fused = fused && terms[k].c_im == 0.0 && (terms[k].x_mask | terms[k].z_mask) != 0;
    if (fused) {
        std::vector<qi_gate> step_gates;
        if (order == 1) {
            trotter_step_gates(terms, nterms, dt, false, step_gates);
        } else {
            trotter_step_gates(terms, nterms, 0.5 * dt, false, step_gates);
            trotter_step_gates(terms, nterms, 0.5 * dt, true, step_gates);
        }
        qi_exec_opts o{};
        o.mode = QI_EXEC_TILED;
        o.use_graph = 1;
        for (uint64_t step = 0; step < steps; step++)
            if (!(e = apply_gates(S, S.n, step_gates.data(), step_gates.size(), &o, true)).ok()) QI_RET(e);
        QI_RET(Status{});  // last_plan_steps: the tile passes of one step
    }
    // streaming kernels: one pass per non-diagonal term, one per diagonal run
    std::vector<TrotterPiece> plan;
    std::vector<double> tab;
    for (TrotterPiece &p : trotter_pieces(terms, nterms, dt, order)) {
        if (p.diag) diag_tables(p.terms, plan, tab);
        else plan.push_back(p);
    }
    Ctx &C = *S.ctx;
    if (!tab.empty()) {
        if (!(e = ensure(C.trot, tab.size() * sizeof(double))).ok()) QI_RET(e);
        hipError_t he = hipMemcpyAsync(C.trot.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, C.stream);
        if (he != hipSuccess) QI_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(he)));
    }
    C.last_plan_steps = plan.size();
    const double2 *dtab = (const double2 *)C.trot.p;
    for (uint64_t step = 0; step < steps; step++)
        for (const TrotterPiece &p : plan) {
            if (p.diag) e = launch_diag_tables(S, p.pass, dtab + p.table_off, C.stream);
            else e = pauli_exp(S, p.terms[0].first, p.terms[0].second);
            if (!e.ok()) QI_RET(e);
        }
    QI_RET(device_check(S.ctx->stream));
}

qi_status qi_inner_product(qi_state_t a, qi_state_t b, double *re, double *im, qi_error *err) {
    QI_NULLCHK(a); QI_NULLCHK(b); QI_NULLCHK(re); QI_NULLCHK(im);
    Status e = check_pair(a->s, b->s);
    if (!e.ok()) QI_RET(e);
    QI_LOCK(a->s.ctx);
    double r[2];
    if (!(e = reduce_state(a->s, &b->s, RK_INNER, r, a->s.ctx->stream)).ok()) QI_RET(e);
    *re = r[0]; *im = r[1];
    QI_RET(Status{});
}

qi_status qi_norm_sqr(qi_state_t s, double *out, qi_error *err) {
    QI_NULLCHK(s); QI_NULLCHK(out);
    QI_LOCK(s->s.ctx);
    QI_RET(norm_sqr(s->s, out));
}

qi_status qi_normalise(qi_state_t s, qi_error *err) {
    QI_NULLCHK(s);
    QI_LOCK(s->s.ctx);
    Status e = normalise(s->s);
    if (!e.ok()) QI_RET(e);
    QI_RET(device_check(s->s.ctx->stream));
}

// The normalised inner product <a/|a| | b/|b|> of fs_fidelity / fs_dist
// (state.rs:470-498), with the reference's error order: each state's
// normalise (ZeroNorm) before the inner product's qubit-count check
// (state.rs:890-897).
static Status normalised_inner(State &a, State &b, double *ir, double *ii) {
    double na, nb, r[2];
    Status e;
    if (!(e = norm_sqr(a, &na)).ok()) return e;
    if (std::sqrt(na) == 0.0) return make_err(QI_ERR_ZERO_NORM);
    if (!(e = norm_sqr(b, &nb)).ok()) return e;
    if (std::sqrt(nb) == 0.0) return make_err(QI_ERR_ZERO_NORM);
    if (!(e = check_pair(a, b)).ok()) return e;
    if (!(e = reduce_state(a, &b, RK_INNER, r, a.ctx->stream)).ok()) return e;
    const double scale = std::sqrt(na) * std::sqrt(nb);
    *ir = r[0] / scale;
    *ii = r[1] / scale;
    return Status{};
}

qi_status qi_fidelity(qi_state_t a, qi_state_t b, double *out, qi_error *err) {
    // fs_fidelity, state.rs:492-498: |<a/|a| | b/|b|>|^2
    QI_NULLCHK(a); QI_NULLCHK(b); QI_NULLCHK(out);
    if (a->s.ctx != b->s.ctx) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "states on different contexts"));
    QI_LOCK(a->s.ctx);
    double ir, ii;
    Status e = normalised_inner(a->s, b->s, &ir, &ii);
    if (!e.ok()) QI_RET(e);
    *out = ir * ir + ii * ii;
    QI_RET(Status{});
}

qi_status qi_fs_dist(qi_state_t a, qi_state_t b, double *out, qi_error *err) {
    // fs_dist, state.rs:470-478: acos(|<a/|a| | b/|b|>|), unclamped as in the
    // reference (a modulus rounded above 1 gives NaN there too)
    QI_NULLCHK(a); QI_NULLCHK(b); QI_NULLCHK(out);
    if (a->s.ctx != b->s.ctx) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "states on different contexts"));
    QI_LOCK(a->s.ctx);
    double ir, ii;
    Status e = normalised_inner(a->s, b->s, &ir, &ii);
    if (!e.ok()) QI_RET(e);
    *out = std::acos(std::hypot(ir, ii));  // Complex::norm is hypot
    QI_RET(Status{});
}

qi_status qi_state_scale(qi_state_t s, double re, double im, qi_error *err) {
    QI_NULLCHK(s);
    QI_LOCK(s->s.ctx);
    Status e = launch_scale(s->s, re, im, s->s.ctx->stream);
    if (!e.ok()) QI_RET(e);
    QI_RET(device_check(s->s.ctx->stream));
}

qi_status qi_state_axpy(qi_state_t y, double re, double im, qi_state_t x, qi_error *err) {
    QI_NULLCHK(y); QI_NULLCHK(x);
    Status e = check_pair(y->s, x->s);
    if (!e.ok()) QI_RET(e);
    QI_LOCK(y->s.ctx);
    if (!(e = launch_axpy(y->s, re, im, x->s, y->s.ctx->stream)).ok()) QI_RET(e);
    QI_RET(device_check(y->s.ctx->stream));
}

static qi_status addsub(qi_state_t a, qi_state_t b, qi_state_t out, bool sub, qi_error *err) {
    QI_NULLCHK(a); QI_NULLCHK(b); QI_NULLCHK(out);
    Status e = check_pair(a->s, b->s);  // the reference panics on a qubit-count mismatch
    if (!e.ok()) QI_RET(e);
    if (!(e = check_pair(a->s, out->s)).ok()) QI_RET(e);
    QI_LOCK(a->s.ctx);
    if (!(e = launch_addsub(out->s, a->s, b->s, sub, a->s.ctx->stream)).ok()) QI_RET(e);
    QI_RET(device_check(a->s.ctx->stream));
}

qi_status qi_state_add(qi_state_t a, qi_state_t b, qi_state_t out, qi_error *err) { return addsub(a, b, out, false, err); }
qi_status qi_state_sub(qi_state_t a, qi_state_t b, qi_state_t out, qi_error *err) { return addsub(a, b, out, true, err); }

qi_status qi_measure_probs(qi_state_t s, const uint64_t *qubits, uint32_t m, double *probs, qi_error *err) {
    QI_NULLCHK(s); QI_NULLCHK(probs);
    State &S = s->s;
    std::vector<uint64_t> all;
    if (m == 0) { for (uint32_t i = 0; i < S.n; i++) all.push_back(i); qubits = all.data(); m = S.n; }
    QI_NULLCHK(qubits);
    Status e = check_measure(S.n, qubits, m);
    if (!e.ok()) QI_RET(e);
    QI_LOCK(S.ctx);
    double *hm = nullptr;
    if (!(e = measure_probs_device(S, qubits, m, &hm, S.ctx->stream)).ok()) QI_RET(e);
    hipError_t he = hipMemcpyAsync(probs, hm, (size_t(1) << m) * sizeof(double), hipMemcpyDeviceToHost, S.ctx->stream);
    if (he != hipSuccess) QI_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(he)));
    QI_RET(device_check(S.ctx->stream));
}

qi_status qi_measure(qi_state_t s, qi_basis basis, const double *custom_u, const uint64_t *qubits, uint32_t m,
                     double uniform, qi_state_t out_state, uint64_t *outcome, qi_error *err) {
    QI_NULLCHK(s); QI_NULLCHK(out_state); QI_NULLCHK(outcome);
    if (m && !qubits) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "qubits is NULL"));
    if (basis < QI_BASIS_COMPUTATIONAL || basis > QI_BASIS_CUSTOM) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "basis"));
    if (basis == QI_BASIS_CUSTOM && !custom_u) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "custom basis is NULL"));
    Status e = check_pair(s->s, out_state->s);
    if (!e.ok()) QI_RET(e);
    QI_LOCK(s->s.ctx);
    if (!(e = measure_one(s->s, basis, custom_u, qubits, m, uniform, out_state->s, outcome)).ok()) QI_RET(e);
    QI_RET(device_check(s->s.ctx->stream));
}

qi_status qi_measure_n(qi_state_t s, qi_basis basis, const double *custom_u, const uint64_t *qubits, uint32_t m,
                       const double *uniforms, size_t shots, uint64_t *outcomes, qi_error *err) {
    QI_NULLCHK(s);
    if (shots == 0) QI_RET(make_err(QI_ERR_INVALID_NUMBER_OF_MEASUREMENTS, 0));  // state.rs:756-758
    QI_NULLCHK(outcomes);
    if (m && !qubits) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "qubits is NULL"));
    if (basis < QI_BASIS_COMPUTATIONAL || basis > QI_BASIS_CUSTOM) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "basis"));
    if (basis == QI_BASIS_CUSTOM && !custom_u) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "custom basis is NULL"));
    State &S = s->s;
    std::vector<uint64_t> all;
    if (m == 0) { for (uint32_t i = 0; i < S.n; i++) all.push_back(i); qubits = all.data(); m = S.n; }
    Status e = check_measure(S.n, qubits, m);
    if (!e.ok()) QI_RET(e);
    if (basis == QI_BASIS_CUSTOM && !(e = check_unitary(custom_u)).ok()) QI_RET(e);
    for (size_t k = 0; uniforms && k < shots; k++)
        if (!(uniforms[k] >= 0.0 && uniforms[k] < 1.0)) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "uniform draw outside [0,1)"));
    QI_LOCK(S.ctx);
    std::vector<double> drawn;
    if (!uniforms) {  // one draw of the context's stream per shot (rand::rng(), state.rs:601-602)
        drawn.resize(shots);
        for (double &u : drawn) u = splitmix_uniform(S.ctx->rng);
        uniforms = drawn.data();
    }
    // every shot measures the same state: one histogram and scan, many searches
    TmpState w;
    State *src = &S;
    if (basis != QI_BASIS_COMPUTATIONAL) {
        if (!(e = alloc_state(*S.ctx, S.n, S.prec, w.s)).ok()) QI_RET(e);
        if (!(e = launch_copy(w.s, S, S.ctx->stream)).ok()) QI_RET(e);
        if (!(e = basis_in(w.s, basis, custom_u, qubits, m)).ok()) QI_RET(e);
        src = &w.s;
    }
    QI_RET(measure_sample_collapse(*src, *src, qubits, m, uniforms, shots, outcomes, false, S.ctx->stream));
}

qi_status qi_measure_collapse(qi_state_t src, qi_state_t dst, const uint64_t *qubits, uint32_t m, uint64_t outcome,
                              qi_error *err) {
    QI_NULLCHK(src); QI_NULLCHK(dst);
    Status e = check_pair(src->s, dst->s);
    if (!e.ok()) QI_RET(e);
    State &S = src->s;
    std::vector<uint64_t> all;
    if (m == 0) { for (uint32_t i = 0; i < S.n; i++) all.push_back(i); qubits = all.data(); m = S.n; }
    if (!(e = check_measure(S.n, qubits, m)).ok()) QI_RET(e);
    if (m < 64 && outcome >= (uint64_t(1) << m)) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "outcome out of range"));
    // the outcome's bits must agree on repeated qubits
    for (uint32_t a = 0; a < m; a++)
        for (uint32_t b = a + 1; b < m; b++)
            if (qubits[a] == qubits[b] && (((outcome >> a) ^ (outcome >> b)) & 1))
                QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "outcome inconsistent on repeated qubits"));
    QI_LOCK(S.ctx);
    if (!(e = collapse_outcome(S, dst->s, qubits, m, outcome, S.ctx->stream)).ok()) QI_RET(e);
    QI_RET(check_normalised(dst->s));
}

qi_status qi_measure_result_state(qi_state_t src, qi_basis basis, const double *custom_u, const uint64_t *qubits,
                                  uint32_t m, uint64_t outcome, qi_state_t dst, qi_error *err) {
    QI_NULLCHK(src); QI_NULLCHK(dst);
    if (m && !qubits) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "qubits is NULL"));
    if (basis < QI_BASIS_COMPUTATIONAL || basis > QI_BASIS_CUSTOM) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "basis"));
    if (basis == QI_BASIS_CUSTOM && !custom_u) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "custom basis is NULL"));
    Status e = check_pair(src->s, dst->s);
    if (!e.ok()) QI_RET(e);
    State &S = src->s;
    std::vector<uint64_t> all;
    if (m == 0) { for (uint32_t i = 0; i < S.n; i++) all.push_back(i); qubits = all.data(); m = S.n; }
    if (!(e = check_measure(S.n, qubits, m)).ok()) QI_RET(e);
    if (basis == QI_BASIS_CUSTOM && !(e = check_unitary(custom_u)).ok()) QI_RET(e);
    if (m < 64 && outcome >= (uint64_t(1) << m)) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "outcome out of range"));
    QI_LOCK(S.ctx);
    if (!(e = measured_state(S, basis, custom_u, qubits, m, outcome, dst->s)).ok()) QI_RET(e);
    QI_RET(device_check(S.ctx->stream));
}

qi_status qi_param_circuit_create(qi_ctx_t ctx, uint32_t n, qi_prec prec, const qi_gate *gates, size_t ngates,
                                  const int32_t *angle_slots, const qi_exec_opts *opts, qi_param_circuit_t *out,
                                  qi_error *err) {
    QI_NULLCHK(ctx); QI_NULLCHK(out);
    if (ngates && !gates) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "gates is NULL"));
    if (n == 0) QI_RET(make_err(QI_ERR_INVALID_NUMBER_OF_QUBITS, 0));
    if (n > QI_MAX_QUBITS) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "more qubits than the engine supports"));
    if (prec != QI_C64 && prec != QI_C128) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "precision"));
    std::unique_ptr<qi_param_circuit_s> pc(new qi_param_circuit_s());
    pc->ctx = &ctx->c; pc->n = n; pc->prec = prec;
    pc->gates.assign(gates, gates + ngates);
    pc->slots.assign(ngates, -1);
    Status e;
    for (size_t g = 0; g < ngates; g++) {
        if (!is_operator(gates[g].kind)) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, g, 0, 0, 0, "parametric circuits hold operator gates only"));
        if (!(e = validate_gate(n, gates[g])).ok()) QI_RET(e);
        const int32_t sl = angle_slots ? angle_slots[g] : -1;
        if (sl < 0) continue;
        const int k = gates[g].kind;
        if (k != QI_GATE_RX && k != QI_GATE_RY && k != QI_GATE_RZ && k != QI_GATE_P)
            QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, g, 0, 0, 0, "only RX, RY, RZ and P gates take an angle slot"));
        if (sl >= (1 << 24)) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, g, 0, 0, 0, "angle slot out of range"));
        pc->slots[g] = sl;
        if ((size_t)sl >= pc->angles.size()) pc->angles.resize((size_t)sl + 1, 0.0);
    }
    // the angle table starts from the gates' own angles (first gate of each slot)
    std::vector<char> seen(pc->angles.size(), 0);
    for (size_t g = 0; g < ngates; g++)
        if (pc->slots[g] >= 0 && !seen[(size_t)pc->slots[g]]) {
            seen[(size_t)pc->slots[g]] = 1;
            pc->angles[(size_t)pc->slots[g]] = gates[g].params[0];
        }
    qi_exec_opts o{};
    o.mode = QI_EXEC_AUTO;
    if (opts) o = *opts;
    int mode = o.mode;
    if (mode == QI_EXEC_AUTO) mode = n >= 12 ? QI_EXEC_TILED : QI_EXEC_PER_GATE;
    if (mode != QI_EXEC_TILED && mode != QI_EXEC_PER_GATE)
        QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "parametric circuits run tiled or per gate"));
    State shape;
    shape.ctx = &ctx->c; shape.n = n; shape.prec = prec;
    if (!(e = build_plan(shape, pc->gates.data(), ngates, mode, o, pc->plan, pc->slots.data())).ok()) QI_RET(e);
    if (!pc->plan.mats.empty()) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "dense blocks are not parametric"));
    for (const ExecStep &st : pc->plan.steps) pc->graph_ok = pc->graph_ok && !(st.type == ExecStep::GATE && st.L.slot >= 0);
    patch_param_plan(*pc);
    {
        QI_LOCK(&ctx->c);
        if (!pc->plan.reg_ops.empty() && !(e = ensure(pc->blob, pc->plan.reg_ops.size() * sizeof(RegOp))).ok()) QI_RET(e);
    }
    *out = pc.release();
    QI_RET(Status{});
}

qi_status qi_param_circuit_destroy(qi_param_circuit_t pc) {
    if (!pc) return QI_OK;
    {
        QI_LOCK(pc->ctx);
        (void)hipStreamSynchronize(pc->ctx->stream);
        if (pc->graph) (void)hipGraphExecDestroy(pc->graph);
        if (pc->blob.p) (void)hipFree(pc->blob.p);
    }
    delete pc;
    return QI_OK;
}

qi_status qi_param_circuit_num_slots(qi_param_circuit_t pc, uint64_t *nslots, qi_error *err) {
    QI_NULLCHK(pc); QI_NULLCHK(nslots);
    *nslots = pc->angles.size();
    QI_RET(Status{});
}

qi_status qi_param_circuit_set_angles(qi_param_circuit_t pc, const double *angles, size_t nangles, qi_error *err) {
    QI_NULLCHK(pc);
    if (nangles != pc->angles.size())
        QI_RET(make_err(QI_ERR_MISMATCHED_PARAMETERS, pc->angles.size(), nangles));  // MismatchedNumberOfParameters{expected, actual}
    if (nangles) QI_NULLCHK(angles);
    QI_LOCK(pc->ctx);
    std::copy(angles, angles + nangles, pc->angles.begin());
    patch_param_plan(*pc);
    QI_RET(Status{});
}

qi_status qi_param_circuit_run(qi_param_circuit_t pc, qi_state_t s, qi_error *err) {
    QI_NULLCHK(pc); QI_NULLCHK(s);
    State &S = s->s;
    if (S.n != pc->n) QI_RET(make_err(QI_ERR_INVALID_NUMBER_OF_QUBITS, S.n));  // circuit.rs:161-163
    if (S.prec != pc->prec) QI_RET(make_err(QI_ERR_PRECISION_MISMATCH));
    if (S.ctx != pc->ctx) QI_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "state on another context"));
    QI_LOCK(pc->ctx);
    Ctx &C = *pc->ctx;
    hipStream_t st = C.stream;
    Status e;
    if (pc->dirty && !pc->plan.reg_ops.empty()) {
        // the previous run may still read the op table: the stream orders the copy after it
        hipError_t he = hipMemcpyAsync(pc->blob.p, pc->plan.reg_ops.data(), pc->plan.reg_ops.size() * sizeof(RegOp),
                                       hipMemcpyHostToDevice, st);
        if (he == hipSuccess) he = hipStreamSynchronize(st);  // the host table may change before the next run
        if (he != hipSuccess) QI_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(he)));
    }
    pc->dirty = false;
    QI_HIP_TRY_C(hipEventRecord(C.ev0, st));
    if (!pc->graph_ok) {
        if (!(e = launch_plan(S, pc->plan, pc->gates.data(), pc->blob.p)).ok()) QI_RET(e);
    } else {
        if (!pc->graph || pc->graph_state != S.d) {
            if (pc->graph) { (void)hipGraphExecDestroy(pc->graph); pc->graph = nullptr; }
            hipGraph_t graph = nullptr;
            QI_HIP_TRY_C(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            e = launch_plan(S, pc->plan, pc->gates.data(), pc->blob.p);
            hipError_t ce = hipStreamEndCapture(st, &graph);
            if (!e.ok()) { if (graph) (void)hipGraphDestroy(graph); QI_RET(e); }
            if (ce != hipSuccess) QI_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(ce)));
            hipError_t ie = hipGraphInstantiate(&pc->graph, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) { pc->graph = nullptr; QI_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(ie))); }
            pc->graph_state = S.d;
        }
        QI_HIP_TRY_C(hipGraphLaunch(pc->graph, st));
    }
    QI_HIP_TRY_C(hipEventRecord(C.ev1, st));
    if (!(e = device_check(st)).ok()) QI_RET(e);
    float ms = 0.f;
    QI_HIP_TRY_C(hipEventElapsedTime(&ms, C.ev0, C.ev1));
    C.last_exec_ms = ms;
    C.last_plan_steps = pc->plan.steps.size();
    QI_RET(Status{});
}

qi_status qi_last_exec_ms(qi_ctx_t ctx, double *ms, qi_error *err) {
    QI_NULLCHK(ctx); QI_NULLCHK(ms);
    *ms = ctx->c.last_exec_ms;
    QI_RET(Status{});
}

qi_status qi_last_exec_steps(qi_ctx_t ctx, uint64_t *steps, qi_error *err) {
    QI_NULLCHK(ctx); QI_NULLCHK(steps);
    *steps = ctx->c.last_plan_steps;
    QI_RET(Status{});
}

}  // extern "C"
