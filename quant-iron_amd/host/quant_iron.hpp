// quant_iron.hpp — C++ host mirror of quant-iron's state-vector API over the
// C ABI (include/qiron.h). Header-only; link with libqiron.so.
//
// Names and argument order follow the reference (src/components/state.rs,
// gate.rs, pauli_string.rs, circuit.rs, algorithms/time_evolution.rs): every
// State method returns a new State, errors are thrown as quant_iron::Error
// carrying the reference's variant and payload (src/errors.rs:4-97).
#pragma once

#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/qiron.h"

namespace quant_iron {

using C64 = std::complex<double>;

class Error : public std::runtime_error {
  public:
    explicit Error(const qi_error &e)
        : std::runtime_error(e.msg), code(static_cast<qi_status>(e.code)), a(e.a), b(e.b), coefficient(e.re, e.im) {}
    qi_status code;
    uint64_t a, b;
    C64 coefficient;
};

inline void check(qi_status st, const qi_error &e) {
    if (st != QI_OK) throw Error(e);
}

class Context {
  public:
    explicit Context(int device = 0) {
        qi_error e{};
        check(qi_ctx_create(device, &h_, &e), e);
    }
    ~Context() { qi_ctx_destroy(h_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    qi_ctx_t get() const { return h_; }
    static Context &global() {
        static Context c(0);
        return c;
    }

  private:
    qi_ctx_t h_ = nullptr;
};

enum class Precision { C64 = QI_C64, C128 = QI_C128 };

struct MeasurementBasis {
    qi_basis kind = QI_BASIS_COMPUTATIONAL;
    C64 u[2][2] = {{1, 0}, {0, 1}};
    static MeasurementBasis computational() { return {}; }
    static MeasurementBasis x() { MeasurementBasis b; b.kind = QI_BASIS_X; return b; }
    static MeasurementBasis y() { MeasurementBasis b; b.kind = QI_BASIS_Y; return b; }
    static MeasurementBasis custom(const C64 (&m)[2][2]) {
        MeasurementBasis b;
        b.kind = QI_BASIS_CUSTOM;
        for (int r = 0; r < 2; r++)
            for (int c = 0; c < 2; c++) b.u[r][c] = m[r][c];
        return b;
    }
    void flat(double *m8) const {
        for (int k = 0; k < 4; k++) { m8[2 * k] = u[k / 2][k % 2].real(); m8[2 * k + 1] = u[k / 2][k % 2].imag(); }
    }
};

class State;

struct MeasurementResult {
    MeasurementBasis basis;
    std::vector<size_t> indices;
    std::vector<uint8_t> outcomes;
    std::shared_ptr<State> new_state;
};

inline qi_gate make_gate(qi_gate_kind kind, std::vector<size_t> targets, std::vector<size_t> controls = {},
                         std::vector<double> params = {}, const C64 (*m)[2] = nullptr) {
    qi_gate g;
    std::memset(&g, 0, sizeof g);
    g.kind = kind;
    g.ntargets = (uint32_t)targets.size();
    for (size_t i = 0; i < targets.size() && i < QI_MAX_TARGETS; i++) g.targets[i] = targets[i];
    if (controls.size() > QI_MAX_CONTROLS) throw std::invalid_argument("too many controls");
    g.ncontrols = (uint32_t)controls.size();
    for (size_t i = 0; i < controls.size(); i++) g.controls[i] = controls[i];
    for (size_t i = 0; i < params.size() && i < 3; i++) g.params[i] = params[i];
    if (m)
        for (int k = 0; k < 4; k++) { g.matrix[2 * k] = m[k / 2][k % 2].real(); g.matrix[2 * k + 1] = m[k / 2][k % 2].imag(); }
    return g;
}

class Gate {
  public:
    explicit Gate(const qi_gate &g) : g_(g) {}
    const qi_gate &raw() const { return g_; }
    State apply(const State &s) const;
    static Gate h_gate(size_t q) { return Gate(make_gate(QI_GATE_H, {q})); }
    static Gate x_gate(size_t q) { return Gate(make_gate(QI_GATE_X, {q})); }
    static Gate y_gate(size_t q) { return Gate(make_gate(QI_GATE_Y, {q})); }
    static Gate z_gate(size_t q) { return Gate(make_gate(QI_GATE_Z, {q})); }
    static Gate rx_gate(size_t q, double a) { return Gate(make_gate(QI_GATE_RX, {q}, {}, {a})); }
    static Gate ry_gate(size_t q, double a) { return Gate(make_gate(QI_GATE_RY, {q}, {}, {a})); }
    static Gate rz_gate(size_t q, double a) { return Gate(make_gate(QI_GATE_RZ, {q}, {}, {a})); }
    static Gate p_gate(size_t q, double a) { return Gate(make_gate(QI_GATE_P, {q}, {}, {a})); }
    static Gate cnot_gate(size_t target, size_t control) { return Gate(make_gate(QI_GATE_CNOT, {target}, {control})); }
    static Gate swap_gate(size_t a, size_t b) { return Gate(make_gate(QI_GATE_SWAP, {a, b})); }
    static Gate toffoli_gate(size_t target, std::vector<size_t> controls) {
        return Gate(make_gate(QI_GATE_TOFFOLI, {target}, std::move(controls)));
    }
    static std::vector<Gate> p_controlled_gates(std::vector<size_t> targets, std::vector<size_t> controls, double angle) {
        std::vector<Gate> out;
        for (size_t t : targets) out.emplace_back(make_gate(QI_GATE_P, {t}, controls, {angle}));
        return out;
    }

  private:
    qi_gate g_;
};

class State {
  public:
    State(uint32_t n, Precision p, Context &ctx = Context::global()) : ctx_(&ctx) {
        qi_error e{};
        check(qi_state_create(ctx.get(), n, static_cast<qi_prec>(p), &h_, &e), e);
    }
    ~State() { qi_state_destroy(h_); }
    State(const State &o) : ctx_(o.ctx_) {
        qi_error e{};
        check(qi_state_clone(o.h_, &h_, &e), e);
    }
    State(State &&o) noexcept : ctx_(o.ctx_), h_(o.h_) { o.h_ = nullptr; }
    State &operator=(State o) { std::swap(h_, o.h_); std::swap(ctx_, o.ctx_); return *this; }

    // constructors (state.rs:99-300)
    static State init(uint32_t n, qi_init k, uint64_t basis = 0, Precision p = Precision::C128) {
        if (n == 0) { qi_error e{}; e.code = QI_ERR_INVALID_NUMBER_OF_QUBITS; std::strcpy(e.msg, "Invalid number of qubits: 0"); throw Error(e); }
        State s(n, p);
        qi_error e{};
        check(qi_state_init(s.h_, k, basis, &e), e);
        return s;
    }
    static State new_zero(uint32_t n, Precision p = Precision::C128) { return init(n, QI_INIT_ZERO, 0, p); }
    static State new_basis_n(uint32_t n, uint64_t i, Precision p = Precision::C128) { return init(n, QI_INIT_BASIS, i, p); }
    static State new_plus(uint32_t n, Precision p = Precision::C128) { return init(n, QI_INIT_PLUS, 0, p); }
    static State new_minus(uint32_t n, Precision p = Precision::C128) { return init(n, QI_INIT_MINUS, 0, p); }
    static State new_ghz(uint32_t n, Precision p = Precision::C128) { return init(n, QI_INIT_GHZ, 0, p); }
    static State new_(const std::vector<C64> &v, Precision p = Precision::C128) {
        size_t len = v.size();
        qi_error e{};
        if (len == 0 || (len & (len - 1))) {
            e.code = QI_ERR_INVALID_NUMBER_OF_QUBITS;
            e.a = len ? (uint64_t)std::floor(std::log2((double)len)) : 0;
            std::snprintf(e.msg, sizeof e.msg, "Invalid number of qubits: %llu", (unsigned long long)e.a);
            throw Error(e);
        }
        State s((uint32_t)__builtin_ctzll(len), p);
        check(qi_state_upload(s.h_, reinterpret_cast<const double *>(v.data()), 1, &e), e);
        return s;
    }

    uint32_t num_qubits() const { return qi_state_num_qubits(h_); }
    qi_state_t handle() const { return h_; }
    std::vector<C64> state_vector() const {
        std::vector<C64> v(size_t(1) << num_qubits());
        qi_error e{};
        check(qi_state_download(h_, reinterpret_cast<double *>(v.data()), &e), e);
        return v;
    }

    State &apply_inplace(const qi_gate &g) {
        qi_error e{};
        check(qi_apply_gate(h_, &g, &e), e);
        return *this;
    }
    State op(qi_gate_kind k, std::vector<size_t> t, std::vector<size_t> c = {}, std::vector<double> p = {},
             const C64 (*m)[2] = nullptr) const {
        State s(*this);
        s.apply_inplace(make_gate(k, std::move(t), std::move(c), std::move(p), m));
        return s;
    }
    State multi(qi_gate_kind k, const std::vector<size_t> &qs, std::vector<size_t> c = {}, std::vector<double> p = {},
                const C64 (*m)[2] = nullptr) const {
        State s(*this);
        for (size_t q : qs) s.apply_inplace(make_gate(k, {q}, c, p, m));
        return s;
    }

    // gate methods (state.rs:1019-2345)
    State h(size_t i) const { return op(QI_GATE_H, {i}); }
    State h_multi(const std::vector<size_t> &q) const { return multi(QI_GATE_H, q); }
    State ch_multi(const std::vector<size_t> &t, const std::vector<size_t> &c) const { return multi(QI_GATE_H, t, c); }
    State x(size_t i) const { return op(QI_GATE_X, {i}); }
    State cx_multi(const std::vector<size_t> &t, const std::vector<size_t> &c) const { return multi(QI_GATE_X, t, c); }
    State y(size_t i) const { return op(QI_GATE_Y, {i}); }
    State z(size_t i) const { return op(QI_GATE_Z, {i}); }
    State cz_multi(const std::vector<size_t> &t, const std::vector<size_t> &c) const { return multi(QI_GATE_Z, t, c); }
    State i(size_t q) const { return op(QI_GATE_I, {q}); }
    State s(size_t q) const { return op(QI_GATE_S, {q}); }
    State t(size_t q) const { return op(QI_GATE_T, {q}); }
    State s_dag(size_t q) const { return op(QI_GATE_SDAG, {q}); }
    State t_dag(size_t q) const { return op(QI_GATE_TDAG, {q}); }
    State p(size_t q, double a) const { return op(QI_GATE_P, {q}, {}, {a}); }
    State cp_multi(const std::vector<size_t> &t, const std::vector<size_t> &c, double a) const {
        return multi(QI_GATE_P, t, c, {a});
    }
    State rx(size_t q, double a) const { return op(QI_GATE_RX, {q}, {}, {a}); }
    State rx_multi(const std::vector<size_t> &q, double a) const { return multi(QI_GATE_RX, q, {}, {a}); }
    State crx_multi(const std::vector<size_t> &t, const std::vector<size_t> &c, double a) const {
        return multi(QI_GATE_RX, t, c, {a});
    }
    State ry(size_t q, double a) const { return op(QI_GATE_RY, {q}, {}, {a}); }
    State rz(size_t q, double a) const { return op(QI_GATE_RZ, {q}, {}, {a}); }
    State unitary(size_t q, const C64 (&m)[2][2]) const { return op(QI_GATE_U2, {q}, {}, {}, m); }
    State cnot(size_t control, size_t target) const { return op(QI_GATE_CNOT, {target}, {control}); }
    State swap(size_t a, size_t b) const { return op(QI_GATE_SWAP, {a, b}); }
    State cswap(size_t a, size_t b, const std::vector<size_t> &c) const { return op(QI_GATE_SWAP, {a, b}, c); }
    State matchgate(size_t t, double th, double p1, double p2) const { return op(QI_GATE_MATCHGATE, {t}, {}, {th, p1, p2}); }
    State toffoli(size_t c1, size_t c2, size_t t) const { return op(QI_GATE_TOFFOLI, {t}, {c1, c2}); }

    // reductions (state.rs:470-498, 890-945)
    C64 inner_product(const State &o) const {
        double re, im;
        qi_error e{};
        check(qi_inner_product(h_, o.h_, &re, &im, &e), e);
        return {re, im};
    }
    State normalise() const {
        State s(*this);
        qi_error e{};
        check(qi_normalise(s.h_, &e), e);
        return s;
    }
    double fs_fidelity(const State &o) const {
        double f;
        qi_error e{};
        check(qi_fidelity(h_, o.h_, &f, &e), e);
        return f;
    }
    double fs_dist(const State &o) const {
        double d;
        qi_error e{};
        check(qi_fs_dist(h_, o.h_, &d, &e), e);
        return d;
    }
    State operator*(C64 c) const {
        State s(*this);
        qi_error e{};
        check(qi_state_scale(s.h_, c.real(), c.imag(), &e), e);
        return s;
    }
    // impl Add / Sub for State (state.rs:2773-2800, 2836-2862)
    State operator+(const State &o) const {
        State s(num_qubits(), precision());
        qi_error e{};
        check(qi_state_add(h_, o.h_, s.h_, &e), e);
        return s;
    }
    State operator-(const State &o) const {
        State s(num_qubits(), precision());
        qi_error e{};
        check(qi_state_sub(h_, o.h_, s.h_, &e), e);
        return s;
    }

    // measurement (state.rs:525-784), uniform draws injected
    MeasurementResult measure(const MeasurementBasis &b, const std::vector<size_t> &qubits, double uniform) const {
        std::vector<uint64_t> q(qubits.begin(), qubits.end());
        double m8[8];
        b.flat(m8);
        auto out = std::make_shared<State>(num_qubits(), precision());
        uint64_t o = 0;
        qi_error e{};
        check(qi_measure(h_, b.kind, m8, q.empty() ? nullptr : q.data(), (uint32_t)q.size(), uniform, out->h_, &o, &e), e);
        MeasurementResult r;
        r.basis = b;
        if (qubits.empty()) for (size_t k = 0; k < num_qubits(); k++) r.indices.push_back(k);
        else r.indices = qubits;
        for (size_t k = 0; k < r.indices.size(); k++) r.outcomes.push_back((uint8_t)((o >> k) & 1));
        r.new_state = out;
        return r;
    }
    // measure_n outcomes (bit b = qubits[b]) for injected uniforms
    std::vector<uint64_t> measure_n_outcomes(const MeasurementBasis &b, const std::vector<size_t> &qubits,
                                             const std::vector<double> &uniforms) const {
        std::vector<uint64_t> q(qubits.begin(), qubits.end()), out(uniforms.size());
        double m8[8];
        b.flat(m8);
        qi_error e{};
        check(qi_measure_n(h_, b.kind, m8, q.empty() ? nullptr : q.data(), (uint32_t)q.size(), uniforms.data(),
                           uniforms.size(), out.data(), &e), e);
        return out;
    }
    // State::measure_n (state.rs:750-784): one MeasurementResult per shot, the
    // collapsed state computed once per distinct outcome
    std::vector<MeasurementResult> measure_n(const MeasurementBasis &b, const std::vector<size_t> &qubits,
                                             const std::vector<double> &uniforms) const {
        auto outs = measure_n_outcomes(b, qubits, uniforms);
        std::vector<uint64_t> q(qubits.begin(), qubits.end());
        double m8[8];
        b.flat(m8);
        std::vector<MeasurementResult> res;
        std::vector<std::pair<uint64_t, std::shared_ptr<State>>> seen;
        for (uint64_t o : outs) {
            MeasurementResult r;
            r.basis = b;
            if (qubits.empty()) for (size_t k = 0; k < num_qubits(); k++) r.indices.push_back(k);
            else r.indices = qubits;
            for (size_t k = 0; k < r.indices.size(); k++) r.outcomes.push_back((uint8_t)((o >> k) & 1));
            for (auto &p : seen) if (p.first == o) r.new_state = p.second;
            if (!r.new_state) {
                r.new_state = std::make_shared<State>(num_qubits(), precision());
                qi_error e{};
                check(qi_measure_result_state(h_, b.kind, m8, q.empty() ? nullptr : q.data(), (uint32_t)q.size(), o,
                                              r.new_state->h_, &e), e);
                seen.emplace_back(o, r.new_state);
            }
            res.push_back(std::move(r));
        }
        return res;
    }

    Precision precision() const { return static_cast<Precision>(qi_state_precision(h_)); }

    // PartialEq (state.rs:2348-2372): |dre|, |dim| <= f32::EPSILON
    bool operator==(const State &o) const {
        if (num_qubits() != o.num_qubits()) return false;
        auto a = state_vector(), b = o.state_vector();
        const double eps = 1.1920928955078125e-07;
        for (size_t k = 0; k < a.size(); k++)
            if (std::abs(a[k].real() - b[k].real()) > eps || std::abs(a[k].imag() - b[k].imag()) > eps) return false;
        return true;
    }

  private:
    Context *ctx_;
    qi_state_t h_ = nullptr;
};

inline State Gate::apply(const State &s) const {
    State o(s);
    o.apply_inplace(g_);
    return o;
}

class PauliString {
  public:
    explicit PauliString(C64 c = 1.0) : coefficient_(c) {}
    PauliString &with_op(size_t q, char op) {
        if (ops_.count(q)) throw std::logic_error("Duplicate Pauli string operator for qubit");
        ops_[q] = op;
        return *this;
    }
    C64 coefficient() const { return coefficient_; }
    qi_pauli_term term() const {
        qi_pauli_term t{0, 0, coefficient_.real(), coefficient_.imag()};
        for (auto &kv : ops_) {
            if (kv.second == 'X' || kv.second == 'Y') t.x_mask |= uint64_t(1) << kv.first;
            if (kv.second == 'Z' || kv.second == 'Y') t.z_mask |= uint64_t(1) << kv.first;
        }
        return t;
    }
    State apply(const State &s) const {
        State out(s);
        qi_pauli_term t = term();
        qi_error e{};
        check(qi_pauli_apply(s.handle(), out.handle(), &t, &e), e);
        return out;
    }
    State apply_normalised(const State &s) const {
        State out(s);
        qi_pauli_term t = term();
        qi_error e{};
        check(qi_pauli_apply_normalised(s.handle(), out.handle(), &t, &e), e);
        return out;
    }
    State apply_exp_factor(const State &s, C64 f) const {
        State out(s);
        qi_pauli_term t = term();
        qi_error e{};
        check(qi_pauli_exp(out.handle(), &t, f.real(), f.imag(), &e), e);
        return out;
    }
    State apply_exp(const State &s) const { return apply_exp_factor(s, 1.0); }
    State apply_exp_neg_i_dt(const State &s, double dt) const {
        State out(s);
        qi_pauli_term t = term();
        qi_error e{};
        check(qi_pauli_exp_neg_i_dt(out.handle(), &t, dt, &e), e);
        return out;
    }

  private:
    C64 coefficient_;
    std::map<size_t, char> ops_;
};

class SumOp {
  public:
    explicit SumOp(std::vector<PauliString> terms = {}) : terms_(std::move(terms)) {}
    std::vector<qi_pauli_term> raw() const {
        std::vector<qi_pauli_term> r;
        for (auto &t : terms_) r.push_back(t.term());
        return r;
    }
    State apply(const State &s) const {
        State out(s);
        auto r = raw();
        qi_error e{};
        check(qi_sumop_apply(s.handle(), out.handle(), r.data(), r.size(), &e), e);
        return out;
    }
    C64 expectation_value(const State &s) const {
        auto r = raw();
        double re, im;
        qi_error e{};
        check(qi_sumop_expectation(s.handle(), r.data(), r.size(), &re, &im, &e), e);
        return {re, im};
    }
    size_t num_terms() const { return terms_.size(); }

  private:
    std::vector<PauliString> terms_;
};

class Circuit {
  public:
    explicit Circuit(uint32_t n) : n_(n) {}
    void add_gate(const Gate &g) { gates_.push_back(g.raw()); }
    void add_gates(const std::vector<Gate> &gs) { for (auto &g : gs) gates_.push_back(g.raw()); }
    State execute(const State &s, const qi_exec_opts *opts = nullptr) const {
        State out(s);
        qi_error e{};
        check(qi_apply_gates(out.handle(), n_, gates_.data(), gates_.size(), opts, &e), e);
        return out;
    }
    size_t size() const { return gates_.size(); }

  private:
    uint32_t n_;
    std::vector<qi_gate> gates_;
};

enum class TrotterOrder { First = 1, Second = 2 };

inline State trotter_evolve_state(const SumOp &h, const State &s, double dt, uint64_t steps, TrotterOrder o) {
    State out(s);
    auto r = h.raw();
    qi_error e{};
    check(qi_trotter_evolve(out.handle(), r.data(), r.size(), dt, steps, static_cast<int>(o), &e), e);
    return out;
}

}  // namespace quant_iron
