// C++ mirror tests: reference known-answer tests restated against the C++ host
// API (quant-iron_amd/host/quant_iron.hpp), in the reference's own style
// (construct, apply, compare with State's f32-epsilon PartialEq).
// Exit status 0 = all passed. Built by the build step, run by a GPU test.
#include <cmath>
#include <cstdio>
#include <functional>
#include <vector>

#include "../../quant-iron_amd/host/quant_iron.hpp"

using namespace quant_iron;

static int failures = 0;
#define CHECK(cond)                                                                 \
    do {                                                                            \
        if (!(cond)) { std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond); failures++; } \
    } while (0)

template <typename F> static void expect_error(F f, qi_status code, uint64_t a, uint64_t b) {
    try {
        f();
        std::printf("FAIL: no error (expected code %d)\n", (int)code);
        failures++;
    } catch (const Error &e) {
        if (e.code != code || e.a != a || e.b != b) {
            std::printf("FAIL: error %d(%llu,%llu) expected %d(%llu,%llu)\n", (int)e.code, (unsigned long long)e.a,
                        (unsigned long long)e.b, (int)code, (unsigned long long)a, (unsigned long long)b);
            failures++;
        }
    }
}

int main() {
    const double PI = std::acos(-1.0), r2 = 1.0 / std::sqrt(2.0);
    // h(|0>) = |+>, h(|1>) = |-> (operator_tests.rs:17-122)
    CHECK(State::new_zero(1).h(0) == State::new_plus(1));
    CHECK(State::new_basis_n(1, 1).h(0) == State::new_minus(1));
    // rx(|0>) = cos(t/2)|0> - i sin(t/2)|1> (operator_tests.rs:1251-1275)
    {
        double th = PI / 2.5;
        State got = State::new_zero(1).rx(0, th);
        State exp = State::new_(std::vector<C64>{{std::cos(th / 2), 0}, {0, -std::sin(th / 2)}});
        CHECK(got == exp);
    }
    // CNOT(control 0, target 1)|01> = |11> (operator_tests.rs:1986-2068)
    CHECK(State::new_basis_n(2, 1).cnot(0, 1) == State::new_basis_n(2, 3));
    // SWAP, Toffoli
    CHECK(State::new_basis_n(2, 1).swap(0, 1) == State::new_basis_n(2, 2));
    CHECK(State::new_basis_n(3, 3).toffoli(0, 1, 2) == State::new_basis_n(3, 7));
    // S^dag S = I, T T = S
    {
        State p = State::new_plus(3).rx(1, 0.3);
        CHECK(p.s(2).s_dag(2) == p);
        CHECK(p.t(0).t(0) == p.s(0));
    }
    // PauliString (2+2i) X0 Y1 |11> = (2-2i)|00> (pauli_string_tests.rs:96-107)
    {
        PauliString ps(C64(2, 2));
        ps.with_op(0, 'X').with_op(1, 'Y');
        State got = ps.apply(State::new_basis_n(2, 3));
        CHECK(got == State::new_basis_n(2, 0) * C64(2, -2));
    }
    // exp(-i (pi/2) 0.5 Z0 X1)|00> (pauli_string_tests.rs:408-438)
    {
        PauliString ps(C64(0.5 * PI, 0));
        ps.with_op(0, 'Z').with_op(1, 'X');
        State got = ps.apply_exp_neg_i_dt(State::new_zero(2), 0.5);
        CHECK(got == State::new_(std::vector<C64>{{r2, 0}, {0, 0}, {0, -r2}, {0, 0}}));
    }
    // <11| 2X0 + 3Y1 + 4Z1 |11> = -4 (pauli_string_tests.rs:348-364)
    {
        PauliString a(2.0), b(3.0), c(4.0);
        a.with_op(0, 'X'); b.with_op(1, 'Y'); c.with_op(1, 'Z');
        C64 e = SumOp({a, b, c}).expectation_value(State::new_basis_n(2, 3));
        CHECK(std::abs(e - C64(-4.0, 0)) < 1e-12);
    }
    // Circuit::execute H H |00> = |++> and the qubit-count check (circuit_tests.rs:84-107)
    {
        Circuit cz(2);
        cz.add_gate(Gate::h_gate(0));
        cz.add_gate(Gate::h_gate(1));
        CHECK(cz.execute(State::new_zero(2)) == State::new_plus(2));
        expect_error([&] { cz.execute(State::new_zero(1)); }, QI_ERR_INVALID_NUMBER_OF_QUBITS, 1, 0);
    }
    // Trotter, H = X0 + Y1 from |10>, 3 first-order steps (time_evolution_tests.rs)
    {
        PauliString x0(1.0), y1(1.0);
        x0.with_op(0, 'X'); y1.with_op(1, 'Y');
        State got = trotter_evolve_state(SumOp({x0, y1}), State::new_basis_n(2, 2), 0.1, 3, TrotterOrder::First);
        double t = 0.3;
        std::vector<C64> q0{{std::cos(t), 0}, {0, -std::sin(t)}}, q1{{-std::sin(t), 0}, {std::cos(t), 0}}, v;
        for (auto &hi : q1) for (auto &lo : q0) v.push_back(hi * lo);
        CHECK(got == State::new_(v));
    }
    // measurement with an injected uniform: |+> measured with u = 0.25 -> 0, 0.75 -> 1
    {
        auto r0 = State::new_plus(1).measure(MeasurementBasis::computational(), {0}, 0.25);
        auto r1 = State::new_plus(1).measure(MeasurementBasis::computational(), {0}, 0.75);
        CHECK(r0.outcomes[0] == 0 && *r0.new_state == State::new_zero(1));
        CHECK(r1.outcomes[0] == 1 && *r1.new_state == State::new_basis_n(1, 1));
        auto outs = State::new_plus(2).measure_n_outcomes(MeasurementBasis::computational(), {}, {0.1, 0.3, 0.6, 0.9});
        CHECK(outs == (std::vector<uint64_t>{0, 1, 2, 3}));
        // per-shot results: shot k collapses |++> onto basis state k
        auto shots = State::new_plus(2).measure_n(MeasurementBasis::computational(), {}, {0.1, 0.3, 0.6, 0.9});
        for (uint64_t k = 0; k < 4; k++)
            CHECK(shots[k].outcomes[0] == (k & 1) && shots[k].outcomes[1] == (k >> 1) &&
                  *shots[k].new_state == State::new_basis_n(2, k));
    }
    // Fubini-Study metrics (state_tests.rs:225-295) and State Add / Sub
    {
        CHECK(std::abs(State::new_plus(1).fs_fidelity(State::new_basis_n(1, 1)) - 0.5) < 1e-15);
        CHECK(std::abs(State::new_plus(2).fs_dist(State::new_plus(2))) < 1e-15);
        auto p = State::new_plus(1), m = State::new_minus(1);
        // |+> + |-> = sqrt(2)|0>, |+> - |-> = sqrt(2)|1>
        CHECK((p + m) == State::new_zero(1) * std::sqrt(2.0));
        CHECK((p - m) == State::new_basis_n(1, 1) * std::sqrt(2.0));
    }
    // error variants (operator_tests.rs:2513-2703)
    expect_error([] { State::new_zero(2).h(2); }, QI_ERR_INVALID_QUBIT_INDEX, 2, 2);
    expect_error([] { State::new_zero(2).cnot(0, 0); }, QI_ERR_OVERLAPPING_CONTROL_AND_TARGET, 0, 0);
    expect_error([] { State::new_zero(3).matchgate(2, 1.0, 0, 0); }, QI_ERR_INVALID_QUBIT_INDEX, 2, 3);
    expect_error([] { State::new_zero(0); }, QI_ERR_INVALID_NUMBER_OF_QUBITS, 0, 0);
    expect_error([] { State::new_(std::vector<C64>{{1, 0}, {1, 0}}); }, QI_ERR_STATE_NOT_NORMALISED, 0, 0);
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
