"""Write tests/golden/kats.json: known-answer cases of the reference's own tests.

Each case restates one assertion of the reference's test suite as data: the
initial state, the gates applied, and the expected amplitudes computed here
from the closed form that the reference test states (e.g. the comment
"rx(|0>) = cos(theta/2)|0> - i*sin(theta/2)|1>", operator_tests.rs:1251).
Expected values never come from the oracle or the engine, so they pin both.
The reference compares with State's PartialEq (|dre|, |dim| <= f32::EPSILON,
state.rs:2348-2372); the tests here use that tolerance.

Run:  python tests/golden/make_kats.py   (rewrites kats.json)
"""
from __future__ import annotations

import cmath
import json
import math
from pathlib import Path

PI = math.pi
R2 = 1.0 / math.sqrt(2.0)
OUT = Path(__file__).with_name("kats.json")


def basis(n, i):
    v = [0j] * (1 << n)
    v[i] = 1 + 0j
    return v


def plus(n):
    a = 1.0 / math.sqrt(1 << n)
    return [complex(a, 0)] * (1 << n)


def minus(n):
    a = 1.0 / math.sqrt(1 << n)
    return [complex(a if bin(i).count("1") % 2 == 0 else -a, 0) for i in range(1 << n)]


def kron(a, b):  # |a> (high qubits) tensor |b> (low qubits), state.rs:818-820
    return [x * y for x in a for y in b]


def g(kind, targets, controls=(), params=(), matrix=None):
    d = {"kind": kind, "targets": list(targets), "controls": list(controls), "params": list(params)}
    if matrix is not None:
        d["matrix"] = [[z.real, z.imag] for row in matrix for z in row]
    return d


CASES = []


def case(name, cite, init, gates, expected):
    CASES.append({"name": name, "cite": cite, "n": int(math.log2(len(init))),
                  "init": [[z.real, z.imag] for z in init], "gates": gates,
                  "expected": [[z.real, z.imag] for z in expected]})


# ---- Hadamard (operator_tests.rs:17-122)
case("h_zero", "operator_tests.rs:17-122", basis(1, 0), [g("H", [0])], [R2, R2])
case("h_one", "operator_tests.rs:17-122", basis(1, 1), [g("H", [0])], [R2, -R2])
case("h_plus", "operator_tests.rs:17-122", plus(1), [g("H", [0])], [1, 0])
case("h_minus", "operator_tests.rs:17-122", minus(1), [g("H", [0])], [0, 1])
case("h_multi_00", "operator_tests.rs:17-122", basis(2, 0), [g("H", [0]), g("H", [1])], plus(2))
case("ch_c0_t1_11", "operator_tests.rs:17-122", basis(2, 3), [g("H", [1], [0])], kron([R2, -R2], [0, 1]))
case("ch_c0_t1_10", "operator_tests.rs:17-122", basis(2, 2), [g("H", [1], [0])], basis(2, 2))
# ---- Pauli X/Y/Z (operator_tests.rs:124-486)
case("x_zero", "operator_tests.rs:124-238", basis(1, 0), [g("X", [0])], [0, 1])
case("x_plus", "operator_tests.rs:124-238", plus(1), [g("X", [0])], plus(1))
case("cx_c1_t0_10", "operator_tests.rs:124-238", basis(2, 2), [g("X", [0], [1])], basis(2, 3))
case("y_zero", "operator_tests.rs:240-357", basis(1, 0), [g("Y", [0])], [0, 1j])
case("y_one", "operator_tests.rs:240-357", basis(1, 1), [g("Y", [0])], [-1j, 0])
case("y_plus", "operator_tests.rs:240-357", plus(1), [g("Y", [0])], [-1j * R2, 1j * R2])
case("z_one", "operator_tests.rs:359-486", basis(1, 1), [g("Z", [0])], [0, -1])
case("z_plus", "operator_tests.rs:359-486", plus(1), [g("Z", [0])], minus(1))
case("cz_c0_t1_11", "operator_tests.rs:359-486", basis(2, 3), [g("Z", [1], [0])], [0, 0, 0, -1])
case("i_plus", "operator_tests.rs:529-580", plus(2), [g("I", [1])], plus(2))
# ---- phase gates (operator_tests.rs:582-1248)
case("s_one", "operator_tests.rs:582-718", basis(1, 1), [g("S", [0])], [0, 1j])
case("s_plus", "operator_tests.rs:582-718", plus(1), [g("S", [0])], [R2, 1j * R2])
case("t_one", "operator_tests.rs:720-856", basis(1, 1), [g("T", [0])], [0, cmath.exp(1j * PI / 4)])
case("t_minus", "operator_tests.rs:720-856", minus(1), [g("T", [0])], [R2, -R2 * cmath.exp(1j * PI / 4)])
case("sdag_one", "operator_tests.rs:858-984", basis(1, 1), [g("SDAG", [0])], [0, -1j])
case("tdag_plus", "operator_tests.rs:986-1106", plus(1), [g("TDAG", [0])], [R2, R2 * cmath.exp(-1j * PI / 4)])
case("ctdag_c0_t1_+1", "operator_tests.rs:1018-1022", kron(plus(1), basis(1, 1)), [g("TDAG", [1], [0])],
     kron([R2, R2 * cmath.exp(-1j * PI / 4)], basis(1, 1)))
th = PI / 2.5
case("p_plus", "operator_tests.rs:1108-1131", plus(1), [g("P", [0], [], [th])], [R2, R2 * cmath.exp(1j * th)])
case("p_minus", "operator_tests.rs:1108-1131", minus(1), [g("P", [0], [], [th])], [R2, -R2 * cmath.exp(1j * th)])
case("p_multi_11", "operator_tests.rs:1147", basis(2, 3), [g("P", [0], [], [th]), g("P", [1], [], [th])],
     [0, 0, 0, cmath.exp(2j * th)])
case("cp_c0_t1_+1", "operator_tests.rs:1151-1155", kron(plus(1), basis(1, 1)), [g("P", [1], [0], [PI / 2])],
     kron([R2, R2 * cmath.exp(1j * PI / 2)], basis(1, 1)))
case("cp_c1_t0_+0", "operator_tests.rs:1158-1160", kron(plus(1), basis(1, 0)), [g("P", [0], [1], [-PI / 2])],
     kron(plus(1), basis(1, 0)))
# ---- rotations (operator_tests.rs:1250-1750)
c, s = math.cos(th / 2), math.sin(th / 2)
case("rx_zero", "operator_tests.rs:1251-1275", basis(1, 0), [g("RX", [0], [], [th])], [c, -1j * s])
case("rx_one", "operator_tests.rs:1251-1275", basis(1, 1), [g("RX", [0], [], [th])], [-1j * s, c])
case("rx_plus", "operator_tests.rs:1251-1275", plus(1), [g("RX", [0], [], [th])],
     [cmath.exp(-1j * th / 2) * R2, cmath.exp(-1j * th / 2) * R2])
case("rx_multi_00", "operator_tests.rs:1280-1286", basis(2, 0), [g("RX", [0], [], [th]), g("RX", [1], [], [th])],
     [c * c, -1j * c * s, -1j * c * s, -s * s])
case("crx_c0_t1_11_pi", "operator_tests.rs:1288-1292", basis(2, 3), [g("RX", [1], [0], [PI])],
     [0, -1j, 0, math.cos(PI / 2)])
t2 = PI / 1.5
c2, s2 = math.cos(t2 / 2), math.sin(t2 / 2)
case("ry_zero", "operator_tests.rs:1428-1600", basis(1, 0), [g("RY", [0], [], [t2])], [c2, s2])
case("ry_one", "operator_tests.rs:1428-1600", basis(1, 1), [g("RY", [0], [], [t2])], [-s2, c2])
t3 = PI / 2.2
case("rz_zero", "operator_tests.rs:1602-1750", basis(1, 0), [g("RZ", [0], [], [t3])], [cmath.exp(-1j * t3 / 2), 0])
case("rz_plus", "operator_tests.rs:1602-1750", plus(1), [g("RZ", [0], [], [t3])],
     [R2 * cmath.exp(-1j * t3 / 2), R2 * cmath.exp(1j * t3 / 2)])
case("crz_c1_t0_11", "operator_tests.rs:1602-1750", basis(2, 3), [g("RZ", [0], [1], [t3])],
     [0, 0, 0, cmath.exp(1j * t3 / 2)])
# ---- Unitary2 / ry_phase (operator_tests.rs:1752-1984)
u = [[complex(R2, 0), complex(0, R2)], [complex(0, R2), complex(R2, 0)]]
case("u2_zero", "operator_tests.rs:1752-1845", basis(1, 0), [g("U2", [0], matrix=u)], [R2, 1j * R2])
tp, ph = PI / 3, PI / 5
e = cmath.exp(1j * ph)
ryp = [[complex(math.cos(tp / 2), 0), -e * math.sin(tp / 2)], [complex(math.sin(tp / 2), 0), e * math.cos(tp / 2)]]
case("ry_phase_one", "operator_tests.rs:1847-1955", basis(1, 1), [g("U2", [0], matrix=ryp)],
     [-e * math.sin(tp / 2), e * math.cos(tp / 2)])
case("ry_phase_pi_pi_is_x", "operator.rs:2130-2133 (U(pi,pi) = X)", basis(1, 0),
     [g("U2", [0], matrix=[[complex(math.cos(PI / 2), 0), -cmath.exp(1j * PI) * math.sin(PI / 2)],
                           [complex(math.sin(PI / 2), 0), cmath.exp(1j * PI) * math.cos(PI / 2)]])],
     [math.cos(PI / 2), 1])
# ---- CNOT / SWAP / Toffoli / Matchgate (operator_tests.rs:1986-2491)
case("cnot_c0_t1_01", "operator_tests.rs:1986-2068", basis(2, 1), [g("CNOT", [1], [0])], basis(2, 3))
case("cnot_c0_t1_10", "operator_tests.rs:1986-2068", basis(2, 2), [g("CNOT", [1], [0])], basis(2, 2))
case("cnot_bell", "operator_tests.rs:1986-2068", basis(2, 0), [g("H", [0]), g("CNOT", [1], [0])],
     [R2, 0, 0, R2])
case("swap_01", "operator_tests.rs:2070-2212", basis(2, 1), [g("SWAP", [0, 1])], basis(2, 2))
case("swap_3q_02", "operator_tests.rs:2070-2212", basis(3, 1), [g("SWAP", [0, 2])], basis(3, 4))
case("cswap_c2_on", "operator_tests.rs:2070-2212", basis(3, 5), [g("SWAP", [0, 1], [2])], basis(3, 6))
case("cswap_c2_off", "operator_tests.rs:2070-2212", basis(3, 1), [g("SWAP", [0, 1], [2])], basis(3, 1))
case("toffoli_011", "operator_tests.rs:2391-2491", basis(3, 3), [g("TOFFOLI", [2], [0, 1])], basis(3, 7))
case("toffoli_001", "operator_tests.rs:2391-2491", basis(3, 1), [g("TOFFOLI", [2], [0, 1])], basis(3, 1))
mt, p1, p2 = PI, 0.3, 0.9
case("matchgate_pi_01", "operator_tests.rs:2214-2388", basis(2, 1), [g("MATCHGATE", [0], [], [mt, p1, p2])],
     [0, math.cos(PI / 2), 1, 0])
# CPU form of the Matchgate (operator.rs:993-1004): on (|01>, |10>) the matrix
# [[c, -e^{i phi1} s], [s, e^{i phi1} c]], |11> -> e^{i phi2} |11>; the phase
# difference phi1 != phi2 is what the reference's theta = pi test cannot see.
mt = PI / 3
cm, sm, e1, e2 = math.cos(mt / 2), math.sin(mt / 2), cmath.exp(1j * 0.7), cmath.exp(1j * 1.1)
a = plus(2)
case("matchgate_cpu_form", "operator.rs:980-1006", a, [g("MATCHGATE", [0], [], [mt, 0.7, 1.1])],
     [a[0], cm * a[1] - e1 * sm * a[2], sm * a[1] + e1 * cm * a[2], a[3] * e2])
# ---- Pauli strings (pauli_string_tests.rs)
CASES.append({"name": "pauli_apply_x0y1_11", "cite": "pauli_string_tests.rs:96-107", "n": 2, "op": "pauli_apply",
              "init": [[z.real, z.imag] for z in basis(2, 3)], "term": [3, 2, 2.0, 2.0],
              "expected": [[2.0, -2.0], [0, 0], [0, 0], [0, 0]]})
CASES.append({"name": "pauli_exp_neg_i_dt_z0x1", "cite": "pauli_string_tests.rs:408-438", "n": 2,
              "op": "pauli_exp_neg_i_dt", "dt": 0.5,
              "init": [[z.real, z.imag] for z in basis(2, 0)], "term": [2, 1, 0.5 * PI, 0.0],
              "expected": [[R2, 0], [0, 0], [0, -R2], [0, 0]]})
CASES.append({"name": "sumop_apply_2x0_3y1_11", "cite": "pauli_string_tests.rs:307-320", "n": 2, "op": "sumop_apply",
              "init": [[z.real, z.imag] for z in basis(2, 3)], "terms": [[1, 0, 2.0, 0.0], [2, 2, 3.0, 0.0]],
              "expected": [[0, 0], [0, -3.0], [2.0, 0], [0, 0]]})
CASES.append({"name": "sumop_expectation_minus4", "cite": "pauli_string_tests.rs:348-364", "n": 2,
              "op": "expectation", "init": [[z.real, z.imag] for z in basis(2, 3)],
              "terms": [[1, 0, 2.0, 0.0], [2, 2, 3.0, 0.0], [0, 2, 4.0, 0.0]], "expected": [[-4.0, 0.0]]})
# ---- Trotter (time_evolution_tests.rs:14-207): H = X0 + Y1 from |10>, dt = 0.1;
# X0 and Y1 commute, so each step is exactly exp(-i dt X0) exp(-i dt Y1)
for steps, order in ((1, 1), (1, 2), (3, 1), (3, 2)):
    t = 0.1 * steps
    q0 = [math.cos(t), -1j * math.sin(t)]       # exp(-i t X)|0>
    q1 = [-math.sin(t), math.cos(t)]            # exp(-i t Y)|1>
    CASES.append({"name": f"trotter_x0_y1_steps{steps}_order{order}", "cite": "time_evolution_tests.rs:14-207",
                  "n": 2, "op": "trotter", "dt": 0.1, "steps": steps, "order": order,
                  "init": [[z.real, z.imag] for z in basis(2, 2)],
                  "terms": [[1, 0, 1.0, 0.0], [2, 2, 1.0, 0.0]],
                  "expected": [[complex(z).real, complex(z).imag] for z in kron(q1, q0)]})
# ---- Gate::PauliString normalises (gate_tests.rs:124-186): 2 X0 X1 |00> -> |11>
CASES.append({"name": "gate_pauli_string_normalised", "cite": "gate_tests.rs:124-186", "n": 2, "op": "gates",
              "init": [[z.real, z.imag] for z in basis(2, 0)],
              "gates": [{"kind": "PAULI_STRING", "targets": [], "controls": [], "params": [2.0, 0.0],
                         "x_mask": 3, "z_mask": 0}],
              "expected": [[0, 0], [0, 0], [0, 0], [1, 0]]})
# ---- Circuit::execute (circuit_tests.rs:84-107): H (x) H |00> = |++>
case("circuit_hh", "circuit_tests.rs:84-107", basis(2, 0), [g("H", [0]), g("H", [1])], plus(2))
# ---- QFT closed form: U = R F R (F = DFT with omega = e^{+2 pi i / N}, R = bit reversal)
for nq, k in ((3, 5), (4, 6)):
    N = 1 << nq
    rev = lambda x: int(format(x, f"0{nq}b")[::-1], 2)  # noqa: E731
    kk = rev(k)
    exp = [0j] * N
    for j in range(N):
        exp[rev(j)] = cmath.exp(2j * PI * j * kk / N) / math.sqrt(N)
    gates = []
    for i in range(nq):
        gates.append(g("H", [i]))
        den = 2.0
        for kq in range(1, nq - i):
            gates.append(g("P", [i], [i + kq], [PI / den]))
            den *= 2.0
    for i in range(nq // 2):
        gates.append(g("SWAP", [i, nq - 1 - i]))
    case(f"qft{nq}_basis{k}", "subroutine.rs:90-112", basis(nq, k), gates, exp)
# ---- constructors (state_tests.rs)
CASES.append({"name": "new_minus_2", "cite": "state.rs:240-275", "n": 2, "op": "init", "kind": 3,
              "expected": [[z.real, z.imag] for z in minus(2)]})
CASES.append({"name": "new_ghz_3", "cite": "state.rs:277-300", "n": 3, "op": "init", "kind": 4,
              "expected": [[R2 if i in (0, 7) else 0.0, 0.0] for i in range(8)]})
CASES.append({"name": "new_plus_3", "cite": "state.rs:225-237", "n": 3, "op": "init", "kind": 2,
              "expected": [[z.real, z.imag] for z in plus(3)]})

# ---- normalise and the Fubini-Study metrics (state_tests.rs:149-166, 225-295).
# The reference builds some of these states with a num_qubits field that does
# not match their length (e.g. two amplitudes labelled 2 qubits); the length is
# what the arithmetic uses, so here a state of 2^k amplitudes is k qubits.
def vec(v):
    return [[complex(z).real, complex(z).imag] for z in v]


CASES.append({"name": "normalise_3_4", "cite": "state_tests.rs:149-156", "n": 1, "op": "normalise",
              "init": vec([3.0, 4.0]), "expected": vec([0.6, 0.8]), "bit_exact": True})
METRICS = [
    {"name": "fs_dist_zero", "cite": "state_tests.rs:225-229", "metric": "fs_dist",
     "a": vec(plus(2)), "b": vec(plus(2)), "expected": 0.0},
    {"name": "fs_dist_pi_over_4", "cite": "state_tests.rs:231-244", "metric": "fs_dist",
     "a": vec([1.0, 1j]), "b": vec([0.0, 1.0]), "expected": PI / 4.0},
    {"name": "fs_fidelity_zero", "cite": "state_tests.rs:271-278", "metric": "fs_fidelity",
     "a": vec([1.0, 0.0]), "b": vec([0.0, 1.0]), "expected": 0.0},
    {"name": "fs_fidelity_one", "cite": "state_tests.rs:280-283", "metric": "fs_fidelity",
     "a": vec(plus(2)), "b": vec(plus(2)), "expected": 1.0},
    {"name": "fs_fidelity_half", "cite": "state_tests.rs:285-294", "metric": "fs_fidelity",
     "a": vec(plus(1)), "b": vec(basis(1, 1)), "expected": 0.5},
]
# errors: (variant code, payload a); the empty state of state_tests.rs:248-252
# (no amplitudes) cannot be built in the engine (State::new rejects it too)
METRIC_ERRORS = [
    {"name": "fs_dist_qubit_count", "cite": "state_tests.rs:254-259", "metric": "fs_dist",
     "a": vec(plus(2)), "b": vec(plus(1)), "code": 3, "a_payload": 2},
    {"name": "fs_dist_zero_norm", "cite": "state_tests.rs:261-267", "metric": "fs_dist",
     "a": vec([0.0, 0.0]), "b": vec(basis(2, 1)), "code": 14, "a_payload": 0},
    {"name": "fs_fidelity_zero_norm_other", "cite": "state.rs:492-498", "metric": "fs_fidelity",
     "a": vec(basis(2, 1)), "b": vec([0.0, 0.0, 0.0, 0.0]), "code": 14, "a_payload": 0},
    {"name": "normalise_zero_norm", "cite": "state_tests.rs:159-166", "metric": "normalise",
     "a": vec([0.0, 0.0]), "b": None, "code": 14, "a_payload": 0},
]

# ---- error cases: (gate, n, expected variant, payload) from operator_tests.rs:2545-2703 etc.
ERRORS = [
    {"name": "h_bad_target", "cite": "operator_tests.rs:2513-2545", "n": 2, "gate": g("H", [2]),
     "code": 4, "a": 2, "b": 2},
    {"name": "rx_bad_target", "cite": "operator_tests.rs:2513-2545", "n": 2, "gate": g("RX", [2], [], [1.0]),
     "code": 4, "a": 2, "b": 2},
    {"name": "h_bad_control", "cite": "operator.rs:236-238", "n": 2, "gate": g("H", [0], [5]),
     "code": 4, "a": 5, "b": 2},
    {"name": "cnot_overlap", "cite": "operator_tests.rs:2690-2700", "n": 2, "gate": g("CNOT", [0], [0]),
     "code": 2, "a": 0, "b": 0},
    {"name": "cnot_two_controls", "cite": "operator.rs:674-679", "n": 3, "gate": g("CNOT", [0], [1, 2]),
     "code": 3, "a": 2, "b": 0},
    {"name": "cnot_no_targets", "cite": "operator.rs:221-223", "n": 3, "gate": g("CNOT", [], [1]),
     "code": 3, "a": 0, "b": 0},
    {"name": "toffoli_same_controls", "cite": "operator.rs:1070-1072", "n": 3, "gate": g("TOFFOLI", [0], [1, 1]),
     "code": 3, "a": 2, "b": 0},
    {"name": "swap_same_targets", "cite": "operator.rs:249-268", "n": 3, "gate": g("SWAP", [1, 1]),
     "code": 4, "a": 1, "b": 3},
    {"name": "swap_one_target", "cite": "operator.rs:221-223", "n": 3, "gate": g("SWAP", [1]),
     "code": 3, "a": 1, "b": 0},
    {"name": "matchgate_last_qubit", "cite": "operator.rs:903-905", "n": 3, "gate": g("MATCHGATE", [2], [], [1, 0, 0]),
     "code": 4, "a": 2, "b": 3},
    {"name": "u2_non_unitary", "cite": "operator.rs:2092-2118", "n": 1,
     "gate": g("U2", [0], matrix=[[1, 1], [0, 1]]), "code": 6, "a": 0, "b": 0},
    {"name": "pauli_evolution_complex_coeff", "cite": "pauli_string.rs:281-287", "n": 2,
     "gate": {"kind": "PAULI_EVOLUTION", "targets": [], "controls": [], "params": [1.0, 1.0, 0.5],
              "x_mask": 2, "z_mask": 1}, "code": 15, "a": 0, "b": 0},
    {"name": "pauli_string_bad_qubit", "cite": "pauli_string_tests.rs:122-132", "n": 1,
     "gate": {"kind": "PAULI_STRING", "targets": [], "controls": [], "params": [2.0, 2.0],
              "x_mask": 3, "z_mask": 2}, "code": 4, "a": 1, "b": 1},
    {"name": "measure_too_many_qubits", "cite": "state.rs:548-550", "n": 2,
     "gate": {"kind": "MEASURE", "targets": [], "controls": [0, 1, 0], "params": [0.5]}, "code": 3, "a": 2, "b": 0},
    {"name": "measure_bad_qubit", "cite": "state.rs:552-556", "n": 2,
     "gate": {"kind": "MEASURE", "targets": [], "controls": [3], "params": [0.5]}, "code": 4, "a": 3, "b": 2},
]

if __name__ == "__main__":
    OUT.write_text(json.dumps({"cases": CASES, "errors": ERRORS, "metrics": METRICS,
                               "metric_errors": METRIC_ERRORS}, indent=1))
    print(f"wrote {len(CASES)} cases, {len(ERRORS)} error cases, {len(METRICS)} metric cases and "
          f"{len(METRIC_ERRORS)} metric error cases to {OUT}")
