"""Pin the oracle independently of its own formulas: dense-matrix brute force
(numpy Kronecker products of the textbook gate matrices), scipy.linalg.expm for
exp(P), closed forms and invariants. A dropped term, a wrong sign, a swapped
index or a transposed operand in the oracle fails one of these."""
import math

import numpy as np
import pytest
from scipy.linalg import expm

import oracle as O
from kat_util import gates_array, random_gate, random_state, random_terms, term_array, to_gate

I2 = np.eye(2, dtype=complex)
PX = np.array([[0, 1], [1, 0]], dtype=complex)
PY = np.array([[0, -1j], [1j, 0]], dtype=complex)
PZ = np.array([[1, 0], [0, -1]], dtype=complex)


def mat_1q(d):
    k, p = d["kind"], d.get("params", [0, 0, 0])
    h = p[0] / 2
    return {
        "H": np.array([[1, 1], [1, -1]]) / math.sqrt(2), "X": PX, "Y": PY, "Z": PZ, "I": I2,
        "S": np.diag([1, 1j]), "T": np.diag([1, np.exp(1j * math.pi / 4)]), "SDAG": np.diag([1, -1j]),
        "TDAG": np.diag([1, np.exp(-1j * math.pi / 4)]), "P": np.diag([1, np.exp(1j * p[0])]),
        "RX": np.array([[math.cos(h), -1j * math.sin(h)], [-1j * math.sin(h), math.cos(h)]]),
        "RY": np.array([[math.cos(h), -math.sin(h)], [math.sin(h), math.cos(h)]]),
        "RZ": np.diag([np.exp(-1j * h), np.exp(1j * h)]),
        "CNOT": PX, "TOFFOLI": PX,
    }.get(k) if k != "U2" else np.array([complex(a, b) for a, b in d["matrix"]]).reshape(2, 2)


def dense(n, d):
    """Full 2^n x 2^n matrix of one gate, built column by column from basis states."""
    N = 1 << n
    M = np.zeros((N, N), dtype=complex)
    k = d["kind"]
    cm = sum(1 << c for c in d.get("controls", []))
    for col in range(N):
        if k == "SWAP":
            a, b = d["targets"]
            on = (col & cm) == cm
            j = col
            if on and ((col >> a) & 1) != ((col >> b) & 1):
                j = col ^ (1 << a) ^ (1 << b)
            M[j, col] = 1
        elif k == "MATCHGATE":
            q1 = d["targets"][0]
            q2 = q1 + 1
            th, p1, p2 = d["params"]
            c, s = math.cos(th / 2), math.sin(th / 2)
            e1, e2 = np.exp(1j * p1), np.exp(1j * p2)
            b1, b2 = (col >> q1) & 1, (col >> q2) & 1
            base = col & ~(1 << q1) & ~(1 << q2)
            i01, i10, i11 = base | (1 << q1), base | (1 << q2), base | (1 << q1) | (1 << q2)
            # the reference tests controls on i01 for the rotation, on i11 for the phase
            rot_on = (i01 & cm) == cm
            ph_on = (i11 & cm) == cm
            if (b1, b2) == (1, 0):
                if rot_on:
                    M[i01, col] += c; M[i10, col] += s
                else:
                    M[col, col] = 1
            elif (b1, b2) == (0, 1):
                if rot_on:
                    M[i01, col] += -e1 * s; M[i10, col] += e1 * c
                else:
                    M[col, col] = 1
            elif (b1, b2) == (1, 1):
                M[col, col] = e2 if ph_on else 1
            else:
                M[col, col] = 1
        else:
            t = d["targets"][0]
            U = mat_1q(d)
            if (col & cm) != cm:
                M[col, col] = 1
                continue
            b = (col >> t) & 1
            for r in (0, 1):
                M[(col & ~(1 << t)) | (r << t), col] += U[r, b]
    return M


@pytest.mark.parametrize("seed", range(40))
def test_every_gate_kind_against_dense_matrix(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 6))
    psi = random_state(n, rng)
    for _ in range(6):
        d = random_gate(n, rng)
        got = O.apply_gate(psi, to_gate(d))
        exp = dense(n, d) @ psi
        assert np.max(np.abs(got - exp)) < 1e-13, d
        psi = got


def pauli_dense(n, x, z):
    M = np.array([[1]], dtype=complex)
    for q in reversed(range(n)):
        xb, zb = (x >> q) & 1, (z >> q) & 1
        P = PY if xb and zb else PX if xb else PZ if zb else I2
        M = np.kron(M, P)
    return M


@pytest.mark.parametrize("seed", range(12))
def test_pauli_apply_exp_expectation_dense(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 6))
    psi = random_state(n, rng)
    terms = random_terms(n, 5, rng)
    H = np.zeros((1 << n, 1 << n), dtype=complex)
    for t in terms:
        c = complex(t["c_re"], t["c_im"])
        P = pauli_dense(n, int(t["x_mask"]), int(t["z_mask"]))
        np.testing.assert_allclose(O.pauli_apply(psi, t), c * (P @ psi), atol=1e-13)
        f = complex(0.3, -0.7)
        np.testing.assert_allclose(O.pauli_exp(psi, t, f), expm(c * f * P) @ psi, atol=1e-12)
        H += c * P
    np.testing.assert_allclose(O.sumop_apply(psi, terms), H @ psi, atol=1e-12)
    assert abs(O.expectation(psi, terms) - np.vdot(psi, H @ psi)) < 1e-12


def test_exp_of_hermitian_pauli_is_cos_minus_i_sin():
    # exp(-i t P) = cos t I - i sin t P for a Pauli string P (P^2 = I)
    rng = np.random.default_rng(7)
    n = 4
    psi = random_state(n, rng)
    t = term_array([(0b1010, 0b0110, 0.8, 0.0)])[0]
    P = pauli_dense(n, 0b1010, 0b0110)
    dt = 0.37
    exp = math.cos(0.8 * dt) * psi - 1j * math.sin(0.8 * dt) * (P @ psi)
    np.testing.assert_allclose(O.pauli_exp_neg_i_dt(psi, t, dt), exp, atol=1e-13)


def test_trotter_matches_product_of_exponentials_and_commuting_limit():
    rng = np.random.default_rng(11)
    n = 4
    psi = random_state(n, rng)
    terms = random_terms(n, 6, rng)
    terms["c_im"] = 0.0
    dt, steps = 0.05, 3
    U1 = np.eye(1 << n, dtype=complex)
    for t in terms:
        U1 = expm(-1j * dt * t["c_re"] * pauli_dense(n, int(t["x_mask"]), int(t["z_mask"]))) @ U1
    np.testing.assert_allclose(O.trotter(psi, terms, dt, steps, 1), np.linalg.matrix_power(U1, steps) @ psi,
                               atol=1e-12)
    U2 = np.eye(1 << n, dtype=complex)
    for t in list(terms) + list(terms)[::-1]:
        U2 = expm(-1j * dt / 2 * t["c_re"] * pauli_dense(n, int(t["x_mask"]), int(t["z_mask"]))) @ U2
    np.testing.assert_allclose(O.trotter(psi, terms, dt, steps, 2), np.linalg.matrix_power(U2, steps) @ psi,
                               atol=1e-12)
    # all-Z Hamiltonian: terms commute, Trotter is exact
    zt = term_array([(0, 0b0011, 0.4, 0.0), (0, 0b1100, -0.9, 0.0), (0, 0b0100, 0.3, 0.0)])
    Hz = sum(t["c_re"] * pauli_dense(n, 0, int(t["z_mask"])) for t in zt)
    np.testing.assert_allclose(O.trotter(psi, zt, 0.2, 5, 1), expm(-1j * 1.0 * Hz) @ psi, atol=1e-12)


def test_invariants():
    rng = np.random.default_rng(3)
    psi = random_state(3, rng)
    G = lambda k, t, c=(), p=(0,): to_gate({"kind": k, "targets": t, "controls": list(c), "params": list(p)})  # noqa: E731
    hh = O.apply_gate(O.apply_gate(psi, G("H", [1])), G("H", [1]))
    np.testing.assert_allclose(hh, psi, atol=1e-15)
    ss = O.apply_gate(O.apply_gate(psi, G("S", [2])), G("SDAG", [2]))
    np.testing.assert_allclose(ss, psi, atol=1e-15)
    tt = O.apply_gate(O.apply_gate(psi, G("T", [0])), G("T", [0]))
    np.testing.assert_allclose(tt, O.apply_gate(psi, G("S", [0])), atol=1e-15)
    sw = O.apply_gate(psi, G("SWAP", [0, 2]))
    c3 = O.apply_gates(psi, gates_array([{"kind": "CNOT", "targets": [2], "controls": [0]},
                                         {"kind": "CNOT", "targets": [0], "controls": [2]},
                                         {"kind": "CNOT", "targets": [2], "controls": [0]}]))
    np.testing.assert_allclose(sw, c3, atol=1e-15)
    for k in ("H", "RX", "RY", "RZ", "P", "Y"):
        out = O.apply_gate(psi, G(k, [1], [], [1.3]))
        assert abs(np.linalg.norm(out) - 1) < 1e-14


def test_measurement_brute_force_and_sequential_rule():
    rng = np.random.default_rng(5)
    n = 5
    psi = random_state(n, rng)
    qs = [3, 0, 4]
    probs = O.measure_probs(psi, qs)
    bf = np.zeros(8)
    for i, a in enumerate(psi):
        o = sum(((i >> q) & 1) << b for b, q in enumerate(qs))
        bf[o] += abs(a) ** 2
    np.testing.assert_allclose(probs, bf, atol=1e-15)
    # duplicate measured qubits: inconsistent outcomes have probability 0
    pd = O.measure_probs(psi, [1, 1])
    assert pd[1] == 0 and pd[2] == 0 and abs(pd[0] + pd[3] - 1) < 1e-14
    cum = np.cumsum(bf / bf.sum())
    for u in rng.random(200):
        want = int(np.searchsorted(cum, u, side="right"))
        want = min(want, 7)
        assert O.sample_outcome(bf, u) == want
        o, post = O.measure(psi, 0, qs, u)
        assert o == want
        keep = np.array([sum(((i >> q) & 1) << b for b, q in enumerate(qs)) == o for i in range(1 << n)])
        exp = np.where(keep, psi, 0)
        np.testing.assert_allclose(post, exp / np.linalg.norm(exp), atol=1e-14)
    # last bin when u is beyond the final cumulative value (state.rs:614-617)
    assert O.sample_outcome(np.array([0.5, 0.5 - 1e-17]), 0.9999999999999999) == 1


def test_measurement_bases():
    rng = np.random.default_rng(9)
    n = 3
    psi = random_state(n, rng)
    Hd = np.array([[1, 1], [1, -1]]) / math.sqrt(2)
    for basis, U in ((1, Hd), (2, Hd @ np.diag([1, -1j]))):
        o, post = O.measure(psi, basis, [1], 0.37)
        V = np.kron(np.kron(I2, U), I2)
        w = V @ psi
        p1 = np.sum(np.abs(w[[2, 3, 6, 7]]) ** 2)
        want = 0 if 0.37 < (1 - p1) / 1.0 else 1
        assert o == want
        keep = np.array([((i >> 1) & 1) == o for i in range(8)])
        col = np.where(keep, w, 0)
        col /= np.linalg.norm(col)
        back = np.kron(np.kron(I2, U.conj().T), I2) @ col
        np.testing.assert_allclose(post, back, atol=1e-13)


def test_qft_closed_form():
    n = 4
    N = 1 << n
    rev = [int(format(x, f"0{n}b")[::-1], 2) for x in range(N)]
    F = np.array([[np.exp(2j * math.pi * j * k / N) for k in range(N)] for j in range(N)]) / math.sqrt(N)
    R = np.zeros((N, N))
    for x in range(N):
        R[rev[x], x] = 1
    gates = []
    for i in range(n):
        gates.append({"kind": "H", "targets": [i]})
        den = 2.0
        for k in range(1, n - i):
            gates.append({"kind": "P", "targets": [i], "controls": [i + k], "params": [math.pi / den]})
            den *= 2.0
    for i in range(n // 2):
        gates.append({"kind": "SWAP", "targets": [i, n - 1 - i]})
    rng = np.random.default_rng(1)
    psi = random_state(n, rng)
    np.testing.assert_allclose(O.apply_gates(psi, gates_array(gates)), R @ F @ R @ psi, atol=1e-13)


def test_reference_structured_baseline_equals_oracle():
    rng = np.random.default_rng(21)
    n = 8
    psi = random_state(n, rng)
    gates = gates_array([random_gate(n, rng) for _ in range(40)])
    a = O.apply_gates(psi, gates)
    b = O.reference_structured_apply(psi, gates, threads=3)
    np.testing.assert_allclose(a, b, atol=1e-13)
    c, dt = O.reference_structured_timed(psi, gates, threads=3, keep=True)  # the benchmark's timed form
    np.testing.assert_array_equal(b, c)
    assert dt > 0


def _dense_basis_measure(psi, n, qubits, U, u):
    """Measurement in the basis of U on `qubits` from dense algebra: w = (x)U psi,
    the outcome by the sequential-CDF rule on the marginal of w, the projector,
    renormalisation and (x)U^dagger back (state.rs:706-728)."""
    V = np.eye(1, dtype=complex)
    for q in reversed(range(n)):  # qubit q is bit q: the highest qubit is the leftmost factor
        V = np.kron(V, U if q in qubits else I2)
    w = V @ psi
    outcome_of = np.array([sum(((i >> q) & 1) << b for b, q in enumerate(qubits)) for i in range(1 << n)])
    probs = np.bincount(outcome_of, weights=np.abs(w) ** 2, minlength=1 << len(qubits))
    cum = 0.0
    o = len(probs) - 1
    for i, p in enumerate(probs):  # first bin with u < C_i, else the last (state.rs:603-619)
        cum += p / probs.sum()
        if u < cum:
            o = i
            break
    col = np.where(outcome_of == o, w, 0)
    col = col / np.linalg.norm(col)
    return o, V.conj().T @ col


def test_custom_basis_measurement_matches_dense_algebra():
    """MeasurementBasis::Custom(U) (state.rs:706-728) with a general complex
    unitary on two measured qubits of four: outcome and collapsed state against
    the dense product, projector and adjoint."""
    from kat_util import random_unitary2

    rng = np.random.default_rng(77)
    n = 4
    for trial in range(6):
        psi = random_state(n, rng)
        U = random_unitary2(rng)
        qs = [3, 1] if trial % 2 else [0, 2]
        for u in (0.02, 0.31, 0.64, 0.97):
            o, post = O.measure(psi, 3, qs, u, U)
            o_d, post_d = _dense_basis_measure(psi, n, qs, U, u)
            assert o == o_d, (trial, u)
            np.testing.assert_allclose(post, post_d, atol=1e-13)
    # a basis that is not unitary is rejected (Unitary2::new, operator.rs:2092-2118)
    with pytest.raises(O.OracleError) as ei:
        O.measure(random_state(2, rng), 3, [0], 0.5, np.array([[1, 1], [0, 1]]))
    assert ei.value.code == 6


def test_fs_metrics_against_numpy():
    """fs_fidelity = |<a|b>|^2 / (|a|^2 |b|^2) and fs_dist = acos of its root,
    computed here with numpy's vdot on unnormalised random vectors."""
    rng = np.random.default_rng(5)
    for n in (1, 3, 7):
        a = random_state(n, rng) * 1.7
        b = random_state(n, rng) * 0.4
        f = abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)
        assert abs(O.fs_fidelity(a, b) - f) < 1e-14
        assert abs(O.fs_dist(a, b) - math.acos(math.sqrt(f))) < 1e-12


def test_state_arithmetic_against_numpy():
    """State Add / Sub / Mul (state.rs:2687-2862): one IEEE rounding per
    component, the same as numpy's complex add, subtract and multiply."""
    rng = np.random.default_rng(6)
    a, b = random_state(6, rng), random_state(6, rng)
    np.testing.assert_array_equal(O.state_add(a, b), a + b)
    np.testing.assert_array_equal(O.state_sub(a, b), a - b)
    c = complex(0.3, -1.25)
    # (ar*cr - ai*ci, ar*ci + ai*cr), the num-complex product
    exp = (a.real * c.real - a.imag * c.imag) + 1j * (a.real * c.imag + a.imag * c.real)
    np.testing.assert_array_equal(O.state_scale(a, c), exp)


def test_sample_cdf_is_the_sequential_cumulative_sum():
    rng = np.random.default_rng(8)
    p = rng.random(1000)
    cdf = O.sample_cdf(p)
    total = 0.0
    for x in p:
        total += x
    cum, exp = 0.0, []
    for x in p:
        cum += x / total
        exp.append(cum)
    np.testing.assert_array_equal(cdf, np.array(exp))
    for u in rng.random(100):
        assert O.sample_outcome(p, u) == int(np.argmax(u < cdf)) if u < cdf[-1] else len(p) - 1
