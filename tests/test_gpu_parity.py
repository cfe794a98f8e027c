"""Parity of the HIP engine with the CPU oracle and the reference KATs.

Bars: complex128 amplitudes within 1e-12 of the oracle (the streaming and
tiled kernels use the reference's operand order and usually agree to the
bit); complex64 amplitudes within 1e-5 per amplitude and state fidelity
above 1 - 1e-9; measurement outcomes identical for identical uniforms; error
variants and payloads identical."""
import math

import numpy as np
import pytest

import oracle as O
import qiron
from kat_util import (F32_EPS, ONE_Q, cvec, gates_array, load_kats, random_gate, random_state, random_terms,
                      random_unitary2, term_array, to_gate)
from qiron import _lib as L
from qiron import workloads as W

pytestmark = pytest.mark.gpu

KATS = load_kats()


def upload(vec, prec):
    return qiron.State.from_vector_unchecked(vec, prec)


def fidelity(a, b):
    return abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


def check(got, exp, prec, what=""):
    d = float(np.max(np.abs(got - exp))) if got.size else 0.0
    if prec == "c128":
        assert d <= 1e-12, f"{what}: max |diff| {d:.3e}"
    else:
        assert d <= 1e-5, f"{what}: max |diff| {d:.3e}"
        assert fidelity(got, exp) >= 1 - 1e-9, what


def run_engine_case(case, prec):
    op = case.get("op", "gates")
    if op == "init":
        kind = case["kind"]
        s = qiron.State._init(case["n"], kind, 0, prec)
        return s.state_vector
    s = upload(cvec(case["init"]), prec)
    if op == "gates":
        s.apply_gates_inplace(gates_array(case["gates"]))
        return s.state_vector
    if op == "pauli_apply":
        t = term_array([case["term"]])[0]
        ps = qiron.PauliString(complex(t["c_re"], t["c_im"]))
        return _ps_from_term(t).apply(s).state_vector
    if op == "pauli_exp_neg_i_dt":
        return _ps_from_term(term_array([case["term"]])[0]).apply_exp_neg_i_dt(s, case["dt"]).state_vector
    if op == "sumop_apply":
        return _sumop(case["terms"]).apply(s).state_vector
    if op == "expectation":
        return np.array([_sumop(case["terms"]).expectation_value(s)])
    if op == "trotter":
        return qiron.trotter_evolve_state(_sumop(case["terms"]), s, case["dt"], case["steps"], case["order"]).state_vector
    if op == "normalise":
        return s.normalise().state_vector
    raise ValueError(op)


def _ps_from_term(t):
    ops = {}
    x, z = int(t["x_mask"]), int(t["z_mask"])
    for q in range(64):
        xb, zb = (x >> q) & 1, (z >> q) & 1
        if xb or zb:
            ops[q] = "Y" if xb and zb else ("X" if xb else "Z")
    return qiron.PauliString(complex(t["c_re"], t["c_im"]), ops)


def _sumop(ts):
    return qiron.SumOp([_ps_from_term(t) for t in term_array(ts)])


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("case", KATS["cases"], ids=[c["name"] for c in KATS["cases"]])
def test_reference_kats_on_engine(ctx, case, prec):
    got = run_engine_case(case, prec)
    exp = cvec(case["expected"])
    tol = 1e-12 if prec == "c128" else 2 * F32_EPS
    assert np.max(np.abs(got - exp)) <= tol, case["cite"]


@pytest.mark.parametrize("case", KATS["errors"], ids=[c["name"] for c in KATS["errors"]])
def test_error_variants_and_payloads(ctx, case):
    s = qiron.State.new_zero(case["n"])
    before = s.state_vector
    with pytest.raises(qiron.QuantIronError) as ei:
        s.apply_gate_inplace(qiron.Gate(to_gate(case["gate"])))
    e = ei.value
    assert (e.code, e.a, e.b) == (case["code"], case["a"], case["b"]), (case["cite"], str(e))
    with pytest.raises(O.OracleError) as eo:
        O.apply_gate(O.init(case["n"], 0), to_gate(case["gate"]))
    assert (eo.value.code, eo.value.a, eo.value.b) == (e.code, e.a, e.b)
    np.testing.assert_array_equal(s.state_vector, before)  # a rejected gate changes nothing


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("n", [2, 5, 9, 13, 17])
def test_random_gates_match_oracle(ctx, n, prec):
    rng = np.random.default_rng(1000 + n)
    psi = random_state(n, rng)
    s = upload(psi, prec)
    ref = psi
    for k in range(30):
        d = random_gate(n, rng, max_controls=3)
        s.apply_gate_inplace(qiron.Gate(to_gate(d)))
        ref = O.apply_gate(ref, to_gate(d))
    check(s.state_vector, ref, prec, f"n={n}")


def test_c128_gates_are_bit_exact_with_the_oracle(ctx):
    """Same operand order, no FMA contraction, host-computed constants: the
    streaming kernels reproduce the oracle's (the reference's) rounding."""
    rng = np.random.default_rng(77)
    n = 10
    psi = random_state(n, rng)
    s = upload(psi, "c128")
    ref = psi
    for _ in range(60):
        d = random_gate(n, rng, max_controls=2)
        s.apply_gate_inplace(qiron.Gate(to_gate(d)))
        ref = O.apply_gate(ref, to_gate(d))
    got = s.state_vector
    # equal up to the sign of zero
    assert np.array_equal(got.real, ref.real) and np.array_equal(got.imag, ref.imag)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_streaming_gates_at_every_qubit_0_placement(ctx, prec):
    """The complex64 kernels move two amplitudes per 16-byte unit, so qubit 0
    (inside the unit) as target, control, swap partner or bystander takes a
    different code path from the higher qubits; each placement against the
    oracle, on states of 1 to 5 qubits."""
    rng = np.random.default_rng(4242)
    placements = [  # (n, targets, controls)
        (1, [0], []), (2, [0], []), (2, [1], []), (2, [0], [1]), (2, [1], [0]),
        (4, [0], [2]), (4, [2], [0]), (4, [3], [0, 1]), (4, [1], [2, 3]), (5, [4], []),
    ]
    swaps = [(2, [0, 1], []), (4, [0, 2], []), (4, [3, 0], [1]), (4, [1, 3], [0]), (4, [2, 3], []), (5, [1, 4], [2])]
    cases = []
    for kind in ONE_Q:
        for n, t, c in placements:
            cases.append(dict(kind=kind, targets=t, controls=c))
    cases += [dict(kind="SWAP", targets=t, controls=c) for n, t, c in swaps]
    for n in (1, 2, 4, 5):
        psi = random_state(n, rng)
        s = upload(psi, prec)
        ref = psi
        for d in cases:
            if max(d["targets"] + d["controls"]) >= n:
                continue
            d = dict(d, params=[float(rng.uniform(0, 2 * np.pi)) for _ in range(3)])
            if d["kind"] == "U2":
                u = random_unitary2(rng)
                d["matrix"] = [[z.real, z.imag] for z in u.reshape(-1)]
            s.apply_gate_inplace(qiron.Gate(to_gate(d)))
            ref = O.apply_gate(ref, to_gate(d))
            got = s.state_vector
            if prec == "c128":
                assert np.array_equal(got.real, ref.real) and np.array_equal(got.imag, ref.imag), (n, d)
            else:
                check(got, ref, prec, f"n={n} {d['kind']} t={d['targets']} c={d['controls']}")
                ref = got.astype(np.complex128)  # compare each gate from the engine's own state


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_streaming_low_targets_lane_exchange(ctx, prec):
    """A one-target gate whose target lands on index bit 0..2 once the
    controls are removed (partner 16-64 bytes away) takes the lane-exchange
    kernel: every kind with controls below, above and on qubit 0, against the
    oracle on 8- and 14-qubit states (the latter spans several workgroups)."""
    rng = np.random.default_rng(777)
    placements = [  # (targets, controls)
        ([0], []), ([1], []), ([2], []), ([3], []), ([4], []), ([0], [1]), ([1], [0]), ([2], [0]), ([3], [0, 1]),
        ([3], [1]), ([2], [5]), ([1], [7]), ([4], [0, 1]), ([4], [1, 2]), ([3], [0, 7]), ([2], [0, 1, 3]),
        ([5], [0, 2, 4]), ([6], [1, 2, 3, 4]),
    ]
    for n in (8, 14):
        psi = random_state(n, rng)
        s = upload(psi, prec)
        ref = psi
        for kind in ONE_Q:
            for t, c in placements:
                d = dict(kind=kind, targets=t, controls=c, params=[float(rng.uniform(0, 2 * np.pi)) for _ in range(3)])
                if kind == "U2":
                    d["matrix"] = [[z.real, z.imag] for z in random_unitary2(rng).reshape(-1)]
                s.apply_gate_inplace(qiron.Gate(to_gate(d)))
                ref = O.apply_gate(ref, to_gate(d))
                got = s.state_vector
                if prec == "c128":
                    assert np.array_equal(got.real, ref.real) and np.array_equal(got.imag, ref.imag), (n, d)
                else:
                    check(got, ref, prec, f"n={n} {kind} t={t} c={c}")
                    ref = got.astype(np.complex128)


MODES = [
    dict(mode=L.EXEC_PER_GATE),
    dict(mode=L.EXEC_TILED, tile_qubits=10),
    dict(mode=L.EXEC_TILED, tile_qubits=12),
    dict(mode=L.EXEC_FUSED_MFMA, block_qubits=5),
    dict(mode=L.EXEC_FUSED_MFMA, block_qubits=6),
    dict(mode=L.EXEC_AUTO, use_graph=True),
]


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("opts", MODES, ids=["pergate", "tiled10", "tiled12", "fused5", "fused6", "auto_graph"])
def test_circuit_execute_modes_match_oracle(ctx, opts, prec):
    n = 14
    gates = W.random_circuit(n=n, depth=6, seed=123)
    rng = np.random.default_rng(5)
    psi = random_state(n, rng)
    ref = O.apply_gates(psi, gates)
    s = upload(psi, prec)
    s.apply_gates_inplace(gates, **opts)
    got = s.state_vector
    if opts["mode"] == L.EXEC_FUSED_MFMA and prec == "c128":
        assert np.max(np.abs(got - ref)) <= 1e-12
    else:
        check(got, ref, prec, str(opts))
    if opts.get("use_graph"):  # a replay of the cached graph gives the same result
        s2 = upload(psi, prec)
        s2.apply_gates_inplace(gates, **opts)
        s2b = s2.state_vector
        s2.apply_gates_inplace(gates, **opts)
        check(s2.state_vector, O.apply_gates(ref, gates), prec, "graph replay")
        check(s2b, ref, prec, "graph first")


def test_tiled_handles_controls_diagonals_and_matchgates(ctx):
    rng = np.random.default_rng(9)
    n = 13
    kinds = ["H", "RX", "U2", "RZ", "P", "Z", "S", "CNOT", "TOFFOLI", "SWAP", "MATCHGATE", "Y", "T"]
    ds = [random_gate(n, rng, kinds=kinds, max_controls=3) for _ in range(200)]
    gates = gates_array(ds)
    psi = random_state(n, rng)
    ref = O.apply_gates(psi, gates)
    for K in (10, 11, 12):
        s = upload(psi, "c128")
        s.apply_gates_inplace(gates, mode=L.EXEC_TILED, tile_qubits=K)
        check(s.state_vector, ref, "c128", f"K={K}")
    s = upload(psi, "c64")
    s.apply_gates_inplace(gates, mode=L.EXEC_TILED, tile_qubits=13)
    check(s.state_vector, ref, "c64", "c64 K=13")
    for kb in (5, 6):
        s = upload(psi, "c128")
        s.apply_gates_inplace(gates, mode=L.EXEC_FUSED_MFMA, block_qubits=kb)
        assert np.max(np.abs(s.state_vector - ref)) <= 1e-12


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_tiled_every_width_and_thread_count(ctx, prec):
    """A 23-qubit state with controlled SWAPs, matchgates (streamed between
    one-wave passes, LDS operations otherwise) and relabeling SWAPs that force
    permuted stores and lane/register exchanges; every tile width and thread
    count (one-wave tiles included) against the streaming kernels."""
    rng = np.random.default_rng(23)
    n = 23
    kinds = ["H", "RX", "RY", "U2", "RZ", "P", "Z", "X", "CNOT", "TOFFOLI", "SWAP", "MATCHGATE", "Y"]
    gates = gates_array([random_gate(n, rng, kinds=kinds, max_controls=2) for _ in range(160)])
    psi = random_state(n, rng)
    s = upload(psi, prec)
    s.apply_gates_inplace(gates, mode=L.EXEC_PER_GATE)
    ref = s.state_vector
    kmax = 12 if prec == "c128" else 13
    for K in range(10, kmax + 1):
        for tb in ((6, 8, 9) if K <= (12 if prec == "c64" else 11) else (8, 9)):
            s = upload(psi, prec)
            s.apply_gates_inplace(gates, mode=L.EXEC_TILED, tile_qubits=K, tile_threads_log2=tb)
            got = s.state_vector
            if prec == "c128":
                assert np.max(np.abs(got - ref)) <= 1e-12, (K, tb)
            else:
                assert fidelity(got, ref) >= 1 - 1e-9 and np.max(np.abs(got - ref)) <= 1e-5, (K, tb)


def test_plan_and_graph_caches_follow_the_gate_list(ctx):
    """Repeated gate lists reuse the cached plan (and graph); a different list
    in between, with or without a graph, must not leave a stale op table."""
    rng = np.random.default_rng(31)
    n = 21
    kinds = ["H", "RX", "RY", "RZ", "X", "Y", "Z", "P", "CNOT", "SWAP", "TOFFOLI"]
    a = gates_array([random_gate(n, rng, kinds=kinds, max_controls=2) for _ in range(120)])
    b = gates_array([random_gate(n, rng, kinds=kinds, max_controls=2) for _ in range(90)])
    b2 = b.copy()
    b2["params"][:, 0] += 0.25  # same structure, new angles: a different plan
    psi = random_state(n, rng)
    expect = {}
    for name, g in (("a", a), ("b", b), ("b2", b2)):
        s = upload(psi, "c128")
        s.apply_gates_inplace(g, mode=L.EXEC_PER_GATE)
        expect[name] = s.state_vector
    for seq in ((("a", False), ("a", False), ("b", True), ("a", True), ("a", True), ("b2", False), ("b", True), ("b", True)),):
        for name, graph in seq:
            g = {"a": a, "b": b, "b2": b2}[name]
            s = upload(psi, "c128")
            s.apply_gates_inplace(g, mode=L.EXEC_TILED, use_graph=graph)
            assert np.max(np.abs(s.state_vector - expect[name])) <= 1e-12, (name, graph)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_swaps_relabel_qubits_instead_of_moving_data(ctx, prec):
    """Uncontrolled SWAPs rename qubits for the rest of the circuit: layers of
    RY on every qubit with SWAPs across the two halves of a 20-qubit state take
    the three 10-qubit passes the RY layers need (each tile holds the four
    lowest qubits), with an even number of SWAP layers (nothing to undo) or an
    odd one (the final permutation rides along). Kept as data moves, each SWAP
    layer would need tiles spanning both halves. Both against the oracle."""
    n = 20
    rng = np.random.default_rng(99)
    for layers, max_steps in ((6, 3), (5, 3)):
        ds = []
        for _ in range(layers):
            ds += [{"kind": "RY", "targets": [q], "params": [float(rng.uniform(0, 6.28))]} for q in range(n)]
            ds += [{"kind": "SWAP", "targets": [i, i + n // 2]} for i in range(n // 2)]
        gates = gates_array(ds)
        psi = random_state(n, rng)
        ref = O.apply_gates(psi, gates)
        s = upload(psi, prec)
        s.apply_gates_inplace(gates, mode=L.EXEC_TILED, tile_qubits=10)
        assert s.ctx.last_exec_steps() <= max_steps, (layers, s.ctx.last_exec_steps())
        check(s.state_vector, ref, prec, f"{layers} layers")


def test_unscaled_h_keeps_large_unnormalised_states_finite(ctx):
    """Multiples of H run unscaled with the factors applied later; an
    unnormalised complex64 state of magnitude 1e18 (State * Complex allows it)
    through 400 H gates and rotations stays finite and matches the oracle to
    single-precision relative accuracy."""
    n = 14
    rng = np.random.default_rng(8)
    ds = []
    for _ in range(40):
        ds += [{"kind": "H", "targets": [q]} for q in range(4, 14)]
        ds.append({"kind": "RY", "targets": [int(rng.integers(4, 14))], "params": [float(rng.uniform(0, 6.28))]})
    gates = gates_array(ds)
    psi = random_state(n, rng) * 1e18
    ref = O.apply_gates(psi, gates)
    s = upload(psi, "c64")
    s.apply_gates_inplace(gates, mode=L.EXEC_TILED, tile_qubits=12)
    got = s.state_vector
    assert np.all(np.isfinite(got))
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))


def _inverse(gates):
    """The inverse gate list of a circuit of H, X, Y, Z, RX, RY, RZ, P, CNOT, CZ,
    SWAP and Toffoli: reversed, rotation and phase angles negated."""
    inv = gates[::-1].copy()
    rot = np.isin(inv["kind"], [L.GATE_KINDS[k] for k in ("RX", "RY", "RZ", "P")])
    inv["params"][rot, 0] = -inv["params"][rot, 0]
    return inv


@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_random_circuit_then_inverse_restores_the_state_at_26_qubits(ctx, prec):
    """A size-independent check of the tiled executor at a large size: the
    depth-40 random circuit (relabeled SWAPs, shear rotations, unscaled H,
    controlled ops on every kind of tile position) followed by its inverse
    gives back |+...+>; the norm stays 1 throughout."""
    n = 26
    gates = W.random_circuit(n, 40, seed=4321)
    s = qiron.State.new_plus(n, prec)
    s.apply_gates_inplace(gates, mode=L.EXEC_TILED, use_graph=True)
    assert abs(s.norm_sqr() - 1.0) <= (1e-12 if prec == "c128" else 1e-4)
    s.apply_gates_inplace(_inverse(gates), mode=L.EXEC_TILED)
    got = s.state_vector
    d = float(np.max(np.abs(got - 2.0 ** (-n / 2))))
    assert d <= (1e-12 if prec == "c128" else 2e-6), d


def test_circuit_with_measurement_and_pauli_gates(ctx):
    """Gate::Measurement / PauliString / PauliTimeEvolution inside Circuit::execute."""
    n = 12
    rng = np.random.default_rng(2)
    ds = [random_gate(n, rng) for _ in range(20)]
    ds.append({"kind": "MEASURE", "targets": [], "controls": [3, 7], "params": [0.41]})
    ds += [random_gate(n, rng) for _ in range(10)]
    ds.append({"kind": "PAULI_STRING", "targets": [], "controls": [], "params": [2.0, 1.0], "x_mask": 0b101, "z_mask": 0b110})
    ds.append({"kind": "PAULI_EVOLUTION", "targets": [], "controls": [], "params": [0.7, 0.0, 0.3], "x_mask": 0b1001, "z_mask": 0b1})
    ds.append({"kind": "MEASURE", "targets": [], "controls": [], "params": [0.77], "basis": 1})
    gates = gates_array(ds)
    psi = random_state(n, rng)
    ref = O.apply_gates(psi, gates)
    for mode in (L.EXEC_PER_GATE, L.EXEC_TILED):
        s = upload(psi, "c128")
        s.apply_gates_inplace(gates, mode=mode, tile_qubits=10)
        check(s.state_vector, ref, "c128", f"mode {mode}")


def test_circuit_qubit_count_mismatch(ctx):
    s = qiron.State.new_zero(3)
    with pytest.raises(qiron.QuantIronError) as ei:
        s.apply_gates_inplace(gates_array([{"kind": "H", "targets": [0]}]), n_circuit=2)
    assert ei.value.variant == "InvalidNumberOfQubits" and ei.value.a == 3  # circuit.rs:161-163


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_pauli_sumop_expectation_match_oracle(ctx, prec):
    rng = np.random.default_rng(31)
    n = 11
    psi = random_state(n, rng)
    terms = random_terms(n, 12, rng, max_weight=5)
    for t in terms:
        ps = _ps_from_term(t)
        s = upload(psi, prec)
        check(ps.apply(s).state_vector, O.pauli_apply(psi, t), prec, "apply")
        check(ps.apply_normalised(s).state_vector, O.pauli_apply(psi, t, normalised=True), prec, "normalised")
        f = complex(0.2, -0.45)
        check(ps.apply_exp_factor(s, f).state_vector, O.pauli_exp(psi, t, f), prec, "exp")
    sop = qiron.SumOp([_ps_from_term(t) for t in terms])
    s = upload(psi, prec)
    check(sop.apply(s).state_vector, O.sumop_apply(psi, terms), prec, "sumop apply")
    e, eo = sop.expectation_value(s), O.expectation(psi, terms)
    assert abs(e - eo) <= (1e-12 if prec == "c128" else 1e-5), (e, eo)
    real = term_array([(int(t["x_mask"]), int(t["z_mask"]), float(t["c_re"]), 0.0) for t in terms])
    for t in real:
        check(_ps_from_term(t).apply_exp_neg_i_dt(s, 0.3).state_vector, O.pauli_exp_neg_i_dt(psi, t, 0.3), prec, "dt")


def test_inner_product_norm_fidelity(ctx):
    rng = np.random.default_rng(4)
    a, b = random_state(12, rng), random_state(12, rng)
    sa, sb = upload(a, "c128"), upload(b, "c128")
    assert abs(sa.inner_product(sb) - O.inner_product(a, b)) < 1e-14
    assert abs(sa.norm_sqr() - 1) < 1e-14
    assert abs(sa.fs_fidelity(sb) - O.fidelity(a, b)) < 1e-14
    z = upload(np.zeros(8, dtype=complex), "c128")
    with pytest.raises(qiron.QuantIronError) as ei:
        z.normalise()
    assert ei.value.variant == "ZeroNorm"


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_measurement_matches_oracle(ctx, prec):
    rng = np.random.default_rng(12)
    n = 10
    psi = random_state(n, rng)
    s = upload(psi, prec)
    for qs in ([], [0], [9, 2, 5], [4, 4], [1, 3, 1, 0]):
        np.testing.assert_allclose(s.measure_probs(qs), O.measure_probs(psi, qs), atol=1e-12 if prec == "c128" else 1e-6)
    us = rng.random(400)
    qs = [7, 1, 3]
    got = s.measure_outcomes(qs, us)
    ref = O.measure_n(psi, 0, qs, us)
    assert np.array_equal(got, ref)
    for basis, custom in ((1, None), (2, None), (3, np.array([[0.6, 0.8j], [0.8j, 0.6]]))):
        b = (qiron.MeasurementBasis.X if basis == 1 else qiron.MeasurementBasis.Y if basis == 2
             else qiron.MeasurementBasis.Custom(custom))
        for u in (0.05, 0.5, 0.93):
            r = s.measure(b, [2, 6], uniform=u)
            o, post = O.measure(psi, basis, [2, 6], u, custom)
            assert sum(bit << k for k, bit in enumerate(r.outcomes)) == o
            check(r.new_state.state_vector, post, prec, f"basis {basis}")


def test_measure_errors(ctx):
    s = qiron.State.new_zero(3)
    with pytest.raises(qiron.QuantIronError) as ei:
        s.measure_n(qiron.MeasurementBasis.Computational, [0], 0, uniforms=[])
    assert ei.value.variant == "InvalidNumberOfMeasurements" and ei.value.a == 0
    with pytest.raises(qiron.QuantIronError) as ei:
        s.measure(qiron.MeasurementBasis.Computational, [0, 1, 2, 0])
    assert ei.value.variant == "InvalidNumberOfQubits" and ei.value.a == 3


@pytest.mark.parametrize("order", [1, 2])
def test_trotter_lattice_matches_oracle(ctx, order):
    terms, xs = W.trotter_2d(2, 3, seed=42)
    n = 6
    psi = O.init(n, 0)
    for q in xs:
        psi = O.apply_gate(psi, to_gate({"kind": "X", "targets": [q]}))
    ref = O.trotter(psi, terms, 0.02, 20, order)
    for prec in ("c128", "c64"):
        s = upload(psi, prec)
        got = qiron.trotter_evolve_state(_sumop_from_arr(terms), s, 0.02, 20, order).state_vector
        check(got, ref, prec, f"trotter {prec}")


@pytest.mark.parametrize("shape,order", [((4, 4), 1), ((4, 4), 2), ((4, 5), 1)])
def test_trotter_fused_steps_match_oracle(ctx, shape, order):
    """From 16 qubits a Trotter step runs as one gate list through the tiled
    executor (the XX, YY, ZZ terms of a bond as one 7-gate block; other terms:
    Y factors turned to X, CNOTs and CZs gathering the string onto one X, RX
    there and the gathering undone; Z strings a CNOT ladder and RZ): the same
    unitary as the per-term cos/sin form up to rounding."""
    terms, xs = W.trotter_2d(*shape, seed=43)
    n = shape[0] * shape[1]
    psi = O.init(n, 0)
    for q in xs:
        psi = O.apply_gate(psi, to_gate({"kind": "X", "targets": [q]}))
    ref = O.trotter(psi, terms, 0.02, 8, order)
    s = upload(psi, "c128")
    got = qiron.trotter_evolve_state(_sumop_from_arr(terms), s, 0.02, 8, order).state_vector
    check(got, ref, "c128", f"trotter c128 {shape}")
    # complex64: each step is ~7 gates per term (about 7000 gate applications in
    # all), so FP32 rounding accumulates past the per-amplitude bar of short
    # circuits; the state must still match to a 1e-7 fidelity deficit
    s = upload(psi, "c64")
    got = qiron.trotter_evolve_state(_sumop_from_arr(terms), s, 0.02, 8, order).state_vector
    d = float(np.max(np.abs(got - ref)))
    f = fidelity(got, ref)
    assert d <= 1e-4 and f >= 1 - 1e-7, (shape, order, d, 1 - f)
    # every rotation of the fused step runs in a norm-preserving form: a
    # rotation with rounded cos / sin entries (det != 1) would drift the norm
    # by ~4e-8 per application, about 1e-5 here
    drift = abs(float(np.vdot(got, got).real) - 1.0)
    assert drift <= 5e-6, (shape, order, drift)


@pytest.mark.parametrize("order", [1, 2])
def test_trotter_fused_mixed_pauli_strings_match_oracle(ctx, order):
    """Strings mixing X, Y and Z factors (weights 1..5, a pivot at any position
    among them) on 17 qubits, tiled executor against the oracle's per-term
    cos/sin product: a wrong conjugation (a CZ for an X factor, a missing
    S / S-dagger pair, the wrong pivot) moves amplitudes by O(theta)."""
    rng = np.random.default_rng(500 + order)
    n = 17
    terms = random_terms(n, 24, rng, max_weight=5)
    terms["c_im"] = 0.0  # the fused path takes Hermitian terms
    terms = terms[(terms["x_mask"] | terms["z_mask"]) != 0]
    assert np.any(terms["x_mask"] & terms["z_mask"]) and np.any(terms["x_mask"] & ~terms["z_mask"])
    psi = random_state(n, rng)
    ref = O.trotter(psi, terms, 0.05, 3, order)
    s = upload(psi, "c128")
    got = qiron.trotter_evolve_state(_sumop_from_arr(terms), s, 0.05, 3, order).state_vector
    check(got, ref, "c128", f"trotter mixed strings order {order}")


@pytest.mark.parametrize("order", [1, 2])
def test_trotter_fused_bond_runs_match_oracle(ctx, order):
    """Runs of XX / YY / ZZ terms on one qubit pair (every subset, repeated
    terms, XX and YY with equal coefficients) between other strings, 16
    qubits, against the oracle's term-by-term product."""
    rng = np.random.default_rng(600 + order)
    n = 16
    ts = []
    subsets = [(1,), (2,), (3,), (1, 2), (1, 3), (2, 3), (1, 2, 3), (1, 1, 2), (3, 1, 2), (1, 2)]
    for i, sub in enumerate(subsets):
        p, q = (int(x) for x in rng.choice(n, size=2, replace=False))
        b = (1 << p) | (1 << q)
        same = i == len(subsets) - 1  # XX and YY with equal coefficients
        c0 = float(rng.normal())
        for k in sub:
            c = c0 if same else float(rng.normal())
            ts.append({1: (b, 0, c, 0.0), 2: (b, b, c, 0.0), 3: (0, b, c, 0.0)}[k])
        r = int(rng.integers(n))
        ts.append((1 << r, 0, float(rng.normal()), 0.0))  # an X term between runs
    terms = term_array(ts)
    psi = random_state(n, rng)
    ref = O.trotter(psi, terms, 0.07, 3, order)
    s = upload(psi, "c128")
    got = qiron.trotter_evolve_state(_sumop_from_arr(terms), s, 0.07, 3, order).state_vector
    check(got, ref, "c128", f"trotter bond runs order {order}")


def _sumop_from_arr(arr):
    return qiron.SumOp([_ps_from_term(t) for t in arr])


def test_config1_full_size_matches_oracle(ctx):
    """20-qubit QFT + 5 rotation/CNOT layers + open Heisenberg energy, complex128."""
    gates, terms = W.config1(n=20)
    psi = O.init(20, 2)
    ref = O.apply_gates(psi, gates)
    e_ref = O.expectation(ref, terms)
    for opts in (dict(mode=L.EXEC_PER_GATE), dict(mode=L.EXEC_TILED), dict(mode=L.EXEC_FUSED_MFMA)):
        s = qiron.State.new_plus(20)
        s.apply_gates_inplace(gates, **opts)
        got = s.state_vector
        assert np.max(np.abs(got - ref)) <= 1e-12, opts
        e = _sumop_from_arr(terms).expectation_value(s)
        assert abs(e - e_ref) <= 1e-12, (opts, e, e_ref)


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_qft_matches_the_fft_closed_form_at_20_qubits(ctx, prec):
    """Subroutine::qft (subroutine.rs:90-112) on the engine, checked with no
    oracle in the loop: the gate list computes R F R, with F the unitary DFT
    (omega = e^{+2 pi i/N}, numpy's ifft times sqrt(N)) and R the bit
    reversal of the index, on a random 20-qubit state, in every executor."""
    n = 20
    N = 1 << n
    gl = W.GateList()
    W.qft_gates(gl, list(range(n)))
    gates = gl.array()
    idx = np.arange(N)
    rev = np.zeros(N, dtype=np.int64)
    for b in range(n):
        rev |= ((idx >> b) & 1) << (n - 1 - b)
    psi = random_state(n, np.random.default_rng(2024))
    exp = (np.fft.ifft(psi[rev]) * math.sqrt(N))[rev]
    for mode in (L.EXEC_PER_GATE, L.EXEC_TILED, L.EXEC_FUSED_MFMA):
        s = upload(psi, prec)
        s.apply_gates_inplace(gates, mode=mode)
        check(s.state_vector, exp, prec, f"qft mode {mode}")


def test_large_state_properties(ctx):
    """26 qubits (complex64): norm preserved and the three executors agree."""
    n = 26
    gates = W.random_circuit(n=n, depth=3, seed=99)
    outs = []
    for opts in (dict(mode=L.EXEC_PER_GATE), dict(mode=L.EXEC_TILED), dict(mode=L.EXEC_FUSED_MFMA, block_qubits=6)):
        s = qiron.State.new_plus(n, "c64")
        s.apply_gates_inplace(gates, **opts)
        assert abs(s.norm_sqr() - 1) < 1e-4
        outs.append(s.state_vector)
    for o in outs[1:]:
        assert fidelity(o, outs[0]) > 1 - 1e-8
