"""State-level operations on the engine against the reference's known answers,
closed forms and the oracle: the Fubini-Study metrics, State arithmetic,
measurement in every basis (outcomes and per-shot collapsed states), the
inverse QFT, and the error paths (zero norm, out of memory, a collapse onto
an outcome of zero probability)."""
import math

import numpy as np
import pytest

import oracle as O
import qiron
from kat_util import cvec, load_kats, random_state, random_unitary2
from qiron import _lib as L
from qiron import workloads as W

pytestmark = pytest.mark.gpu

KATS = load_kats()
EPS64 = float(np.finfo(np.float64).eps)


def upload(vec, prec):
    return qiron.State.from_vector_unchecked(vec, prec)


def engine_metric(name, a, b):
    return a.fs_dist(b) if name == "fs_dist" else a.fs_fidelity(b)


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("case", KATS["metrics"], ids=[c["name"] for c in KATS["metrics"]])
def test_fs_metric_kats(ctx, case, prec):
    """state_tests.rs:225-295: F = 0, 1, 1/2 and D = 0, pi/4 (the reference
    asserts f64::EPSILON; complex64 inputs are exact here, the FP64 reduction
    keeps the bar)."""
    a, b = upload(cvec(case["a"]), prec), upload(cvec(case["b"]), prec)
    got = engine_metric(case["metric"], a, b)
    tol = EPS64 if prec == "c128" else 1e-7
    assert abs(got - case["expected"]) < tol, (case["cite"], got)


@pytest.mark.parametrize("case", KATS["metric_errors"], ids=[c["name"] for c in KATS["metric_errors"]])
def test_fs_metric_error_kats(ctx, case):
    a = upload(cvec(case["a"]), "c128")
    with pytest.raises(qiron.QuantIronError) as ei:
        if case["metric"] == "normalise":
            a.normalise()
        else:
            engine_metric(case["metric"], a, upload(cvec(case["b"]), "c128"))
    assert (ei.value.code, ei.value.a) == (case["code"], case["a_payload"]), case["cite"]


def test_normalise_kat_is_bit_exact(ctx):
    """state_tests.rs:149-156 asserts [3, 4] -> [0.6, 0.8] with assert_eq."""
    case = next(c for c in KATS["cases"] if c["name"] == "normalise_3_4")
    got = upload(cvec(case["init"]), "c128").normalise().state_vector
    np.testing.assert_array_equal(got, cvec(case["expected"]))


def test_fs_metrics_match_oracle_on_random_states(ctx):
    rng = np.random.default_rng(41)
    for n in (3, 12, 17):
        a, b = random_state(n, rng) * 2.5, random_state(n, rng) * 0.3
        sa, sb = upload(a, "c128"), upload(b, "c128")
        assert abs(sa.fs_fidelity(sb) - O.fs_fidelity(a, b)) < 1e-14
        assert abs(sa.fs_dist(sb) - O.fs_dist(a, b)) < 1e-12


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_state_add_sub_mul_are_the_ieee_elementwise_results(ctx, prec):
    """impl Add / Sub / Mul for State (state.rs:2687-2862) are one rounding per
    component: the engine must equal numpy's elementwise arithmetic in the same
    precision bit for bit, and the oracle's in FP64."""
    dt = np.complex128 if prec == "c128" else np.complex64
    rng = np.random.default_rng(7)
    for n in (1, 5, 16):
        a, b = random_state(n, rng).astype(dt), (random_state(n, rng) * 3.0).astype(dt)
        sa, sb = upload(a, prec), upload(b, prec)
        np.testing.assert_array_equal((sa + sb).state_vector, (a + b).astype(np.complex128))
        np.testing.assert_array_equal((sa - sb).state_vector, (a - b).astype(np.complex128))
        np.testing.assert_array_equal((sa + sa).state_vector, (2 * a).astype(np.complex128))
        if prec == "c128":
            np.testing.assert_array_equal((sa + sb).state_vector, O.state_add(a, b))
            np.testing.assert_array_equal((sa - sb).state_vector, O.state_sub(a, b))
            c = complex(-0.75, 0.5)
            np.testing.assert_array_equal((sa * c).state_vector, O.state_scale(a, c))
            np.testing.assert_array_equal((2.5 * sa).state_vector, O.state_scale(a, 2.5))
        # inputs are left unchanged (value semantics)
        np.testing.assert_array_equal(sa.state_vector, a.astype(np.complex128))
    # Sum is a left fold of Add (state.rs:2802-2834)
    xs = [upload(random_state(4, rng).astype(dt), prec) for _ in range(3)]
    v = [x.state_vector.astype(dt) for x in xs]
    np.testing.assert_array_equal(qiron.State.sum(xs).state_vector, ((v[0] + v[1]) + v[2]).astype(np.complex128))
    with pytest.raises(ValueError, match="different numbers of qubits"):
        _ = xs[0] + upload(random_state(5, rng), prec)
    with pytest.raises(ValueError, match="empty"):
        qiron.State.sum([])


def test_measurement_many_bins_is_exact_away_from_bin_boundaries(ctx):
    """2^17 outcome bins (the chunked histogram and scan): every outcome equals
    the oracle's unless u lies within 1e-12 of one of the oracle's cumulative
    boundaries, and then the two outcomes are neighbouring bins."""
    rng = np.random.default_rng(13)
    n = 18
    psi = random_state(n, rng)
    s = upload(psi, "c128")
    qs = list(range(17))
    probs = O.measure_probs(psi, qs)
    np.testing.assert_allclose(s.measure_probs(qs), probs, atol=1e-14)
    cdf = O.sample_cdf(probs)
    us = np.concatenate([rng.random(400), cdf[rng.integers(0, cdf.size - 1, 20)]])  # some draws on a boundary
    got = s.measure_outcomes(qs, us)
    ref = O.measure_n(psi, 0, qs, us)
    for u, g, r in zip(us, got, ref):
        if g != r:
            assert abs(int(g) - int(r)) == 1, (u, g, r)
            assert abs(u - cdf[min(g, r)]) <= 1e-12, (u, g, r)
    assert np.mean(got == ref) >= 1 - 20 / us.size
    assert np.array_equal(got[:400], ref[:400])  # random draws: no boundary within reach
    got_all = s.measure_outcomes([], us[:50])
    assert np.array_equal(got_all, O.measure_n(psi, 0, [], us[:50]))


BASES = [(0, None), (1, None), (2, None), (3, "random")]


def _basis(code, custom):
    return {0: qiron.MeasurementBasis.Computational, 1: qiron.MeasurementBasis.X,
            2: qiron.MeasurementBasis.Y}.get(code) or qiron.MeasurementBasis.Custom(custom)


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("code,custom", BASES, ids=["computational", "x", "y", "custom"])
def test_measure_n_every_basis_matches_oracle(ctx, code, custom, prec):
    """measure_n (state.rs:750-784) in each basis: outcomes identical to the
    oracle's for identical uniforms, and each shot's collapsed state equal to
    the oracle's State::measure result for that shot's uniform."""
    rng = np.random.default_rng(100 + code)
    n = 9
    psi = random_state(n, rng)
    U = random_unitary2(rng) if custom else None
    s = upload(psi, prec)
    qs = [5, 0, 7]
    us = rng.random(64)
    shots = s.measure_n(_basis(code, U), qs, len(us), uniforms=us)
    ref = O.measure_n(psi, code, qs, us, U)
    got = np.array([sum(b << k for k, b in enumerate(r.outcomes)) for r in shots], dtype=np.uint64)
    assert np.array_equal(got, ref)
    for r, u in list(zip(shots, us))[:12]:
        o, post = O.measure(psi, code, qs, u, U)
        d = float(np.max(np.abs(r.new_state.state_vector - post)))
        assert d <= (1e-12 if prec == "c128" else 2e-6), (u, d)
        assert r.indices == qs


def test_engine_draws_follow_the_seeded_stream(ctx):
    """With no injected uniforms the engine draws from the context's
    SplitMix64 stream: after seeding, the outcomes equal those of the same
    draws injected, and a measurement gate with no uniform draws anew each run."""
    rng = np.random.default_rng(3)
    psi = random_state(6, rng)
    s = upload(psi, "c128")
    ctx.seed(1234)
    drawn = s.measure_n(qiron.MeasurementBasis.Computational, [0, 3, 4], 200)
    us = W.SplitMix64(1234).uniforms(200)
    ref = O.measure_n(psi, 0, [0, 3, 4], us)
    assert [sum(b << k for k, b in enumerate(r.outcomes)) for r in drawn] == [int(x) for x in ref]
    ctx.seed(99)
    r = s.measure(qiron.MeasurementBasis.X, [2])
    o, post = O.measure(psi, 1, [2], W.SplitMix64(99).u())
    assert r.outcomes == [o]
    np.testing.assert_allclose(r.new_state.state_vector, post, atol=1e-12)
    # a Gate::Measurement without a uniform: each execution draws the next value
    ctx.seed(5)
    g = qiron.Gate.new_measurement([0, 1, 2, 3, 4, 5])
    u5 = W.SplitMix64(5).uniforms(2)
    for k in range(2):
        t = upload(psi, "c128")
        t.apply_gate_inplace(g)
        o, post = O.measure(psi, 0, list(range(6)), u5[k])
        np.testing.assert_allclose(t.state_vector, post, atol=1e-12)


def test_collapse_onto_a_zero_probability_outcome_is_rejected_at_any_size(ctx):
    """An outcome of probability 0 leaves an all-zero state, which State::new
    rejects (state.rs:117-122); complex64 states of 23+ qubits included."""
    for n, prec in ((3, "c128"), (23, "c64"), (23, "c128")):
        s = qiron.State.new_zero(n, prec)
        out = qiron.State._alloc(n, prec)
        with pytest.raises(qiron.QuantIronError) as ei:
            err = L.QiError()
            qs = np.array([0], dtype=np.uint64)
            qiron.core._check(s.lib.qi_measure_collapse(s._h, out._h, L.u64ptr(qs), 1, 1, qiron.core.C.byref(err)), err)
        assert ei.value.variant == "StateVectorNotNormalised", (n, prec)


def test_out_of_memory_is_reported(ctx):
    """A state larger than the device (2^36 complex64 amplitudes = 512 GiB)
    returns the engine's out-of-memory status, and the context stays usable."""
    with pytest.raises(qiron.QuantIronError) as ei:
        qiron.State.new_zero(36, "c64")
    assert ei.value.variant == "OutOfMemory"
    s = qiron.State.new_plus(10, "c64")
    assert abs(s.norm_sqr() - 1.0) < 1e-6


def _rev(n):
    idx = np.arange(1 << n)
    rev = np.zeros(1 << n, dtype=np.int64)
    for b in range(n):
        rev |= ((idx >> b) & 1) << (n - 1 - b)
    return rev


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_iqft_matches_the_inverse_dft_closed_form_at_20_qubits(ctx, prec):
    """Subroutine::iqft (subroutine.rs:125-160) computes R F^-1 R (F^-1 = numpy's
    fft / sqrt(N)), and iqft after qft is the identity, in every executor."""
    n = 20
    N = 1 << n
    rev = _rev(n)
    psi = random_state(n, np.random.default_rng(2025))
    exp = (np.fft.fft(psi[rev]) / math.sqrt(N))[rev]
    iqft = qiron.Subroutine.iqft(list(range(n)), n)
    qft = qiron.Subroutine.qft(list(range(n)), n)
    tol = 1e-12 if prec == "c128" else 1e-5
    for mode in (L.EXEC_PER_GATE, L.EXEC_TILED, L.EXEC_FUSED_MFMA):
        s = upload(psi, prec)
        iqft.execute_inplace(s, mode=mode)
        assert float(np.max(np.abs(s.state_vector - exp))) <= tol, mode
        qft.execute_inplace(s, mode=mode)
        assert float(np.max(np.abs(s.state_vector - psi))) <= tol, mode
