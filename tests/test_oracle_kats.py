"""The oracle against the reference's own known-answer tests (tests/golden/kats.json)."""
import numpy as np
import pytest

import oracle as O
from kat_util import F32_EPS, cvec, gates_array, load_kats, term_array, to_gate

KATS = load_kats()
CASES = KATS["cases"]


def run_oracle(case):
    op = case.get("op", "gates")
    if op == "init":
        return O.init(case["n"], case["kind"])
    psi = cvec(case["init"])
    if op == "gates":
        return O.apply_gates(psi, gates_array(case["gates"]))
    if op == "pauli_apply":
        return O.pauli_apply(psi, term_array([case["term"]])[0])
    if op == "pauli_exp_neg_i_dt":
        return O.pauli_exp_neg_i_dt(psi, term_array([case["term"]])[0], case["dt"])
    if op == "sumop_apply":
        return O.sumop_apply(psi, term_array(case["terms"]))
    if op == "expectation":
        return np.array([O.expectation(psi, term_array(case["terms"]))])
    if op == "trotter":
        return O.trotter(psi, term_array(case["terms"]), case["dt"], case["steps"], case["order"])
    if op == "normalise":
        return O.normalise(psi)
    raise ValueError(op)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_matches_reference_kat(case):
    got = run_oracle(case)
    exp = cvec(case["expected"])
    # the reference's PartialEq tolerance, and the FP64 closed form to 1e-12
    assert np.all(np.abs(got.real - exp.real) <= F32_EPS) and np.all(np.abs(got.imag - exp.imag) <= F32_EPS)
    assert np.max(np.abs(got - exp)) <= 1e-12, case["cite"]
    if case.get("bit_exact"):
        np.testing.assert_array_equal(got, exp)


ERRS = KATS["errors"]


@pytest.mark.parametrize("case", ERRS, ids=[c["name"] for c in ERRS])
def test_oracle_error_variants(case):
    psi = O.init(case["n"], 0)
    with pytest.raises(O.OracleError) as ei:
        O.apply_gate(psi, to_gate(case["gate"]))
    e = ei.value
    assert (e.code, e.a, e.b) == (case["code"], case["a"], case["b"]), case["cite"]


METRICS = KATS["metrics"]


def oracle_metric(name, a, b):
    return O.fs_dist(a, b) if name == "fs_dist" else O.fs_fidelity(a, b)


@pytest.mark.parametrize("case", METRICS, ids=[c["name"] for c in METRICS])
def test_oracle_fs_metrics(case):
    """fs_fidelity / fs_dist known answers (state_tests.rs:225-295); the
    reference asserts them within f64::EPSILON."""
    got = oracle_metric(case["metric"], cvec(case["a"]), cvec(case["b"]))
    assert abs(got - case["expected"]) < np.finfo(np.float64).eps, case["cite"]


MERRS = KATS["metric_errors"]


@pytest.mark.parametrize("case", MERRS, ids=[c["name"] for c in MERRS])
def test_oracle_fs_metric_errors(case):
    with pytest.raises(O.OracleError) as ei:
        if case["metric"] == "normalise":
            O.normalise(cvec(case["a"]))
        else:
            oracle_metric(case["metric"], cvec(case["a"]), cvec(case["b"]))
    assert (ei.value.code, ei.value.a) == (case["code"], case["a_payload"]), case["cite"]
