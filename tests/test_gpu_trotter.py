"""Trotter steps on the streaming kernels with diagonal-run fusion: runs of
diagonal terms (Z strings) that commute through the terms between them are
applied in one pass, and repeated exponentials of one string merge. Both are
exact rewrites of the term-by-term product (time_evolution.rs:45-115): the
results match the oracle's per-term chain to rounding."""
import numpy as np
import pytest

import oracle as O
import qiron
from kat_util import random_state, term_array

pytestmark = pytest.mark.gpu


def ising(n, jz, hz, hx, zz_first=True):
    zz = [((0, (1 << i) | (1 << (i + 1)), -jz, 0.0)) for i in range(n - 1)]
    z = [(0, 1 << i, -hz * (1 + 0.1 * i), 0.0) for i in range(n)]
    x = [(1 << i, 0, -hx * (1 - 0.05 * i), 0.0) for i in range(n)]
    return term_array(zz + z + x) if zz_first else term_array([t for i in range(n) for t in
                                                               ([zz[i]] if i < n - 1 else []) + [z[i], x[i]]])


def run(psi, terms, dt, steps, order, prec):
    s = qiron.State.from_vector_unchecked(psi, prec)
    qiron.trotter_evolve_terms_inplace(s, terms, dt, steps, order)
    return s.state_vector, s.ctx.last_exec_steps()


@pytest.mark.parametrize("order", [1, 2])
def test_diagonal_runs_fuse_into_one_pass(ctx, order):
    """Transverse-field Ising with all ZZ and Z terms first: one diagonal pass,
    then one pass per X term (second order: the middle X terms merge)."""
    n = 12
    terms = ising(n, 1.0, 0.7, 0.9)
    psi = random_state(n, np.random.default_rng(50 + order))
    ref = O.trotter(psi, terms, 0.05, 6, order)
    got, passes = run(psi, terms, 0.05, 6, order, "c128")
    assert float(np.max(np.abs(got - ref))) <= 1e-12
    # order 1: 1 diagonal + n X passes; order 2: forward (1 + n), the two copies of
    # the last X term merge, backward (n - 1) X passes + 1 diagonal
    assert passes == (1 + n if order == 1 else 2 * n + 1), passes
    got64, _ = run(psi, terms, 0.05, 6, order, "c64")
    assert float(np.max(np.abs(got64 - ref))) <= 5e-6


def test_interleaved_terms_fuse_only_where_they_commute(ctx):
    """ZZ, Z and X terms interleaved site by site: Z_i moves back past X_{i-1}
    (different qubits) to join ZZ_{i-1,i}'s run only where the strings commute;
    the result still equals the term-by-term product."""
    n = 10
    terms = ising(n, 0.8, 0.6, 1.1, zz_first=False)
    psi = random_state(n, np.random.default_rng(7))
    ref = O.trotter(psi, terms, 0.07, 5, 1)
    got, passes = run(psi, terms, 0.07, 5, 1, "c128")
    assert float(np.max(np.abs(got - ref))) <= 1e-12
    assert passes < terms.size  # some diagonal terms shared a pass


def test_complex_coefficients_and_wide_strings(ctx):
    """Non-Hermitian coefficients (apply_exp_factor with a complex exponent) and
    Z strings wider than one table, against the oracle."""
    n = 13
    rng = np.random.default_rng(8)
    ts = [(0, (1 << 13) - 1, 0.3, 0.2),              # a 13-qubit Z string: no table fits, streamed alone
          (0, 0b1011, 0.5, -0.1), (0, 0b110000000, -0.4, 0.0), (0, 0, 0.1, 0.05),  # incl. the identity string
          (0b100, 0b100, 0.2, 0.0), (0, 0b1001000000001, 0.25, 0.3)]
    terms = term_array(ts)
    psi = random_state(n, rng)
    for order in (1, 2):
        ref = O.trotter(psi, terms, 0.1, 3, order)
        got, _ = run(psi, terms, 0.1, 3, order, "c128")
        assert float(np.max(np.abs(got - ref))) <= 1e-12, order
