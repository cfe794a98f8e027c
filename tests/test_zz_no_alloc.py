"""qi_sumop_expectation_noalloc (no allocation: static device scratch) against
qi_sumop_expectation through the C ABI: identical bits on two calls in a row,
the second with a smaller state and fewer terms, in both precisions."""
import ctypes as C

import numpy as np
import pytest

import qiron
from kat_util import random_state, random_terms
from qiron import _lib as L

pytestmark = pytest.mark.gpu


def both(state, terms):
    lib = state.lib
    arr = np.ascontiguousarray(terms)
    out = []
    for fn in (lib.qi_sumop_expectation, lib.qi_sumop_expectation_noalloc):
        re, im = C.c_double(), C.c_double()
        err = L.QiError()
        rc = fn(state._h, arr.ctypes.data_as(C.c_void_p), arr.size, C.byref(re), C.byref(im), C.byref(err))
        assert rc == 0, err.msg
        out.append((re.value, im.value))
    return out


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_noalloc_expectation_equals_the_allocating_entry_point(ctx, prec):
    rng = np.random.default_rng(4242)
    for n, nt in ((18, 300), (9, 17)):  # the second call: smaller state, fewer terms
        psi = random_state(n, rng)
        s = qiron.State.from_vector_unchecked(psi, prec)
        terms = random_terms(n, nt, rng, max_weight=6)
        (a_re, a_im), (b_re, b_im) = both(s, terms)
        assert (a_re, a_im) == (b_re, b_im), (n, nt, a_re - b_re, a_im - b_im)


def test_noalloc_expectation_limits(ctx):
    s = qiron.State.new_plus(6, "c128")
    lib = s.lib
    rng = np.random.default_rng(1)
    big = np.ascontiguousarray(random_terms(6, 513, rng))
    re, im = C.c_double(), C.c_double()
    err = L.QiError()
    rc = lib.qi_sumop_expectation_noalloc(s._h, big.ctypes.data_as(C.c_void_p), big.size, C.byref(re), C.byref(im),
                                          C.byref(err))
    assert rc == 17 and b"maximum" in err.msg  # InvalidArgument, the limit named
    empty = np.zeros(0, dtype=L.TERM_DTYPE)
    rc = lib.qi_sumop_expectation_noalloc(s._h, empty.ctypes.data_as(C.c_void_p), 0, C.byref(re), C.byref(im),
                                          C.byref(err))
    assert rc == 0 and (re.value, im.value) == (0.0, 0.0)  # an empty SumOp has expectation 0
