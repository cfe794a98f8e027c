"""The 30-qubit depth-40 random circuit at its full size, and the memory
ceiling (34 qubits complex64, 33 qubits complex128, 128 GiB states).

At these sizes the oracle is too slow to run, so the checks are
size-independent properties: the tiled executor against the per-gate
streaming executor in complex128 (two independent kernel families) within
1e-12 per amplitude; the complex64 fidelity deficit against complex128; and a
circuit run forward by one executor and inverted by the other, which must
return |+...+> exactly (a no-op executor fails it)."""
import time

import numpy as np
import pytest

import qiron
from qiron import _lib as L
from qiron import workloads as W

pytestmark = pytest.mark.gpu

N30 = 30


def _max_abs_diff(a: "qiron.State", b: "qiron.State") -> float:
    """max_i |a_i - b_i|: a - b on the device, then the maximum over host chunks."""
    d = (a - b).state_vector
    m = 0.0
    for k in range(0, d.size, 1 << 26):
        m = max(m, float(np.max(np.abs(d[k:k + (1 << 26)]))))
    return m


def _x_sum(n):
    return qiron.SumOp([qiron.PauliString(1.0, {q: "X"}) for q in range(n)])


def test_30q_random_circuit_c128_tiled_equals_per_gate(ctx):
    gates = W.random_circuit(n=N30, depth=40, seed=W.SEED_BASE + 2)
    t0 = time.perf_counter()
    a = qiron.State.new_plus(N30, "c128")
    a.apply_gates_inplace(gates, mode=L.EXEC_TILED)
    b = qiron.State.new_plus(N30, "c128")
    b.apply_gates_inplace(gates, mode=L.EXEC_PER_GATE)
    d = _max_abs_diff(a, b)
    print(f"30q c128 tiled vs per-gate: max|diff| = {d:.3e} ({time.perf_counter() - t0:.1f} s)", flush=True)
    assert d <= 1e-12
    assert abs(a.norm_sqr() - 1.0) < 1e-12


def test_30q_random_circuit_c64_fidelity_deficit(ctx):
    """complex64 amplitudes through the tiled executor against the complex128
    run of the same circuit: 1 - |<psi64|psi128>|^2 (stated in the design notes)."""
    gates = W.random_circuit(n=N30, depth=40, seed=W.SEED_BASE + 2)
    s64 = qiron.State.new_plus(N30, "c64")
    s64.apply_gates_inplace(gates, mode=L.EXEC_TILED)
    wide = qiron.State.from_vector_unchecked(s64.state_vector, "c128")
    del s64
    s128 = qiron.State.new_plus(N30, "c128")
    s128.apply_gates_inplace(gates, mode=L.EXEC_TILED)
    deficit = 1.0 - wide.fs_fidelity(s128)
    e64 = _x_sum(N30).expectation_value(wide).real
    e128 = _x_sum(N30).expectation_value(s128).real
    print(f"30q depth-40: c64 fidelity deficit vs c128 = {deficit:.3e}, |d<sum X>| = {abs(e64 - e128):.3e}", flush=True)
    assert deficit <= 1e-9
    assert abs(e64 - e128) <= 1e-5


@pytest.mark.parametrize("n,prec", [(34, "c64"), (33, "c128")])
def test_memory_ceiling_forward_tiled_inverse_per_gate(ctx, n, prec):
    """One layer of the random circuit on a 128 GiB state: forward through the
    tiled executor, inverse through the per-gate kernels; <sum_q X_q> = n holds
    only for |+...+> (every <X_q> <= 1), so both executors must have acted."""
    gates = W.random_circuit(n=n, depth=1, seed=W.SEED_BASE + 2)
    s = qiron.State.new_plus(n, prec)
    s.apply_gates_inplace(gates, mode=L.EXEC_TILED)
    mid = _x_sum(n).expectation_value(s).real
    s.apply_gates_inplace(W.inverse_gates(gates), mode=L.EXEC_PER_GATE)
    e = _x_sum(n).expectation_value(s).real
    norm = s.norm_sqr()
    print(f"{n}q {prec}: <sum X> after the layer {mid:.6f}, after its inverse {e:.12f} (n = {n}), norm {norm:.12f}",
          flush=True)
    assert abs(mid - n) > 0.5  # the layer moved the state
    tol = 1e-9 if prec == "c128" else 2e-4
    assert abs(e - n) <= tol * n and abs(norm - 1.0) <= tol
