"""Parametric circuits (Gate::Parametric over Parameter<N>, gate.rs:107-114,
parameter.rs:12-68): planned once, re-run with new angle tables, against the
oracle on the same gates with the angles substituted."""
import math

import numpy as np
import pytest

import oracle as O
import qiron
from qiron import _lib as L
from qiron import workloads as W

pytestmark = pytest.mark.gpu


def param_ansatz(n, layers):
    """RY/RZ layers (one slot per rotation), a controlled RZ sharing slot 0, H
    and CNOT ring gates without slots: (gates, slots, number of slots)."""
    gl = W.GateList(8 * n * layers + 8)
    slots, k = [], 0
    for _ in range(layers):
        for kind in ("RY", "RZ", "RX"):
            for q in range(n):
                gl.add(kind, [q], [], [0.0])
                slots.append(k)
                k += 1
        gl.add("H", [0])
        slots.append(-1)
        gl.add("RZ", [n - 1], [0], [0.0])  # controlled, shares slot 0
        slots.append(0)
        gl.add("P", [1], [], [0.0])
        slots.append(k)
        k += 1
        for q in range(n):
            gl.add("CNOT", [(q + 1) % n], [q])
            slots.append(-1)
    return gl.array(), np.array(slots, dtype=np.int32), k


def substituted(gates, slots, angles):
    g = gates.copy()
    on = slots >= 0
    g["params"][on, 0] = angles[slots[on]]
    return g


@pytest.mark.parametrize("prec", ["c128", "c64"])
@pytest.mark.parametrize("n,mode", [(8, L.EXEC_AUTO), (14, L.EXEC_TILED), (20, L.EXEC_TILED), (13, L.EXEC_PER_GATE)])
def test_angle_tables_rerun_one_plan(ctx, n, mode, prec):
    gates, slots, ns = param_ansatz(n, 2)
    pc = qiron.ParametricCircuit(n, gates, slots, prec, mode=mode)
    assert pc.num_slots == ns
    rng = W.SplitMix64(300 + n)
    tol = 1e-12 if prec == "c128" else 2e-6
    # all-zero angles first: no parametric op may have been folded away or dropped
    tables = [np.zeros(ns)] + [2 * math.pi * rng.uniforms(ns) for _ in range(3)]
    s = qiron.State.new_zero(n, prec)
    for a in tables:
        pc.set_angles(a)
        s.copy_from(qiron.State.new_zero(n, prec))
        pc.run(s)
        ref = O.apply_gates(O.init(n, 0), substituted(gates, slots, a))
        d = float(np.max(np.abs(s.state_vector - ref)))
        assert d <= tol, (n, mode, prec, d)
    # the same table on a second state (a new capture) and again on the first (a replay)
    t = pc.execute(qiron.State.new_zero(n, prec))
    assert float(np.max(np.abs(t.state_vector - ref))) <= tol
    s.copy_from(qiron.State.new_zero(n, prec))
    pc.run(s)
    assert float(np.max(np.abs(s.state_vector - ref))) <= tol


def test_parametric_program_from_qasm(ctx):
    n = 12
    text = "OPENQASM 3.0;\nqubit[12] q;\ninput float a;\ninput float b;\n" + \
           "".join(f"ry(a) q[{i}];\nrz(b) q[{i}];\n" for i in range(n)) + \
           "".join(f"cx q[{i}], q[{i + 1}];\n" for i in range(n - 1))
    pc = qiron.ParametricCircuit.from_qasm(text, "c128")
    assert pc.num_slots == 2
    for a, b in ((0.3, -1.2), (2.5, 0.7)):
        pc.set_angles([a, b])
        s = pc.execute(qiron.State.new_zero(n, "c128"))
        ref = O.apply_gates(O.init(n, 0), qiron.core.parse_qasm(text, [a, b])[1])
        assert float(np.max(np.abs(s.state_vector - ref))) <= 1e-12


def test_parametric_errors(ctx):
    gates, slots, ns = param_ansatz(6, 1)
    pc = qiron.ParametricCircuit(6, gates, slots, "c128")
    with pytest.raises(qiron.QuantIronError) as ei:
        pc.set_angles(np.zeros(ns + 1))
    assert ei.value.variant == "MismatchedNumberOfParameters" and (ei.value.a, ei.value.b) == (ns, ns + 1)
    with pytest.raises(qiron.QuantIronError) as ei:
        pc.run(qiron.State.new_zero(7, "c128"))
    assert ei.value.variant == "InvalidNumberOfQubits" and ei.value.a == 7
    bad = slots.copy()
    bad[list(slots).index(-1)] = 0  # the H gate cannot take an angle slot
    with pytest.raises(qiron.QuantIronError) as ei:
        qiron.ParametricCircuit(6, gates, bad, "c128")
    assert ei.value.variant == "InvalidArgument"
