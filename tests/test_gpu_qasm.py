"""OpenQASM 3.0 programs executed by the engine: the exporter's programs and
parametrised ansatz programs in the batched engine, against the oracle."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import qiron
from kat_util import random_state
from qiron import _lib as L
from qiron import workloads as W

pytestmark = pytest.mark.gpu

QASM = Path(__file__).resolve().parent / "golden" / "qasm"


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_exporter_program_runs_like_the_oracle(ctx, prec):
    """exporter_random.qasm read by the engine's reader, its measurement
    draws fixed, run by every executor against the oracle on a random state."""
    n, arr, _, _ = qiron.core.parse_qasm((QASM / "exporter_random.qasm").read_text())
    meas = arr["kind"] == L.GATE_KINDS["MEASURE"]
    arr["params"][meas, 0] = np.linspace(0.11, 0.89, int(meas.sum()))
    psi = random_state(n, np.random.default_rng(17))
    ref = O.apply_gates(psi, arr)
    for mode in (L.EXEC_PER_GATE, L.EXEC_TILED, L.EXEC_FUSED_MFMA):
        s = qiron.State.from_vector_unchecked(psi, prec)
        s.apply_gates_inplace(arr, mode=mode)
        d = float(np.max(np.abs(s.state_vector - ref)))
        assert d <= (1e-12 if prec == "c128" else 2e-6), (mode, d)


def _ansatz_qasm(n, layers):
    """The hardware-efficient ansatz of the batched benchmarks as a program with
    one `input` parameter per rotation (RY then RZ on each qubit, CNOT ring)."""
    lines = ["OPENQASM 3.0;", 'include "stdgates.inc";', f"qubit[{n}] q;"]
    lines += [f"input float[64] t{k};" for k in range(2 * n * layers)]
    k = 0
    for _ in range(layers):
        for kind in ("ry", "rz"):
            for q in range(n):
                lines.append(f"{kind}(t{k}) q[{q}];")
                k += 1
        for q in range(n):
            lines.append(f"cx q[{q}], q[{(q + 1) % n}];")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_batch_from_qasm_equals_the_gate_array_batch(ctx, prec):
    """Parametrised programs added with CircuitBatch.add_qasm read the batch's
    angle table through their slots: energies equal one-circuit oracle runs."""
    rng = W.SplitMix64(77)
    b = qiron.CircuitBatch(prec)
    specs, off = [], 0
    for n, layers in ((6, 2), (11, 3), (15, 2)):
        th = 2 * math.pi * rng.uniforms(2 * n * layers)
        terms = W.heisenberg_open_1d(n, 0.9, 1.1, 1.3)
        b.add_qasm(_ansatz_qasm(n, layers), terms, slot_offset=off)
        specs.append((n, layers, th, terms, off))
        off += th.size
    b.set_angles(np.concatenate([s[2] for s in specs]))
    e = b.run()
    for (n, layers, th, terms, _), got in zip(specs, e):
        ref = O.expectation(O.apply_gates(O.init(n, 0), W.ansatz_circuit(n, layers, th)), terms)
        assert abs(got - ref) <= (1e-12 if prec == "c128" else 2e-5), (n, got, ref)
