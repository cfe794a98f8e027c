"""Shared helpers: JSON gate records -> qi_gate, random inputs, comparisons."""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

from qiron import _lib as L

GOLDEN = Path(__file__).resolve().parent / "golden"
F32_EPS = float(np.finfo(np.float32).eps)


def load_kats():
    return json.loads((GOLDEN / "kats.json").read_text())


def to_gate(d) -> L.QiGate:
    g = L.QiGate()
    g.kind = L.GATE_KINDS[d["kind"]]
    g.ntargets = len(d.get("targets", []))
    for i, t in enumerate(d.get("targets", [])[: L.MAX_TARGETS]):
        g.targets[i] = t
    g.ncontrols = len(d.get("controls", []))
    for i, c in enumerate(d.get("controls", [])):
        g.controls[i] = c
    for i, p in enumerate(d.get("params", [])):
        g.params[i] = p
    if "matrix" in d:
        for i, (re, im) in enumerate(d["matrix"]):
            g.matrix[2 * i] = re
            g.matrix[2 * i + 1] = im
    g.x_mask = d.get("x_mask", 0)
    g.z_mask = d.get("z_mask", 0)
    g.basis = d.get("basis", 0)
    return g


def gates_array(ds) -> np.ndarray:
    arr = np.zeros(len(ds), dtype=L.GATE_DTYPE)
    raw = arr.view(np.uint8).reshape(len(ds), L.GATE_DTYPE.itemsize)
    for i, d in enumerate(ds):
        raw[i, :] = np.frombuffer(bytes(to_gate(d)), dtype=np.uint8)
    return arr


def term_array(ts) -> np.ndarray:
    arr = np.zeros(len(ts), dtype=L.TERM_DTYPE)
    for i, t in enumerate(ts):
        arr[i] = tuple(t)
    return arr


def cvec(pairs) -> np.ndarray:
    a = np.asarray(pairs, dtype=np.float64)
    return (a[:, 0] + 1j * a[:, 1]).astype(np.complex128)


def random_state(n: int, rng: np.random.Generator) -> np.ndarray:
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return (v / np.linalg.norm(v)).astype(np.complex128)


def unitary_ok(u) -> bool:
    """Unitary2::new's acceptance test (operator.rs:2092-2118, tolerance 2 eps)."""
    tol = 2 * np.finfo(np.float64).eps
    a, b, c, d = (complex(x) for x in np.asarray(u).reshape(-1))
    n0 = (a.real * a.real + a.imag * a.imag) + (b.real * b.real + b.imag * b.imag)
    n1 = (c.real * c.real + c.imag * c.imag) + (d.real * d.real + d.imag * d.imag)
    x = a * c.conjugate() + b * d.conjugate()
    return abs(n0 - 1) <= tol and abs(n1 - 1) <= tol and (x.real * x.real + x.imag * x.imag) <= tol * tol


def random_unitary2(rng):
    """A random 2x2 unitary that passes Unitary2::new."""
    while True:
        z = rng.normal(size=(2, 2)) + 1j * rng.normal(size=(2, 2))
        q, r = np.linalg.qr(z)
        q = q * (np.diag(r) / np.abs(np.diag(r)))
        if unitary_ok(q):
            return q


ONE_Q = ["H", "X", "Y", "Z", "I", "S", "T", "SDAG", "TDAG", "P", "RX", "RY", "RZ", "U2"]


def random_gate(n: int, rng, kinds=None, max_controls: int = 2) -> dict:
    """A valid random gate of the operator kinds on n qubits."""
    kinds = kinds or ONE_Q + ["CNOT", "SWAP", "TOFFOLI", "MATCHGATE"]
    while True:
        k = kinds[rng.integers(len(kinds))]
        need = {"CNOT": 2, "SWAP": 2, "TOFFOLI": 3, "MATCHGATE": 2}.get(k, 1)
        if n >= need:
            break
    perm = list(rng.permutation(n))
    d = {"kind": k, "params": [float(rng.uniform(0, 2 * np.pi)) for _ in range(3)]}
    if k == "CNOT":
        d["targets"], d["controls"] = [int(perm[0])], [int(perm[1])]
    elif k == "TOFFOLI":
        d["targets"], d["controls"] = [int(perm[0])], [int(perm[1]), int(perm[2])]
    elif k == "SWAP":
        nc = int(rng.integers(0, min(max_controls, n - 2) + 1))
        d["targets"], d["controls"] = [int(perm[0]), int(perm[1])], [int(x) for x in perm[2:2 + nc]]
    elif k == "MATCHGATE":
        q1 = int(rng.integers(0, n - 1))
        rest = [q for q in perm if q != q1]
        nc = int(rng.integers(0, min(max_controls, n - 2) + 1))
        d["targets"], d["controls"] = [q1], [int(x) for x in rest[:nc]]
    else:
        nc = int(rng.integers(0, min(max_controls, n - 1) + 1))
        d["targets"], d["controls"] = [int(perm[0])], [int(x) for x in perm[1:1 + nc]]
        if k == "U2":
            u = random_unitary2(rng)
            d["matrix"] = [[z.real, z.imag] for z in u.reshape(-1)]
    return d


def random_terms(n: int, count: int, rng, max_weight: int = 4) -> np.ndarray:
    ts = []
    for _ in range(count):
        w = int(rng.integers(0, min(max_weight, n) + 1))
        qs = rng.choice(n, size=w, replace=False)
        x = z = 0
        for q in qs:
            p = int(rng.integers(3))
            if p in (0, 1):
                x |= 1 << int(q)
            if p in (1, 2):
                z |= 1 << int(q)
        ts.append((x, z, float(rng.normal()), float(rng.normal()) if rng.random() < 0.5 else 0.0))
    return term_array(ts)
