"""Seeded synthetic workloads: the circuits and Hamiltonians the benchmarks and
tests run, built as qi_gate / qi_pauli_term arrays.

Random numbers come from SplitMix64 (Steele, Lea & Flood, 2014): draw k of a
stream seeded with s is mix(s + k * 0x9E3779B97F4A7C15), u = (x >> 11) * 2^-53.
The draw order of every builder is fixed and documented at the builder.
Hamiltonian conventions follow src/models/heisenberg.rs (-1/2 J per bond, term
order XX, YY, ZZ per bond) with the wrap-around bonds left out (open chains).
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

from . import _lib as L

GAMMA = np.uint64(0x9E3779B97F4A7C15)
SEED_BASE = 0x5EED0000


class SplitMix64:
    """Counter-based SplitMix64 stream of uniforms in [0, 1)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.k = 0

    def uniforms(self, count: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            ks = np.arange(self.k + 1, self.k + 1 + count, dtype=np.uint64)
            z = self.state + ks * GAMMA
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        self.k += count
        return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def u(self) -> float:
        return float(self.uniforms(1)[0])


def splitmix_at(seed: int, draws: np.ndarray) -> np.ndarray:
    """Uniforms of the given draw numbers (1-based) of the stream `seed`."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + np.asarray(draws, dtype=np.uint64) * GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


K = L.GATE_KINDS


class GateList:
    """Append-only qi_gate array builder."""

    def __init__(self, capacity: int = 1024):
        self.arr = np.zeros(capacity, dtype=L.GATE_DTYPE)
        self.n = 0

    def add(self, kind: str, targets=(), controls=(), params=()):
        if self.n == self.arr.size:
            self.arr = np.concatenate([self.arr, np.zeros(self.arr.size, dtype=L.GATE_DTYPE)])
        g = self.arr[self.n]
        g["kind"] = K[kind]
        g["ntargets"] = len(targets)
        for i, t in enumerate(targets):
            g["targets"][i] = t
        g["ncontrols"] = len(controls)
        for i, c in enumerate(controls):
            g["controls"][i] = c
        for i, p in enumerate(params):
            g["params"][i] = p
        self.n += 1

    def array(self) -> np.ndarray:
        return self.arr[: self.n].copy()


def qft_gates(gl: GateList, qubits: List[int]):
    """Subroutine::qft (subroutine.rs:90-112): H, CP(pi/2^k) target q_i control q_{i+k}, swaps."""
    n = len(qubits)
    for i in range(n):
        gl.add("H", [qubits[i]])
        den = 2.0
        for k in range(1, n - i):
            gl.add("P", [qubits[i]], [qubits[i + k]], [math.pi / den])
            den *= 2.0
    for i in range(n // 2):
        gl.add("SWAP", [qubits[i], qubits[n - 1 - i]])


def heisenberg_open_1d(n: int, jx: float, jy: float, jz: float, h=None) -> np.ndarray:
    """-1/2 (Jx XX + Jy YY + Jz ZZ) on bonds (i, i+1), i < n-1, then -1/2 h_i Z_i
    per site when fields are given; per-site order as heisenberg.rs:68-95."""
    terms = []
    for i in range(n):
        if i + 1 < n:
            b = (1 << i) | (1 << (i + 1))
            terms.append((b, 0, -0.5 * jx, 0.0))
            terms.append((b, b, -0.5 * jy, 0.0))
            terms.append((0, b, -0.5 * jz, 0.0))
        if h is not None:
            terms.append((0, 1 << i, -0.5 * float(h[i]), 0.0))
    arr = np.zeros(len(terms), dtype=L.TERM_DTYPE)
    for k, t in enumerate(terms):
        arr[k] = t
    return arr


def heisenberg_open_2d(rows: int, cols: int, jx: float, jy: float, jz: float, h) -> np.ndarray:
    """2-D anisotropic Heisenberg, site = r*cols + c, term order of heisenberg.rs:160-213
    (field, vertical XX/YY/ZZ, horizontal XX/YY/ZZ), open boundaries."""
    terms = []
    for s in range(rows * cols):
        r, c = divmod(s, cols)
        terms.append((0, 1 << s, -0.5 * float(h[s]), 0.0))
        for nb in ([(r + 1) * cols + c] if r + 1 < rows else []) + ([r * cols + c + 1] if c + 1 < cols else []):
            b = (1 << s) | (1 << nb)
            terms.append((b, 0, -0.5 * jx, 0.0))
            terms.append((b, b, -0.5 * jy, 0.0))
            terms.append((0, b, -0.5 * jz, 0.0))
    arr = np.zeros(len(terms), dtype=L.TERM_DTYPE)
    for k, t in enumerate(terms):
        arr[k] = t
    return arr


# ---------------------------------------------------------------- config 1 --
def config1(n: int = 20, layers: int = 5, seed: int = SEED_BASE + 1) -> Tuple[np.ndarray, np.ndarray]:
    """QFT on all qubits of |+...+>, then `layers` x [RX, RY, RZ on every qubit,
    CNOT i -> i+1]; open 1-D Heisenberg with fields. Draws: angles in order
    (layer, gate kind, qubit), then Jx, Jy, Jz, then h_0..h_{n-1}."""
    rng = SplitMix64(seed)
    gl = GateList(4096)
    qft_gates(gl, list(range(n)))
    for _ in range(layers):
        for kind in ("RX", "RY", "RZ"):
            angles = 2.0 * math.pi * rng.uniforms(n)
            for q in range(n):
                gl.add(kind, [q], [], [angles[q]])
        for i in range(n - 1):
            gl.add("CNOT", [i + 1], [i])
    j = 0.5 + rng.uniforms(3)
    h = 2.0 * rng.uniforms(n) - 1.0
    return gl.array(), heisenberg_open_1d(n, j[0], j[1], j[2], h)


# ---------------------------------------------------------------- config 2 --
ONE_Q = ("H", "X", "Y", "Z", "RX", "RY", "RZ", "P")


def random_circuit(n: int = 30, depth: int = 40, seed: int = SEED_BASE + 2, toffolis: int = 2) -> np.ndarray:
    """Per layer: for q = 0..n-1 a kind floor(8u) of {H,X,Y,Z,RX,RY,RZ,P} (an
    angle 2*pi*u drawn right after a rotation kind); a Fisher-Yates shuffle of
    the qubits (u for i = n-1..1) pairs perm[2k], perm[2k+1] for k < n // 2,
    each a kind floor(3u) of {CNOT a->b, CZ, SWAP}; then `toffolis` Toffolis,
    each on three distinct qubits from a 3-step partial shuffle."""
    rng = SplitMix64(seed)
    gl = GateList(depth * (n + n // 2 + toffolis) + 8)
    for _ in range(depth):
        for q in range(n):
            kind = ONE_Q[min(7, int(8 * rng.u()))]
            if kind in ("RX", "RY", "RZ", "P"):
                gl.add(kind, [q], [], [2.0 * math.pi * rng.u()])
            else:
                gl.add(kind, [q])
        perm = list(range(n))
        for i in range(n - 1, 0, -1):
            j = min(i, int(rng.u() * (i + 1)))
            perm[i], perm[j] = perm[j], perm[i]
        for k in range(n // 2):
            a, b = perm[2 * k], perm[2 * k + 1]
            kind = min(2, int(3 * rng.u()))
            if kind == 0:
                gl.add("CNOT", [b], [a])
            elif kind == 1:
                gl.add("Z", [b], [a])
            else:
                gl.add("SWAP", [a, b])
        for _t in range(toffolis):
            p = list(range(n))
            for i in range(3):
                j = i + min(n - 1 - i, int(rng.u() * (n - i)))
                p[i], p[j] = p[j], p[i]
            gl.add("TOFFOLI", [p[2]], [p[0], p[1]])
    return gl.array()


def inverse_gates(gates: np.ndarray) -> np.ndarray:
    """The inverse of a circuit of H, X, Y, Z, CNOT, CZ, SWAP, Toffoli (each its
    own inverse) and RX, RY, RZ, P (inverted by negating the angle): the list
    reversed with rotation and phase angles negated."""
    bad = np.isin(gates["kind"], [L.GATE_KINDS[k] for k in ("S", "T", "SDAG", "TDAG", "U2", "MATCHGATE")])
    if bad.any():
        raise ValueError("inverse_gates covers self-inverse gates and RX/RY/RZ/P only")
    inv = gates[::-1].copy()
    rot = np.isin(inv["kind"], [L.GATE_KINDS[k] for k in ("RX", "RY", "RZ", "P")])
    inv["params"][rot, 0] = -inv["params"][rot, 0]
    return inv


def minimal_touch_bytes(gates: np.ndarray, n: int, amp_bytes: int) -> float:
    """Bytes a gate must move at the least: dense or permutation 1-qubit gates and
    RZ read and write the state (2S); Z/S/T/P touch the target-1 half (S);
    controls halve the touched amplitudes per control; SWAP touches half (S)."""
    S = float(amp_bytes) * (1 << n)
    total = 0.0
    for g in gates:
        k = int(g["kind"])
        nc = int(g["ncontrols"])
        if k in (K["Z"], K["S"], K["T"], K["SDAG"], K["TDAG"], K["P"]):
            b = S
        elif k == K["I"]:
            b = 0.0
        elif k == K["SWAP"]:
            b = S
        elif k in (K["CNOT"], K["TOFFOLI"]):
            b = 2 * S
        else:
            b = 2 * S
        total += b / (2 ** nc)
    return total


# ---------------------------------------------------------------- config 3 --
def ansatz_circuit(n: int, layers: int, thetas: np.ndarray) -> np.ndarray:
    """Hardware-efficient ansatz: per layer RY(theta) on every qubit, RZ(theta)
    on every qubit, then CNOT i -> (i+1) mod n for i = 0..n-1."""
    gl = GateList(layers * 3 * n + 4)
    t = 0
    for _ in range(layers):
        for q in range(n):
            gl.add("RY", [q], [], [thetas[t]]); t += 1
        for q in range(n):
            gl.add("RZ", [q], [], [thetas[t]]); t += 1
        for i in range(n):
            gl.add("CNOT", [(i + 1) % n], [i])
    return gl.array()


def batched_ansatz(count: int = 4096, seed: int = SEED_BASE + 3, nmin: int = 10, nmax: int = 22,
                   lmin: int = 4, lmax: int = 8):
    """Per circuit: n = nmin + floor((nmax-nmin+1) u), L = lmin + floor((lmax-lmin+1) u),
    2 n L angles 2*pi*u (layer by layer, RY block then RZ block), Jx, Jy, Jz = 0.5 + u;
    open 1-D Heisenberg without field. Returns a list of (n, L, thetas, J)."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        n = nmin + min(nmax - nmin, int((nmax - nmin + 1) * rng.u()))
        layers = lmin + min(lmax - lmin, int((lmax - lmin + 1) * rng.u()))
        thetas = 2.0 * math.pi * rng.uniforms(2 * n * layers)
        j = 0.5 + rng.uniforms(3)
        out.append((n, layers, thetas, j))
    return out


# ---------------------------------------------------------------- config 4 --
def trotter_2d(rows: int = 5, cols: int = 5, seed: int = SEED_BASE + 4):
    """5x5 open lattice, Jx=1.0, Jy=0.8, Jz=1.2, h_i = u - 0.5; initial state
    |0..0> with X on sites where (r + c) is even. Returns (terms, x_sites)."""
    rng = SplitMix64(seed)
    h = rng.uniforms(rows * cols) - 0.5
    terms = heisenberg_open_2d(rows, cols, 1.0, 0.8, 1.2, h)
    xs = [r * cols + c for r in range(rows) for c in range(cols) if (r + c) % 2 == 0]
    return terms, xs
