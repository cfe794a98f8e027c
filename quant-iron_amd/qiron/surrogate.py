"""Learned screening of ansatz parameters.

An MLP surrogate of the energy E(theta, J) of a hardware-efficient ansatz
(RY/RZ layers with a CNOT ring) under an open 1-D Heisenberg SumOp is trained
on energies produced by the batched simulator itself, using the engine's own
HIP training kernels (qi_mlp_*). It ranks candidate parameter sets; the top
fraction is re-evaluated exactly by the batched simulator, which stays the
judge of every accepted value.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib as L
from . import workloads as W
from .core import CircuitBatch, Context, QuantIronError, _check, default_context

# the qi_mlp_* entry points
L.SIGNATURES.update({
    "qi_mlp_create": [L.P, C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(L.P), L.ERR],
    "qi_mlp_destroy": [L.P],
    "qi_mlp_num_params": [L.P, L.U64P, L.ERR],
    "qi_mlp_get_params": [L.P, C.POINTER(C.c_float), L.ERR],
    "qi_mlp_set_params": [L.P, C.POINTER(C.c_float), L.ERR],
    "qi_mlp_gradients": [L.P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_size_t, C.POINTER(C.c_float),
                         C.POINTER(C.c_float), L.ERR],
    "qi_mlp_set_data": [L.P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_size_t, L.ERR],
    "qi_mlp_train_epoch": [L.P, C.c_uint32, C.c_float, C.c_uint64, C.POINTER(C.c_float), L.ERR],
    "qi_mlp_predict": [L.P, C.POINTER(C.c_float), C.c_size_t, C.POINTER(C.c_float), L.ERR],
    "qi_mlp_predict_ansatz": [L.P, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, C.POINTER(C.c_float), L.ERR],
    "qi_mlp_predict_ansatz_local": [L.P, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, L.DP, C.c_double, L.DP,
                                    C.POINTER(C.c_float), L.ERR],
})


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MLP:
    """n_in -> H -> H -> H -> 1 with GELU; BF16 weights, FP32 accumulation."""

    def __init__(self, n_in: int, hidden: int = 512, seed: int = 1, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.lib = L.load()
        for name in ("qi_mlp_create", "qi_mlp_destroy", "qi_mlp_num_params", "qi_mlp_get_params", "qi_mlp_set_params",
                     "qi_mlp_gradients", "qi_mlp_set_data", "qi_mlp_train_epoch", "qi_mlp_predict",
                     "qi_mlp_predict_ansatz", "qi_mlp_predict_ansatz_local"):
            f = getattr(self.lib, name)
            f.argtypes = L.SIGNATURES[name]
            f.restype = C.c_int
        self.n_in, self.hidden = n_in, hidden
        self.k_in = (n_in + 15) // 16 * 16
        self._h = C.c_void_p()
        err = L.QiError()
        _check(self.lib.qi_mlp_create(self.ctx.handle, n_in, hidden, seed, C.byref(self._h), C.byref(err)), err)
        n = C.c_uint64()
        _check(self.lib.qi_mlp_num_params(self._h, C.byref(n), C.byref(err)), err)
        self.nparam = n.value

    def __del__(self):
        try:
            if self._h:
                self.lib.qi_mlp_destroy(self._h)
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.empty(self.nparam, dtype=np.float32)
        err = L.QiError()
        _check(self.lib.qi_mlp_get_params(self._h, _fp(out), C.byref(err)), err)
        return out

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        err = L.QiError()
        _check(self.lib.qi_mlp_set_params(self._h, _fp(p), C.byref(err)), err)

    def gradients(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(self.nparam, dtype=np.float32)
        loss = C.c_float()
        err = L.QiError()
        _check(self.lib.qi_mlp_gradients(self._h, _fp(x), _fp(t), x.shape[0], _fp(g), C.byref(loss), C.byref(err)), err)
        return float(loss.value), g

    def set_data(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        err = L.QiError()
        _check(self.lib.qi_mlp_set_data(self._h, _fp(x), _fp(t), x.shape[0], C.byref(err)), err)

    def train_epoch(self, batch: int, lr: float, seed: int) -> float:
        loss = C.c_float()
        err = L.QiError()
        _check(self.lib.qi_mlp_train_epoch(self._h, batch, lr, seed, C.byref(loss), C.byref(err)), err)
        return float(loss.value)

    def predict(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty(x.shape[0], dtype=np.float32)
        err = L.QiError()
        _check(self.lib.qi_mlp_predict(self._h, _fp(x), x.shape[0], _fp(y), C.byref(err)), err)
        return y

    def predict_stream(self, seed: int, draw0: int, count: int, nang: int) -> np.ndarray:
        """Predictions for `count` candidates drawn on the device from the
        SplitMix64 stream (seed, draw0) as sample_params would draw them."""
        y = np.empty(count, dtype=np.float32)
        err = L.QiError()
        _check(self.lib.qi_mlp_predict_ansatz(self._h, seed, draw0, count, nang, _fp(y), C.byref(err)), err)
        return y

    def last_ms(self) -> float:
        return self.ctx.last_exec_ms()


# ----------------------------------------------------------- ansatz energies --
class AnsatzEnergies:
    """Exact energies of one ansatz shape (n qubits, L layers) for many
    (theta, J) at once: a CircuitBatch of `slots` circuits sharing one planned
    program; each chunk rewrites the angle table and the term coefficients
    and replays the captured graph."""

    def __init__(self, n: int, layers: int, slots: int = 4096, precision: str = "c64"):
        self.n, self.layers, self.slots, self.precision = n, layers, slots, precision
        self.nang = 2 * n * layers
        proto = W.ansatz_circuit(n, layers, np.zeros(self.nang))
        rot = np.where(np.isin(proto["kind"], [L.GATE_KINDS["RY"], L.GATE_KINDS["RZ"]]))[0]
        assert rot.size == self.nang
        terms = W.heisenberg_open_1d(n, 1.0, 1.0, 1.0)
        self.nterms = terms.size
        self.batch = CircuitBatch(precision)
        for c in range(slots):
            sl = np.full(proto.size, -1, dtype=np.int32)
            sl[rot] = c * self.nang + np.arange(self.nang)
            self.batch.add(n, proto, terms, sl)
        self.device_seconds = 0.0

    def coeffs(self, J: np.ndarray) -> np.ndarray:
        """Term coefficients of the open chain: per bond -Jx/2, -Jy/2, -Jz/2."""
        nb = self.n - 1
        return np.tile((-0.5 * J)[:, None, :], (1, nb, 1)).reshape(-1).astype(np.complex128)

    def energies(self, thetas: np.ndarray, J: np.ndarray) -> np.ndarray:
        """thetas [count x 2nL] (layer by layer, RY block then RZ block), J [count x 3]."""
        count = thetas.shape[0]
        out = np.empty(count)
        for c0 in range(0, count, self.slots):
            th = thetas[c0:c0 + self.slots]
            jj = J[c0:c0 + self.slots]
            k = th.shape[0]
            if k < self.slots:  # pad the last chunk with copies of its last row
                th = np.concatenate([th, np.repeat(th[-1:], self.slots - k, axis=0)])
                jj = np.concatenate([jj, np.repeat(jj[-1:], self.slots - k, axis=0)])
            self.batch.set_angles(th.reshape(-1))
            self.batch.set_term_coeffs(self.coeffs(jj))
            e = self.batch.run(use_graph=True)
            self.device_seconds += self.batch.info()["last_ms"] * 1e-3
            out[c0:c0 + k] = e.real[:k]
        return out


def features(thetas: np.ndarray, J: np.ndarray) -> np.ndarray:
    """Network inputs: each angle as (cos theta, sin theta), so the periodic
    parameters stay continuous, then J - 1."""
    return np.concatenate([np.cos(thetas), np.sin(thetas), J - 1.0], axis=1).astype(np.float32)


def n_features(nang: int) -> int:
    return 2 * nang + 3


def sample_params(rng: W.SplitMix64, count: int, nang: int):
    th = 2.0 * math.pi * rng.uniforms(count * nang).reshape(count, nang)
    J = 0.5 + rng.uniforms(count * 3).reshape(count, 3)
    return th, J


def params_at(seed: int, draw0: int, count: int, nang: int, idx: np.ndarray):
    """The (theta, J) that sample_params draws for candidates `idx` of a block
    of `count` starting after draw `draw0` of the stream `seed`."""
    idx = np.asarray(idx, dtype=np.uint64)
    j = np.arange(nang, dtype=np.uint64)
    th = 2.0 * math.pi * W.splitmix_at(seed, np.uint64(draw0) + idx[:, None] * np.uint64(nang) + j + np.uint64(1))
    q = np.arange(3, dtype=np.uint64)
    J = 0.5 + W.splitmix_at(seed, np.uint64(draw0 + count * nang) + idx[:, None] * np.uint64(3) + q + np.uint64(1))
    return th.reshape(len(idx), nang), J.reshape(len(idx), 3)


def ansatz_batch(specs, precision: str = "c64", ctx: Optional[Context] = None) -> CircuitBatch:
    """One CircuitBatch over (n, L, thetas, J) specs: every (n, L) shape is one
    shared program whose rotations take their angles from the batch's table."""
    b = CircuitBatch(precision, ctx)
    protos = {}
    angles = []
    slot = 0
    for n, layers, th, J in specs:
        if (n, layers) not in protos:
            proto = W.ansatz_circuit(n, layers, np.zeros(2 * n * layers))
            rot = np.where(np.isin(proto["kind"], [L.GATE_KINDS["RY"], L.GATE_KINDS["RZ"]]))[0]
            protos[(n, layers)] = (proto, rot)
        proto, rot = protos[(n, layers)]
        sl = np.full(proto.size, -1, dtype=np.int32)
        sl[rot] = slot + np.arange(rot.size)
        b.add(n, proto, W.heisenberg_open_1d(n, *J), sl)
        angles.append(np.asarray(th, dtype=np.float64))
        slot += rot.size
    b.set_angles(np.concatenate(angles))
    return b


@dataclass
class ScreenResult:
    candidates: int
    kept: int
    surrogate_seconds: float
    exact_seconds: float
    best_exact: float
    mean_exact_kept: float
    mean_exact_random: float
    surrogate_mae_kept: float
    max_abs_err_kept: float


def screen_stream(mlp: MLP, sim: AnsatzEnergies, seed: int, draw0: int, count: int, keep_frac: float, t_mean: float,
                  t_std: float, rng_np: np.random.Generator, random_baseline: bool = True) -> ScreenResult:
    """Screen `count` candidates generated on the device: rank by the
    surrogate, regenerate the top fraction's parameters on the host from their
    indices and evaluate them exactly; optionally evaluate a random subset of
    the same size (selection without the surrogate)."""
    nang = sim.nang
    t0 = time.perf_counter()
    pred = mlp.predict_stream(seed, draw0, count, nang).astype(np.float64) * t_std + t_mean
    k = max(1, int(round(keep_frac * count)))
    top = np.argpartition(pred, k - 1)[:k]
    t1 = time.perf_counter()
    th, J = params_at(seed, draw0, count, nang, top)
    exact = sim.energies(th, J)
    t2 = time.perf_counter()
    mean_rnd = float("nan")
    if random_baseline:
        rnd = rng_np.choice(count, size=k, replace=False)
        thr, Jr = params_at(seed, draw0, count, nang, rnd)
        mean_rnd = float(sim.energies(thr, Jr).mean())
    err = np.abs(pred[top] - exact)
    return ScreenResult(count, k, t1 - t0, t2 - t1, float(exact.min()), float(exact.mean()), mean_rnd,
                        float(err.mean()), float(err.max()))


def screen(mlp: MLP, sim: AnsatzEnergies, thetas: np.ndarray, J: np.ndarray, keep_frac: float, t_mean: float,
           t_std: float, rng_np: np.random.Generator) -> ScreenResult:
    """Rank all candidates with the surrogate, re-evaluate the top fraction
    exactly; also evaluate a random subset of the same size (the method
    without the surrogate) for comparison."""
    t0 = time.perf_counter()
    pred = mlp.predict(features(thetas, J)).astype(np.float64) * t_std + t_mean
    t1 = time.perf_counter()
    k = max(1, int(round(keep_frac * thetas.shape[0])))
    top = np.argsort(pred)[:k]
    exact = sim.energies(thetas[top], J[top])
    t2 = time.perf_counter()
    rnd = rng_np.choice(thetas.shape[0], size=k, replace=False)
    exact_rnd = sim.energies(thetas[rnd], J[rnd])
    err = np.abs(pred[top] - exact)
    return ScreenResult(thetas.shape[0], k, t1 - t0, t2 - t1, float(exact.min()), float(exact.mean()),
                        float(exact_rnd.mean()), float(err.mean()), float(err.max()))
