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

    def predict_local(self, seed: int, draw0: int, count: int, theta0: np.ndarray, sigma: float,
                      J: np.ndarray) -> np.ndarray:
        """Predictions for `count` candidates theta0 + sigma * z drawn on the
        device (z standard normal from the SplitMix64 stream (seed, draw0),
        see local_params), couplings J fixed."""
        th0 = np.ascontiguousarray(theta0, dtype=np.float64)
        jj = np.ascontiguousarray(J, dtype=np.float64)
        y = np.empty(count, dtype=np.float32)
        err = L.QiError()
        _check(self.lib.qi_mlp_predict_ansatz_local(
            self._h, seed, draw0, count, th0.size, th0.ctypes.data_as(L.DP), float(sigma), jj.ctypes.data_as(L.DP),
            _fp(y), C.byref(err)), err)
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


def local_params(seed: int, draw0: int, nang: int, theta0: np.ndarray, sigma: float, idx: np.ndarray) -> np.ndarray:
    """The angles qi_mlp_predict_ansatz_local draws for candidates `idx`:
    theta0[j] + sigma * z_j with one Box-Muller pair per two angles: for
    p = j // 2, u1 and u2 are the draws d + 1 and d + 2 of the stream `seed`,
    d = draw0 + 2 (c P + p), P = ceil(nang / 2), rho = sqrt(-2 ln(1 - u1)),
    z_2p = rho cos(2 pi u2), z_2p+1 = rho sin(2 pi u2)."""
    idx = np.asarray(idx, dtype=np.uint64)
    pairs = (nang + 1) // 2
    p = np.arange(pairs, dtype=np.uint64)
    d = np.uint64(draw0) + np.uint64(2) * (idx[:, None] * np.uint64(pairs) + p)
    u1 = W.splitmix_at(seed, d + np.uint64(1))
    u2 = W.splitmix_at(seed, d + np.uint64(2))
    rho = np.sqrt(-2.0 * np.log(1.0 - u1))
    z = np.empty((idx.size, 2 * pairs))
    z[:, 0::2] = rho * np.cos(2.0 * math.pi * u2)
    z[:, 1::2] = rho * np.sin(2.0 * math.pi * u2)
    return np.asarray(theta0, dtype=np.float64)[None, :] + sigma * z[:, :nang]


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
    sd_random: float = float("nan")
    heldout_r2: float = float("nan")  # on the random subset (not trained on)
    best_theta: Optional[np.ndarray] = None
    best_J: Optional[np.ndarray] = None
    surrogate_device_seconds: float = float("nan")  # the candidate screen's kernels alone


def screen_stream(mlp: MLP, sim: AnsatzEnergies, seed: int, draw0: int, count: int, keep_frac: float, t_mean: float,
                  t_std: float, rng_np: np.random.Generator, random_baseline: bool = True) -> ScreenResult:
    """Screen `count` candidates generated on the device: rank by the
    surrogate, regenerate the top fraction's parameters on the host from their
    indices and evaluate them exactly; optionally evaluate a random subset of
    the same size (selection without the surrogate)."""
    nang = sim.nang
    t0 = time.perf_counter()
    pred = mlp.predict_stream(seed, draw0, count, nang).astype(np.float64) * t_std + t_mean
    dev = mlp.last_ms() * 1e-3
    k = max(1, int(round(keep_frac * count)))
    top = np.argpartition(pred, k - 1)[:k]
    t1 = time.perf_counter()
    th, J = params_at(seed, draw0, count, nang, top)
    exact = sim.energies(th, J)
    t2 = time.perf_counter()
    mean_rnd = sd_rnd = r2 = float("nan")
    if random_baseline:
        rnd = rng_np.choice(count, size=k, replace=False)
        thr, Jr = params_at(seed, draw0, count, nang, rnd)
        er = sim.energies(thr, Jr)
        mean_rnd, sd_rnd = float(er.mean()), float(er.std())
        r2 = float(1.0 - np.mean((pred[rnd] - er) ** 2) / np.var(er))
    err = np.abs(pred[top] - exact)
    b = int(np.argmin(exact))
    return ScreenResult(count, k, t1 - t0, t2 - t1, float(exact.min()), float(exact.mean()), mean_rnd,
                        float(err.mean()), float(err.max()), sd_rnd, r2, th[b].copy(), J[b].copy(), dev)


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


# ------------------------------------------------------- optimisation rounds --
@dataclass
class RoundResult:
    round: int
    centre_energy: float       # exact energy at the round's centre
    train_samples: int
    heldout_r2: float          # on the random control set (unbiased, never trained on)
    mean_exact_kept: float     # top keep_frac by the surrogate, evaluated exactly
    mean_exact_random: float   # the same number of candidates chosen at random
    sd_random: float
    gain_sd: float             # (random mean - kept mean) / sd of the random set
    best_exact: float          # best exact energy among the kept candidates
    surrogate_mae_kept: float
    surrogate_max_err_kept: float
    baseline_best: float       # the chain without the network: best energy so far
    seconds: dict
    centre_after: float = float("nan")  # the network chain's best energy after the round


def optimise_local(mlp: MLP, sim: AnsatzEnergies, J: np.ndarray, theta: np.ndarray, rounds: int, sigma: float,
                   n_local: int, candidates: int, keep_frac: float, epochs: int, lr: float, seed: int,
                   batch: int = 512) -> list:
    """Surrogate-screened local search. Each round: exact energies of n_local
    points theta + sigma z around the centre (the training data, from the
    batched simulator), a few epochs of training from the current weights,
    the network ranks `candidates` local candidates drawn on the device, the
    top keep_frac are evaluated exactly. A random keep_frac of the same
    candidates is evaluated too: held-out R^2 and the selection gain are
    measured on it. The best of the round's exact energies (training points,
    kept and random candidates) becomes the next centre if it improves on it.
    A second chain without the network spends the same number of exact
    evaluations per round (n_local + 2 kept) on random local candidates and
    moves to its best; its best energy is reported beside the network's."""
    nang = sim.nang
    J = np.asarray(J, dtype=np.float64)
    rng = np.random.default_rng(seed)
    k = max(1, int(round(keep_frac * candidates)))
    centre = np.asarray(theta, dtype=np.float64).copy()
    e_centre = float(sim.energies(centre[None, :], J[None, :])[0])
    base_centre, base_best = centre.copy(), e_centre
    out = []
    for r in range(rounds):
        t = {}
        t0 = time.perf_counter()
        th = centre[None, :] + sigma * rng.standard_normal((n_local, nang))
        Jl = np.repeat(J[None, :], n_local, axis=0)
        E = sim.energies(th, Jl)
        t["train_data"] = time.perf_counter() - t0
        mu, sd = float(E.mean()), float(E.std())
        t0 = time.perf_counter()
        mlp.set_data(features(th, Jl), ((E - mu) / sd).astype(np.float32))
        for ep in range(epochs):
            mlp.train_epoch(batch, lr, seed=seed * 1000 + r * 100 + ep)
        t["train"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        cseed = seed * 7919 + r
        pred = mlp.predict_local(cseed, 0, candidates, centre, sigma, J).astype(np.float64) * sd + mu
        top = np.argpartition(pred, k - 1)[:k]
        t["screen"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        th_k = local_params(cseed, 0, nang, centre, sigma, top)
        E_k = sim.energies(th_k, np.repeat(J[None, :], k, axis=0))
        t["exact_kept"] = time.perf_counter() - t0
        rnd = rng.choice(candidates, size=k, replace=False)
        th_r = local_params(cseed, 0, nang, centre, sigma, rnd)
        E_r = sim.energies(th_r, np.repeat(J[None, :], k, axis=0))
        r2 = float(1.0 - np.mean((pred[rnd] - E_r) ** 2) / np.var(E_r))
        err = np.abs(pred[top] - E_k)
        best = int(np.argmin(E_k))
        # the chain without the network, same number of exact evaluations
        nb = n_local + 2 * k
        thb = base_centre[None, :] + sigma * rng.standard_normal((nb, nang))
        Eb = sim.energies(thb, np.repeat(J[None, :], nb, axis=0))
        if Eb.min() < base_best:
            base_best, base_centre = float(Eb.min()), thb[int(np.argmin(Eb))].copy()
        out.append(RoundResult(r, e_centre, n_local, r2, float(E_k.mean()), float(E_r.mean()), float(E_r.std()),
                               float((E_r.mean() - E_k.mean()) / max(E_r.std(), 1e-300)), float(E_k[best]),
                               float(err.mean()), float(err.max()), base_best, t))
        pool_e = np.concatenate([E, E_k, E_r])
        j = int(np.argmin(pool_e))
        if pool_e[j] < e_centre:
            centre = np.concatenate([th, th_k, th_r])[j].copy()
            e_centre = float(pool_e[j])
        out[-1].centre_after = e_centre
    return out
