"""Benchmark: 30-qubit random circuit (depth 40) on one MI355X.

One step = Circuit::execute of the 1 880-gate circuit (per layer: a random gate
from {H,X,Y,Z,RX,RY,RZ,P} on each of 30 qubits, 15 disjoint CNOT/CZ/SWAP, 2
Toffolis) on a device-resident 2^30-amplitude state. The headline value is
whole-job gates/s of the register-tiled executor with complex64 amplitudes;
complex128 and the other executors are reported beside it.

Prints ONE JSON line (rank 0). Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "quant-iron_amd"))

METRIC = "gates/s + HBM GB/s @30q; ansatz energies/s FP32+surrogate vs FP64"
HBM_PEAK_GBPS = 8000.0
VALU_CLOCK_HZ = 2.4e9  # MI355X peak engine clock
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def run_circuit(qiron, L, n, prec, gates, opts, steps, warmup):
    s = qiron.State.new_plus(n, prec)
    ctx = s.ctx
    for _ in range(warmup):
        s.apply_gates_inplace(gates, **opts)
    ctx.synchronize()
    t0 = time.perf_counter()
    dev_ms = 0.0
    for _ in range(steps):
        s.apply_gates_inplace(gates, **opts)
        dev_ms += ctx.last_exec_ms()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    nsteps = ctx.last_exec_steps()
    norm = s.norm_sqr()
    del s
    return wall, dev_ms, nsteps, norm


def host_cpu():
    """The host's CPU model and core counts (the baseline states what it ran on)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "nproc": os.cpu_count(), "usable": usable}


def cpu_baseline(qiron_workloads, n: int, depth: int, budget_s: float, threads: int):
    """The reference-structured CPU path (clone per gate, update list + serial
    scatter, `threads` workers) on the first gates of the benchmark circuit
    itself (n qubits, complex128 as the reference computes), timed directly:
    gates are run until about budget_s seconds of the C call have passed (at
    least one). Returns (gates done, seconds)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    gates = qiron_workloads.random_circuit(n=n, depth=depth, seed=qiron_workloads.SEED_BASE + 2)
    psi = O.init(n, 2)
    done, spent = 0, 0.0
    while done < gates.size and (done == 0 or spent + spent / done <= budget_s):
        _, dt = O.reference_structured_timed(psi, gates[done:done + 1], threads)
        spent += dt
        done += 1
    del psi
    return done, spent


def bench_config1(qiron, L, W, args):
    """20-qubit complex128: QFT on |+..+>, 5 rotation/CNOT-ladder layers, the
    energy of an open Heisenberg chain with fields; GPU against the oracle
    (amplitudes and energy) and the reference-structured CPU path."""
    import numpy as np

    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as O

    gates, terms = W.config1()
    n = 20
    s = qiron.State.new_plus(n, "c128")
    s.apply_gates_inplace(gates)  # warm-up (plans, loads code)
    s = qiron.State.new_plus(n, "c128")
    s.ctx.synchronize()
    t0 = time.perf_counter()
    s.apply_gates_inplace(gates)
    e = qiron.expectation_terms(s, terms)
    s.ctx.synchronize()
    t_gpu = time.perf_counter() - t0
    psi = s.state_vector
    t0 = time.perf_counter()
    ref = O.apply_gates(O.init(n, 2), gates)
    t_oracle = time.perf_counter() - t0
    e_ref = O.expectation(ref, terms)
    out = dict(gates=int(gates.size), gpu_seconds=t_gpu, energy=e.real, energy_oracle=e_ref.real,
               max_abs_amp_diff=float(np.max(np.abs(psi - ref))), abs_dE=abs(e.real - e_ref.real),
               oracle_seconds=t_oracle)
    if not args.no_cpu:
        threads = min(host_cpu()["usable"], 16)
        rs, dt = O.reference_structured_timed(O.init(n, 2), gates, threads, keep=True)
        out["reference_structured_seconds"] = dt
        out["reference_structured_threads"] = threads
        out["reference_structured_max_diff"] = float(np.max(np.abs(rs - ref)))
    log("config 1 (20q c128):", out)
    return out


def bench_trotter(qiron, W, args):
    """5x5 open 2-D Heisenberg, 200 first-order steps of dt = 0.02 from the Neel
    state, complex64 beside complex128: energy drift and fidelity every 20 steps."""
    import numpy as np

    terms, xs = W.trotter_2d()
    n = 25
    states = {}
    for prec in ("c64", "c128"):
        st = qiron.State.new_zero(n, prec)
        for q in xs:
            st.apply_gate_inplace(qiron.Gate.x_gate(q))
        states[prec] = st
    e0 = {p: qiron.expectation_terms(st, terms).real for p, st in states.items()}
    rows, t_c64 = [], 0.0
    for chunk in range(10):
        for prec, st in states.items():
            st.ctx.synchronize()
            t0 = time.perf_counter()
            qiron.trotter_evolve_terms_inplace(st, terms, 0.02, 20)
            st.ctx.synchronize()
            if prec == "c64":
                t_c64 += time.perf_counter() - t0
                step_passes = st.ctx.last_exec_steps()  # launches of one Trotter step
        a, b = states["c64"].state_vector, states["c128"].state_vector
        na, nb = float(np.vdot(a, a).real), float(np.vdot(b, b).real)
        fid = float(abs(np.vdot(b, a)) ** 2 / (na * nb))
        e = {p: qiron.expectation_terms(st, terms).real for p, st in states.items()}
        # complex64 rounding moves the norm; the energy of the normalised state isolates the state error
        rows.append(dict(step=20 * (chunk + 1), fidelity_c64_vs_c128=fid, norm_sq_c64=na,
                         drift_c64=e["c64"] - e0["c64"], drift_c128=e["c128"] - e0["c128"],
                         dE_c64_vs_c128=e["c64"] - e["c128"],
                         dE_normalised_c64_vs_c128=e["c64"] / na - e["c128"] / nb))
    apps = 200 * terms.size
    S = (1 << n) * 8
    # HBM rate of the passes: each pass reads and writes the state once
    out = dict(terms=int(terms.size), steps=200, exp_applications=apps, c64_seconds=t_c64,
               c64_exp_per_s=apps / t_c64, c64_passes_per_step=step_passes,
               c64_pass_GBps=2 * S * step_passes * 200 / t_c64 / 1e9, checkpoints=rows)
    log(f"trotter 25q: {apps / t_c64:.0f} exp(P)/s (c64), {step_passes} passes per step at "
        f"{out['c64_pass_GBps']:.0f} GB/s, final fidelity {rows[-1]['fidelity_c64_vs_c128']:.9f}, "
        f"drift c64 {rows[-1]['drift_c64']:.2e}")
    return out


def bench_measure(qiron, W):
    """measure_n on the 30-qubit state after the benchmark circuit's first
    layer: 2^20 shots of 20 measured qubits (one marginal histogram over 2^20
    bins in FP64, one cumulative scan, one sampling pass), complex64."""
    import numpy as np

    n, m, shots = 30, 20, 1 << 20
    s = qiron.State.new_plus(n, "c64")
    s.apply_gates_inplace(W.random_circuit(n=n, depth=1, seed=W.SEED_BASE + 2))
    u = W.SplitMix64(W.SEED_BASE + 9).uniforms(shots)
    qs = list(range(m))
    s.measure_outcomes(qs, u[:16])  # first launches, untimed
    s.ctx.synchronize()
    t0 = time.perf_counter()
    outs = s.measure_outcomes(qs, u)
    s.ctx.synchronize()
    dt = time.perf_counter() - t0
    del s
    out = dict(qubits=n, measured=m, shots=shots, seconds=dt, shots_per_s=shots / dt,
               distinct_outcomes=int(np.unique(outs).size))
    log("measure_n:", out)
    return out


def bench_ceiling(qiron, L, W):
    """Memory ceiling: one layer of the random circuit at 34 qubits complex64 and
    33 qubits complex128 (128 GiB states), tiled executor."""
    out = {}
    for n, prec, ab in ((34, "c64", 8), (33, "c128", 16)):
        gates = W.random_circuit(n=n, depth=1, seed=W.SEED_BASE + 2)
        try:
            s = qiron.State.new_plus(n, prec)
            s.apply_gates_inplace(gates, mode=L.EXEC_TILED)
            s.ctx.synchronize()
            t0 = time.perf_counter()
            s.apply_gates_inplace(gates, mode=L.EXEC_TILED)
            s.ctx.synchronize()
            dt = time.perf_counter() - t0
            passes = s.ctx.last_exec_steps()
            norm = s.norm_sqr()
            del s
            S = (1 << n) * ab
            out[f"{n}q_{prec}"] = dict(gates=int(gates.size), seconds=dt, gates_per_s=gates.size / dt, passes=passes,
                                       pass_GBps=2 * S * passes / dt / 1e9, norm=norm, state_GiB=S / 2 ** 30)
        except Exception as ex:  # report, do not hide
            out[f"{n}q_{prec}"] = dict(error=str(ex))
        log(f"ceiling {n}q {prec}:", out[f"{n}q_{prec}"])
    return out


def bench_energies(qiron, L, W, args):
    """Batched ansatz energies (4096 circuits, 10..22 qubits, 4..8 layers, each
    with its own open Heisenberg chain) in complex64 and complex128, against
    the plain FP64 GPU path: one circuit at a time, per-gate kernels in
    complex128, then the SumOp expectation."""
    import numpy as np

    from qiron import surrogate as S

    specs = W.batched_ansatz(args.batch_circuits)
    count = len(specs)
    res, E = {}, {}
    for prec in ("c64", "c128"):
        t0 = time.perf_counter()
        b = S.ansatz_batch(specs, prec)
        fixed = getattr(args, f"schedule_{prec}")
        if fixed:  # a given schedule instead of the autotuner (profiles of the timed runs alone)
            k, ks = (int(x) for x in fixed.split(","))
            b.set_schedule(k, ks)
            b.run(use_graph=True)
            tuned_ms = b.info()["last_ms"]
        else:
            k, ks, tuned_ms = b.autotune(trials=2)  # plans, captures and times every schedule
        setup = time.perf_counter() - t0
        walls = []
        for _ in range(args.steps):
            t1 = time.perf_counter()
            e = b.run(use_graph=True)
            walls.append(time.perf_counter() - t1)
        info = b.info()
        wall = min(walls)
        res[f"batch_{prec}"] = dict(energies_per_s=count / wall, wall_ms=wall * 1e3, device_ms=info["last_ms"],
                                    setup_s=setup, levels=info["levels"], lds_resident=info["lds_resident"],
                                    schedule={"executor": b.executor(), "tile_qubits": k, "lds_resident_max_qubits": ks,
                                              "tuned_ms": tuned_ms})
        log(f"batched {prec}: {count} energies in {wall * 1e3:.1f} ms ({count / wall:.0f}/s), device "
            f"{info['last_ms']:.1f} ms, schedule {b.executor()} k={k} ks={ks}, setup {setup:.1f} s")
        E[prec] = e.real
        del b
    dE = float(np.max(np.abs(E["c64"] - E["c128"])))
    res["max_abs_dE_fp32_vs_fp64"] = dE
    log(f"max |E_fp32 - E_fp64| over {count} circuits: {dE:.2e}")
    # the plain FP64 path on every k-th circuit
    stride = max(1, count // args.plain_circuits)
    t_plain, worst, n_plain = 0.0, 0.0, 0
    for idx in range(0, count, stride):
        n, layers, th, J = specs[idx]
        gates = W.ansatz_circuit(n, layers, th)
        terms = W.heisenberg_open_1d(n, *J)
        t0 = time.perf_counter()
        s = qiron.State.new_zero(n, "c128")
        s.apply_gates_inplace(gates, mode=L.EXEC_PER_GATE)
        e = qiron.expectation_terms(s, terms)
        t_plain += time.perf_counter() - t0
        del s
        worst = max(worst, abs(e.real - E["c128"][idx]))
        n_plain += 1
    plain = n_plain / t_plain
    res["plain_fp64"] = dict(energies_per_s=plain, circuits=n_plain, seconds=t_plain, max_abs_dE_vs_batched_fp64=worst)
    res["batched_fp32_vs_plain_fp64"] = res["batch_c64"]["energies_per_s"] / plain
    log(f"plain FP64 path: {n_plain} circuits (every {stride}th) in {t_plain:.2f} s = {plain:.1f}/s; "
        f"max |dE| vs batched FP64 {worst:.1e}; batched FP32 is {res['batched_fp32_vs_plain_fp64']:.0f}x")
    return res, plain


def bench_surrogate(W, args):
    """The learned screen on the 16-qubit, 6-layer ansatz (192 angles + 3
    couplings). Round 0: train on 1M batched-simulator energies at uniform
    angles and couplings, rank 1M candidates drawn on the device, re-evaluate
    the top 1% exactly. Then local optimisation rounds from the best point
    found: each round trains further on exact energies around the centre and
    screens 1M local candidates. Skill is measured on a random 1% of each
    round's candidates (held-out R^2, kept-vs-random gain in sd units), and a
    chain without the network that spends the same exact evaluations runs
    beside it."""
    import numpy as np

    from qiron import surrogate as S

    n, layers = 16, 6
    sim = S.AnsatzEnergies(n, layers, slots=4096, precision="c64")
    nang = sim.nang
    fwd_flop = 2.0 * (S.n_features(nang) * 512 + 2 * 512 * 512 + 512)  # per sample: 3 hidden layers and the head
    rng = W.SplitMix64(W.SEED_BASE + 5)
    t0 = time.perf_counter()
    th, J = S.sample_params(rng, args.train_samples, nang)
    E = sim.energies(th, J)
    t_gen = time.perf_counter() - t0
    mu, sd = float(E.mean()), float(E.std())
    net = S.MLP(S.n_features(nang), 512, seed=7)
    t0 = time.perf_counter()
    net.set_data(S.features(th, J), ((E - mu) / sd).astype(np.float32))
    del th
    t_up = time.perf_counter() - t0
    losses, t_train, dev_train = [], 0.0, 0.0
    for ep in range(args.epochs):
        t1 = time.perf_counter()
        losses.append(net.train_epoch(1024, 1e-3, seed=ep))
        t_train += time.perf_counter() - t1
        dev_train += net.last_ms() * 1e-3
    seed, draw0, cand = W.SEED_BASE + 6, 0, args.candidates
    net.predict_stream(seed + 1, 0, 4096, nang)  # first launches of the screening kernels, untimed
    res = S.screen_stream(net, sim, seed, draw0, cand, 0.01, mu, sd, np.random.default_rng(3))
    dev_screen = res.surrogate_device_seconds
    train_tflops = 3.0 * fwd_flop * args.epochs * args.train_samples / dev_train * 1e-12
    screen_tflops = fwd_flop * cand / dev_screen * 1e-12
    round0 = dict(train_samples=args.train_samples, gen_seconds=t_gen, gen_energies_per_s=args.train_samples / t_gen,
                  upload_seconds=t_up, epochs=args.epochs, train_seconds=t_train, train_device_seconds=dev_train,
                  train_TFLOPs=train_tflops, loss=[round(x, 4) for x in losses], candidates=cand, kept=res.kept,
                  screen_device_seconds=dev_screen, screen_TFLOPs=screen_tflops,
                  screened_candidates_per_s=cand / (res.surrogate_seconds + res.exact_seconds),
                  heldout_r2=res.heldout_r2, mean_exact_kept=res.mean_exact_kept,
                  mean_exact_random=res.mean_exact_random, sd_random=res.sd_random,
                  gain_sd=(res.mean_exact_random - res.mean_exact_kept) / res.sd_random,
                  best_exact=res.best_exact, surrogate_mae_kept=res.surrogate_mae_kept,
                  surrogate_max_err_kept=res.max_abs_err_kept)
    log(f"surrogate round 0 (uniform): R^2 {res.heldout_r2:.3f}, kept {res.mean_exact_kept:.4f} vs random "
        f"{res.mean_exact_random:.4f} (sd {res.sd_random:.4f}); training {train_tflops:.0f} TFLOP/s, "
        f"screening {screen_tflops:.0f} TFLOP/s")
    # the uniform-data network has no skill at this depth (R^2 about 0): the
    # local rounds start from fresh weights and learn from local data only
    net = S.MLP(S.n_features(nang), 512, seed=8)
    rounds = S.optimise_local(net, sim, res.best_J, res.best_theta, rounds=args.local_rounds, sigma=0.1,
                              n_local=65536, candidates=cand, keep_frac=0.01, epochs=8, lr=1e-3, seed=17)
    local = []
    for r in rounds:
        local.append(dict(round=r.round + 1, centre_energy=r.centre_energy, heldout_r2=r.heldout_r2,
                          mean_exact_kept=r.mean_exact_kept, mean_exact_random=r.mean_exact_random,
                          sd_random=r.sd_random, gain_sd=r.gain_sd, best_kept=r.best_exact,
                          surrogate_mae_kept=r.surrogate_mae_kept, surrogate_max_err_kept=r.surrogate_max_err_kept,
                          best_with_network=r.centre_after, baseline_best=r.baseline_best, seconds={k: round(v, 3) for k, v in r.seconds.items()}))
        log(f"surrogate round {r.round + 1} (local, sigma 0.1): centre {r.centre_energy:.4f}, R^2 {r.heldout_r2:.3f}, "
            f"kept {r.mean_exact_kept:.4f} vs random {r.mean_exact_random:.4f} = {r.gain_sd:.2f} sd; "
            f"best {r.centre_after:.4f}, chain without the network {r.baseline_best:.4f}")
    final = rounds[-1].centre_after
    return dict(qubits=n, layers=layers, round0=round0, local_rounds=local,
                local_gain_sd_mean=float(np.mean([r.gain_sd for r in rounds])),
                local_heldout_r2_mean=float(np.mean([r.heldout_r2 for r in rounds])),
                best_energy_with_network=final, best_energy_without_network=rounds[-1].baseline_best)


def profiled_traffic(kernel: str, path: str):
    """HBM bytes per dispatch of `kernel` from a committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE runs of this same workload, scaled by a plain
    copy's counters); None when the summary does not cover this kernel."""
    full = os.path.join(os.path.dirname(os.path.abspath(__file__)), path)
    if not os.path.exists(full):
        return None
    lines = open(full).read().splitlines()
    for i, line in enumerate(lines):
        if line.startswith(kernel) and i + 2 < len(lines) and "traffic_bytes_per_dispatch:" in lines[i + 2]:
            return int(lines[i + 2].split("traffic_bytes_per_dispatch:")[1].split()[0])
    return None


def round_trip(qiron, W, n, prec, gates, opts):
    """1 - |<+...+| U^-1 U |+...+>|^2 at the full size: the circuit, then its
    inverse, through the same executor (a size-independent check of the path
    being timed; FP64 reduction on the device)."""
    s = qiron.State.new_plus(n, prec)
    s.apply_gates_inplace(gates, **opts)
    s.apply_gates_inplace(W.inverse_gates(gates), **opts)
    ref = qiron.State.new_plus(n, prec)
    f = s.fs_fidelity(ref)
    del s, ref
    return 1.0 - f


def profiled_valu(kernel: str, path: str):
    """wave64 VALU instructions per dispatch of `kernel` from a committed
    rocprofv3 --pmc summary (SQ_INSTS_VALU) of this same workload; None when
    the summary is missing or covers another kernel."""
    full = os.path.join(os.path.dirname(os.path.abspath(__file__)), path)
    if not os.path.exists(full):
        return None
    text = open(full).read()
    if kernel not in text.splitlines()[0]:
        return None
    for line in text.splitlines():
        if line.startswith("SQ_INSTS_VALU"):
            return float(line.split()[1])
    return None


def single_gates(qiron, L, W, ctx, n, results):
    """One H gate on the middle qubit of an n-qubit state, streaming kernel:
    the HBM-roofline case (reads and writes the whole state once)."""
    for prec, ab in (("c64", 8), ("c128", 16)):
        S = (1 << n) * ab
        s = qiron.State.new_plus(n, prec)
        one = W.GateList(4)
        one.add("H", [n // 2])
        arr = one.array()
        ts = []
        for rep in range(6):
            s.apply_gates_inplace(arr, mode=L.EXEC_PER_GATE)
            if rep:
                ts.append(ctx.last_exec_ms())
        del s
        ms = min(ts)
        gbps = 2 * S / (ms * 1e-3) / 1e9
        log(f"{prec} single H gate at {n}q: {ms:.3f} ms, {gbps:.0f} GB/s ({gbps / HBM_PEAK_GBPS:.1%} of 8 TB/s)")
        results[f"{prec}_single_gate"] = dict(ms=ms, GBps=gbps, frac=gbps / HBM_PEAK_GBPS,
                                              ms_all=[round(t, 4) for t in ts])
        # the same gate on every target qubit (the stride between partners changes the rate)
        s = qiron.State.new_plus(n, prec)
        fr = []
        for t in range(n):
            g = W.GateList(4)
            g.add("H", [t])
            a = g.array()
            tt = []
            for rep in range(4):
                s.apply_gates_inplace(a, mode=L.EXEC_PER_GATE)
                if rep:
                    tt.append(ctx.last_exec_ms())
            fr.append(2 * S / (min(tt) * 1e-3) / 1e9 / HBM_PEAK_GBPS)
        del s
        srt = sorted(fr)
        results[f"{prec}_single_gate"]["every_target"] = dict(
            frac_of_peak=[round(x, 4) for x in fr], median=srt[n // 2], min=srt[0], max=srt[-1],
            at_least_80pct=sum(x >= 0.8 for x in fr))
        log(f"{prec} H on each of {n} targets: median {srt[n // 2]:.1%}, range {srt[0]:.1%}-{srt[-1]:.1%}, "
            f"{sum(x >= 0.8 for x in fr)} targets at >= 80%")


def copy_calibration(qiron, ctx, n=29, reps=5):
    """A plain device copy of known size (the library's state copy kernel,
    2^29 complex64 amplitudes = 4 GiB read and 4 GiB written per launch): the
    reference that scales the HBM byte counters of a profiled run."""
    src = qiron.State.new_plus(n, "c64")
    dst = qiron.State.new_plus(n, "c64")
    nbytes = (1 << n) * 8
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_from(src)
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / reps
    log(f"copy of {nbytes} bytes: {dt * 1e3:.3f} ms per copy (wall), {2 * nbytes / dt / 1e9:.0f} GB/s (read + write)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--only", default="", choices=["", "tiled", "tiled128", "gate", "copy", "surrogate", "ceiling",
                                                   "energies", "trotter"],
                    help="profiling runs: only the headline executor (complex64, or complex128 with tiled128), only "
                         "the single-gate kernels, only a plain copy of known size (the reference for HBM byte "
                         "counters), only the surrogate, only the 128 GiB memory-ceiling runs, only the batched "
                         "energies, or only the Trotter evolution")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-qubits", type=int, default=0, help="tile width of the headline executor (0: default)")
    ap.add_argument("--batch-circuits", type=int, default=4096)
    ap.add_argument("--plain-circuits", type=int, default=64, help="circuits timed on the plain FP64 path")
    ap.add_argument("--schedule-c64", default="", help="batched energies: 'k,ks' instead of autotuning (complex64)")
    ap.add_argument("--schedule-c128", default="", help="the same for complex128")
    ap.add_argument("--train-samples", type=int, default=1 << 20)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--local-rounds", type=int, default=5, help="local optimisation rounds after the uniform round")
    ap.add_argument("--candidates", type=int, default=1 << 20)
    ap.add_argument("--no-energies", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip configs 1 and 4 and the memory-ceiling runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: the state-vector engine here is single-GPU "
              f"(--gpus {args.gpus} is not supported; sharding is described in the design notes)", file=sys.stderr)
        sys.exit(2)

    import qiron
    from qiron import _lib as L
    from qiron import workloads as W

    ctx = qiron.default_context()
    info = ctx.info()
    log("device", info)
    n, depth = args.qubits, args.depth
    gates = W.random_circuit(n=n, depth=depth, seed=W.SEED_BASE + 2)
    G = int(gates.size)
    log(f"circuit: {G} gates on {n} qubits")

    if args.only == "copy":
        copy_calibration(qiron, ctx)
        return
    results = {}
    if args.only == "gate":
        single_gates(qiron, L, W, ctx, n, results)
        print(json.dumps(results), flush=True)
        return
    if args.only == "surrogate":
        print(json.dumps(bench_surrogate(W, args)), flush=True)
        return
    if args.only == "energies":
        eres, _ = bench_energies(qiron, L, W, args)
        print(json.dumps(eres), flush=True)
        return
    if args.only == "ceiling":
        print(json.dumps(bench_ceiling(qiron, L, W)), flush=True)
        return
    if args.only == "trotter":
        print(json.dumps(bench_trotter(qiron, W, args)), flush=True)
        return

    tiled = dict(mode=L.EXEC_TILED, use_graph=True, tile_qubits=args.tile_qubits)
    if args.only == "tiled128":
        wall, dev_ms, passes, norm = run_circuit(qiron, L, n, "c128", gates, tiled, args.steps, args.warmup)
        pass_ms = dev_ms / args.steps / max(passes, 1)
        log(f"c128 tiled: {wall / args.steps * 1e3:.1f} ms/step, {G * args.steps / wall:.1f} gates/s, {passes} passes, "
            f"{pass_ms:.3f} ms/pass, {2 * (1 << n) * 16 / (pass_ms * 1e-3) / 1e9:.0f} GB/s per pass, norm {norm:.6f}")
        return
    # the library's complex64 defaults: 4096-amplitude tiles on one wave (512 threads for 8192 amplitudes)
    tile_k = args.tile_qubits or 12
    tile_tb = 9 if tile_k >= 13 else 6
    # ---- headline: tiled executor, complex64
    wall, dev_ms, passes, norm = run_circuit(qiron, L, n, "c64", gates, tiled, args.steps, args.warmup)
    ms_step = wall / args.steps * 1e3
    value = G * args.steps / wall
    S64 = (1 << n) * 8
    pass_ms = dev_ms / args.steps / max(passes, 1)
    achieved = 2 * S64 / (pass_ms * 1e-3) / 1e9
    log(f"c64 tiled: {ms_step:.1f} ms/step, {value:.1f} gates/s, {passes} passes, {pass_ms:.3f} ms/pass, "
        f"{achieved:.0f} GB/s per pass, norm {norm:.6f}")
    results["c64_tiled"] = dict(gates_per_s=value, ms_per_step=ms_step, passes=passes, pass_ms=pass_ms,
                                pass_GBps=achieved, norm=norm)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                "kernel": f"k_wave_pass<float, {tile_k}>" if tile_tb == 6 else f"k_reg_pass<float, {tile_k}, {tile_tb}>",
                "algorithmic_bytes_per_launch": 2 * S64}
    if n == 30 and depth == 40:
        src = "profiles/k_wave_pass_c64_30q_random_circuit.txt"
        roofline["traffic"] = profiled_traffic(roofline["kernel"], src)
        if roofline["traffic"] is not None:
            roofline["traffic_source"] = src
        # the second resource of the pass: VALU issue (wave64 instruction = 4 cycles of a 16-lane SIMD)
        vsrc = "profiles/k_wave_pass_valu_c64_30q_random_circuit.txt"
        valu = profiled_valu(roofline["kernel"], vsrc)
        if valu is not None:
            valu_ms = valu * 4 / (info["cu_count"] * 4) / VALU_CLOCK_HZ * 1e3
            roofline["valu"] = {"instructions_per_launch": valu, "issue_ms": round(valu_ms, 3),
                                "issue_frac": round(valu_ms / pass_ms, 4), "source": vsrc}

    if not args.only:
        rt = round_trip(qiron, W, n, "c64", gates, tiled)
        results["c64_tiled"]["round_trip_infidelity"] = rt
        log(f"c64 tiled round trip (circuit, then its inverse) at {n}q: 1 - fidelity {rt:.2e}")
        single_gates(qiron, L, W, ctx, n, results)
        # ---- complex128 tiled
        wall, dev_ms, passes, norm = run_circuit(qiron, L, n, "c128", gates, tiled, 1, 1)
        results["c128_tiled"] = dict(gates_per_s=G / wall, ms_per_step=wall * 1e3, passes=passes,
                                     pass_GBps=2 * (1 << n) * 16 * passes / (dev_ms * 1e-3) / 1e9)
        results["c128_tiled"]["round_trip_infidelity"] = round_trip(qiron, W, n, "c128", gates, tiled)
        log(f"c128 tiled: {wall * 1e3:.1f} ms/step {G / wall:.1f} gates/s, round trip 1 - fidelity "
            f"{results['c128_tiled']['round_trip_infidelity']:.2e}")
        # ---- per-gate streaming and fused MFMA on the first 5 layers
        sub = gates[: (G // depth) * 5]
        for name, opts, prec in (("c64_pergate", dict(mode=L.EXEC_PER_GATE), "c64"),
                                 ("c128_pergate", dict(mode=L.EXEC_PER_GATE), "c128"),
                                 ("c64_fused6", dict(mode=L.EXEC_FUSED_MFMA, block_qubits=6), "c64"),
                                 ("c64_fused5", dict(mode=L.EXEC_FUSED_MFMA, block_qubits=5), "c64"),
                                 ("c128_fused6", dict(mode=L.EXEC_FUSED_MFMA, block_qubits=6), "c128")):
            wall, dev_ms, steps_, norm = run_circuit(qiron, L, n, prec, sub, opts, 1, 1)
            r = dict(gates_per_s=sub.size / wall, ms=wall * 1e3, launches=steps_)
            if "fused" in name:
                kb = opts["block_qubits"]
                flops = 8.0 * (1 << n) * (1 << kb) * steps_
                tf = flops / (dev_ms * 1e-3) / 1e12
                peak = FP32_PEAK_TFLOPS if prec == "c64" else FP64_PEAK_TFLOPS
                r.update(mfma_TFLOPs=tf, mfma_util=tf / peak)
            results[name] = r
            log(name, r)

    if not args.only and not args.no_energies:
        eres, plain_eps = bench_energies(qiron, L, W, args)
        results["energies"] = eres
        results["surrogate"] = bench_surrogate(W, args)
    if not args.only and not args.no_extras:
        results["config1"] = bench_config1(qiron, L, W, args)
        results["trotter"] = bench_trotter(qiron, W, args)
        results["measure"] = bench_measure(qiron, W)
        results["ceiling"] = bench_ceiling(qiron, L, W)

    cpu = None
    if not args.no_cpu and not args.only:
        host = host_cpu()
        threads = min(host["usable"], 16)
        done, dt = cpu_baseline(W, n, depth, args.cpu_seconds, threads)
        gps = done / dt
        cpu = {"value": round(gps, 6), "unit": "gates/s", "cores": threads, "kind": "port",
               "sample": f"reference-structured CPU path (clone per gate, update list + serial scatter, {threads} "
                         f"threads), complex128, the first {done} gates of this {n}-qubit circuit run directly "
                         f"({dt:.1f} s of CPU path time, host buffer copies excluded); host {host['model']}, "
                         f"nproc {host['nproc']}, {host['usable']} usable",
               "host": host}
        log("cpu baseline", cpu)

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "gates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"random circuit, {n} qubits, depth {depth} ({G} gates), complex64 amplitudes, "
                               f"register-tiled executor", "qubits": n, "depth": depth, "gates": G, "seed": W.SEED_BASE + 2},
        "roofline": roofline, "cpu_baseline": cpu, "device": info["name"], "extra": results,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
