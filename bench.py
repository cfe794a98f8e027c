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


def cpu_baseline(qiron_workloads, n_cpu: int, budget_s: float, threads: int):
    """The reference-structured CPU path (clone per gate, update list + serial
    scatter, `threads` workers) on the first layer of the same circuit family
    at n_cpu qubits, complex128, for about budget_s seconds."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np

    import oracle as O

    gates = qiron_workloads.random_circuit(n=n_cpu, depth=2, seed=qiron_workloads.SEED_BASE + 2)
    psi = O.init(n_cpu, 2)
    done = 0
    t0 = time.perf_counter()
    while done < gates.size and time.perf_counter() - t0 < budget_s:
        psi = O.reference_structured_apply(psi, gates[done:done + 1], threads)
        done += 1
    dt = time.perf_counter() - t0
    return done, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--only", default="", help="run only the headline executor (for profiling)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile-qubits", type=int, default=0, help="tile width of the headline executor (0: default)")
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

    tiled = dict(mode=L.EXEC_TILED, use_graph=True, tile_qubits=args.tile_qubits)
    results = {}
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
                "kernel": "k_reg_pass<float,13>", "algorithmic_bytes_per_launch": 2 * S64}

    if not args.only:
        # ---- single 30-qubit gate application (streaming kernel): the HBM-roofline target
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
            results[f"{prec}_single_gate"] = dict(ms=ms, GBps=gbps, frac=gbps / HBM_PEAK_GBPS)
        # ---- complex128 tiled
        wall, dev_ms, passes, norm = run_circuit(qiron, L, n, "c128", gates, tiled, 1, 1)
        results["c128_tiled"] = dict(gates_per_s=G / wall, ms_per_step=wall * 1e3, passes=passes,
                                     pass_GBps=2 * (1 << n) * 16 * passes / (dev_ms * 1e-3) / 1e9)
        log(f"c128 tiled: {wall * 1e3:.1f} ms/step {G / wall:.1f} gates/s")
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

    cpu = None
    if not args.no_cpu and not args.only:
        threads = min(os.cpu_count() or 1, 16)
        n_cpu = 26
        done, dt = cpu_baseline(W, n_cpu, args.cpu_seconds, threads)
        gps = done / dt
        cpu = {"value": round(gps * 2.0 ** (n_cpu - n), 6), "unit": "gates/s", "cores": threads, "kind": "port",
               "sample": f"reference-structured CPU path (clone per gate, update list + serial scatter, {threads} "
                         f"threads), complex128, first {done} gates of the same circuit family at {n_cpu} qubits in "
                         f"{dt:.1f} s = {gps:.3f} gates/s; value scaled by 2^({n_cpu}-{n}) to {n} qubits "
                         f"(cost per gate is linear in the state size)"}
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
