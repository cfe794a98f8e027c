"""The register-tiled planner's op tables, checked on the CPU.

tests/cpp/tile_emu.cpp plans a gate list with the engine's planner
(qi::plan_tiled from libqiron.so) and executes every step with the device
kernels' per-op semantics in complex128. Its output must equal the oracle's
on the same input up to rounding: a wrong op form, a stale layout after a
lane/register exchange, a diagonal folded across a non-diagonal op or a
wrong store layout all move amplitudes by O(1). These run without a GPU and
cover every tile width and thread count, both precisions' op forms,
relabelled SWAPs, streamed matchgates and parametric rewrites."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import qiron.workloads as W
from kat_util import gates_array, random_gate, random_state

ROOT = Path(__file__).resolve().parent.parent
LIBDIR = ROOT / "quant-iron_amd" / "qiron"


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    if not (LIBDIR / "libqiron.so").exists():
        pytest.fail("libqiron.so is not built (run __graft_entry__.build())")
    exe = tmp_path_factory.mktemp("emu") / "tile_emu"
    cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           str(ROOT / "tests" / "cpp" / "tile_emu.cpp"), "-o", str(exe), f"-L{LIBDIR}", "-lqiron",
           f"-Wl,-rpath,{LIBDIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


@pytest.fixture(scope="module")
def emu_san(tmp_path_factory):
    """The planner (tile_plan.cpp, host_common.cpp: plain C++) and the CPU model
    built with the address and undefined-behaviour sanitizers, any finding
    fatal: out-of-range layout / op-table indexing or integer overflow in the
    planner aborts the run."""
    exe = tmp_path_factory.mktemp("emu_san") / "tile_emu_san"
    csrc = ROOT / "quant-iron_amd" / "csrc"
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           str(csrc / "tile_plan.cpp"), str(csrc / "host_common.cpp"), str(ROOT / "tests" / "cpp" / "tile_emu.cpp"),
           "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def run_emu(emu, tmp_path, gates, psi, n, prec, K, TB, angles=None, oop=False, full=False):
    g, i, o = tmp_path / "g.bin", tmp_path / "in.bin", tmp_path / "out.bin"
    g.write_bytes(np.ascontiguousarray(gates).tobytes())
    i.write_bytes(np.ascontiguousarray(psi, dtype=np.complex128).tobytes())
    args = [str(emu), str(g), str(n), "128" if prec == "c128" else "64", str(K), str(TB), str(i), str(o)]
    if angles is not None:
        a = tmp_path / "angles.bin"
        a.write_bytes(np.asarray(angles, dtype=np.float64).tobytes())
        args += ["--angles", str(a)]
    if oop:
        args.append("--oop")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run(args, capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    out = np.frombuffer(o.read_bytes(), dtype=np.complex128)
    steps, moves = (int(x) for x in r.stdout.split()[:2])
    return (out, steps, moves) if full else (out, steps)


# (K, TB) pairs the executor accepts: one-wave tiles (TB 6) up to 12 / 11
# tile qubits, multi-wave LDS tiles (TB 8, 9) up to 13 / 12.
SHAPES = {"c64": [(10, 6), (11, 6), (12, 6), (12, 8), (13, 9)], "c128": [(10, 6), (11, 6), (11, 8), (12, 9)]}
CASES = [(p, k, tb) for p in ("c64", "c128") for (k, tb) in SHAPES[p]]


@pytest.mark.parametrize("prec,K,TB", CASES, ids=[f"{p}-K{k}-T{t}" for p, k, t in CASES])
def test_random_circuit_plans_execute_like_the_oracle(emu, tmp_path, prec, K, TB):
    n = 16
    gates = W.random_circuit(n=n, depth=8, seed=321 + K + TB)
    psi = random_state(n, np.random.default_rng(K * 10 + TB))
    got, steps = run_emu(emu, tmp_path, gates, psi, n, prec, K, TB)
    ref = O.apply_gates(psi, gates)
    assert np.max(np.abs(got - ref)) <= 1e-11, steps


@pytest.mark.parametrize("prec,K,TB", CASES, ids=[f"{p}-K{k}-T{t}" for p, k, t in CASES])
def test_controlled_gates_and_matchgates_plan_correctly(emu, tmp_path, prec, K, TB):
    """Every operator kind with up to three controls (lane-, register- and
    outside-tile controls, controlled SWAPs, matchgates)."""
    n = 15
    rng = np.random.default_rng(1000 + K * 10 + TB)
    gates = gates_array([random_gate(n, rng, max_controls=3) for _ in range(240)])
    psi = random_state(n, rng)
    got, _ = run_emu(emu, tmp_path, gates, psi, n, prec, K, TB)
    assert np.max(np.abs(got - O.apply_gates(psi, gates))) <= 1e-11


def test_thirty_qubit_circuit_shape_on_twenty_qubits(emu, tmp_path):
    """The 30-qubit depth-40 circuit's gate mix at 20 qubits, one-wave 12-qubit
    tiles (the complex64 headline plan: exchanges, shears, unscaled H, SWAP
    relabelling and the final permutation)."""
    n = 20
    gates = W.random_circuit(n=n, depth=40, seed=W.SEED_BASE + 2)
    psi = random_state(n, np.random.default_rng(7))
    got, steps = run_emu(emu, tmp_path, gates, psi, n, "c64", 12, 6)
    ref = O.apply_gates(psi, gates)
    assert np.max(np.abs(got - ref)) <= 1e-10, steps


@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_parametric_ops_rewritten_for_new_angles(emu, tmp_path, prec):
    """A plan built with every RX / RY / RZ / P parametric, its ops rewritten
    by param_op_coeffs for new angles, equals the oracle run at those angles."""
    n = 14
    gates = W.random_circuit(n=n, depth=10, seed=99)
    kinds = {int(W.K[k]) for k in ("RX", "RY", "RZ", "P")}
    idx = [i for i, g in enumerate(gates) if int(g["kind"]) in kinds]
    rng = np.random.default_rng(11)
    angles = rng.uniform(-2 * np.pi, 2 * np.pi, size=len(idx))
    moved = gates.copy()
    for a, i in zip(angles, idx):
        moved[i]["params"][0] = a
    psi = random_state(n, rng)
    got, _ = run_emu(emu, tmp_path, gates, psi, n, prec, 12 if prec == "c64" else 11, 6, angles=angles)
    assert np.max(np.abs(got - O.apply_gates(psi, moved))) <= 1e-11


@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_out_of_place_plan_with_final_permutation(emu, tmp_path, prec):
    """With a second buffer, the SWAP-relabelled plan moves the state to it in
    its first pass and back through the final qubit permutation in one last
    pass (base bits permuted at the store): the result is the oracle's, and the
    plan is shorter than the in-place one."""
    n = 20
    gates = W.random_circuit(n=n, depth=40, seed=W.SEED_BASE + 2)
    psi = random_state(n, np.random.default_rng(9))
    K = 10  # 10-qubit tiles: the in-place final permutation needs several passes at 20 qubits
    got, steps, moves = run_emu(emu, tmp_path, gates, psi, n, prec, K, 6, oop=True, full=True)
    _, steps_inplace, moves_inplace = run_emu(emu, tmp_path, gates, psi, n, prec, K, 6, full=True)
    assert moves == 2 and moves_inplace == 0
    assert steps < steps_inplace, (steps, steps_inplace)
    assert np.max(np.abs(got - O.apply_gates(psi, gates))) <= 1e-10


SAN_CASES = [("c64", 12, 6, False, 15), ("c128", 11, 6, False, 15), ("c64", 13, 9, False, 15),
             ("c64", 10, 6, True, 15), ("c128", 10, 6, True, 15), ("c64", 12, 6, True, 20)]


@pytest.mark.parametrize("prec,K,TB,oop,n", SAN_CASES,
                         ids=[f"{p}-K{k}-T{t}{'-oop' if o else ''}-n{n}" for p, k, t, o, n in SAN_CASES])
def test_planner_under_sanitizers(emu_san, tmp_path, prec, K, TB, oop, n):
    """Plans of a random circuit with SWAP relabelling, controlled gates and
    matchgates (one-wave and LDS tiles, in place and out of place; at 20
    qubits with the tile qubit search and its pair exchanges) run with
    AddressSanitizer and UndefinedBehaviorSanitizer and still match the
    oracle."""
    rng = np.random.default_rng(77 + K + TB)
    gates = np.concatenate([W.random_circuit(n=n, depth=6, seed=5 + K),
                            gates_array([random_gate(n, rng, max_controls=2) for _ in range(60)])])
    psi = random_state(n, rng)
    got, _ = run_emu(emu_san, tmp_path, gates, psi, n, prec, K, TB, oop=oop)
    assert np.max(np.abs(got - O.apply_gates(psi, gates))) <= 1e-10
