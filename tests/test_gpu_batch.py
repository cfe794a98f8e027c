"""The batched engine (one graph launch for many circuits) against the oracle
run one circuit at a time."""
import numpy as np
import pytest

import oracle as O
import qiron
from kat_util import ONE_Q, gates_array, random_gate, random_terms
from qiron import workloads as W

pytestmark = pytest.mark.gpu


def build(batch, specs):
    """specs: (n, layers, thetas, J) from workloads.batched_ansatz; gate angles
    come from slots into one shared table."""
    angles = []
    refs = []
    for (n, layers, thetas, j) in specs:
        gates = W.ansatz_circuit(n, layers, thetas)
        terms = W.heisenberg_open_1d(n, j[0], j[1], j[2])
        slots = np.full(gates.size, -1, dtype=np.int32)
        rot = np.where(np.isin(gates["kind"], [10, 11, 12]))[0]
        slots[rot] = len(angles) + np.arange(rot.size)
        angles.extend(gates["params"][rot, 0].tolist())
        batch.add(n, gates, terms, slots)
        refs.append((n, gates, terms))
    batch.set_angles(np.array(angles))
    return np.array(angles), refs


def oracle_energies(refs):
    out = []
    for n, gates, terms in refs:
        psi = O.apply_gates(O.init(n, 0), gates)
        out.append(O.expectation(psi, terms))
    return np.array(out)


@pytest.mark.parametrize("executor", ["wave", "lds"])
@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_batch_energies_match_one_at_a_time_oracle(ctx, prec, executor):
    specs = W.batched_ansatz(count=48, seed=77, nmin=3, nmax=17, lmin=1, lmax=4)
    b = qiron.CircuitBatch(prec)
    b.set_executor(executor)
    assert b.executor() == executor
    angles, refs = build(b, specs)
    e = b.run(use_graph=True)
    ref = oracle_energies(refs)
    tol = 1e-11 if prec == "c128" else 2e-5
    assert np.max(np.abs(e - ref)) <= tol, np.max(np.abs(e - ref))
    info = b.info()
    assert info["circuits"] == 48 and 0 < info["lds_resident"] < 48
    # a replay of the captured graph gives the same energies
    e2 = b.run(use_graph=True)
    assert np.array_equal(e, e2)
    # new angles through the table, same graph
    rng = np.random.default_rng(3)
    new = rng.uniform(0, 2 * np.pi, size=angles.size)
    b.set_angles(new)
    e3 = b.run(use_graph=True)
    k = 0
    refs2 = []
    for n, gates, terms in refs:
        g2 = gates.copy()
        rot = np.where(np.isin(g2["kind"], [10, 11, 12]))[0]
        g2["params"][rot, 0] = new[k:k + rot.size]
        k += rot.size
        refs2.append((n, g2, terms))
    ref3 = oracle_energies(refs2)
    assert np.max(np.abs(e3 - ref3)) <= tol


def test_batch_angle_table_of_another_size_replans(ctx):
    """A table with extra (unused) slots changes the size the per-slot
    trigonometry kernel covers: the batch is planned again and gives the same
    energies; a same-size table afterwards replays the graph."""
    specs = W.batched_ansatz(count=12, seed=5, nmin=3, nmax=15, lmin=1, lmax=3)
    b = qiron.CircuitBatch("c128")
    angles, refs = build(b, specs)
    e = b.run(use_graph=True)
    b.set_angles(np.concatenate([angles, np.full(7, 1.25)]))
    e2 = b.run(use_graph=True)
    assert np.max(np.abs(e2 - e)) <= 1e-13
    b.set_angles(np.concatenate([angles, np.full(7, -0.5)]))
    e3 = b.run(use_graph=True)
    assert np.max(np.abs(e3 - e)) <= 1e-13
    assert np.max(np.abs(e - oracle_energies(refs))) <= 1e-11


def test_batch_rejects_invalid_gates(ctx):
    b = qiron.CircuitBatch("c64")
    g = W.GateList()
    g.add("CNOT", [0], [0])
    with pytest.raises(qiron.QuantIronError) as ei:
        b.add(3, g.array(), W.heisenberg_open_1d(3, 1, 1, 1))
    assert ei.value.variant == "OverlappingControlAndTargetQubits"


@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_every_schedule_and_the_autotuned_one_match_the_oracle(ctx, prec):
    specs = W.batched_ansatz(count=24, seed=78, nmin=9, nmax=16, lmin=1, lmax=3)
    b = qiron.CircuitBatch(prec)
    _, refs = build(b, specs)
    ref = oracle_energies(refs)
    tol = 1e-11 if prec == "c128" else 2e-5
    kmax = 12 if prec == "c128" else 13
    for executor in ("lds", "wave"):
        b.set_executor(executor)
        for k in range(10, kmax + 1):
            for ks in (k, kmax):
                b.set_schedule(k, ks)
                e = b.run(use_graph=True)
                assert np.max(np.abs(e - ref)) <= tol, (executor, k, ks, np.max(np.abs(e - ref)))
                assert b.schedule() == (k, ks)
    k, ks, ms = b.autotune(trials=1)
    assert 10 <= k <= ks <= kmax and ms > 0
    assert b.schedule() == (k, ks) and b.executor() in ("wave", "lds")
    e = b.run(use_graph=True)
    assert np.max(np.abs(e - ref)) <= tol


def test_schedule_out_of_range_is_rejected(ctx):
    b = qiron.CircuitBatch("c128")
    with pytest.raises(qiron.QuantIronError):
        b.set_schedule(13, 13)
    with pytest.raises(qiron.QuantIronError):
        b.set_schedule(12, 11)


@pytest.mark.parametrize("prec", ["c64", "c128"])
def test_one_wave_passes_on_general_circuits_match_the_oracle(ctx, prec):
    """Circuits of every operator kind with up to two controls on 12..15
    qubits, RX / RY / RZ / P (controlled or not) reading the angle table, and
    random Pauli terms with complex coefficients: the one-wave executor and
    the LDS one both give the oracle's energies, before and after new angles.
    Uncontrolled SWAPs make one-wave plans with in-place permutation passes;
    circuits with controlled SWAPs or matchgates (streamed gates in a one-wave
    plan) run LDS tiles inside the same batch."""
    rng = np.random.default_rng(91)
    rot_kinds = [W.K[k] for k in ("RX", "RY", "RZ", "P")]

    def circuits(angle_rng):
        out, angles = [], []
        for c in range(12):
            n = 12 + c % 4
            kinds = ONE_Q + ["CNOT", "TOFFOLI"] + (["SWAP", "MATCHGATE"] if c % 3 == 0 else [])
            if c % 3 == 1:
                kinds = kinds + ["SWAP"]
            g = gates_array([random_gate(n, np.random.default_rng(1000 + c), kinds=kinds, max_controls=2)
                             for _ in range(150)])
            if c % 3 == 1:  # uncontrolled SWAPs only (the relabelled form)
                g = g[~((g["kind"] == W.K["SWAP"]) & (g["ncontrols"] > 0))]
            rot = np.where(np.isin(g["kind"], rot_kinds))[0]
            slots = np.full(g.size, -1, np.int32)
            slots[rot] = len(angles) + np.arange(rot.size)
            g["params"][rot, 0] = angle_rng.uniform(-2 * np.pi, 2 * np.pi, size=rot.size)
            angles.extend(g["params"][rot, 0].tolist())
            out.append((n, g, slots))
        # edge shapes: no gates at all, and uncontrolled SWAPs only (a plan of
        # permutation passes alone)
        out.append((13, gates_array([]), np.zeros(0, np.int32)))
        sw = gates_array([{"kind": "SWAP", "targets": [int(a), int(b)], "controls": []}
                          for a, b in np.random.default_rng(7).integers(0, 12, size=(40, 2)) if a != b])
        out.append((12, sw, np.full(sw.size, -1, np.int32)))
        return out, np.array(angles)

    circs, angles = circuits(rng)
    terms = [random_terms(n, 10, np.random.default_rng(50 + i)) for i, (n, _, _) in enumerate(circs)]
    b = qiron.CircuitBatch(prec)
    for (n, g, slots), t in zip(circs, terms):
        b.add(n, g, t, slots)
    b.set_angles(angles)
    tol = 1e-11 if prec == "c128" else 5e-5
    ref = oracle_energies([(n, g, t) for (n, g, _), t in zip(circs, terms)])
    for executor in ("wave", "lds"):
        b.set_executor(executor)
        e = b.run(use_graph=True)
        assert np.max(np.abs(e - ref)) <= tol, (executor, np.max(np.abs(e - ref)))
    # new angles through the table: one upload and a graph replay
    circs2, angles2 = circuits(np.random.default_rng(92))
    b.set_angles(angles2)
    ref2 = oracle_energies([(n, g, t) for (n, g, _), t in zip(circs2, terms)])
    e2 = b.run(use_graph=True)
    assert np.max(np.abs(e2 - ref2)) <= tol
    b.set_executor("wave")
    e3 = b.run(use_graph=True)
    assert np.max(np.abs(e3 - ref2)) <= tol
    # new term coefficients (complex) without replanning: the grouped term
    # table keeps its order, only the signed coefficients change
    crng = np.random.default_rng(93)
    terms2 = []
    for t in terms:
        t2 = t.copy()
        t2["c_re"] = crng.normal(size=t.size)
        t2["c_im"] = crng.normal(size=t.size) * (crng.random(t.size) < 0.5)
        terms2.append(t2)
    b.set_term_coeffs(np.concatenate([t2["c_re"] + 1j * t2["c_im"] for t2 in terms2]))
    ref4 = oracle_energies([(n, g, t2) for (n, g, _), t2 in zip(circs2, terms2)])
    for executor in ("wave", "lds"):
        b.set_executor(executor)
        e4 = b.run(use_graph=True)
        assert np.max(np.abs(e4 - ref4)) <= tol, (executor, np.max(np.abs(e4 - ref4)))


def test_full_size_ansatz_circuits_match_the_per_gate_path(ctx):
    """Ansatz circuits at the batch's largest sizes (20..22 qubits, 8 layers:
    one-wave plans with the tile qubit search), in both precisions, against
    the per-gate complex128 path and the SumOp expectation (each checked
    against the oracle at smaller sizes): 1e-11 in complex128, 1e-5 in
    complex64."""
    from qiron import _lib as L
    specs = [(n, 8, np.random.default_rng(40 + n).uniform(0, 2 * np.pi, 2 * n * 8), (0.7, 1.1, 1.3))
             for n in (20, 21, 22)]
    ref = []
    for n, layers, th, j in specs:
        s = qiron.State.new_zero(n, "c128")
        s.apply_gates_inplace(W.ansatz_circuit(n, layers, th), mode=L.EXEC_PER_GATE)
        ref.append(qiron.expectation_terms(s, W.heisenberg_open_1d(n, *j)))
        del s
    ref = np.array(ref)
    for prec, tol in (("c128", 1e-11), ("c64", 1e-5)):
        b = qiron.CircuitBatch(prec)
        build(b, specs)
        assert b.executor() == "wave"
        e = b.run(use_graph=True)
        assert np.max(np.abs(e - ref)) <= tol, (prec, np.max(np.abs(e - ref)))
