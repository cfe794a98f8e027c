"""OpenQASM 3.0 input (qi_qasm_parse): the exporter's programs read back into the
gate records they came from, standard syntax, modifiers and errors. The reader
is host code in the engine library, so these run without a GPU.

Semantics are pinned against dense matrices built here from the OpenQASM 3
definitions (U(θ, φ, λ), sx, cu, ctrl / negctrl / inv), applied to random
states by the oracle: a wrong decomposition, a dropped phase or a swapped
control fails them."""
import json
import math
import zlib
from pathlib import Path

import numpy as np
import pytest

import oracle as O
import qiron
from kat_util import random_state
from qiron import _lib as L

QASM = Path(__file__).resolve().parent / "golden" / "qasm"
CASES = json.loads((QASM / "qasm_cases.json").read_text())
KIND_NAME = {v: k for k, v in L.GATE_KINDS.items()}


def parse(text, params=None):
    return qiron.core.parse_qasm(text, params)


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_exporter_programs_read_back_to_their_gates(case):
    n, arr, slots, npar = parse((QASM / case["file"]).read_text())
    assert n == case["qubits"] and npar == 0 and np.all(slots == -1)
    exp = case["gates"]
    assert arr.size == len(exp), case["cite"]
    for g, e in zip(arr, exp):
        kind = KIND_NAME[int(g["kind"])]
        assert kind == e["kind"], (kind, e)
        if kind == "MEASURE":
            assert int(g["basis"]) == e["basis"]
            assert [int(x) for x in g["controls"][: int(g["ncontrols"])]] == e["qubits"]
            assert math.isnan(float(g["params"][0]))  # drawn when the gate runs
            continue
        assert [int(x) for x in g["targets"][: int(g["ntargets"])]] == e["targets"]
        assert [int(x) for x in g["controls"][: int(g["ncontrols"])]] == e["controls"]
        for k, p in enumerate(e["params"]):
            assert float(g["params"][k]) == p  # printed with full precision (or 3 decimals for U)


def test_exporter_test_circuit_keeps_its_invalid_gate():
    """The circuit of circuit_tests.rs:135-157 has ctdag with target 3 and
    control 3: reading it succeeds, running it fails validation with the
    reference's OverlappingControlAndTargetQubits(3, 3), before any gate runs."""
    n, arr, _, _ = parse((QASM / "circuit_tests_to_qasm.qasm").read_text())
    with pytest.raises(O.OracleError) as ei:
        O.apply_gates(O.init(n, 0), arr)
    assert (ei.value.code, ei.value.a, ei.value.b) == (2, 3, 3)


I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)


def u_matrix(th, ph, la):
    """OpenQASM 3 U(θ, φ, λ)."""
    c, s = math.cos(th / 2), math.sin(th / 2)
    return np.array([[c, -np.exp(1j * la) * s], [np.exp(1j * ph) * s, np.exp(1j * (ph + la)) * c]])


def dense_1q(n, m, target, controls=(), neg=()):
    """The 2^n matrix of a (multi-)controlled one-qubit gate; `neg` controls fire on 0."""
    N = 1 << n
    M = np.zeros((N, N), dtype=complex)
    for col in range(N):
        on = all((col >> c) & 1 for c in controls) and all(not (col >> c) & 1 for c in neg)
        if not on:
            M[col, col] = 1
            continue
        b = (col >> target) & 1
        for nb in (0, 1):
            row = (col & ~(1 << target)) | (nb << target)
            M[row, col] += m[nb, b]
    return M


def run(text, psi, params=None):
    _, arr, _, _ = parse(text, params)
    return O.apply_gates(psi, arr)


HEAD = 'OPENQASM 3.0;\ninclude "stdgates.inc";\nqubit[3] q;\n'


@pytest.mark.parametrize("stmt,mat", [
    ("U(0.7, -1.3, 2.1) q[1];", dense_1q(3, u_matrix(0.7, -1.3, 2.1), 1)),
    ("u3(0.7, -1.3, 2.1) q[1];", dense_1q(3, u_matrix(0.7, -1.3, 2.1), 1)),
    ("u2(0.4, 1.9) q[2];", dense_1q(3, u_matrix(math.pi / 2, 0.4, 1.9), 2)),
    ("sx q[0];", dense_1q(3, np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]]) / 2, 0)),
    ("inv @ sx q[0];", dense_1q(3, np.array([[1 - 1j, 1 + 1j], [1 + 1j, 1 - 1j]]) / 2, 0)),
    ("cu(0.3, 0.5, -0.8, 1.1) q[2], q[0];",
     dense_1q(3, np.diag([1, np.exp(1.1j)]), 2) @ dense_1q(3, u_matrix(0.3, 0.5, -0.8), 0, [2])),
    ("ctrl @ U(0.7, -1.3, 2.1) q[0], q[2];", dense_1q(3, u_matrix(0.7, -1.3, 2.1), 2, [0])),
    ("inv @ U(0.7, -1.3, 2.1) q[1];", dense_1q(3, u_matrix(0.7, -1.3, 2.1).conj().T, 1)),
    ("negctrl @ x q[0], q[1];", dense_1q(3, X, 1, [], [0])),
    ("ctrl(2) @ rz(0.9) q[0], q[1], q[2];",
     dense_1q(3, np.diag([np.exp(-0.45j), np.exp(0.45j)]), 2, [0, 1])),
    ("cp(pi/3) q[1], q[0];", dense_1q(3, np.diag([1, np.exp(1j * math.pi / 3)]), 0, [1])),
    ("ccx q[2], q[0], q[1];", dense_1q(3, X, 1, [2, 0])),
    ("inv @ s q[2];", dense_1q(3, np.diag([1, -1j]), 2)),
    ("inv @ tdg q[2];", dense_1q(3, np.diag([1, np.exp(1j * math.pi / 4)]), 2)),
    ("ry(2 * pi / 5 - 0.25) q[0];", dense_1q(3, np.array([[math.cos(math.pi / 5 - 0.125), -math.sin(math.pi / 5 - 0.125)],
                                                           [math.sin(math.pi / 5 - 0.125), math.cos(math.pi / 5 - 0.125)]]), 0)),
])
def test_standard_gates_match_their_definitions(stmt, mat):
    psi = random_state(3, np.random.default_rng(zlib.crc32(stmt.encode())))
    np.testing.assert_allclose(run(HEAD + stmt, psi), mat @ psi, atol=1e-13)


def test_broadcast_registers_and_the_exporter_control_form():
    psi = random_state(4, np.random.default_rng(3))
    text = 'OPENQASM 3;\nqubit[2] a;\nqubit[2] b;\nh a;\ncx a, b;\nctrl(1) @ a[1] z b[0] // exporter form\n'
    H = np.array([[1, 1], [1, -1]]) / math.sqrt(2)
    M = dense_1q(4, H, 0) @ psi
    M = dense_1q(4, H, 1) @ M
    M = dense_1q(4, X, 2, [0]) @ M
    M = dense_1q(4, X, 3, [1]) @ M
    M = dense_1q(4, np.diag([1, -1]), 2, [1]) @ M
    np.testing.assert_allclose(run(text, psi), M, atol=1e-13)


def test_measurements_group_by_register_and_basis():
    text = HEAD + "bit[3] c;\nbit[1] d;\nc[0] = measure q[0];\nc[1] = measure q[2];\nd[0] = xmeasure q[1];\n" \
                  "measure q[1] -> c[2];\nh q[0];\nc = measure q;\n"
    _, arr, _, _ = parse(text)
    meas = [(int(g["basis"]), [int(x) for x in g["controls"][: int(g["ncontrols"])]]) for g in arr
            if int(g["kind"]) == L.GATE_KINDS["MEASURE"]]
    assert meas == [(0, [0, 2]), (1, [1]), (0, [1]), (0, [0, 1, 2])]


def test_circuit_parameters_become_angle_slots():
    text = HEAD + "input float[64] theta;\ninput angle phi;\nry(theta) q[0];\nrz(phi) q[1];\nrx(0.5) q[2];\nry(theta) q[2];\n"
    n, arr, slots, npar = parse(text, [0.25, -1.5])
    assert (n, npar) == (3, 2)
    assert list(slots) == [0, 1, -1, 0]
    assert [float(g["params"][0]) for g in arr] == [0.25, -1.5, 0.5, 0.25]
    c = qiron.Circuit.from_qasm(text, [0.25, -1.5])
    assert c.num_qubits == 3 and len(c.gates) == 4


@pytest.mark.parametrize("text,line,what", [
    (HEAD + "h q[3];", 4, "out of range"),
    (HEAD + "foo q[0];", 4, "unknown gate"),
    (HEAD + "\nrx q[0];", 5, "takes 1 arguments"),
    (HEAD + "cx q[0];", 4, "wrong number of qubits"),
    (HEAD + "input float t;\nry(2 * t) q[0];", 5, "whole angle"),
    (HEAD + "reset q[0];", 4, "not supported"),
    (HEAD + "gate g a { h a; }", 4, "not supported"),
    (HEAD + "bit[2] c;\nc = measure q;", 5, "differ in number"),
    (HEAD + "h q[0] x q[1];", 4, "expected ';'"),
])
def test_errors_name_the_line(text, line, what):
    with pytest.raises(qiron.QuantIronError) as ei:
        parse(text)
    assert ei.value.variant == "InvalidInputValue" and ei.value.a == line and what in str(ei.value), str(ei.value)


@pytest.fixture(scope="module")
def qasm_san(tmp_path_factory):
    """The reader (csrc/qasm.cpp, plain C++) with a small driver, built with the
    address and undefined-behaviour sanitizers (any finding fatal)."""
    import subprocess
    root = Path(__file__).resolve().parent.parent
    exe = tmp_path_factory.mktemp("qasm_san") / "qasm_san"
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", str(root / "quant-iron_amd" / "csrc" / "qasm.cpp"),
           str(root / "tests" / "cpp" / "qasm_san.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def test_reader_under_sanitizers_on_truncated_and_corrupted_programs(qasm_san, tmp_path):
    """Every prefix of an exporter program (cut at each byte of its first 600
    and at each line end) and 400 single-byte corruptions parse to a result or
    an error, with AddressSanitizer and UndefinedBehaviorSanitizer clean; the
    intact programs give the library's own gate counts."""
    import os
    import subprocess
    rng = np.random.default_rng(17)
    texts = [(QASM / c["file"]).read_bytes() for c in CASES]
    files = []

    def add(data):
        p = tmp_path / f"p{len(files)}.qasm"
        p.write_bytes(data)
        files.append(p)

    for t in texts:
        add(t)
    base = texts[-1]
    cuts = set(range(0, min(600, len(base)))) | {i + 1 for i, b in enumerate(base) if b == ord("\n")}
    for c in sorted(cuts):
        add(base[:c])
    alphabet = b";()[]{}@,.-+*/^0123456789abcxyzq \n\t\"'#"
    for _ in range(400):
        t = bytearray(texts[int(rng.integers(len(texts)))])
        i = int(rng.integers(len(t)))
        t[i] = alphabet[int(rng.integers(len(alphabet)))] if rng.random() < 0.8 else int(rng.integers(256))
        add(bytes(t))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(qasm_san)] + [str(p) for p in files], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    out = r.stdout.splitlines()
    assert len(out) == len(files)
    for line, t in zip(out[:len(texts)], texts):
        n, arr, _, _ = parse(t.decode())
        assert line == f"ok {arr.size} {n} 0", line
    print(f"{sum(l.startswith('ok') for l in out)} parsed, {sum(l.startswith('err') for l in out)} rejected")
