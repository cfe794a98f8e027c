"""The C-ABI library loads and exports every entry point include/qiron.h
declares, and the ctypes/numpy record layouts equal the C structs. CPU only:
no compute call is made here."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from qiron import _lib as L
from qiron import surrogate  # noqa: F401  (registers the MLP entry points' signatures)

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "qiron.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:qi_status|uint32_t|int32_t|void\s*\*)\s*(qi_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_boundary():
    names = declared_functions()
    assert len(names) >= 30
    for must in ("qi_apply_gate", "qi_apply_gates", "qi_sumop_expectation", "qi_pauli_exp", "qi_measure_n"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(L.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each
    assert set(declared_functions()) <= set(L.SIGNATURES)


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "qiron.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu\\n\", sizeof(qi_gate), sizeof(qi_pauli_term),"
        " sizeof(qi_error), sizeof(qi_exec_opts), offsetof(qi_gate, params), offsetof(qi_gate, x_mask));return 0;}\n")
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", str(HEADER.parent), str(src), "-o", str(exe)], check=True)
    out = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == C.sizeof(L.QiGate) == L.GATE_DTYPE.itemsize
    assert out[1] == C.sizeof(L.QiPauliTerm) == L.TERM_DTYPE.itemsize
    assert out[2] == C.sizeof(L.QiError)
    assert out[3] == C.sizeof(L.QiExecOpts)
    assert out[4] == L.QiGate.params.offset == L.GATE_DTYPE.fields["params"][1]
    assert out[5] == L.QiGate.x_mask.offset == L.GATE_DTYPE.fields["x_mask"][1]


def test_engine_has_no_cpu_path():
    """The product package never imports or loads the oracle, and the library
    carries device code for gfx950."""
    pkg = ROOT / "quant-iron_amd"
    for f in list(pkg.rglob("*.py")) + list(pkg.rglob("*.hip")) + list(pkg.rglob("*.h")):
        t = f.read_text()
        assert "import oracle" not in t and "liboracle" not in t and "qi_oracle" not in t, f
    blob = L.LIB_PATH.read_bytes()
    assert b"gfx950" in blob


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(OSError):
        L.load(tmp_path / "libqiron.so")
