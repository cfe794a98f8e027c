"""The C++ host mirror (quant-iron_amd/host/quant_iron.hpp) runs the reference
KATs in C++ against the engine."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parent.parent / "quant-iron_amd" / "host" / "test_state_api"


def test_cpp_host_api_kats():
    assert BIN.exists(), "build the package first (__graft_entry__.build)"
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
