import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "quant-iron_amd", ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


@pytest.fixture(scope="session")
def ctx():
    """The engine context on device 0. GPU tests fail loudly if the HIP library
    or the device is missing: there is no CPU path to fall back to."""
    import qiron
    c = qiron.default_context()
    print(f"[gpu] {c.info()}", flush=True)
    return c
