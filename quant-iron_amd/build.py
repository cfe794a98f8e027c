"""Build the gfx950 engine library (libqiron.so) from csrc/ with hipcc.

Each .hip file is compiled to an object in parallel, then linked into
quant-iron_amd/qiron/libqiron.so (in-tree, so it travels with the sources).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "qiron" / "libqiron.so"
BUILD = PKG / "_build"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off: hipcc must not emit v_pk_*_f32
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


# Per-file flags. tiled.hip keeps a thread's share of the tile in registers
# across a loop over op kinds: common-code sinking would merge element
# updates into stores through selected addresses (the tile would live in
# scratch memory), and structurizing the uniform op switch would give each
# path its own copy of the tile registers (whole-tile moves at every join).
FILE_FLAGS = {"tiled.hip": ["-mllvm", "-simplifycfg-sink-common=false", "-mllvm", "-structurizecfg-skip-uniform-regions"]}


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(src.name, []), "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{r.stderr[-6000:]}")
    return obj


HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", "-D__HIP_PLATFORM_AMD__",
              "-I/opt/rocm/include"]


def _compile_host(src: Path) -> Path:
    """Host-only sources (no device code): plain g++."""
    obj = BUILD / (src.stem + ".o")
    r = subprocess.run(["g++", *HOST_FLAGS, "-c", str(src), "-o", str(obj)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed on {src.name}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    host_srcs = sorted(CSRC.glob("*.cpp"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs) + len(host_srcs))) as ex:
        objs = list(ex.map(_compile, srcs)) + list(ex.map(_compile_host, host_srcs))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    build_host_tests(verbose)
    return OUT


HOST_TEST_SRC = PKG.parent / "tests" / "cpp" / "test_state_api.cpp"
HOST_TEST_BIN = PKG / "host" / "test_state_api"


def build_host_tests(verbose: bool = False) -> Path:
    """The C++ host-API test program (plain g++ against libqiron.so)."""
    cmd = ["g++", "-O1", "-std=c++17", "-I", str(PKG.parent / "include"), str(HOST_TEST_SRC), "-o", str(HOST_TEST_BIN),
           "-L", str(OUT.parent), "-lqiron", "-Wl,-rpath,$ORIGIN/../qiron"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host test build failed:\n{r.stderr[-4000:]}")
    if verbose:
        print(f"built {HOST_TEST_BIN}")
    return HOST_TEST_BIN


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
