"""ctypes binding of the C ABI declared in include/qiron.h.

The engine library is loaded from this directory (libqiron.so, built by
quant-iron_amd/build.py). There is no fallback: if the library is missing or a
symbol is absent, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libqiron.so"

MAX_TARGETS = 4
MAX_CONTROLS = 62

# qi_status values (errors.rs:4-97 order, then engine-only codes)
STATUS_NAMES = {
    0: "Ok",
    1: "InvalidNumberOfMeasurements",
    2: "OverlappingControlAndTargetQubits",
    3: "InvalidNumberOfQubits",
    4: "InvalidQubitIndex",
    5: "StateVectorNotNormalised",
    6: "NonUnitaryMatrix",
    7: "InvalidNumberOfInputs",
    8: "MismatchedNumberOfParameters",
    9: "UnknownError",
    10: "DeviceError",
    11: "GpuContextLockError",
    12: "CircuitMacroError",
    13: "InvalidInputValue",
    14: "ZeroNorm",
    15: "InvalidPauliStringCoefficient",
    16: "OutOfMemory",
    17: "InvalidArgument",
    18: "PrecisionMismatch",
}

# qi_gate_kind
GATE_KINDS = dict(H=0, X=1, Y=2, Z=3, I=4, S=5, T=6, SDAG=7, TDAG=8, P=9, RX=10, RY=11, RZ=12, U2=13,
                  CNOT=14, SWAP=15, MATCHGATE=16, TOFFOLI=17, MEASURE=18, PAULI_STRING=19,
                  PAULI_EVOLUTION=20)

C64, C128 = 0, 1
INIT_ZERO, INIT_BASIS, INIT_PLUS, INIT_MINUS, INIT_GHZ = range(5)
BASIS_COMPUTATIONAL, BASIS_X, BASIS_Y, BASIS_CUSTOM = range(4)
EXEC_PER_GATE, EXEC_TILED, EXEC_FUSED_MFMA, EXEC_AUTO = range(4)


class QiError(C.Structure):
    _fields_ = [("code", C.c_int32), ("a", C.c_uint64), ("b", C.c_uint64), ("re", C.c_double),
                ("im", C.c_double), ("msg", C.c_char * 256)]


class QiGate(C.Structure):
    _fields_ = [("kind", C.c_int32), ("ntargets", C.c_uint32), ("ncontrols", C.c_uint32),
                ("basis", C.c_uint32), ("targets", C.c_uint64 * MAX_TARGETS),
                ("controls", C.c_uint64 * MAX_CONTROLS), ("params", C.c_double * 3),
                ("matrix", C.c_double * 8), ("x_mask", C.c_uint64), ("z_mask", C.c_uint64)]


class QiPauliTerm(C.Structure):
    _fields_ = [("x_mask", C.c_uint64), ("z_mask", C.c_uint64), ("c_re", C.c_double), ("c_im", C.c_double)]


class QiExecOpts(C.Structure):
    _fields_ = [("mode", C.c_int32), ("tile_qubits", C.c_int32), ("block_qubits", C.c_int32),
                ("use_graph", C.c_int32), ("tile_threads_log2", C.c_int32)]


# numpy mirrors, for building large gate lists without per-gate ctypes objects
GATE_DTYPE = np.dtype([("kind", "<i4"), ("ntargets", "<u4"), ("ncontrols", "<u4"), ("basis", "<u4"),
                       ("targets", "<u8", (MAX_TARGETS,)), ("controls", "<u8", (MAX_CONTROLS,)),
                       ("params", "<f8", (3,)), ("matrix", "<f8", (8,)), ("x_mask", "<u8"),
                       ("z_mask", "<u8")], align=True)
TERM_DTYPE = np.dtype([("x_mask", "<u8"), ("z_mask", "<u8"), ("c_re", "<f8"), ("c_im", "<f8")], align=True)
assert GATE_DTYPE.itemsize == C.sizeof(QiGate), (GATE_DTYPE.itemsize, C.sizeof(QiGate))
assert TERM_DTYPE.itemsize == C.sizeof(QiPauliTerm)

P = C.c_void_p
ERR = C.POINTER(QiError)
D = C.c_double
U64 = C.c_uint64
U32 = C.c_uint32
I32 = C.c_int32
DP = C.POINTER(C.c_double)
U64P = C.POINTER(C.c_uint64)

# (name, argtypes) of every entry point of include/qiron.h
SIGNATURES = {
    "qi_ctx_create": [C.c_int, C.POINTER(P), ERR],
    "qi_ctx_destroy": [P],
    "qi_ctx_synchronize": [P, ERR],
    "qi_ctx_seed": [P, U64, ERR],
    "qi_ctx_info": [P, C.c_char_p, C.c_size_t, U64P, C.POINTER(C.c_int), ERR],
    "qi_state_create": [P, U32, C.c_int, C.POINTER(P), ERR],
    "qi_state_destroy": [P],
    "qi_state_clone": [P, C.POINTER(P), ERR],
    "qi_state_copy": [P, P, ERR],
    "qi_state_num_qubits": [P],
    "qi_state_precision": [P],
    "qi_state_device_ptr": [P],
    "qi_state_init": [P, C.c_int, U64, ERR],
    "qi_state_upload": [P, DP, C.c_int, ERR],
    "qi_state_download": [P, DP, ERR],
    "qi_apply_gate": [P, C.c_void_p, ERR],
    "qi_apply_gates": [P, U32, C.c_void_p, C.c_size_t, C.POINTER(QiExecOpts), ERR],
    "qi_pauli_apply": [P, P, C.POINTER(QiPauliTerm), ERR],
    "qi_pauli_apply_normalised": [P, P, C.POINTER(QiPauliTerm), ERR],
    "qi_pauli_exp": [P, C.POINTER(QiPauliTerm), D, D, ERR],
    "qi_pauli_exp_neg_i_dt": [P, C.POINTER(QiPauliTerm), D, ERR],
    "qi_sumop_apply": [P, P, C.c_void_p, C.c_size_t, ERR],
    "qi_sumop_expectation": [P, C.c_void_p, C.c_size_t, DP, DP, ERR],
    "qi_sumop_expectation_noalloc": [P, C.c_void_p, C.c_size_t, DP, DP, ERR],
    "qi_trotter_evolve": [P, C.c_void_p, C.c_size_t, D, U64, C.c_int, ERR],
    "qi_inner_product": [P, P, DP, DP, ERR],
    "qi_norm_sqr": [P, DP, ERR],
    "qi_normalise": [P, ERR],
    "qi_fidelity": [P, P, DP, ERR],
    "qi_fs_dist": [P, P, DP, ERR],
    "qi_state_scale": [P, D, D, ERR],
    "qi_state_axpy": [P, D, D, P, ERR],
    "qi_state_add": [P, P, P, ERR],
    "qi_state_sub": [P, P, P, ERR],
    "qi_measure_probs": [P, U64P, U32, DP, ERR],
    "qi_measure": [P, C.c_int, DP, U64P, U32, D, P, U64P, ERR],
    "qi_measure_n": [P, C.c_int, DP, U64P, U32, DP, C.c_size_t, U64P, ERR],
    "qi_measure_collapse": [P, P, U64P, U32, U64, ERR],
    "qi_measure_result_state": [P, C.c_int, DP, U64P, U32, U64, P, ERR],
    "qi_qasm_parse": [C.c_char_p, C.c_size_t, DP, C.c_size_t, C.c_void_p, C.POINTER(C.c_int32), C.c_size_t,
                      C.POINTER(C.c_size_t), C.POINTER(U32), C.POINTER(U32), ERR],
    "qi_param_circuit_create": [P, U32, C.c_int, C.c_void_p, C.c_size_t, C.POINTER(C.c_int32),
                                C.POINTER(QiExecOpts), C.POINTER(P), ERR],
    "qi_param_circuit_destroy": [P],
    "qi_param_circuit_num_slots": [P, U64P, ERR],
    "qi_param_circuit_set_angles": [P, DP, C.c_size_t, ERR],
    "qi_param_circuit_run": [P, P, ERR],
    "qi_batch_create": [P, C.c_int, C.POINTER(P), ERR],
    "qi_batch_destroy": [P],
    "qi_batch_add": [P, U32, C.c_void_p, C.c_size_t, C.POINTER(C.c_int32), C.c_void_p, C.c_size_t, ERR],
    "qi_batch_set_angles": [P, DP, C.c_size_t, ERR],
    "qi_batch_set_term_coeffs": [P, DP, C.c_size_t, ERR],
    "qi_batch_run": [P, DP, C.c_int, ERR],
    "qi_batch_info": [P, U64P, U64P, U64P, DP, ERR],
    "qi_batch_set_schedule": [P, C.c_uint32, C.c_uint32, ERR],
    "qi_batch_get_schedule": [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), ERR],
    "qi_batch_autotune": [P, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), DP, ERR],
    "qi_batch_set_executor": [P, C.c_int, ERR],
    "qi_batch_get_executor": [P, C.POINTER(C.c_int), ERR],
    "qi_last_exec_ms": [P, DP, ERR],
    "qi_last_exec_steps": [P, U64P, ERR],
}

RESTYPES = {"qi_state_num_qubits": U32, "qi_state_precision": I32, "qi_state_device_ptr": P}

_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libqiron.so and declare every exported signature."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise OSError(f"{p} is missing: build it with quant-iron_amd/build.py (hipcc --offload-arch=gfx950)")
    lib = C.CDLL(str(p))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)  # AttributeError if the symbol is missing
        fn.argtypes = args
        fn.restype = RESTYPES.get(name, C.c_int)
    if path is None:
        _lib = lib
    return lib


def dptr(a: np.ndarray):
    return a.ctypes.data_as(DP)


def u64ptr(a: np.ndarray):
    return a.ctypes.data_as(U64P)
