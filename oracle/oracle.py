"""ctypes loader of the CPU FP64 oracle (liboracle.so, built from qi_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
benchmark's CPU-baseline leg, never by the engine. Arrays are complex128
numpy vectors (interleaved re, im doubles, qubit k = bit k of the index).
"""
from __future__ import annotations

import ctypes as C
import time
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
SRC = HERE / "qi_oracle.c"


def build() -> Path:
    """gcc -O2 -ffp-contract=off: no FMA contraction, so the per-gate arithmetic
    rounds exactly as the reference's num-complex expressions do."""
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-Wall", "-o", str(LIB), str(SRC), "-lm",
           "-lpthread"]
    subprocess.run(cmd, check=True)
    return LIB


class OracleError(Exception):
    def __init__(self, code, a, b, re, im, msg):
        self.code, self.a, self.b, self.coefficient = code, a, b, complex(re, im)
        super().__init__(f"{code}: {msg}")


_lib = None


def _qiron_lib():
    import sys
    pkg = HERE.parent / "quant-iron_amd"
    if str(pkg) not in sys.path:
        sys.path.insert(0, str(pkg))
    from qiron import _lib as L  # struct layouts only (no engine calls)
    return L


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists() or LIB.stat().st_mtime < SRC.stat().st_mtime:
            build()
        _lib = C.CDLL(str(LIB))
        L = _qiron_lib()
        P, ERR = C.c_void_p, C.POINTER(L.QiError)
        DP, U64P = C.POINTER(C.c_double), C.POINTER(C.c_uint64)
        sig = {
            "or_validate_gate": [C.c_uint32, P, ERR],
            "or_apply_gate": [DP, C.c_uint32, P, ERR],
            "or_apply_gates": [DP, C.c_uint32, C.c_uint32, P, C.c_size_t, ERR],
            "or_state_init": [DP, C.c_uint32, C.c_int, C.c_uint64, ERR],
            "or_state_check_norm": [DP, C.c_uint32, ERR],
            "or_inner_product": [DP, DP, C.c_uint32, DP, DP],
            "or_norm_sqr": [DP, C.c_uint32],
            "or_normalise": [DP, C.c_uint32, ERR],
            "or_fidelity": [DP, DP, C.c_uint32, DP, ERR],
            "or_pauli_apply": [DP, DP, C.c_uint32, P, ERR],
            "or_pauli_apply_normalised": [DP, DP, C.c_uint32, P, ERR],
            "or_pauli_exp": [DP, C.c_uint32, P, C.c_double, C.c_double, ERR],
            "or_pauli_exp_neg_i_dt": [DP, C.c_uint32, P, C.c_double, ERR],
            "or_sumop_apply": [DP, DP, C.c_uint32, P, C.c_size_t, ERR],
            "or_sumop_expectation": [DP, C.c_uint32, P, C.c_size_t, DP, DP, ERR],
            "or_trotter_evolve": [DP, C.c_uint32, P, C.c_size_t, C.c_double, C.c_uint64, C.c_int, ERR],
            "or_measure_probs": [DP, C.c_uint32, U64P, C.c_uint32, DP, ERR],
            "or_measure": [DP, C.c_uint32, C.c_int, DP, U64P, C.c_uint32, C.c_double, DP, U64P, ERR],
            "or_measure_n": [DP, C.c_uint32, C.c_int, DP, U64P, C.c_uint32, DP, C.c_size_t, U64P, ERR],
            "or_measure_collapse": [DP, DP, C.c_uint32, U64P, C.c_uint32, C.c_uint64, ERR],
            "or_sample_outcome": [DP, C.c_uint64, C.c_double],
            "rf_apply_gates": [C.POINTER(DP), C.c_uint32, P, C.c_size_t, C.c_int, ERR],
            "or_fs_metric": [DP, C.c_uint32, DP, C.c_uint32, C.c_int, DP, ERR],
            "or_state_add": [DP, DP, DP, C.c_uint32, C.c_int],
            "or_state_scale": [DP, C.c_uint32, C.c_double, C.c_double],
            "or_sample_cdf": [DP, C.c_uint64, DP],
        }
        for name, args in sig.items():
            f = getattr(_lib, name)
            f.argtypes = args
            f.restype = C.c_int
        _lib.or_norm_sqr.restype = C.c_double
        _lib.or_inner_product.restype = None
        _lib.or_sample_outcome.restype = C.c_uint64
        for name in ("or_state_add", "or_state_scale", "or_sample_cdf"):
            getattr(_lib, name).restype = None
    return _lib


def _dp(a):
    return a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))


def _err():
    return _qiron_lib().QiError()


def _check(rc, e):
    if rc != 0:
        raise OracleError(rc, e.a, e.b, e.re, e.im, e.msg.decode(errors="replace"))


def _vec(psi):
    return np.ascontiguousarray(np.array(psi, dtype=np.complex128))


def n_of(psi) -> int:
    return int(psi.size).bit_length() - 1


def init(n: int, kind: int, basis: int = 0) -> np.ndarray:
    psi = np.zeros(1 << n, dtype=np.complex128)
    e = _err()
    _check(lib().or_state_init(_dp(psi), n, kind, basis, C.byref(e)), e)
    return psi


def validate_gate(n: int, gate) -> None:
    e = _err()
    _check(lib().or_validate_gate(n, C.byref(gate), C.byref(e)), e)


def apply_gate(psi, gate) -> np.ndarray:
    out = _vec(psi)
    e = _err()
    _check(lib().or_apply_gate(_dp(out), n_of(out), C.byref(gate), C.byref(e)), e)
    return out


def apply_gates(psi, gates: np.ndarray, n_circuit=None) -> np.ndarray:
    out = _vec(psi)
    n = n_of(out)
    e = _err()
    _check(lib().or_apply_gates(_dp(out), n, n if n_circuit is None else n_circuit,
                                gates.ctypes.data_as(C.c_void_p), gates.size, C.byref(e)), e)
    return out


def reference_structured_apply(psi, gates: np.ndarray, threads: int) -> np.ndarray:
    """The rayon-structured CPU path (clone per gate, update list, serial scatter)."""
    n = n_of(psi)
    buf = (C.c_double * (2 * psi.size)).from_buffer_copy(_vec(psi).tobytes())
    # the C side frees and replaces the buffer per gate: hand it a malloc'd copy
    libc = C.CDLL(None)
    libc.malloc.restype = C.c_void_p
    libc.malloc.argtypes = [C.c_size_t]
    libc.free.argtypes = [C.c_void_p]
    p = libc.malloc(psi.size * 16)
    C.memmove(p, buf, psi.size * 16)
    pp = C.cast(p, C.POINTER(C.c_double))
    e = _err()
    rc = lib().rf_apply_gates(C.byref(pp), n, gates.ctypes.data_as(C.c_void_p), gates.size, threads, C.byref(e))
    out = np.ctypeslib.as_array(pp, shape=(2 * psi.size,)).copy().view(np.complex128)
    libc.free(C.cast(pp, C.c_void_p))
    _check(rc, e)
    return out


def reference_structured_timed(psi, gates: np.ndarray, threads: int, keep: bool = False):
    """reference_structured_apply with only the C call timed (the host-side
    buffer copies of this wrapper stay outside): (result or None, seconds)."""
    n = n_of(psi)
    src = _vec(psi)
    libc = C.CDLL(None)
    libc.malloc.restype = C.c_void_p
    libc.malloc.argtypes = [C.c_size_t]
    libc.free.argtypes = [C.c_void_p]
    p = libc.malloc(src.size * 16)
    if not p:
        raise MemoryError("host buffer for the reference-structured path")
    C.memmove(p, src.ctypes.data, src.size * 16)
    pp = C.cast(p, C.POINTER(C.c_double))
    e = _err()
    t0 = time.perf_counter()
    rc = lib().rf_apply_gates(C.byref(pp), n, gates.ctypes.data_as(C.c_void_p), gates.size, threads, C.byref(e))
    dt = time.perf_counter() - t0
    out = np.ctypeslib.as_array(pp, shape=(2 * src.size,)).copy().view(np.complex128) if keep and rc == 0 else None
    libc.free(C.cast(pp, C.c_void_p))
    _check(rc, e)
    return out, dt


def _term(t):
    L = _qiron_lib()
    return L.QiPauliTerm(int(t["x_mask"]), int(t["z_mask"]), float(t["c_re"]), float(t["c_im"])) \
        if not isinstance(t, L.QiPauliTerm) else t


def pauli_apply(psi, term, normalised=False) -> np.ndarray:
    src = _vec(psi)
    out = np.empty_like(src)
    t = _term(term)
    e = _err()
    f = lib().or_pauli_apply_normalised if normalised else lib().or_pauli_apply
    _check(f(_dp(src), _dp(out), n_of(src), C.byref(t), C.byref(e)), e)
    return out


def pauli_exp(psi, term, factor=1.0) -> np.ndarray:
    out = _vec(psi)
    t = _term(term)
    f = complex(factor)
    e = _err()
    _check(lib().or_pauli_exp(_dp(out), n_of(out), C.byref(t), f.real, f.imag, C.byref(e)), e)
    return out


def pauli_exp_neg_i_dt(psi, term, dt) -> np.ndarray:
    out = _vec(psi)
    t = _term(term)
    e = _err()
    _check(lib().or_pauli_exp_neg_i_dt(_dp(out), n_of(out), C.byref(t), float(dt), C.byref(e)), e)
    return out


def sumop_apply(psi, terms: np.ndarray) -> np.ndarray:
    src = _vec(psi)
    out = np.empty_like(src)
    e = _err()
    _check(lib().or_sumop_apply(_dp(src), _dp(out), n_of(src), terms.ctypes.data_as(C.c_void_p), terms.size,
                                C.byref(e)), e)
    return out


def expectation(psi, terms: np.ndarray) -> complex:
    src = _vec(psi)
    re, im = C.c_double(), C.c_double()
    e = _err()
    _check(lib().or_sumop_expectation(_dp(src), n_of(src), terms.ctypes.data_as(C.c_void_p), terms.size,
                                      C.byref(re), C.byref(im), C.byref(e)), e)
    return complex(re.value, im.value)


def trotter(psi, terms: np.ndarray, dt: float, steps: int, order: int = 1) -> np.ndarray:
    out = _vec(psi)
    e = _err()
    _check(lib().or_trotter_evolve(_dp(out), n_of(out), terms.ctypes.data_as(C.c_void_p), terms.size, float(dt),
                                   int(steps), int(order), C.byref(e)), e)
    return out


def inner_product(a, b) -> complex:
    a, b = _vec(a), _vec(b)
    re, im = C.c_double(), C.c_double()
    lib().or_inner_product(_dp(a), _dp(b), n_of(a), C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def norm_sqr(a) -> float:
    a = _vec(a)
    return lib().or_norm_sqr(_dp(a), n_of(a))


def normalise(a) -> np.ndarray:
    out = _vec(a)
    e = _err()
    _check(lib().or_normalise(_dp(out), n_of(out), C.byref(e)), e)
    return out


def fidelity(a, b) -> float:
    a, b = _vec(a), _vec(b)
    v = C.c_double()
    e = _err()
    _check(lib().or_fidelity(_dp(a), _dp(b), n_of(a), C.byref(v), C.byref(e)), e)
    return v.value


def fs_fidelity(a, b) -> float:
    """fs_fidelity of states of any sizes, with the reference's error order."""
    return _fs_metric(a, b, 0)


def fs_dist(a, b) -> float:
    return _fs_metric(a, b, 1)


def _fs_metric(a, b, dist):
    a, b = _vec(a), _vec(b)
    v = C.c_double()
    e = _err()
    _check(lib().or_fs_metric(_dp(a), n_of(a), _dp(b), n_of(b), dist, C.byref(v), C.byref(e)), e)
    return v.value


def state_add(a, b) -> np.ndarray:
    a, b = _vec(a), _vec(b)
    out = np.empty_like(a)
    lib().or_state_add(_dp(a), _dp(b), _dp(out), n_of(a), 0)
    return out


def state_sub(a, b) -> np.ndarray:
    a, b = _vec(a), _vec(b)
    out = np.empty_like(a)
    lib().or_state_add(_dp(a), _dp(b), _dp(out), n_of(a), 1)
    return out


def state_scale(a, c) -> np.ndarray:
    out = _vec(a)
    c = complex(c)
    lib().or_state_scale(_dp(out), n_of(out), c.real, c.imag)
    return out


def sample_cdf(probs) -> np.ndarray:
    p = np.ascontiguousarray(np.asarray(probs, dtype=np.float64))
    out = np.empty_like(p)
    lib().or_sample_cdf(p.ctypes.data_as(C.POINTER(C.c_double)), p.size, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def check_norm(a) -> None:
    a = _vec(a)
    e = _err()
    _check(lib().or_state_check_norm(_dp(a), n_of(a), C.byref(e)), e)


def _mat(custom):
    if custom is None:
        return None
    m = np.asarray(custom, dtype=np.complex128).reshape(-1)
    out = np.empty(8)
    out[0::2], out[1::2] = m.real, m.imag
    return out


def measure_probs(psi, qubits) -> np.ndarray:
    src = _vec(psi)
    q = np.asarray(list(qubits), dtype=np.uint64)
    m = q.size if q.size else n_of(src)
    out = np.empty(1 << m)
    e = _err()
    _check(lib().or_measure_probs(_dp(src), n_of(src), q.ctypes.data_as(C.POINTER(C.c_uint64)) if q.size else None,
                                  q.size, out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(e)), e)
    return out


def measure(psi, basis: int, qubits, uniform: float, custom=None):
    src = _vec(psi)
    out = np.empty_like(src)
    q = np.asarray(list(qubits), dtype=np.uint64)
    o = C.c_uint64()
    m = _mat(custom)
    e = _err()
    _check(lib().or_measure(_dp(src), n_of(src), basis, m.ctypes.data_as(C.POINTER(C.c_double)) if m is not None else None,
                            q.ctypes.data_as(C.POINTER(C.c_uint64)) if q.size else None, q.size, float(uniform),
                            _dp(out), C.byref(o), C.byref(e)), e)
    return o.value, out


def measure_n(psi, basis: int, qubits, uniforms, custom=None) -> np.ndarray:
    src = _vec(psi)
    q = np.asarray(list(qubits), dtype=np.uint64)
    u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64))
    outs = np.zeros(max(u.size, 1), dtype=np.uint64)
    m = _mat(custom)
    e = _err()
    _check(lib().or_measure_n(_dp(src), n_of(src), basis,
                              m.ctypes.data_as(C.POINTER(C.c_double)) if m is not None else None,
                              q.ctypes.data_as(C.POINTER(C.c_uint64)) if q.size else None, q.size,
                              u.ctypes.data_as(C.POINTER(C.c_double)), u.size,
                              outs.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(e)), e)
    return outs[: u.size]


def collapse(psi, qubits, outcome: int) -> np.ndarray:
    src = _vec(psi)
    out = np.empty_like(src)
    q = np.asarray(list(qubits), dtype=np.uint64)
    e = _err()
    _check(lib().or_measure_collapse(_dp(src), _dp(out), n_of(src), q.ctypes.data_as(C.POINTER(C.c_uint64)) if q.size else None,
                                     q.size, int(outcome), C.byref(e)), e)
    return out


def sample_outcome(probs, u: float) -> int:
    p = np.ascontiguousarray(np.asarray(probs, dtype=np.float64))
    return int(lib().or_sample_outcome(p.ctypes.data_as(C.POINTER(C.c_double)), p.size, float(u)))
