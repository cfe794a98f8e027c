"""Host-side mirror of quant-iron's state-vector API over the C ABI.

Names, argument order and error behaviour follow the reference:
  State               src/components/state.rs (constructors :99-373, gates :1019-2345,
                      measure :525-784, inner_product :890-917, normalise :924-945)
  Gate                src/components/gate.rs (:13-122 and the constructor helpers)
  Circuit / Builder   src/circuit.rs (:34-202, builder :288-1742)
  Subroutine          src/subroutine.rs (:90-160)
  PauliString, SumOp  src/components/pauli_string.rs (:13-507)
  MeasurementBasis    src/components/measurement.rs (:15-90)
Every operation returns a new State (value semantics, as in Rust); the
`*_inplace` methods expose the engine's in-place form. All arithmetic runs
in the HIP engine (libqiron.so); nothing here computes amplitudes.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Dict, Iterable, List, Optional, Sequence

import numpy as np

from . import _lib as L


class QuantIronError(Exception):
    """An `Error` variant of the reference (errors.rs:4-97) with its payload."""

    def __init__(self, code: int, a: int = 0, b: int = 0, re: float = 0.0, im: float = 0.0, msg: str = ""):
        self.code = code
        self.variant = L.STATUS_NAMES.get(code, f"code{code}")
        self.a, self.b = a, b
        self.coefficient = complex(re, im)
        super().__init__(f"{self.variant}: {msg}")

    @property
    def payload(self):
        v = self.variant
        if v in ("InvalidQubitIndex", "OverlappingControlAndTargetQubits", "InvalidNumberOfInputs"):
            return (self.a, self.b)
        if v in ("InvalidNumberOfQubits", "InvalidNumberOfMeasurements", "InvalidInputValue"):
            return (self.a,)
        if v == "InvalidPauliStringCoefficient":
            return (self.coefficient,)
        return ()


def _check(rc: int, err: L.QiError):
    if rc != 0:
        raise QuantIronError(rc, err.a, err.b, err.re, err.im, err.msg.decode(errors="replace"))


# ------------------------------------------------------------------ context --
class Context:
    """One device context (stream, scratch buffers, mutex)."""

    def __init__(self, device: int = 0):
        self.lib = L.load()
        self._h = C.c_void_p()
        err = L.QiError()
        _check(self.lib.qi_ctx_create(device, C.byref(self._h), C.byref(err)), err)

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        err = L.QiError()
        _check(self.lib.qi_ctx_synchronize(self._h, C.byref(err)), err)

    def seed(self, seed: int):
        """Seeds the stream that measurements without injected uniforms draw
        from (the reference's rand::rng(), state.rs:601-602)."""
        err = L.QiError()
        _check(self.lib.qi_ctx_seed(self._h, int(seed) & 0xFFFFFFFFFFFFFFFF, C.byref(err)), err)

    def info(self):
        name = C.create_string_buffer(256)
        hbm = C.c_uint64()
        cus = C.c_int()
        err = L.QiError()
        _check(self.lib.qi_ctx_info(self._h, name, 256, C.byref(hbm), C.byref(cus), C.byref(err)), err)
        return {"name": name.value.decode(), "hbm_bytes": hbm.value, "cu_count": cus.value}

    def last_exec_ms(self) -> float:
        v = C.c_double()
        err = L.QiError()
        _check(self.lib.qi_last_exec_ms(self._h, C.byref(v), C.byref(err)), err)
        return v.value

    def last_exec_steps(self) -> int:
        v = C.c_uint64()
        err = L.QiError()
        _check(self.lib.qi_last_exec_steps(self._h, C.byref(v), C.byref(err)), err)
        return v.value

    def close(self):
        if self._h:
            self.lib.qi_ctx_destroy(self._h)
            self._h = C.c_void_p()


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


PRECISIONS = {"c64": L.C64, "c128": L.C128}


# ------------------------------------------------------------------- gates --
def _gate(kind: str, targets: Sequence[int] = (), controls: Sequence[int] = (), params: Sequence[float] = (),
          matrix=None, basis: int = 0, x_mask: int = 0, z_mask: int = 0) -> L.QiGate:
    g = L.QiGate()
    g.kind = L.GATE_KINDS[kind]
    g.ntargets = len(targets)
    for i, t in enumerate(list(targets)[: L.MAX_TARGETS]):
        g.targets[i] = int(t)
    if len(controls) > L.MAX_CONTROLS:
        raise QuantIronError(17, msg="too many controls")
    g.ncontrols = len(controls)
    for i, c in enumerate(controls):
        g.controls[i] = int(c)
    for i, p in enumerate(params):
        g.params[i] = float(p)
    if matrix is not None:
        m = np.asarray(matrix, dtype=np.complex128).reshape(2, 2)
        for i, v in enumerate(m.reshape(-1)):
            g.matrix[2 * i] = v.real
            g.matrix[2 * i + 1] = v.imag
    g.basis = basis
    g.x_mask, g.z_mask = x_mask, z_mask
    return g


class MeasurementBasis:
    """MeasurementBasis (measurement.rs:76-86)."""

    Computational = ("Computational", None)
    X = ("X", None)
    Y = ("Y", None)

    @staticmethod
    def Custom(matrix):
        return ("Custom", np.asarray(matrix, dtype=np.complex128).reshape(2, 2))


def _basis_code(basis):
    name, mat = basis
    return {"Computational": L.BASIS_COMPUTATIONAL, "X": L.BASIS_X, "Y": L.BASIS_Y, "Custom": L.BASIS_CUSTOM}[name], mat


def _mat8(mat):
    if mat is None:
        return None
    m = np.asarray(mat, dtype=np.complex128).reshape(-1)
    out = np.empty(8, dtype=np.float64)
    out[0::2] = m.real
    out[1::2] = m.imag
    return out


class Gate:
    """`Gate` (gate.rs:13-52); wraps one qi_gate record."""

    def __init__(self, raw: L.QiGate, label: str = ""):
        self.raw = raw
        self.label = label

    def apply(self, state: "State") -> "State":
        """Gate::apply (gate.rs:99-122): returns a new state."""
        out = state.clone()
        out.apply_gate_inplace(self)
        return out

    # constructor helpers of gate.rs (host-side only)
    @staticmethod
    def h_gate(q): return Gate(_gate("H", [q]), "H")
    @staticmethod
    def x_gate(q): return Gate(_gate("X", [q]), "X")
    @staticmethod
    def y_gate(q): return Gate(_gate("Y", [q]), "Y")
    @staticmethod
    def z_gate(q): return Gate(_gate("Z", [q]), "Z")
    @staticmethod
    def i_gate(q): return Gate(_gate("I", [q]), "I")
    @staticmethod
    def s_gate(q): return Gate(_gate("S", [q]), "S")
    @staticmethod
    def t_gate(q): return Gate(_gate("T", [q]), "T")
    @staticmethod
    def s_dag_gate(q): return Gate(_gate("SDAG", [q]), "Sdag")
    @staticmethod
    def t_dag_gate(q): return Gate(_gate("TDAG", [q]), "Tdag")
    @staticmethod
    def p_gate(q, angle): return Gate(_gate("P", [q], params=[angle]), "P")
    @staticmethod
    def rx_gate(q, angle): return Gate(_gate("RX", [q], params=[angle]), "RX")
    @staticmethod
    def ry_gate(q, angle): return Gate(_gate("RY", [q], params=[angle]), "RY")
    @staticmethod
    def rz_gate(q, angle): return Gate(_gate("RZ", [q], params=[angle]), "RZ")
    @staticmethod
    def unitary2_gate(q, matrix): return Gate(_gate("U2", [q], matrix=matrix), "U2")
    @staticmethod
    def cnot_gate(target, control):  # gate.rs: target first (circuit.rs:1071-1075)
        return Gate(_gate("CNOT", [target], [control]), "CNOT")
    @staticmethod
    def swap_gate(q1, q2): return Gate(_gate("SWAP", [q1, q2]), "SWAP")
    @staticmethod
    def toffoli_gate(target, controls): return Gate(_gate("TOFFOLI", [target], list(controls)), "Toffoli")
    @staticmethod
    def matchgate(target, theta, phi1, phi2, controls=()):
        return Gate(_gate("MATCHGATE", [target], list(controls), [theta, phi1, phi2]), "Matchgate")
    @staticmethod
    def operator(kind: str, targets, controls=(), params=(), matrix=None):
        """Gate::Operator(op, targets, controls) for any built-in operator kind."""
        return Gate(_gate(kind, list(targets), list(controls), list(params), matrix), kind)
    @staticmethod
    def controlled(kind: str, targets, controls, params=(), matrix=None):
        """The *_controlled_gates helpers: one gate per target with shared controls."""
        return [Gate(_gate(kind, [t], list(controls), list(params), matrix), kind) for t in targets]
    @staticmethod
    def p_controlled_gates(targets, controls, angle):
        return Gate.controlled("P", targets, controls, [angle])
    @staticmethod
    def new_measurement(qubits, basis=MeasurementBasis.Computational, uniform: Optional[float] = None):
        """Gate::Measurement. `uniform` injects the draw of rand::rng(); None draws
        from the context's stream each time the gate runs (Context.seed)."""
        code, mat = _basis_code(basis)
        uniform = math.nan if uniform is None else uniform
        return Gate(_gate("MEASURE", [], list(qubits), [uniform], mat, basis=code), "Measure")
    @staticmethod
    def pauli_string_gate(ps: "PauliString"):
        x, z = ps.masks()
        c = ps.coefficient
        return Gate(_gate("PAULI_STRING", params=[c.real, c.imag], x_mask=x, z_mask=z), "PauliString")
    @staticmethod
    def pauli_time_evolution_gate(ps: "PauliString", time: float):
        x, z = ps.masks()
        c = ps.coefficient
        return Gate(_gate("PAULI_EVOLUTION", params=[c.real, c.imag, time], x_mask=x, z_mask=z), "PauliEvolution")


# ------------------------------------------------------------- measurement --
@dataclass
class MeasurementResult:
    """MeasurementResult (measurement.rs:15-72). For measure_n results the
    collapsed state is computed by the engine on first access and shared by
    the shots with the same outcome (the reference builds one per shot)."""

    basis: tuple
    indices: List[int]
    outcomes: List[int]
    _state: Optional["State"] = None
    _parent: Optional["State"] = field(default=None, repr=False)
    _cache: Optional[dict] = field(default=None, repr=False)

    @property
    def new_state(self) -> "State":
        if self._state is None:
            o = sum(int(b) << k for k, b in enumerate(self.outcomes))
            if self._cache is not None and o in self._cache:
                self._state = self._cache[o]
            else:
                self._state = self._parent._measured_state(self.basis, self.indices, o)
                if self._cache is not None:
                    self._cache[o] = self._state
        return self._state

    def get_new_state(self) -> "State":
        return self.new_state


# ------------------------------------------------------------------- state --
class State:
    """`State` (state.rs:74-81), device-resident."""

    def __init__(self, handle: C.c_void_p, ctx: Context):
        self._h = handle
        self.ctx = ctx
        self.lib = ctx.lib

    def __del__(self):
        try:
            if self._h:
                self.lib.qi_state_destroy(self._h)
        except Exception:
            pass

    # ---- construction
    @classmethod
    def _alloc(cls, n: int, precision: str = "c128", ctx: Optional[Context] = None) -> "State":
        ctx = ctx or default_context()
        h = C.c_void_p()
        err = L.QiError()
        _check(ctx.lib.qi_state_create(ctx.handle, int(n), PRECISIONS[precision], C.byref(h), C.byref(err)), err)
        return cls(h, ctx)

    @classmethod
    def new(cls, state_vector, precision: str = "c128", ctx=None) -> "State":
        """State::new (state.rs:99-127): length must be a power of two and the norm 1."""
        v = np.ascontiguousarray(np.asarray(state_vector, dtype=np.complex128))
        n_len = v.size
        if n_len == 0:
            raise QuantIronError(3, 0, msg="Invalid number of qubits: 0")
        if n_len & (n_len - 1):
            k = int(math.floor(math.log2(n_len)))
            raise QuantIronError(3, k, msg=f"Invalid number of qubits: {k}")
        n = n_len.bit_length() - 1
        s = cls._alloc(n, precision, ctx)
        err = L.QiError()
        _check(s.lib.qi_state_upload(s._h, v.view(np.float64).ctypes.data_as(L.DP), 1, C.byref(err)), err)
        return s

    @classmethod
    def from_vector_unchecked(cls, state_vector, precision="c128", ctx=None) -> "State":
        v = np.ascontiguousarray(np.asarray(state_vector, dtype=np.complex128))
        s = cls._alloc(v.size.bit_length() - 1, precision, ctx)
        err = L.QiError()
        _check(s.lib.qi_state_upload(s._h, v.view(np.float64).ctypes.data_as(L.DP), 0, C.byref(err)), err)
        return s

    @classmethod
    def _init(cls, n, kind, basis=0, precision="c128", ctx=None):
        if n == 0:
            raise QuantIronError(3, 0, msg="Invalid number of qubits: 0")
        s = cls._alloc(n, precision, ctx)
        err = L.QiError()
        _check(s.lib.qi_state_init(s._h, kind, int(basis), C.byref(err)), err)
        return s

    @classmethod
    def new_zero(cls, n, precision="c128", ctx=None): return cls._init(n, L.INIT_ZERO, 0, precision, ctx)
    @classmethod
    def new_basis_n(cls, n, idx, precision="c128", ctx=None):
        if idx >= (1 << n):  # state.rs:196-198 (checked before the qubit count)
            raise QuantIronError(4, idx, n, msg=f"Invalid qubit index: {idx} for {n} qubits")
        return cls._init(n, L.INIT_BASIS, idx, precision, ctx)
    @classmethod
    def new_plus(cls, n, precision="c128", ctx=None): return cls._init(n, L.INIT_PLUS, 0, precision, ctx)
    @classmethod
    def new_minus(cls, n, precision="c128", ctx=None): return cls._init(n, L.INIT_MINUS, 0, precision, ctx)
    @classmethod
    def new_ghz(cls, n, precision="c128", ctx=None): return cls._init(n, L.INIT_GHZ, 0, precision, ctx)

    # ---- accessors
    @property
    def num_qubits(self) -> int:
        return int(self.lib.qi_state_num_qubits(self._h))

    @property
    def precision(self) -> str:
        return "c128" if self.lib.qi_state_precision(self._h) == L.C128 else "c64"

    @property
    def state_vector(self) -> np.ndarray:
        out = np.empty(1 << self.num_qubits, dtype=np.complex128)
        err = L.QiError()
        _check(self.lib.qi_state_download(self._h, out.view(np.float64).ctypes.data_as(L.DP), C.byref(err)), err)
        return out

    def amplitude(self, n: int) -> complex:
        if n >= (1 << self.num_qubits):
            raise QuantIronError(4, n, self.num_qubits)
        return complex(self.state_vector[n])

    def clone(self) -> "State":
        h = C.c_void_p()
        err = L.QiError()
        _check(self.lib.qi_state_clone(self._h, C.byref(h), C.byref(err)), err)
        return State(h, self.ctx)

    def copy_from(self, src: "State") -> "State":
        """Overwrite this state's amplitudes with src's (same qubit count and
        precision), device to device."""
        err = L.QiError()
        _check(self.lib.qi_state_copy(self._h, src._h, C.byref(err)), err)
        return self

    # ---- in-place engine entry points
    def apply_gate_inplace(self, gate: Gate) -> "State":
        err = L.QiError()
        _check(self.lib.qi_apply_gate(self._h, C.byref(gate.raw), C.byref(err)), err)
        return self

    def apply_gates_inplace(self, gates, n_circuit: Optional[int] = None, mode: int = L.EXEC_AUTO,
                            tile_qubits: int = 0, block_qubits: int = 0, use_graph: bool = False,
                            tile_threads_log2: int = 0) -> "State":
        arr = gates if isinstance(gates, np.ndarray) else gate_array(gates)
        opts = L.QiExecOpts(mode, tile_qubits, block_qubits, 1 if use_graph else 0, tile_threads_log2)
        err = L.QiError()
        n_c = self.num_qubits if n_circuit is None else n_circuit
        _check(self.lib.qi_apply_gates(self._h, n_c, arr.ctypes.data_as(C.c_void_p), arr.size, C.byref(opts),
                                       C.byref(err)), err)
        return self

    def _op(self, kind, targets, controls=(), params=(), matrix=None) -> "State":
        return Gate(_gate(kind, targets, controls, params, matrix)).apply(self)

    def _multi(self, kind, qubits, controls=(), params=(), matrix=None) -> "State":
        # the *_multi loops of state.rs apply one gate per listed qubit, in order
        out = self.clone()
        for q in qubits:
            out.apply_gate_inplace(Gate(_gate(kind, [q], controls, params, matrix)))
        return out

    # ---- single-qubit gates (state.rs:1019-2193)
    def h(self, i): return self._op("H", [i])
    def h_multi(self, qs): return self._multi("H", qs)
    def ch_multi(self, ts, cs): return self._multi("H", ts, cs)
    def x(self, i): return self._op("X", [i])
    def x_multi(self, qs): return self._multi("X", qs)
    def cx_multi(self, ts, cs): return self._multi("X", ts, cs)
    def y(self, i): return self._op("Y", [i])
    def y_multi(self, qs): return self._multi("Y", qs)
    def cy_multi(self, ts, cs): return self._multi("Y", ts, cs)
    def z(self, i): return self._op("Z", [i])
    def z_multi(self, qs): return self._multi("Z", qs)
    def cz_multi(self, ts, cs): return self._multi("Z", ts, cs)
    def i(self, q): return self._op("I", [q])
    def i_multi(self, qs): return self._multi("I", qs)
    def ci_multi(self, ts, cs): return self._multi("I", ts, cs)
    def s(self, i): return self._op("S", [i])
    def s_multi(self, qs): return self._multi("S", qs)
    def cs_multi(self, ts, cs): return self._multi("S", ts, cs)
    def t(self, i): return self._op("T", [i])
    def t_multi(self, qs): return self._multi("T", qs)
    def ct_multi(self, ts, cs): return self._multi("T", ts, cs)
    def s_dag(self, i): return self._op("SDAG", [i])
    def s_dag_multi(self, qs): return self._multi("SDAG", qs)
    def cs_dag_multi(self, ts, cs): return self._multi("SDAG", ts, cs)
    def t_dag(self, i): return self._op("TDAG", [i])
    def t_dag_multi(self, qs): return self._multi("TDAG", qs)
    def ct_dag_multi(self, ts, cs): return self._multi("TDAG", ts, cs)
    def p(self, i, angle): return self._op("P", [i], params=[angle])
    def p_multi(self, qs, angle): return self._multi("P", qs, params=[angle])
    def cp_multi(self, ts, cs, angle): return self._multi("P", ts, cs, [angle])
    def rx(self, i, angle): return self._op("RX", [i], params=[angle])
    def rx_multi(self, qs, angle): return self._multi("RX", qs, params=[angle])
    def crx_multi(self, ts, cs, angle): return self._multi("RX", ts, cs, [angle])
    def ry(self, i, angle): return self._op("RY", [i], params=[angle])
    def ry_multi(self, qs, angle): return self._multi("RY", qs, params=[angle])
    def cry_multi(self, ts, cs, angle): return self._multi("RY", ts, cs, [angle])
    def rz(self, i, angle): return self._op("RZ", [i], params=[angle])
    def rz_multi(self, qs, angle): return self._multi("RZ", qs, params=[angle])
    def crz_multi(self, ts, cs, angle): return self._multi("RZ", ts, cs, [angle])
    def unitary(self, i, u): return self._op("U2", [i], matrix=u)
    def unitary_multi(self, qs, u): return self._multi("U2", qs, matrix=u)
    def cunitary_multi(self, ts, cs, u): return self._multi("U2", ts, cs, matrix=u)
    def ry_phase(self, i, angle, phase): return self._op("U2", [i], matrix=ry_phase_matrix(angle, phase))
    def ry_phase_multi(self, qs, angle, phase): return self._multi("U2", qs, matrix=ry_phase_matrix(angle, phase))
    def cry_phase_gates(self, ts, cs, angle, phase): return self._multi("U2", ts, cs, matrix=ry_phase_matrix(angle, phase))
    def ry_phase_dag(self, i, angle, phase): return self._op("U2", [i], matrix=ry_phase_dag_matrix(angle, phase))
    def ry_phase_dag_multi(self, qs, angle, phase): return self._multi("U2", qs, matrix=ry_phase_dag_matrix(angle, phase))
    def cry_phase_dag_gates(self, ts, cs, angle, phase):
        return self._multi("U2", ts, cs, matrix=ry_phase_dag_matrix(angle, phase))

    # ---- multi-qubit gates (state.rs:2230-2345)
    def cnot(self, control, target): return self._op("CNOT", [target], [control])
    def swap(self, q1, q2): return self._op("SWAP", [q1, q2])
    def cswap(self, t1, t2, controls): return self._op("SWAP", [t1, t2], list(controls))
    def matchgate(self, target, theta, phi1, phi2): return self._op("MATCHGATE", [target], [], [theta, phi1, phi2])
    def cmatchgate(self, target, theta, phi1, phi2, controls):
        return self._op("MATCHGATE", [target], list(controls), [theta, phi1, phi2])
    def toffoli(self, c1, c2, target): return self._op("TOFFOLI", [target], [c1, c2])

    def operate(self, kind: str, targets, controls=(), params=(), matrix=None) -> "State":
        """State::operate with a built-in operator kind."""
        return self._op(kind, list(targets), list(controls), params, matrix)

    # ---- reductions and arithmetic
    def inner_product(self, other: "State") -> complex:
        re, im = C.c_double(), C.c_double()
        err = L.QiError()
        _check(self.lib.qi_inner_product(self._h, other._h, C.byref(re), C.byref(im), C.byref(err)), err)
        return complex(re.value, im.value)

    def norm_sqr(self) -> float:
        v = C.c_double()
        err = L.QiError()
        _check(self.lib.qi_norm_sqr(self._h, C.byref(v), C.byref(err)), err)
        return v.value

    def normalise(self) -> "State":
        out = self.clone()
        err = L.QiError()
        _check(self.lib.qi_normalise(out._h, C.byref(err)), err)
        return out

    def fs_fidelity(self, other: "State") -> float:
        v = C.c_double()
        err = L.QiError()
        _check(self.lib.qi_fidelity(self._h, other._h, C.byref(v), C.byref(err)), err)
        return v.value

    def fs_dist(self, other: "State") -> float:
        """fs_dist (state.rs:470-478): acos(|<a|b>|) of the normalised states."""
        v = C.c_double()
        err = L.QiError()
        _check(self.lib.qi_fs_dist(self._h, other._h, C.byref(v), C.byref(err)), err)
        return v.value

    def __mul__(self, c) -> "State":
        """impl Mul<Complex<f64>> / Mul<f64> for State (state.rs:2687-2771)."""
        out = self.clone()
        c = complex(c)
        err = L.QiError()
        _check(self.lib.qi_state_scale(out._h, c.real, c.imag, C.byref(err)), err)
        return out

    __rmul__ = __mul__

    def _addsub(self, other: "State", fn, verb: str) -> "State":
        if self.num_qubits != other.num_qubits:  # the reference panics with this message
            raise ValueError(f"Cannot {verb} states with different numbers of qubits: "
                             f"{self.num_qubits} != {other.num_qubits}")
        out = State._alloc(self.num_qubits, self.precision, self.ctx)
        err = L.QiError()
        _check(fn(self._h, other._h, out._h, C.byref(err)), err)
        return out

    def __add__(self, other: "State") -> "State":
        """impl Add for State (state.rs:2773-2800)."""
        return self._addsub(other, self.lib.qi_state_add, "add")

    def __sub__(self, other: "State") -> "State":
        """impl Sub for State (state.rs:2836-2862)."""
        return self._addsub(other, self.lib.qi_state_sub, "subtract")

    @staticmethod
    def sum(states: Iterable["State"]) -> "State":
        """impl Sum for State (state.rs:2802-2834): a left fold of Add."""
        it = iter(states)
        try:
            acc = next(it)
        except StopIteration:
            raise ValueError("Cannot sum an empty iterator of State objects: num_qubits for the sum is undefined.")
        for st in it:
            if acc.num_qubits != st.num_qubits:
                raise ValueError(f"Cannot sum states with different numbers of qubits: "
                                 f"{acc.num_qubits} != {st.num_qubits}")
            acc = acc + st
        return acc

    def __eq__(self, other) -> bool:
        """PartialEq (state.rs:2348-2372): |dre|, |dim| <= f32::EPSILON per amplitude."""
        if not isinstance(other, State) or self.num_qubits != other.num_qubits:
            return False
        a, b = self.state_vector, other.state_vector
        eps = float(np.finfo(np.float32).eps)
        return bool(np.all(np.abs(a.real - b.real) <= eps) and np.all(np.abs(a.imag - b.imag) <= eps))

    __hash__ = None

    # ---- measurement (state.rs:525-784): uniforms injected, or drawn by the
    # engine from the context's stream (Context.seed) when not given
    def measure(self, basis=MeasurementBasis.Computational, measured_qubits: Sequence[int] = (),
                uniform: Optional[float] = None) -> MeasurementResult:
        code, mat = _basis_code(basis)
        uniform = math.nan if uniform is None else uniform
        qs = np.asarray(list(measured_qubits), dtype=np.uint64)
        out = State._alloc(self.num_qubits, self.precision, self.ctx)
        o = C.c_uint64()
        m8 = _mat8(mat)
        err = L.QiError()
        _check(self.lib.qi_measure(self._h, code, L.dptr(m8) if m8 is not None else None,
                                   L.u64ptr(qs) if qs.size else None, qs.size, float(uniform), out._h, C.byref(o),
                                   C.byref(err)), err)
        idx = list(measured_qubits) if len(measured_qubits) else list(range(self.num_qubits))
        bits = [(o.value >> b) & 1 for b in range(len(idx))]
        return MeasurementResult(basis, idx, bits, out, self)

    def measure_n(self, basis=MeasurementBasis.Computational, measured_qubits: Sequence[int] = (), n: int = 1,
                  uniforms: Optional[Sequence[float]] = None) -> List[MeasurementResult]:
        """State::measure_n (state.rs:750-784): n shots, each with its outcome
        bits and collapsed state. `uniforms` injects the n draws; None draws
        them from the context's stream."""
        code, mat = _basis_code(basis)
        qs = np.asarray(list(measured_qubits), dtype=np.uint64)
        u = None
        if uniforms is not None:
            u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64))
            if u.size < n:
                raise QuantIronError(17, msg=f"{u.size} uniforms for {n} shots")
        outs = np.zeros(max(int(n), 1), dtype=np.uint64)
        m8 = _mat8(mat)
        err = L.QiError()
        _check(self.lib.qi_measure_n(self._h, code, L.dptr(m8) if m8 is not None else None,
                                     L.u64ptr(qs) if qs.size else None, qs.size,
                                     L.dptr(u) if u is not None else None, int(n), L.u64ptr(outs), C.byref(err)), err)
        idx = list(measured_qubits) if len(measured_qubits) else list(range(self.num_qubits))
        src, cache = self.clone(), {}  # the shots' states derive from this state as it is now
        return [MeasurementResult(basis, idx, [(int(o) >> b) & 1 for b in range(len(idx))], None, src, cache)
                for o in outs[:n]]

    def measure_outcomes(self, measured_qubits: Sequence[int], uniforms, basis=MeasurementBasis.Computational) -> np.ndarray:
        """measure_n outcomes only, as packed integers (bit b = measured_qubits[b])."""
        code, mat = _basis_code(basis)
        qs = np.asarray(list(measured_qubits), dtype=np.uint64)
        u = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float64))
        outs = np.zeros(u.size, dtype=np.uint64)
        m8 = _mat8(mat)
        err = L.QiError()
        _check(self.lib.qi_measure_n(self._h, code, L.dptr(m8) if m8 is not None else None,
                                     L.u64ptr(qs) if qs.size else None, qs.size, L.dptr(u), u.size, L.u64ptr(outs),
                                     C.byref(err)), err)
        return outs

    def measure_probs(self, measured_qubits: Sequence[int] = ()) -> np.ndarray:
        qs = np.asarray(list(measured_qubits), dtype=np.uint64)
        m = qs.size if qs.size else self.num_qubits
        out = np.empty(1 << m, dtype=np.float64)
        err = L.QiError()
        _check(self.lib.qi_measure_probs(self._h, L.u64ptr(qs) if qs.size else None, qs.size, L.dptr(out),
                                         C.byref(err)), err)
        return out

    def _measured_state(self, basis, idx, outcome: int) -> "State":
        """MeasurementResult::new_state of `outcome` (qi_measure_result_state)."""
        code, mat = _basis_code(basis)
        qs = np.asarray(list(idx), dtype=np.uint64)
        m8 = _mat8(mat)
        out = State._alloc(self.num_qubits, self.precision, self.ctx)
        err = L.QiError()
        _check(self.lib.qi_measure_result_state(self._h, code, L.dptr(m8) if m8 is not None else None,
                                                L.u64ptr(qs) if qs.size else None, qs.size, int(outcome), out._h,
                                                C.byref(err)), err)
        return out


def ry_phase_matrix(theta: float, phi: float):
    """Unitary2::from_ry_phase (operator.rs:2140-2156), num-complex arithmetic."""
    c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
    e = complex(math.cos(phi), math.sin(phi))  # exp(i phi) = from_polar(1, phi)
    return [[complex(c, 0.0), -complex(e.real * s, e.imag * s)], [complex(s, 0.0), complex(e.real * c, e.imag * c)]]


def ry_phase_dag_matrix(theta: float, phi: float):
    """Unitary2::from_ry_phase_dagger (operator.rs:2173-2192)."""
    c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
    e = complex(math.cos(-phi), math.sin(-phi))
    return [[complex(c, 0.0), complex(s, 0.0)], [-complex(e.real * s, e.imag * s), complex(e.real * c, e.imag * c)]]


def gate_array(gates: Iterable) -> np.ndarray:
    """Pack Gate objects (or raw QiGate records) into a contiguous qi_gate array."""
    gs = list(gates)
    arr = np.zeros(len(gs), dtype=L.GATE_DTYPE)
    raw = arr.view(np.uint8).reshape(len(gs), L.GATE_DTYPE.itemsize) if gs else None
    for i, g in enumerate(gs):
        r = g.raw if isinstance(g, Gate) else g
        raw[i, :] = np.frombuffer(bytes(r), dtype=np.uint8)
    return arr


# ------------------------------------------------------------ Pauli algebra --
class PauliString:
    """PauliString (pauli_string.rs:13-287): ops {qubit: 'X'|'Y'|'Z'} and a coefficient."""

    def __init__(self, coefficient: complex = 1.0, ops: Optional[Dict[int, str]] = None):
        self.coefficient = complex(coefficient)
        self.ops: Dict[int, str] = dict(ops or {})

    def add_op(self, qubit: int, op: str):
        if qubit in self.ops:
            raise ValueError(f"Duplicate Pauli string operator for qubit: {qubit}")
        self.ops[qubit] = op.upper()

    def with_op(self, qubit: int, op: str) -> "PauliString":
        self.add_op(qubit, op)
        return self

    def __len__(self):
        return len(self.ops)

    def get_targets(self):
        return sorted(self.ops)

    def masks(self):
        x = z = 0
        for q, op in self.ops.items():
            if q >= 64:
                raise QuantIronError(4, q, 0, msg="qubit index beyond the 64-bit mask")
            if op in ("X", "Y"):
                x |= 1 << q
            if op in ("Z", "Y"):
                z |= 1 << q
        return x, z

    def term(self) -> L.QiPauliTerm:
        x, z = self.masks()
        return L.QiPauliTerm(x, z, self.coefficient.real, self.coefficient.imag)

    def hermitian_conjugate(self) -> "PauliString":
        return PauliString(self.coefficient.conjugate(), self.ops)

    def __mul__(self, c):
        return PauliString(self.coefficient * complex(c), self.ops)

    __rmul__ = __mul__

    def __add__(self, other):
        if isinstance(other, PauliString):
            return SumOp([self, other])
        return NotImplemented

    def apply(self, state: State) -> State:
        out = State._alloc(state.num_qubits, state.precision, state.ctx)
        t = self.term()
        err = L.QiError()
        _check(state.lib.qi_pauli_apply(state._h, out._h, C.byref(t), C.byref(err)), err)
        return out

    def apply_normalised(self, state: State) -> State:
        out = State._alloc(state.num_qubits, state.precision, state.ctx)
        t = self.term()
        err = L.QiError()
        _check(state.lib.qi_pauli_apply_normalised(state._h, out._h, C.byref(t), C.byref(err)), err)
        return out

    def apply_exp_factor(self, state: State, factor: complex) -> State:
        out = state.clone()
        t = self.term()
        f = complex(factor)
        err = L.QiError()
        _check(state.lib.qi_pauli_exp(out._h, C.byref(t), f.real, f.imag, C.byref(err)), err)
        return out

    def apply_exp(self, state: State) -> State:
        return self.apply_exp_factor(state, 1.0)

    def apply_exp_neg_i_dt(self, state: State, dt: float) -> State:
        out = state.clone()
        t = self.term()
        err = L.QiError()
        _check(state.lib.qi_pauli_exp_neg_i_dt(out._h, C.byref(t), float(dt), C.byref(err)), err)
        return out

    def to_gates(self):
        return [Gate.operator(op, [q]) for q, op in self.ops.items()]


class SumOp:
    """SumOp (pauli_string.rs:390-507)."""

    def __init__(self, terms: Optional[List[PauliString]] = None):
        self.terms: List[PauliString] = list(terms or [])

    def num_terms(self):
        return len(self.terms)

    def add_term(self, t: PauliString):
        self.terms.append(t)

    def with_term(self, t: PauliString) -> "SumOp":
        self.add_term(t)
        return self

    def __add__(self, other):
        if isinstance(other, SumOp):
            return SumOp(self.terms + other.terms)
        if isinstance(other, PauliString):
            return SumOp(self.terms + [other])
        return NotImplemented

    def __mul__(self, c):
        return SumOp([t * c for t in self.terms])

    def term_array(self) -> np.ndarray:
        arr = np.zeros(len(self.terms), dtype=L.TERM_DTYPE)
        for i, t in enumerate(self.terms):
            x, z = t.masks()
            arr[i] = (x, z, t.coefficient.real, t.coefficient.imag)
        return arr

    def apply(self, state: State) -> State:
        out = State._alloc(state.num_qubits, state.precision, state.ctx)
        arr = self.term_array()
        err = L.QiError()
        _check(state.lib.qi_sumop_apply(state._h, out._h, arr.ctypes.data_as(C.c_void_p), arr.size, C.byref(err)), err)
        return out

    def expectation_value(self, state: State) -> complex:
        arr = self.term_array()
        re, im = C.c_double(), C.c_double()
        err = L.QiError()
        _check(state.lib.qi_sumop_expectation(state._h, arr.ctypes.data_as(C.c_void_p), arr.size, C.byref(re),
                                               C.byref(im), C.byref(err)), err)
        return complex(re.value, im.value)


def expectation_terms(state: State, terms: np.ndarray) -> complex:
    """<psi| sum_k c_k P_k |psi> for a qi_pauli_term array (SumOp::expectation_value)."""
    arr = np.ascontiguousarray(terms)
    re, im = C.c_double(), C.c_double()
    err = L.QiError()
    _check(state.lib.qi_sumop_expectation(state._h, arr.ctypes.data_as(C.c_void_p), arr.size, C.byref(re), C.byref(im),
                                           C.byref(err)), err)
    return complex(re.value, im.value)


# --------------------------------------------------------- time evolution --
FIRST, SECOND = 1, 2


def trotter_evolve_state(hamiltonian: SumOp, state: State, dt: float, num_steps: int, order: int = FIRST) -> State:
    """trotter_evolve_state (time_evolution.rs:140-166)."""
    out = state.clone()
    arr = hamiltonian.term_array()
    err = L.QiError()
    _check(state.lib.qi_trotter_evolve(out._h, arr.ctypes.data_as(C.c_void_p), arr.size, float(dt), int(num_steps),
                                       int(order), C.byref(err)), err)
    return out


def trotter_evolve_terms_inplace(state: State, terms: np.ndarray, dt: float, num_steps: int, order: int = FIRST) -> State:
    """trotter_evolve_state on a qi_pauli_term array, in place on `state`."""
    arr = np.ascontiguousarray(terms)
    err = L.QiError()
    _check(state.lib.qi_trotter_evolve(state._h, arr.ctypes.data_as(C.c_void_p), arr.size, float(dt), int(num_steps),
                                       int(order), C.byref(err)), err)
    return state


def first_order_trotter_step(hamiltonian: SumOp, state: State, dt: float) -> State:
    return trotter_evolve_state(hamiltonian, state, dt, 1, FIRST)


def second_order_trotter_step(hamiltonian: SumOp, state: State, dt: float) -> State:
    return trotter_evolve_state(hamiltonian, state, dt, 1, SECOND)


# ----------------------------------------------------------------- circuits --
class Circuit:
    """Circuit (circuit.rs:34-202)."""

    def __init__(self, num_qubits: int, gates: Optional[List[Gate]] = None):
        self.num_qubits = num_qubits
        self.gates: List[Gate] = list(gates or [])
        self._arr = None

    def add_gate(self, g: Gate):
        self.gates.append(g)
        self._arr = None

    def add_gates(self, gs):
        self.gates.extend(gs)
        self._arr = None

    def array(self) -> np.ndarray:
        if self._arr is None:
            self._arr = gate_array(self.gates)
        return self._arr

    @classmethod
    def from_qasm(cls, text: str, params=None) -> "Circuit":
        """A circuit from OpenQASM 3.0 text (the exporter's output of
        Circuit::to_qasm, circuit.rs:244-278, or standard syntax); `params`
        binds the program's `input` parameters in declaration order."""
        n, arr, _slots, _np = parse_qasm(text, params)
        c = cls(n, [Gate(L.QiGate.from_buffer_copy(arr[i].tobytes())) for i in range(arr.size)])
        c._arr = arr
        return c

    def execute(self, state: State, **opts) -> State:
        """Circuit::execute: a new state; all gates in one engine call."""
        out = state.clone()
        out.apply_gates_inplace(self.array(), n_circuit=self.num_qubits, **opts)
        return out

    def execute_inplace(self, state: State, **opts) -> State:
        return state.apply_gates_inplace(self.array(), n_circuit=self.num_qubits, **opts)

    def trace_execution(self, state: State) -> List[State]:
        """circuit.rs:188-202: every intermediate state."""
        if state.num_qubits != self.num_qubits:
            raise QuantIronError(3, state.num_qubits)
        cur = state.clone()
        states = [cur.clone()]
        for g in self.gates:
            cur.apply_gate_inplace(g)
            states.append(cur.clone())
        return states


def parse_qasm(text: str, params=None):
    """OpenQASM 3.0 text -> (qubit count, qi_gate array, angle slots, parameter
    count) through the engine's reader (qi_qasm_parse). `params` gives the
    values of the program's `input` parameters in declaration order."""
    lib = L.load()
    raw = text.encode()
    vals = None if params is None else np.ascontiguousarray(np.asarray(params, dtype=np.float64))
    ng, nq, npar = C.c_size_t(), C.c_uint32(), C.c_uint32()
    err = L.QiError()
    args = (raw, len(raw), L.dptr(vals) if vals is not None else None, 0 if vals is None else vals.size)
    _check(lib.qi_qasm_parse(*args, None, None, 0, C.byref(ng), C.byref(nq), C.byref(npar), C.byref(err)), err)
    arr = np.zeros(ng.value, dtype=L.GATE_DTYPE)
    slots = np.full(ng.value, -1, dtype=np.int32)
    _check(lib.qi_qasm_parse(*args, arr.ctypes.data_as(C.c_void_p), slots.ctypes.data_as(C.POINTER(C.c_int32)),
                             arr.size, C.byref(ng), C.byref(nq), C.byref(npar), C.byref(err)), err)
    return nq.value, arr, slots, npar.value


class CircuitBuilder:
    """CircuitBuilder (circuit.rs:288-1742): the fluent methods used by the workloads."""

    def __init__(self, num_qubits: int):
        self.num_qubits = num_qubits
        self.gates: List[Gate] = []

    def add_gate(self, g): self.gates.append(g); return self
    def add_gates(self, gs): self.gates.extend(gs); return self
    def h_gate(self, q): return self.add_gate(Gate.h_gate(q))
    def x_gate(self, q): return self.add_gate(Gate.x_gate(q))
    def y_gate(self, q): return self.add_gate(Gate.y_gate(q))
    def z_gate(self, q): return self.add_gate(Gate.z_gate(q))
    def s_gate(self, q): return self.add_gate(Gate.s_gate(q))
    def t_gate(self, q): return self.add_gate(Gate.t_gate(q))
    def p_gate(self, q, a): return self.add_gate(Gate.p_gate(q, a))
    def rx_gate(self, q, a): return self.add_gate(Gate.rx_gate(q, a))
    def ry_gate(self, q, a): return self.add_gate(Gate.ry_gate(q, a))
    def rz_gate(self, q, a): return self.add_gate(Gate.rz_gate(q, a))
    def cp_gates(self, targets, controls, angle): return self.add_gates(Gate.p_controlled_gates(targets, controls, angle))
    def cz_gates(self, targets, controls): return self.add_gates(Gate.controlled("Z", targets, controls))
    def cnot_gate(self, target, control): return self.add_gate(Gate.cnot_gate(target, control))
    def swap_gate(self, q1, q2): return self.add_gate(Gate.swap_gate(q1, q2))
    def toffoli_gate(self, c1, c2, target): return self.add_gate(Gate.toffoli_gate(target, [c1, c2]))
    def measure_gate(self, basis, qubits, uniform=None): return self.add_gate(Gate.new_measurement(qubits, basis, uniform))
    def pauli_string_gate(self, ps): return self.add_gate(Gate.pauli_string_gate(ps))
    def pauli_time_evolution_gate(self, ps, t): return self.add_gate(Gate.pauli_time_evolution_gate(ps, t))
    def build(self) -> Circuit: return Circuit(self.num_qubits, self.gates)
    def build_final(self) -> Circuit: return self.build()


class Subroutine:
    """Subroutine::qft / iqft (subroutine.rs:90-160): gate lists."""

    @staticmethod
    def qft(qubits: Sequence[int], num_qubits: int) -> Circuit:
        b = CircuitBuilder(num_qubits)
        n = len(qubits)
        for i in range(n):
            b.h_gate(qubits[i])
            den = 2.0
            for k in range(1, n - i):
                b.cp_gates([qubits[i]], [qubits[i + k]], math.pi / den)
                den *= 2.0
        for i in range(n // 2):
            b.swap_gate(qubits[i], qubits[n - 1 - i])
        return b.build()

    @staticmethod
    def iqft(qubits: Sequence[int], num_qubits: int) -> Circuit:
        b = CircuitBuilder(num_qubits)
        n = len(qubits)
        for i in range(n // 2):
            b.swap_gate(qubits[i], qubits[n - 1 - i])
        for i in reversed(range(n)):
            if n > i + 1:
                k0 = (n - 1) - i
                den = 2.0 ** k0
                for it in range(k0):
                    k = k0 - it
                    b.cp_gates([qubits[i]], [qubits[i + k]], -math.pi / den)
                    if k > 1:
                        den /= 2.0
            b.h_gate(qubits[i])
        return b.build()


# ------------------------------------------------------ parametric circuits --
class ParametricCircuit:
    """A circuit whose RX/RY/RZ/P gates may read their angle from a table
    (Gate::Parametric over Parameter<N> values, gate.rs:107-114,
    parameter.rs:12-68), planned once and re-run with new angles: set_angles
    rewrites the parametric ops' coefficients, run replays the captured
    launch sequence on a state (qi_param_circuit_* in include/qiron.h)."""

    def __init__(self, num_qubits: int, gates, angle_slots=None, precision: str = "c64", ctx=None,
                 mode: int = L.EXEC_AUTO, tile_qubits: int = 0):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.num_qubits = num_qubits
        self.precision = precision
        arr = gates if isinstance(gates, np.ndarray) else gate_array(gates)
        slots = None
        if angle_slots is not None:
            slots = np.ascontiguousarray(np.asarray(angle_slots, dtype=np.int32))
        opts = L.QiExecOpts(mode, tile_qubits, 0, 1, 0)
        self._h = C.c_void_p()
        err = L.QiError()
        _check(self.lib.qi_param_circuit_create(self.ctx.handle, int(num_qubits), PRECISIONS[precision],
                                                arr.ctypes.data_as(C.c_void_p), arr.size,
                                                slots.ctypes.data_as(C.POINTER(C.c_int32)) if slots is not None else None,
                                                C.byref(opts), C.byref(self._h), C.byref(err)), err)
        n = C.c_uint64()
        _check(self.lib.qi_param_circuit_num_slots(self._h, C.byref(n), C.byref(err)), err)
        self.num_slots = n.value

    @classmethod
    def from_qasm(cls, text: str, precision: str = "c64", ctx=None, **kw) -> "ParametricCircuit":
        """The program's `input` parameters become the angle slots, in declaration order."""
        n, arr, slots, _np = parse_qasm(text)
        return cls(n, arr, slots, precision, ctx, **kw)

    def __del__(self):
        try:
            if self._h:
                self.lib.qi_param_circuit_destroy(self._h)
        except Exception:
            pass

    def set_angles(self, angles) -> None:
        a = np.ascontiguousarray(np.asarray(angles, dtype=np.float64))
        err = L.QiError()
        _check(self.lib.qi_param_circuit_set_angles(self._h, L.dptr(a), a.size, C.byref(err)), err)

    def run(self, state: State) -> State:
        """Applies the circuit to `state` in place."""
        err = L.QiError()
        _check(self.lib.qi_param_circuit_run(self._h, state._h, C.byref(err)), err)
        return state

    def execute(self, state: State) -> State:
        """Circuit::execute: a new state."""
        return self.run(state.clone())


# ------------------------------------------------------------------- batches --
class CircuitBatch:
    """Many parametrised circuits, each with its own SumOp, evaluated to
    energies in one engine launch (qi_batch_* in include/qiron.h)."""

    def __init__(self, precision: str = "c64", ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self._h = C.c_void_p()
        self.count = 0
        err = L.QiError()
        _check(self.lib.qi_batch_create(self.ctx.handle, PRECISIONS[precision], C.byref(self._h), C.byref(err)), err)

    def __del__(self):
        try:
            if self._h:
                self.lib.qi_batch_destroy(self._h)
        except Exception:
            pass

    def add(self, n: int, gates: np.ndarray, terms: np.ndarray, angle_slots: Optional[np.ndarray] = None) -> int:
        g = np.ascontiguousarray(gates)
        t = np.ascontiguousarray(terms)
        slots = None
        if angle_slots is not None:
            slots = np.ascontiguousarray(np.asarray(angle_slots, dtype=np.int32))
        err = L.QiError()
        _check(self.lib.qi_batch_add(self._h, int(n), g.ctypes.data_as(C.c_void_p), g.size,
                                     slots.ctypes.data_as(C.POINTER(C.c_int32)) if slots is not None else None,
                                     t.ctypes.data_as(C.c_void_p), t.size, C.byref(err)), err)
        self.count += 1
        return self.count - 1

    def add_qasm(self, text: str, terms: np.ndarray, slot_offset: int = 0) -> int:
        """Adds a circuit given as OpenQASM 3.0 text; its k-th `input`
        parameter reads entry slot_offset + k of the batch's angle table."""
        n, arr, slots, npar = parse_qasm(text)
        slots = np.where(slots >= 0, slots + int(slot_offset), -1).astype(np.int32)
        return self.add(n, arr, terms, slots if npar else None)

    def set_angles(self, angles) -> None:
        a = np.ascontiguousarray(np.asarray(angles, dtype=np.float64))
        err = L.QiError()
        _check(self.lib.qi_batch_set_angles(self._h, L.dptr(a), a.size, C.byref(err)), err)

    def set_term_coeffs(self, coeffs) -> None:
        """Every circuit's term coefficients (complex), circuit by circuit in add order."""
        c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.complex128))
        err = L.QiError()
        _check(self.lib.qi_batch_set_term_coeffs(self._h, c.view(np.float64).ctypes.data_as(L.DP), c.size,
                                                 C.byref(err)), err)

    def run(self, use_graph: bool = True) -> np.ndarray:
        out = np.zeros(2 * max(self.count, 1), dtype=np.float64)
        err = L.QiError()
        _check(self.lib.qi_batch_run(self._h, L.dptr(out), 1 if use_graph else 0, C.byref(err)), err)
        return out[0::2][: self.count] + 1j * out[1::2][: self.count]

    def set_schedule(self, k: int, ks: int) -> None:
        """Tile width k of the tiled circuits and the largest LDS-resident circuit ks."""
        err = L.QiError()
        _check(self.lib.qi_batch_set_schedule(self._h, k, ks, C.byref(err)), err)

    def schedule(self):
        k, ks = C.c_uint32(), C.c_uint32()
        err = L.QiError()
        _check(self.lib.qi_batch_get_schedule(self._h, C.byref(k), C.byref(ks), C.byref(err)), err)
        return k.value, ks.value

    EXECUTORS = {"lds": 0, "wave": 1}

    def set_executor(self, name: str) -> None:
        """Executor of the circuits above ks qubits: "wave" (default; one-wave
        register passes for 12+ / 11+ qubits) or "lds" (LDS tiles of width k)."""
        err = L.QiError()
        _check(self.lib.qi_batch_set_executor(self._h, self.EXECUTORS[name], C.byref(err)), err)

    def executor(self) -> str:
        e = C.c_int()
        err = L.QiError()
        _check(self.lib.qi_batch_get_executor(self._h, C.byref(e), C.byref(err)), err)
        return {v: k for k, v in self.EXECUTORS.items()}[e.value]

    def autotune(self, trials: int = 2):
        """Times every schedule on this batch and keeps the fastest: (k, ks, ms)."""
        k, ks, ms = C.c_uint32(), C.c_uint32(), C.c_double()
        err = L.QiError()
        _check(self.lib.qi_batch_autotune(self._h, trials, C.byref(k), C.byref(ks), C.byref(ms), C.byref(err)), err)
        return k.value, ks.value, ms.value

    def info(self):
        nc, ns, nl = C.c_uint64(), C.c_uint64(), C.c_uint64()
        ms = C.c_double()
        err = L.QiError()
        _check(self.lib.qi_batch_info(self._h, C.byref(nc), C.byref(ns), C.byref(nl), C.byref(ms), C.byref(err)), err)
        return {"circuits": nc.value, "lds_resident": ns.value, "levels": nl.value, "last_ms": ms.value}
