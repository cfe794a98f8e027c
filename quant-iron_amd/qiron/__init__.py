"""qiron — MI355X (gfx950) state-vector engine for quant-iron's hot path.

The Python layer mirrors the reference's API (State, Gate, Circuit,
PauliString, SumOp, measurement, Trotter evolution) over the C ABI in
include/qiron.h; all amplitude arithmetic runs in the HIP library
libqiron.so next to this file.
"""
from ._lib import load as load_library  # noqa: F401
from .core import (  # noqa: F401
    FIRST,
    SECOND,
    Circuit,
    CircuitBatch,
    CircuitBuilder,
    Context,
    Gate,
    MeasurementBasis,
    MeasurementResult,
    ParametricCircuit,
    PauliString,
    QuantIronError,
    State,
    SumOp,
    Subroutine,
    default_context,
    expectation_terms,
    first_order_trotter_step,
    gate_array,
    ry_phase_dag_matrix,
    ry_phase_matrix,
    second_order_trotter_step,
    trotter_evolve_state,
    trotter_evolve_terms_inplace,
)
