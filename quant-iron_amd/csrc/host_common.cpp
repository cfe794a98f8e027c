// host_common.cpp — host-only code shared by the launch code and the tile
// planner: the error values (errors.rs) and the lowering of a gate record to
// the 2x2 / diagonal / two-target forms the kernels and the planner use. Plain
// C++ (built with g++), so the planner, this file and the CPU model of the tiled
// executor also build with the address and undefined-behaviour sanitizers.
#include <cmath>
#include <cstdio>

#include "qi_internal.h"

namespace qi {

// Display strings of errors.rs:4-97.
Status make_err(int code, uint64_t a, uint64_t b, double re, double im, const char *detail) {
    Status s;
    s.code = code; s.a = a; s.b = b; s.re = re; s.im = im;
    char buf[256];
    switch (code) {
    case QI_ERR_INVALID_NUMBER_OF_MEASUREMENTS: snprintf(buf, sizeof buf, "Invalid number of measurements: %llu", (unsigned long long)a); break;
    case QI_ERR_OVERLAPPING_CONTROL_AND_TARGET:
        snprintf(buf, sizeof buf, "Control qubit index %llu overlaps with target qubit index %llu", (unsigned long long)a, (unsigned long long)b); break;
    case QI_ERR_INVALID_NUMBER_OF_QUBITS: snprintf(buf, sizeof buf, "Invalid number of qubits: %llu", (unsigned long long)a); break;
    case QI_ERR_INVALID_QUBIT_INDEX:
        snprintf(buf, sizeof buf, "Invalid qubit index: %llu for %llu qubits", (unsigned long long)a, (unsigned long long)b); break;
    case QI_ERR_STATE_NOT_NORMALISED: snprintf(buf, sizeof buf, "State vector is not normalised"); break;
    case QI_ERR_NON_UNITARY_MATRIX: snprintf(buf, sizeof buf, "Non-unitary matrix"); break;
    case QI_ERR_UNKNOWN: snprintf(buf, sizeof buf, "An unknown error occurred"); break;
    case QI_ERR_DEVICE: snprintf(buf, sizeof buf, "Device error: %s", detail ? detail : "unknown"); break;
    case QI_ERR_LOCK: snprintf(buf, sizeof buf, "Failed to lock GPU context"); break;
    case QI_ERR_ZERO_NORM: snprintf(buf, sizeof buf, "The state cannot be normalised because it has zero norm."); break;
    case QI_ERR_INVALID_PAULI_COEFFICIENT: snprintf(buf, sizeof buf, "Invalid Pauli String coefficient: %g+%gi", re, im); break;
    case QI_ERR_OUT_OF_MEMORY: snprintf(buf, sizeof buf, "Device out of memory%s%s", detail ? ": " : "", detail ? detail : ""); break;
    case QI_ERR_INVALID_ARGUMENT: snprintf(buf, sizeof buf, "Invalid argument%s%s", detail ? ": " : "", detail ? detail : ""); break;
    case QI_ERR_PRECISION_MISMATCH: snprintf(buf, sizeof buf, "States differ in precision"); break;
    default: snprintf(buf, sizeof buf, "error %d", code); break;
    }
    s.msg = buf;
    return s;
}

Status lower_gate(uint32_t n, const qi_gate &g, Lowered &L) {
    (void)n;
    L = Lowered{};
    L.t0 = (uint32_t)g.targets[0];
    L.t1 = g.ntargets > 1 ? (uint32_t)g.targets[1] : 0;
    for (uint32_t k = 0; k < g.ncontrols; k++) L.ctrl_mask |= uint64_t(1) << g.controls[k];
    const double r2 = 1.0 / std::sqrt(2.0);
    const double h = g.params[0] / 2.0;
    switch (g.kind) {
    case QI_GATE_H: L.kind = LK_H; L.c[0] = r2; break;
    case QI_GATE_X: case QI_GATE_CNOT: case QI_GATE_TOFFOLI: L.kind = LK_X; break;
    case QI_GATE_Y:  // -i_complex = (-0.0, -1.0), i_complex = (0, 1) (operator.rs:489)
        L.kind = LK_Y; L.c[0] = -0.0; L.c[1] = -1.0; L.c[2] = 0.0; L.c[3] = 1.0; break;
    case QI_GATE_Z: L.kind = LK_Z; break;
    case QI_GATE_I: L.kind = LK_NOP; break;
    case QI_GATE_S: L.kind = LK_PHASE; L.c[0] = 0.0; L.c[1] = 1.0; break;
    case QI_GATE_SDAG: L.kind = LK_PHASE; L.c[0] = 0.0; L.c[1] = -1.0; break;
    case QI_GATE_T: L.kind = LK_PHASE; L.c[0] = r2; L.c[1] = r2; break;
    case QI_GATE_TDAG: L.kind = LK_PHASE; L.c[0] = r2; L.c[1] = -r2; break;
    case QI_GATE_P: L.kind = LK_PHASE; L.c[0] = std::cos(g.params[0]); L.c[1] = std::sin(g.params[0]); break;
    case QI_GATE_RX: {  // i*sin, (-i)*sin as Complex * f64 (operator.rs:1756-1757)
        double ch = std::cos(h), sh = std::sin(h);
        L.kind = LK_RX; L.c[0] = ch; L.c[1] = 0.0 * sh; L.c[2] = 1.0 * sh; L.c[3] = -0.0 * sh; L.c[4] = -1.0 * sh;
    } break;
    case QI_GATE_RY: L.kind = LK_RY; L.c[0] = std::cos(h); L.c[1] = std::sin(h); break;
    case QI_GATE_RZ:
        L.kind = LK_RZ; L.c[0] = std::cos(h); L.c[1] = -std::sin(h); L.c[2] = std::cos(h); L.c[3] = std::sin(h); break;
    case QI_GATE_U2: L.kind = LK_U2; for (int q = 0; q < 8; q++) L.c[q] = g.matrix[q]; break;
    case QI_GATE_SWAP: L.kind = LK_SWAP; break;
    case QI_GATE_MATCHGATE: {
        double th = g.params[0] / 2.0, ch = std::cos(th), sh = std::sin(th);
        // exp(i phi) = from_polar(exp(0), phi) (num-complex)
        double e1r = std::exp(0.0) * std::cos(g.params[1]), e1i = std::exp(0.0) * std::sin(g.params[1]);
        double e2r = std::exp(0.0) * std::cos(g.params[2]), e2i = std::exp(0.0) * std::sin(g.params[2]);
        L.kind = LK_MATCH; L.t1 = L.t0 + 1;
        L.c[0] = ch; L.c[1] = sh;
        L.c[2] = e1r * sh; L.c[3] = e1i * sh;  // exp_i_phi1 * sin_theta_half
        L.c[4] = e1r * ch; L.c[5] = e1i * ch;  // exp_i_phi1 * cos_theta_half
        L.c[6] = e2r; L.c[7] = e2i;
    } break;
    default: return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "gate kind has no streaming kernel");
    }
    return Status{};
}

}  // namespace qi
