// CPU model of the register-tiled executor, for tests without a GPU.
//
// Plans a gate list with the engine's own planner (qi::plan_tiled, linked
// from libqiron.so) and executes the plan's steps on the CPU with the
// semantics of the device code: the per-op forms of k_wave_pass /
// k_reg_pass (tiled.hip: pair ops on register and lane bits, diagonal ops,
// single-control X / Z, shears, unscaled H, lane/register exchanges, LDS
// transposes and gates) and the streaming gate kernels for GATE steps.
// Arithmetic is complex128 whatever the plan's precision, so a planner
// mistake (a wrong op form, a stale layout, a fold across a non-diagonal
// op, a wrong store layout) shows as a difference far above rounding when
// the output is compared with the oracle on the same input.
//
//   tile_emu <gates.bin> <n> <prec 64|128> <K> <TB> <in.bin> <out.bin> [--angles a.bin] [--oop]
//
// gates.bin: qi_gate records; in/out: 2^n complex128; --angles: float64
// angle per slot, where gate g's slot is g's index among the RX / RY / RZ / P
// gates; the plan is built with every such gate parametric and its ops are
// rewritten with qi::param_op_coeffs before execution. --oop: the planner may
// use a second state buffer (moving and permuting passes). Prints the plan's
// step count and its number of moving passes. No HIP call is made.
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../quant-iron_amd/csrc/qi_internal.h"
#include "../../quant-iron_amd/csrc/exec_plan.h"

using namespace qi;
using cd = std::complex<double>;

namespace {

template <typename T> std::vector<T> read_all(const char *path) {
    std::vector<T> v;
    FILE *f = fopen(path, "rb");
    if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
    T x;
    while (fread(&x, sizeof x, 1, f) == 1) v.push_back(x);
    fclose(f);
    return v;
}

uint64_t insert_zeros(uint64_t k, uint64_t fixed) {
    while (fixed) {
        const unsigned p = __builtin_ctzll(fixed);
        k = ((k >> p) << (p + 1)) | (k & ((uint64_t(1) << p) - 1));
        fixed &= fixed - 1;
    }
    return k;
}

// The 2x2 matrix of a pair op: m[row][col].
struct M2 { cd m[2][2]; };

M2 pair_matrix(const RegOp &op) {
    const double *c = op.c;
    M2 M{};
    switch (op.kind) {
    case RO_X: case RO_XC1: M.m[0][1] = M.m[1][0] = 1; break;
    case RO_RR: M.m[0][0] = c[0]; M.m[0][1] = c[1]; M.m[1][0] = c[2]; M.m[1][1] = c[3]; break;
    case RO_RXI:  // (a, b) <- (c0 a + i c1 b, i c2 a + c3 b)
        M.m[0][0] = c[0]; M.m[0][1] = cd(0, c[1]); M.m[1][0] = cd(0, c[2]); M.m[1][1] = c[3]; break;
    case RO_Y: M.m[0][1] = cd(0, -1); M.m[1][0] = cd(0, 1); break;
    case RO_GEN:
        M.m[0][0] = cd(c[0], c[1]); M.m[0][1] = cd(c[2], c[3]); M.m[1][0] = cd(c[4], c[5]); M.m[1][1] = cd(c[6], c[7]);
        break;
    default: fprintf(stderr, "not a pair op %d\n", op.kind); exit(3);
    }
    return M;
}

// Lowered gate on the whole state (GATE steps), as the streaming kernels apply it.
void apply_lowered(std::vector<cd> &psi, const Lowered &L) {
    const uint64_t N = psi.size(), cm = L.ctrl_mask, b0 = uint64_t(1) << L.t0, b1 = uint64_t(1) << L.t1;
    const double *c = L.c;
    if (L.kind == LK_SWAP || L.kind == LK_MATCH) {
        // matchgates test their controls on i01 (rotation) and i11 (phase);
        // a control may sit on the second target
        for (uint64_t i = 0; i < N; i++) {
            if (i & (b0 | b1)) continue;
            const uint64_t i01 = i | b0, i10 = i | b1, i11 = i | b0 | b1;
            if (L.kind == LK_SWAP) {
                if ((i & cm) == cm) std::swap(psi[i01], psi[i10]);
                continue;
            }
            if ((i01 & cm) == cm) {
                const cd a01 = psi[i01], a10 = psi[i10];
                psi[i01] = c[0] * a01 - cd(c[2], c[3]) * a10;
                psi[i10] = c[1] * a01 + cd(c[4], c[5]) * a10;
            }
            if ((i11 & cm) == cm) psi[i11] *= cd(c[6], c[7]);
        }
        return;
    }
    M2 M{};
    switch (L.kind) {
    case LK_H: M.m[0][0] = M.m[0][1] = M.m[1][0] = c[0]; M.m[1][1] = -c[0]; break;
    case LK_X: M.m[0][1] = M.m[1][0] = 1; break;
    case LK_Y: M.m[0][1] = cd(c[0], c[1]); M.m[1][0] = cd(c[2], c[3]); break;
    case LK_RX: M.m[0][0] = M.m[1][1] = c[0]; M.m[0][1] = -cd(c[1], c[2]); M.m[1][0] = cd(c[3], c[4]); break;
    case LK_RY: M.m[0][0] = M.m[1][1] = c[0]; M.m[0][1] = -c[1]; M.m[1][0] = c[1]; break;
    case LK_U2:
        M.m[0][0] = cd(c[0], c[1]); M.m[0][1] = cd(c[2], c[3]); M.m[1][0] = cd(c[4], c[5]); M.m[1][1] = cd(c[6], c[7]);
        break;
    case LK_Z: M.m[0][0] = 1; M.m[1][1] = -1; break;
    case LK_PHASE: M.m[0][0] = 1; M.m[1][1] = cd(c[0], c[1]); break;
    case LK_RZ: M.m[0][0] = cd(c[0], c[1]); M.m[1][1] = cd(c[2], c[3]); break;
    default: fprintf(stderr, "lowered kind %d\n", L.kind); exit(3);
    }
    for (uint64_t i = 0; i < N; i++) {
        if ((i & cm) != cm || (i & b0)) continue;
        const cd a = psi[i], b = psi[i | b0];
        psi[i] = M.m[0][0] * a + M.m[0][1] * b;
        psi[i | b0] = M.m[1][0] * a + M.m[1][1] * b;
    }
}

// One tile of a register pass: v[t * E + r] is thread t's register r.
struct Tile {
    int TB, RB, T, E;
    std::vector<cd> v;
    uint8_t lay_t[12], lay_r[8];  // local bit held by each thread / register bit
    cd &at(uint32_t t, uint32_t r) { return v[size_t(t) * E + r]; }
    uint32_t local(uint32_t t, uint32_t r) const {
        uint32_t e = 0;
        for (int i = 0; i < TB; i++) e |= ((t >> i) & 1u) << lay_t[i];
        for (int j = 0; j < RB; j++) e |= ((r >> j) & 1u) << lay_r[j];
        return e;
    }
};

bool ctrl_ok(const RegOp &op, uint32_t t, uint32_t r) { return (t & op.tcm) == op.tcm && (r & op.rcm) == op.rcm; }

void pair_op(Tile &X, const RegOp &op) {
    const M2 M = pair_matrix(op);
    if (op.tgt >= kLaneTarget) {
        const uint32_t bt = 1u << (op.tgt - kLaneTarget);
        for (uint32_t t = 0; t < (uint32_t)X.T; t++) {
            if (t & bt) continue;
            for (uint32_t r = 0; r < (uint32_t)X.E; r++) {
                if (!ctrl_ok(op, t, r)) continue;  // controls never sit on the target bit
                cd &a = X.at(t, r), &b = X.at(t | bt, r);
                const cd na = M.m[0][0] * a + M.m[0][1] * b, nb = M.m[1][0] * a + M.m[1][1] * b;
                a = na; b = nb;
            }
        }
        return;
    }
    const uint32_t br = 1u << op.tgt;
    for (uint32_t t = 0; t < (uint32_t)X.T; t++)
        for (uint32_t r = 0; r < (uint32_t)X.E; r++) {
            if ((r & br) || !ctrl_ok(op, t, r)) continue;
            cd &a = X.at(t, r), &b = X.at(t, r | br);
            const cd na = M.m[0][0] * a + M.m[0][1] * b, nb = M.m[1][0] * a + M.m[1][1] * b;
            a = na; b = nb;
        }
}

void diag_op(Tile &X, const RegOp &op, uint64_t base) {
    for (uint32_t t = 0; t < (uint32_t)X.T; t++)
        for (uint32_t r = 0; r < (uint32_t)X.E; r++) {
            if (!ctrl_ok(op, t, r)) continue;
            const bool one = op.dloc == 0 ? ((r >> op.dbit) & 1u) : op.dloc == 1 ? ((t >> op.dbit) & 1u)
                                                                                  : ((base >> op.dbit) & 1u);
            cd &a = X.at(t, r);
            if (op.kind == RO_Z || op.kind == RO_ZC1) { if (one) a = -a; }
            else if (op.kind == RO_P1) { if (one) a *= cd(op.c[2], op.c[3]); }
            else a *= one ? cd(op.c[2], op.c[3]) : cd(op.c[0], op.c[1]);
        }
}

// p += al q; q += be p; p += al q (the kernels' in-place shear sequence)
void shear(double &p, double &q, double al, double be) { p += al * q; q += be * p; p += al * q; }

void shear_op(Tile &X, const RegOp &op) {
    const uint32_t br = 1u << op.tgt;
    const double al = op.c[0], be = op.c[1];
    for (uint32_t t = 0; t < (uint32_t)X.T; t++)
        for (uint32_t r = 0; r < (uint32_t)X.E; r++) {
            if (r & br) continue;
            cd &a = X.at(t, r), &b = X.at(t, r | br);
            double ax = a.real(), ay = a.imag(), bx = b.real(), by = b.imag();
            if (op.kind == RO_RRS) { shear(ax, bx, al, be); shear(ay, by, al, be); }
            else { shear(ax, by, al, be); shear(ay, bx, -al, -be); }  // RO_RXS
            a = cd(ax, ay); b = cd(bx, by);
        }
}

void hs_op(Tile &X, const RegOp &op) {
    const uint32_t br = 1u << op.tgt;
    for (uint32_t t = 0; t < (uint32_t)X.T; t++)
        for (uint32_t r = 0; r < (uint32_t)X.E; r++) {
            if (r & br) continue;
            cd &a = X.at(t, r), &b = X.at(t, r | br);
            const cd na = a + b, nb = a - b;
            a = na; b = nb;
        }
}

void p1s_op(Tile &X, const RegOp &op) {
    const uint32_t br = 1u << op.dbit;
    for (uint32_t t = 0; t < (uint32_t)X.T; t++)
        for (uint32_t r = 0; r < (uint32_t)X.E; r++) {
            if (!(r & br)) continue;
            cd &a = X.at(t, r);
            double x = a.real(), y = a.imag();
            shear(x, y, op.c[0], op.c[1]);
            a = cd(x, y);
        }
}

// The amplitude at (thread bit i = a, register bit j = b) moves to (b, a).
void xlr_op(Tile &X, const RegOp &op) {
    const uint32_t bt = 1u << op.tgt, br = 1u << op.dbit;
    for (uint32_t t = 0; t < (uint32_t)X.T; t++)
        for (uint32_t r = 0; r < (uint32_t)X.E; r++)
            if (!(t & bt) && (r & br)) std::swap(X.at(t, r), X.at(t | bt, r & ~br));
    std::swap(X.lay_t[op.tgt], X.lay_r[op.dbit]);
}

void xll_op(Tile &X, const RegOp &op) {
    const uint32_t b1 = 1u << op.tgt, b2 = 1u << op.dbit;
    for (uint32_t t = 0; t < (uint32_t)X.T; t++)
        for (uint32_t r = 0; r < (uint32_t)X.E; r++)
            if ((t & b1) && !(t & b2)) std::swap(X.at(t, r), X.at(t ^ (b1 | b2), r));
    std::swap(X.lay_t[op.tgt], X.lay_t[op.dbit]);
}

// Through the tile's local index space (the LDS copy of k_reg_pass).
void to_local(Tile &X, std::vector<cd> &loc) {
    for (uint32_t t = 0; t < (uint32_t)X.T; t++)
        for (uint32_t r = 0; r < (uint32_t)X.E; r++) loc[X.local(t, r)] = X.at(t, r);
}
void from_local(Tile &X, const std::vector<cd> &loc) {
    for (uint32_t t = 0; t < (uint32_t)X.T; t++)
        for (uint32_t r = 0; r < (uint32_t)X.E; r++) X.at(t, r) = loc[X.local(t, r)];
}

void lds_gate(std::vector<cd> &loc, const RegOp &op) {
    const uint32_t b0 = 1u << op.lp0, b1 = 1u << op.lp1;
    const double *c = op.c;
    for (uint32_t l = 0; l < loc.size(); l++) {
        if (l & (b0 | b1)) continue;
        if (op.lkind == LK_SWAP) {
            if (((l | b0) & op.lcm) != op.lcm) continue;
            std::swap(loc[l | b0], loc[l | b1]);
            continue;
        }
        const uint32_t i01 = l | b0, i10 = l | b1, i11 = l | b0 | b1;
        if ((i01 & op.lcm) == op.lcm) {
            const cd a01 = loc[i01], a10 = loc[i10];
            loc[i01] = c[0] * a01 - cd(c[2], c[3]) * a10;
            loc[i10] = c[1] * a01 + cd(c[4], c[5]) * a10;
        }
        if ((i11 & op.lcm) == op.lcm) loc[i11] *= cd(c[6], c[7]);
    }
}

// Loads from src, stores to dst (the same vector for an in-place pass).
void run_pass(const std::vector<cd> &src, std::vector<cd> &dst, int n, const RegPass &p, const std::vector<RegOp> &ops) {
    Tile X;
    X.TB = p.tb; X.RB = p.K - p.tb; X.T = 1 << X.TB; X.E = 1 << X.RB;
    X.v.assign(size_t(1) << p.K, cd());
    std::vector<cd> loc(size_t(1) << p.K);
    const bool wave = p.tb == kLaneBits;
    const uint64_t tiles = uint64_t(1) << (n - p.K);
    for (uint64_t b = 0; b < tiles; b++) {
        const uint64_t base = insert_zeros(b, p.qmask);
        memcpy(X.lay_t, p.lay_t, sizeof X.lay_t);
        memcpy(X.lay_r, p.lay_r, sizeof X.lay_r);
        std::vector<cd> ph(size_t(X.T), cd(1));  // each lane's pending phase (RO_LPH / RO_LPF)
        for (uint32_t t = 0; t < (uint32_t)X.T; t++)
            for (uint32_t r = 0; r < (uint32_t)X.E; r++) {
                uint64_t a = base;
                for (int i = 0; i < X.TB; i++) a |= uint64_t((t >> i) & 1u) << p.qt_in[i];
                for (int j = 0; j < X.RB; j++) a |= uint64_t((r >> j) & 1u) << p.qr_in[j];
                X.at(t, r) = src[a];
            }
        for (int o = 0; o < p.nops; o++) {
            const RegOp &op = ops[size_t(p.op0) + o];
            if (wave && op.kind >= RO_TRANSPOSE) continue;  // one-wave passes hold no LDS ops
            if ((base & op.gcm) != op.gcm) continue;
            switch (op.kind) {
            case RO_X: case RO_RR: case RO_GEN: case RO_RXI: case RO_Y: case RO_XC1: pair_op(X, op); break;
            case RO_Z: case RO_DG: case RO_P1: case RO_ZC1: diag_op(X, op, base); break;
            case RO_RRS: case RO_RXS: shear_op(X, op); break;
            case RO_HS: hs_op(X, op); break;
            case RO_P1S: p1s_op(X, op); break;
            case RO_XLR: xlr_op(X, op); break;
            case RO_XLL: xll_op(X, op); break;
            case RO_LPH:
                for (uint32_t t = 0; t < (uint32_t)X.T; t++) {
                    if ((t & op.tcm) != op.tcm) continue;
                    const bool one = op.dloc == 1 ? ((t >> op.dbit) & 1u) : ((base >> op.dbit) & 1u);
                    ph[t] *= one ? cd(op.c[2], op.c[3]) : cd(op.c[0], op.c[1]);
                }
                break;
            case RO_LPF:
                for (uint32_t t = 0; t < (uint32_t)X.T; t++) {
                    for (uint32_t r = 0; r < (uint32_t)X.E; r++) X.at(t, r) *= ph[t];
                    ph[t] = 1;
                }
                break;
            case RO_TRANSPOSE:
                to_local(X, loc);
                memcpy(X.lay_t, op.lay_t, sizeof X.lay_t);
                memcpy(X.lay_r, op.lay_r, sizeof X.lay_r);
                from_local(X, loc);
                break;
            case RO_LDS:
                to_local(X, loc);
                lds_gate(loc, op);
                from_local(X, loc);
                break;
            case RO_NOP: break;
            default: fprintf(stderr, "op kind %d\n", op.kind); exit(3);
            }
        }
        for (uint32_t t = 0; t < (uint32_t)X.T; t++)
            for (uint32_t r = 0; r < (uint32_t)X.E; r++) {
                uint64_t a = 0;
                for (int i = 0; i < X.TB; i++) a |= uint64_t((t >> i) & 1u) << p.qt_out[i];
                for (int j = 0; j < X.RB; j++) a |= uint64_t((r >> j) & 1u) << p.qr_out[j];
                uint64_t b = base;
                if (p.permute) {  // a permuting pass moves the base bits too
                    b = 0;
                    for (uint64_t m = base; m; m &= m - 1) b |= uint64_t(1) << p.perm[__builtin_ctzll(m)];
                }
                dst[b | (a ^ p.flip)] = X.at(t, r);  // the Pauli frame's X part flips tile bits
            }
    }
}

bool angled(int kind) { return kind == QI_GATE_RX || kind == QI_GATE_RY || kind == QI_GATE_RZ || kind == QI_GATE_P; }

}  // namespace

int main(int argc, char **argv) {
    if (argc < 8) {
        fprintf(stderr, "usage: tile_emu gates.bin n prec K TB in.bin out.bin [--angles a.bin] [--oop]\n");
        return 2;
    }
    const char *angles_path = nullptr;
    bool oop = false;
    for (int a = 8; a < argc; a++) {
        if (!strcmp(argv[a], "--oop")) oop = true;
        else if (!strcmp(argv[a], "--angles") && a + 1 < argc) angles_path = argv[++a];
        else { fprintf(stderr, "argument %s\n", argv[a]); return 2; }
    }
    const std::vector<qi_gate> gates = read_all<qi_gate>(argv[1]);
    const int n = atoi(argv[2]);
    State s;
    s.n = (uint32_t)n;
    s.prec = atoi(argv[3]) == 128 ? QI_C128 : QI_C64;
    const int K = atoi(argv[4]), TB = atoi(argv[5]);
    std::vector<cd> psi = read_all<cd>(argv[6]);
    if (psi.size() != (size_t(1) << n)) { fprintf(stderr, "state size\n"); return 2; }
    std::vector<double> angles;
    if (angles_path) angles = read_all<double>(angles_path);
    std::vector<Lowered> seg;
    int32_t slot = 0;
    for (const qi_gate &g : gates) {
        Lowered L;
        const Status e = lower_gate((uint32_t)n, g, L);
        if (!e.ok()) { fprintf(stderr, "lower_gate: %s\n", e.msg.c_str()); return 3; }
        if (!angles.empty() && angled(g.kind)) L.slot = slot++;
        if (L.kind != LK_NOP) seg.push_back(L);
    }
    ExecPlan plan;
    const Status st = plan_tiled(s, seg, K, TB, plan, oop);
    if (!st.ok()) { fprintf(stderr, "plan_tiled: %s\n", st.msg.c_str()); return 3; }
    if (!angles.empty())
        for (RegOp &op : plan.reg_ops)
            if (op.param > 0) {
                if (op.param > (int)angles.size()) { fprintf(stderr, "slot out of range\n"); return 3; }
                param_op_coeffs(op, angles[size_t(op.param) - 1]);
            }
    // the executor's buffer tracking: a moving pass switches to the other buffer
    std::vector<cd> other(psi.size());
    std::vector<cd> *cur = &psi, *alt = &other;
    int moves = 0;
    for (const ExecStep &step : plan.steps) {
        if (step.type == ExecStep::GATE) apply_lowered(*cur, step.L);
        else if (step.type == ExecStep::REG) {
            if (step.reg.oop) {
                run_pass(*cur, *alt, n, step.reg, plan.reg_ops);
                std::swap(cur, alt);
                moves++;
            } else {
                run_pass(*cur, *cur, n, step.reg, plan.reg_ops);
            }
        } else { fprintf(stderr, "step type %d\n", (int)step.type); return 3; }
    }
    if (cur != &psi) { fprintf(stderr, "plan ends in the second buffer\n"); return 3; }
    FILE *f = fopen(argv[7], "wb");
    if (!f || fwrite(psi.data(), sizeof(cd), psi.size(), f) != psi.size()) { fprintf(stderr, "write\n"); return 2; }
    fclose(f);
    printf("%zu %d\n", plan.steps.size(), moves);
    return 0;
}
