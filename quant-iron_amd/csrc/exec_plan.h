// exec_plan.h — the launch plan of one qi_apply_gates call. Planning runs on
// the host before any launch; the op tables of all steps are uploaded in one
// copy, so the launch sequence itself is pure kernel launches and can be
// captured into a hipGraph and replayed.
#pragma once

#include "qi_internal.h"

namespace qi {

// One gate inside an LDS tile. Local positions index the tile's qubit list.
struct TileOp {
    int32_t kind;        // LoweredKind
    uint8_t lp0, lp1;    // local positions of the targets
    uint8_t outside;     // diagonal gate whose target is outside the tile
    uint8_t gt;          // that target's qubit (tested on the tile base)
    uint32_t lcm;        // controls inside the tile (local bit mask)
    uint32_t pad;        // planner: source gate index; batched engine: angle slot (-1: none)
    uint64_t gcm;        // controls outside the tile (global bit mask)
    double c[8];         // coefficients as in Lowered (phases in c[2], c[3])
    int32_t slot2 = -1;  // batched engine: angle slot of a one-sided phase applied after this op (-1: none)
    int32_t slot3 = -1;  // batched engine, two RY in one sweep: the second RY's angle slot (its qubit in lp1)
    int32_t slot4 = -1;  //   and the slot of its trailing phase (-1: none)
    int32_t pad2 = 0;
};

struct TilePass {
    uint64_t qmask;      // qubits spanned by a tile
    int32_t K;           // tile width (qubits)
    int32_t op0, nops;   // slice of the op table
    uint8_t pos[48];     // qubit of each local bit, ascending
};

// ---- register-resident tile passes (single-state executor) ----
// 2^TB threads (TB = 6, 8 or 9) hold a tile of 2^K amplitudes, 2^(K-TB) per
// thread in registers. A layout assigns each of the K local index bits to a
// thread bit (lane bits 0..5, wave bits 6..TB-1) or a register bit. A 2x2
// gate acts on a target held in registers (pairs inside a thread) or on lanes
// (the partner comes from another lane of the wave); wave bits are never
// targets. With TB = 6 a tile is one wave: no wave bits, no LDS, and layout
// changes are lane/register exchanges (RO_XLR, RO_XLL).
constexpr int kMaxThreadBits = 10;
constexpr int kLaneBits = 6;
constexpr int kLaneTarget = 8;   // RegOp::tgt >= kLaneTarget: lane bit tgt - kLaneTarget

enum RegOpKind : int32_t {
    RO_X = 0,          // swap the pair
    RO_RR,             // real 2x2 [[c0, c1], [c2, c3]] (H, RY)
    RO_GEN,            // complex 2x2, row-major (c0,c1) (c2,c3) / (c4,c5) (c6,c7) (U2)
    RO_RXI,            // [[c0, i c1], [i c2, c3]] with real c (RX)
    RO_Y,              // [[0, -i], [i, 0]]
    RO_Z,              // -1 where the target bit is 1
    RO_DG,             // p0 = (c0, c1) where the target bit is 0, p1 = (c2, c3) where it is 1 (RZ)
    RO_P1,             // p1 = (c2, c3) where the target bit is 1, nothing where it is 0 (S, T, P)
    RO_XC1,            // RO_X with one register-bit control (lp0) and no lane controls (CNOT, Toffoli)
    RO_ZC1,            // RO_Z with one register-bit control (lp0) and no lane controls (CZ)
    RO_RRS,            // rotation [[cos, -sin], [sin, cos]] on a register bit, uncontrolled in the tile, as three
                       // shears per component: a += c0 b, b += c1 a, a += c0 b (c0 = -tan(phi/2), c1 = sin phi)
    RO_RXS,            // RO_RXI rotation [[cos, i sin], [i sin, cos]]: shears (c0, c1) on (a.re, b.im) and
                       // (-c0, -c1) on (a.im, b.re)
    RO_P1S,            // RO_P1 on a register bit, no in-tile control, |phase| <= pi/2: the target-1 amplitudes'
                       // (re, im) rotated by three shears (c0 = -tan(phi/2), c1 = sin phi)
    RO_HS,             // a multiple of [[1, 1], [1, -1]] on a register bit, uncontrolled, unscaled:
                       // (a, b) <- (a + b, a - b); the planner applies the collected factor through an
                       // uncontrolled diagonal
    RO_XLR,            // exchange the local bits of lane bit tgt and register bit dbit
    RO_XLL,            // exchange the local bits of lane bits tgt and dbit
    RO_LPH,            // one-wave tiles: multiply the per-lane pending phase by p0 = (c0, c1) where the target
                       // (lane bit dbit for dloc 1, qubit dbit for dloc 2) is 0 and p1 = (c2, c3) where it is 1,
                       // on the lanes whose thread-index bits tcm are all 1 (a diagonal gate deferred)
    RO_LPF,            // multiply every amplitude by the per-lane pending phase and reset it (lp0 = 1: the
                       // phase is a sign, applied as a sign-bit XOR)
    RO_TRANSPOSE,      // new layout through LDS
    RO_LDS,            // gate on the LDS copy (controlled SWAP, Matchgate)
    RO_NOP             // closes each pass's slice (read ahead, never run)
};

struct RegOp {            // every field dword-sized: the kernel reads ops with scalar loads
    // fields the register kernels read come first (one 32-byte scalar load, then gcm)
    int32_t kind;        // RegOpKind
    int32_t tgt;         // pair ops: register bit, or kLaneTarget + lane bit
    int32_t dloc;        // diagonal ops: 0 register bit (dbit = its index), 1 thread bit (dbit = thread-index bit),
                         //               2 outside the tile (dbit = qubit)
    int32_t dbit;
    int32_t lp0, lp1;    // RO_LDS: local bits of the targets; RO_XC1 / RO_ZC1: lp0 = the control's register bit
    uint32_t rcm;        // controls on register bits (register-bit mask)
    uint32_t tcm;        // controls on thread bits (mask over thread-index bits)
    uint64_t gcm;        // controls outside the tile (qubit mask)
    int32_t lkind;       // RO_LDS: LK_SWAP or LK_MATCH
    int32_t param;       // angle slot + 1 of a parametric op (RO_RR: RY, RO_RXI: RX, RO_DG: RZ, RO_P1: P); 0: none
    uint32_t lcm;        // RO_LDS: controls inside the tile (local-bit mask)
    uint32_t src;        // source gate index
    uint8_t lay_r[8];    // RO_TRANSPOSE: local bit of each register bit
    uint8_t lay_t[12];   //               local bit of each thread bit (also after RO_XLR / RO_XLL)
    uint32_t pad1;
    double c[8];         // coefficients
    float cf[8];         // the same, rounded to float (read by the complex64 kernels)
};

struct RegPass {
    uint64_t qmask;      // qubits spanned by a tile
    int32_t K, op0, nops;
    int32_t sync_store;  // 1: amplitudes are stored to other addresses than they were loaded from
    int32_t tb;          // thread bits (2^tb threads per tile)
    int32_t chunk_log2;  // tile order: 0 = tile b is workgroup b; else see the kernels (set at launch)
    uint8_t lay_r[8];    // layout at load: local bit of each register bit
    uint8_t lay_t[12];   //                 local bit of each thread bit
    uint8_t qr_in[8];    // qubit of each register bit at load
    uint8_t qt_in[12];   // qubit of each thread bit at load
    uint8_t qr_out[8];   // at store (SWAPs relabel qubits, the final layout differs)
    uint8_t qt_out[12];
    uint64_t flip;       // qubits (inside the tile) whose bit the store address flips: the X part of the pass's Pauli frame
    int32_t oop;         // 1: stores go to the other of the state's two buffers (one-wave tiles only)
    int32_t permute;     // 1: the store address's bits outside the tile move to perm[bit] (a qubit permutation)
    uint8_t perm[64];
};

// A dense 2^k x 2^k block applied to tiles of 2^(k + kc) amplitudes.
struct FusedBlock {
    uint64_t bmask;      // block qubits
    uint64_t cmask;      // column qubits (lowest qubits outside the block)
    int32_t kb, kc;      // widths
    int64_t mat_off;     // byte offset of the real-embedded matrix in the blob
    int32_t ngates;      // gates folded into this block
    int32_t order;       // 1: enumerate block rows fastest (block holds qubit 0)
    uint8_t epos[48];    // qubit of each enumeration bit (ascending qubits)
    uint8_t eisb[48];    // 1 if that enumeration bit is a block qubit
};

struct ExecStep {
    enum Type { GATE = 0, TILE = 1, FUSED = 2, HOST = 3, REG = 4 } type = GATE;
    Lowered L;
    TilePass tile{};
    RegPass reg{};
    FusedBlock fused{};
    size_t host_gate = 0;  // index into the caller's gate list (HOST)
    static ExecStep gate(const Lowered &l) { ExecStep s; s.type = GATE; s.L = l; return s; }
    static ExecStep tile_step(const TilePass &p) { ExecStep s; s.type = TILE; s.tile = p; return s; }
    static ExecStep fused_step(const FusedBlock &b) { ExecStep s; s.type = FUSED; s.fused = b; return s; }
    static ExecStep reg_step(const RegPass &p) { ExecStep s; s.type = REG; s.reg = p; return s; }
    static ExecStep host(size_t g) { ExecStep s; s.type = HOST; s.host_gate = g; return s; }
};

struct ExecPlan {
    std::vector<ExecStep> steps;
    std::vector<TileOp> tile_ops;          // LDS passes (batched engine; uploaded by it)
    std::vector<RegOp> reg_ops;            // uploaded first in the blob
    std::vector<unsigned char> mats;       // fused matrices, after the register-pass ops
    size_t mats_offset() const { return (reg_ops.size() * sizeof(RegOp) + 255) & ~size_t(255); }
    size_t blob_bytes() const { return mats_offset() + mats.size(); }
};

// Coefficients of a parametric op (RegOp::param > 0) for the angle theta, in
// the op's form: the same values plan_tiled writes for that gate at that angle.
void param_op_coeffs(RegOp &op, double theta);

int default_tile_qubits(const State &s);
int default_tile_threads_log2(int prec, int K);
int tile_low_qubits(int prec);
// Register-resident passes for one state (ExecStep::REG, plan.reg_ops).
// allow_oop: the plan may use a second state-sized buffer (Ctx::alt): with
// SWAPs relabelled, the first pass then stores to it and a last pass brings the
// state back through any qubit permutation in its store addresses.
Status plan_tiled(const State &s, const std::vector<Lowered> &gates, int K, int TB, ExecPlan &plan,
                  bool allow_oop = false);
// src / dst: the buffers the pass loads from and stores to (equal unless p.oop).
Status launch_reg_pass(State &s, const RegPass &p, const RegOp *dops, hipStream_t st, void *src, void *dst);
// LDS passes over the same pass selection (ExecStep::TILE, plan.tile_ops),
// used by the batched engine.
Status plan_lds_passes(const State &s, const std::vector<Lowered> &gates, int K, ExecPlan &plan);

// Batched one-wave passes (the batched engine's circuits of 12+ / 11+
// qubits): one launch runs pass p of many circuits, one tile per wave.
constexpr int kBatchWaveK64 = 12, kBatchWaveK128 = 11;  // tile qubits (complex64 / complex128)
// Values per angle slot in the parametric-coefficient table the kernel reads
// (the kernel's precision): cos(t/2), -sin(t/2), sin(t/2), cos t, sin t, then
// the shears (a, b) of the rotation by t/2 brought into [-pi/2, pi/2] (a sign,
// a global phase, dropped: for uncontrolled RY whose sign no result sees), 0.
constexpr int kParamRow = 8;
struct WaveTask {
    uint64_t off;       // element offset of the circuit's state in the arena
    uint32_t block0;    // first workgroup of this task in the launch
    uint32_t pass;      // index of its RegPass in the pass table
    int32_t slot_base;  // first angle slot of the circuit (RegOp::param counts from it)
    uint32_t pad;
};
// init: the launch holds the first pass of each of its circuits (the state is
// |0..0>, the pass loads nothing).
Status launch_batch_wave(int prec, bool init, uint32_t blocks, hipStream_t st, void *arena, const WaveTask *tasks,
                         int ntasks, const RegPass *passes, const RegOp *ops, const void *ptab);

Status plan_fused(const State &s, const std::vector<Lowered> &gates, int kb, ExecPlan &plan);
Status launch_fused_block(State &s, const FusedBlock &b, const void *dmat, hipStream_t st);

}  // namespace qi
