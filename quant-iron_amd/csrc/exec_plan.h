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
};

struct TilePass {
    uint64_t qmask;      // qubits spanned by a tile
    int32_t K;           // tile width (qubits)
    int32_t op0, nops;   // slice of the op table
    uint8_t pos[48];     // qubit of each local bit, ascending
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
    enum Type { GATE = 0, TILE = 1, FUSED = 2, HOST = 3 } type = GATE;
    Lowered L;
    TilePass tile{};
    FusedBlock fused{};
    size_t host_gate = 0;  // index into the caller's gate list (HOST)
    static ExecStep gate(const Lowered &l) { ExecStep s; s.type = GATE; s.L = l; return s; }
    static ExecStep tile_step(const TilePass &p) { ExecStep s; s.type = TILE; s.tile = p; return s; }
    static ExecStep fused_step(const FusedBlock &b) { ExecStep s; s.type = FUSED; s.fused = b; return s; }
    static ExecStep host(size_t g) { ExecStep s; s.type = HOST; s.host_gate = g; return s; }
};

struct ExecPlan {
    std::vector<ExecStep> steps;
    std::vector<TileOp> tile_ops;          // uploaded first in the blob
    std::vector<unsigned char> mats;       // fused matrices, after the tile ops
    size_t mats_offset() const { return (tile_ops.size() * sizeof(TileOp) + 255) & ~size_t(255); }
    size_t blob_bytes() const { return mats_offset() + mats.size(); }
};

int default_tile_qubits(const State &s);
Status plan_tiled(const State &s, const std::vector<Lowered> &gates, int K, ExecPlan &plan);
Status launch_tile_pass(State &s, const TilePass &p, const TileOp *dops, hipStream_t st);

Status plan_fused(const State &s, const std::vector<Lowered> &gates, int kb, ExecPlan &plan);
Status launch_fused_block(State &s, const FusedBlock &b, const void *dmat, hipStream_t st);

}  // namespace qi
