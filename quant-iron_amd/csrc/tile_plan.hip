// tile_plan.hip — host planning of cache-blocked passes.
//
// A pass loads tiles of 2^K amplitudes spanned by a qubit set Q, applies a
// group of gates inside each tile and writes the tile back, so one HBM round
// trip carries many gates instead of one. A gate fits a pass when its
// non-diagonal qubits (pair targets, both SWAP / Matchgate targets) lie in Q.
// Controls and diagonal targets outside Q are constant across a tile (the
// tile fixes all other index bits), so they become a per-tile predicate or
// phase. Q always contains the lowest qubits, which makes every tile row a
// contiguous run of 128 bytes in HBM.
//
// Ordering: a gate may move ahead of earlier gates only when they commute in
// the exact sense that each acts diagonally (as a control or a diagonal
// target) on every qubit the other touches; the pass sequence then applies
// the circuit's unitary.
//
// Two emitters share the pass selection: register-resident passes for one
// state (plan_tiled, executed by k_reg_pass) and LDS passes for the batched
// engine (plan_lds_passes).
#include <algorithm>

#include "qi_internal.h"
#include "exec_plan.h"

namespace qi {

int tile_low_qubits(int prec) { return prec == QI_C128 ? 3 : 4; }  // 128-byte rows

int default_tile_qubits(const State &s) { return s.prec == QI_C128 ? 12 : 13; }

namespace {

struct GateSets {
    uint64_t need;  // qubits that must lie in the tile
    uint64_t supp;  // every qubit the gate touches
    uint64_t nd;    // qubits it acts on non-diagonally
};

GateSets gate_sets(const Lowered &L) {
    const uint64_t t0 = uint64_t(1) << L.t0, t1 = uint64_t(1) << L.t1;
    switch (L.kind) {
    case LK_Z: case LK_PHASE: case LK_RZ: return {0, t0 | L.ctrl_mask, 0};
    case LK_SWAP: case LK_MATCH: return {t0 | t1, t0 | t1 | L.ctrl_mask, t0 | t1};
    default: return {t0, t0 | L.ctrl_mask, t0};
    }
}

// Union of skipped gates; a later gate may pass them iff it commutes with
// each of them.
struct Blocked {
    uint64_t supp = 0, nd = 0;
    bool blocks(const GateSets &g) const { return (g.nd & supp) != 0 || (g.supp & nd) != 0; }
    void add(const GateSets &g) { supp |= g.supp; nd |= g.nd; }
};

// Members of the next pass (list order) and the tile's qubit set.
uint64_t select_pass(const std::vector<GateSets> &gs, std::vector<char> &done, size_t first, int K, int n, int low,
                     std::vector<size_t> &members) {
    const uint64_t all = n >= 64 ? ~uint64_t(0) : ((uint64_t(1) << n) - 1);
    uint64_t active = (uint64_t(1) << low) - 1;
    Blocked blk;
    for (size_t g = first; g < gs.size(); g++) {
        if (done[g]) continue;
        if (blk.blocks(gs[g])) blk.add(gs[g]);
        else if (__builtin_popcountll(active | gs[g].need) <= K) {
            active |= gs[g].need;
            members.push_back(g);
            done[g] = 1;
        } else blk.add(gs[g]);
        if ((blk.nd & all) == all) break;  // every later gate is blocked
    }
    uint64_t q = active;  // fill with the lowest unused qubits (longer HBM rows)
    for (int b = 0; __builtin_popcountll(q) < K && b < n; b++) q |= uint64_t(1) << b;
    return q;
}

bool is_relabel(const Lowered &L) { return L.kind == LK_SWAP && L.ctrl_mask == 0; }
bool is_pair(const Lowered &L) { return L.kind <= LK_U2; }

// ---------------------------------------------------------- register passes --
struct RegPlanner {
    int K = 0, RB = 0, L = 0;
    int lbit[64];       // qubit -> local bit (-1: outside the tile)
    int qof[16];        // local bit -> qubit
    uint8_t cr[8], ct[8];  // layout: local bit of each register / thread bit
    std::vector<RegOp> out;

    uint32_t lane_set() const {
        uint32_t m = 0;
        for (int i = 0; i < kLaneBits; i++) m |= 1u << ct[i];
        return m;
    }
    uint32_t io_set() const {
        uint32_t m = 0;
        for (int k = 0; k < L; k++) m |= 1u << lbit[k];
        return m;
    }
    int lane_index(int lb) const {
        for (int i = 0; i < kLaneBits; i++)
            if (ct[i] == lb) return i;
        return -1;
    }
    bool io_ok() const {
        for (int k = 0; k < L; k++)
            if (ct[k] != lbit[k]) return false;
        return true;
    }
    // Lane set of kLaneBits local bits containing `need`: the row qubits'
    // bits join when they fit (so loads and stores need no transpose), then
    // the lowest other bits.
    uint32_t fill_lanes(uint32_t need) const {
        uint32_t s = need;
        if (__builtin_popcount(s | io_set()) <= kLaneBits) s |= io_set();
        for (int b = 0; b < K && __builtin_popcount(s) < kLaneBits; b++) s |= 1u << b;
        return s;
    }
    // Layout with lane bits `lanes` (row qubits first when all are lanes),
    // then the two wave bits and the register bits in ascending order.
    void make(uint32_t lanes, uint8_t *r, uint8_t *t) const {
        int i = 0, j = 0;
        uint32_t used = 0;
        if ((lanes & io_set()) == io_set())
            for (int k = 0; k < L; k++) { t[i++] = (uint8_t)lbit[k]; used |= 1u << lbit[k]; }
        for (int b = 0; b < K; b++)
            if (((lanes >> b) & 1) && !((used >> b) & 1)) { t[i++] = (uint8_t)b; used |= 1u << b; }
        for (int b = 0; b < K && i < kRegThreadBits; b++)
            if (!((used >> b) & 1)) { t[i++] = (uint8_t)b; used |= 1u << b; }
        for (int b = 0; b < K; b++)
            if (!((used >> b) & 1)) r[j++] = (uint8_t)b;
        for (; j < 8; j++) r[j] = 0;
    }
    void transpose_to(uint32_t lanes) {
        uint8_t r[8], t[8];
        make(lanes, r, t);
        if (!memcmp(r, cr, 8) && !memcmp(t, ct, 8)) return;
        RegOp op{};
        op.kind = RO_TRANSPOSE;
        memcpy(op.lay_r, r, 8);
        memcpy(op.lay_t, t, 8);
        out.push_back(op);
        memcpy(cr, r, 8);
        memcpy(ct, t, 8);
    }
    void controls(uint64_t cm, RegOp &op) const {
        for (; cm; cm &= cm - 1) {
            const int q = __builtin_ctzll(cm);
            if (lbit[q] < 0) op.gcm |= uint64_t(1) << q;
            else op.cm |= 1u << lbit[q];
        }
    }
    Status emit(const Lowered &G, uint32_t src) {
        RegOp op{};
        op.src = src;
        switch (G.kind) {
        case LK_Z: case LK_PHASE: case LK_RZ: {
            // p0 = (c0, c1) on |0>, p1 = (c2, c3) on |1>
            if (G.kind == LK_Z) { op.kind = RO_Z; op.c[0] = 1; op.c[2] = -1; }
            else if (G.kind == LK_PHASE) { op.kind = RO_PHASE; op.c[0] = 1; op.c[2] = G.c[0]; op.c[3] = G.c[1]; }
            else { op.kind = RO_RZ; for (int k = 0; k < 4; k++) op.c[k] = G.c[k]; }
            const int lb = lbit[G.t0];
            if (lb < 0) { op.dloc = 2; op.dbit = (int)G.t0; }
            else { op.dloc = 0; op.dbit = lb; }
            controls(G.ctrl_mask, op);
        } break;
        case LK_SWAP: case LK_MATCH: {
            if (is_relabel(G)) {  // no data moves: the two local bits trade qubits
                const int a = lbit[G.t0], b = lbit[G.t1];
                std::swap(lbit[G.t0], lbit[G.t1]);
                qof[a] = (int)G.t1;
                qof[b] = (int)G.t0;
                return Status{};
            }
            op.kind = RO_LDS;
            op.lkind = G.kind;
            for (int k = 0; k < 8; k++) op.c[k] = G.c[k];
            op.lp0 = lbit[G.t0];
            op.lp1 = lbit[G.t1];
            controls(G.ctrl_mask, op);
        } break;
        default: {
            const int i = lane_index(lbit[G.t0]);
            if (i < 0) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "tile planner: target not on a lane");
            op.lane = i;
            switch (G.kind) {
            case LK_X: op.kind = RO_X; break;
            case LK_H: op.kind = RO_H; op.c[0] = G.c[0]; break;
            case LK_RY: op.kind = RO_REAL; op.c[0] = G.c[0]; op.c[1] = G.c[1]; break;
            case LK_Y:  // [[0, (c0,c1)], [(c2,c3), 0]]
                op.kind = RO_GEN;
                op.c[2] = G.c[0]; op.c[3] = G.c[1]; op.c[4] = G.c[2]; op.c[5] = G.c[3];
                break;
            case LK_RX:  // [[c0, -(c1,c2)], [(c3,c4), c0]]
                op.kind = RO_GEN;
                op.c[0] = G.c[0]; op.c[2] = -G.c[1]; op.c[3] = -G.c[2]; op.c[4] = G.c[3]; op.c[5] = G.c[4];
                op.c[6] = G.c[0];
                break;
            default:  // U2, row-major
                op.kind = RO_GEN;
                for (int k = 0; k < 8; k++) op.c[k] = G.c[k];
                break;
            }
            controls(G.ctrl_mask, op);
        } break;
        }
        out.push_back(op);
        return Status{};
    }
};

struct Phase {
    std::vector<size_t> ops;  // indices into the pass members, in execution order
    uint32_t need = 0;        // local bits that must be lane bits
};

}  // namespace

Status plan_tiled(const State &s, const std::vector<Lowered> &gates, int K, ExecPlan &plan) {
    if (K <= 0) K = default_tile_qubits(s);
    const int kmax = s.prec == QI_C128 ? 12 : 13;
    if (K < 10 || K > kmax) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "tile width out of range");
    if ((int)s.n < K) {  // state smaller than a tile: stream gate by gate
        for (const Lowered &L : gates) plan.steps.push_back(ExecStep::gate(L));
        return Status{};
    }
    const int low = tile_low_qubits(s.prec), RB = K - kRegThreadBits, LANES = kLaneBits;
    std::vector<GateSets> gs(gates.size());
    for (size_t g = 0; g < gates.size(); g++) gs[g] = gate_sets(gates[g]);
    std::vector<char> done(gates.size(), 0);
    size_t first = 0;
    while (first < gates.size()) {
        std::vector<size_t> members;
        const uint64_t q = select_pass(gs, done, first, K, (int)s.n, low, members);
        RegPlanner P;
        P.K = K; P.RB = RB; P.L = low;
        for (int b = 0; b < 64; b++) P.lbit[b] = -1;
        RegPass rp{};
        rp.qmask = q;
        rp.K = K;
        for (int b = 0, j = 0; b < (int)s.n; b++)
            if ((q >> b) & 1) { P.lbit[b] = j; P.qof[j] = b; j++; }
        // ---- phases: ops whose targets fit the lane bits, commuting-greedy
        std::vector<Phase> phases;
        {
            int lb[64];
            memcpy(lb, P.lbit, sizeof lb);
            std::vector<char> pd(members.size(), 0);
            size_t left = members.size();
            while (left) {
                Phase ph;
                Blocked blk;
                for (size_t i = 0; i < members.size(); i++) {
                    if (pd[i]) continue;
                    const Lowered &G = gates[members[i]];
                    const GateSets &g = gs[members[i]];
                    if (blk.blocks(g)) { blk.add(g); continue; }
                    const uint32_t need = is_pair(G) ? 1u << lb[G.t0] : 0u;
                    if (__builtin_popcount(ph.need | need) > LANES) { blk.add(g); continue; }
                    ph.need |= need;
                    ph.ops.push_back(i);
                    pd[i] = 1;
                    left--;
                    if (is_relabel(G)) std::swap(lb[G.t0], lb[G.t1]);
                }
                phases.push_back(std::move(ph));
            }
        }
        // ---- layout groups: consecutive phases whose targets fit one lane set
        struct Group { size_t p0, p1; uint32_t need; };
        std::vector<Group> groups;
        for (size_t i = 0; i < phases.size();) {
            uint32_t need = phases[i].need;
            size_t j = i + 1;
            while (j < phases.size() && __builtin_popcount(need | phases[j].need) <= LANES) need |= phases[j++].need;
            groups.push_back({i, j, need});
            i = j;
        }
        // ---- load layout: the first group's, when the row qubits fit beside its targets
        {
            const uint32_t first_need = groups.empty() ? 0u : groups[0].need;
            const uint32_t lanes = __builtin_popcount(first_need | P.io_set()) <= LANES ? P.fill_lanes(first_need)
                                                                                      : P.fill_lanes(0);
            P.make(lanes, P.cr, P.ct);
            memcpy(rp.lay_r, P.cr, 8);
            memcpy(rp.lay_t, P.ct, 8);
            for (int j = 0; j < RB; j++) rp.qr_in[j] = (uint8_t)P.qof[P.cr[j]];
            for (int i = 0; i < kRegThreadBits; i++) rp.qt_in[i] = (uint8_t)P.qof[P.ct[i]];
        }
        for (const Group &gr : groups) {
            if ((gr.need & P.lane_set()) != gr.need) P.transpose_to(P.fill_lanes(gr.need));
            for (size_t p = gr.p0; p < gr.p1; p++)
                for (size_t i : phases[p].ops) {
                    Status e = P.emit(gates[members[i]], (uint32_t)members[i]);
                    if (!e.ok()) return e;
                }
        }
        // ---- store: row qubits back on lanes 0..L-1
        if (!P.io_ok()) {
            P.transpose_to(P.fill_lanes(0));
            if (!P.io_ok()) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "tile planner: store layout");
        }
        for (int j = 0; j < RB; j++) rp.qr_out[j] = (uint8_t)P.qof[P.cr[j]];
        for (int i = 0; i < kRegThreadBits; i++) rp.qt_out[i] = (uint8_t)P.qof[P.ct[i]];
        // SWAP relabels or transposes may send an amplitude to an address another
        // thread loads: the kernel then waits for every load before storing
        rp.sync_store = memcmp(rp.qt_out, rp.qt_in, 8) != 0 || memcmp(rp.qr_out, rp.qr_in, 8) != 0;
        rp.op0 = (int)plan.reg_ops.size();
        rp.nops = (int)P.out.size();
        plan.reg_ops.insert(plan.reg_ops.end(), P.out.begin(), P.out.end());
        RegOp end{};
        end.kind = RO_NOP;
        plan.reg_ops.push_back(end);  // the kernel reads one op ahead
        plan.steps.push_back(ExecStep::reg_step(rp));
        while (first < gates.size() && done[first]) first++;
    }
    return Status{};
}

// ------------------------------------------------------------- LDS passes --
Status plan_lds_passes(const State &s, const std::vector<Lowered> &gates, int K, ExecPlan &plan) {
    if (K <= 0) K = default_tile_qubits(s);
    const int kmax = s.prec == QI_C128 ? 12 : 13;
    if (K < 10 || K > kmax) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "tile width out of range");
    if ((int)s.n < K) {
        for (const Lowered &L : gates) plan.steps.push_back(ExecStep::gate(L));
        return Status{};
    }
    std::vector<GateSets> gs(gates.size());
    for (size_t g = 0; g < gates.size(); g++) gs[g] = gate_sets(gates[g]);
    std::vector<char> done(gates.size(), 0);
    size_t first = 0;
    while (first < gates.size()) {
        std::vector<size_t> members;
        const uint64_t q = select_pass(gs, done, first, K, (int)s.n, tile_low_qubits(s.prec), members);
        TilePass p{};
        p.qmask = q;
        p.K = K;
        uint8_t local[64];
        for (int b = 0, j = 0; b < (int)s.n; b++)
            if ((q >> b) & 1) { local[b] = (uint8_t)j; p.pos[j++] = (uint8_t)b; }
        p.op0 = (int)plan.tile_ops.size();
        for (size_t g : members) {
            const Lowered &L = gates[g];
            TileOp op{};
            op.kind = L.kind;
            for (int c = 0; c < 8; c++) op.c[c] = L.c[c];
            if (L.kind == LK_Z) { op.c[2] = -1.0; op.c[3] = 0.0; }
            if (L.kind == LK_PHASE) { op.c[2] = L.c[0]; op.c[3] = L.c[1]; }
            const uint64_t tb0 = uint64_t(1) << L.t0;
            if ((q & tb0) == 0) { op.outside = 1; op.gt = (uint8_t)L.t0; }
            else op.lp0 = local[L.t0];
            if (L.kind == LK_SWAP || L.kind == LK_MATCH) op.lp1 = local[L.t1];
            op.gcm = L.ctrl_mask & ~q;
            uint32_t lcm = 0;
            for (uint64_t lc = L.ctrl_mask & q; lc; lc &= lc - 1) lcm |= 1u << local[__builtin_ctzll(lc)];
            op.lcm = lcm;
            op.pad = (uint32_t)g;  // source gate (the batched engine maps it to an angle slot)
            plan.tile_ops.push_back(op);
        }
        p.nops = (int)members.size();
        plan.steps.push_back(ExecStep::tile_step(p));
        while (first < gates.size() && done[first]) first++;
    }
    return Status{};
}

}  // namespace qi
