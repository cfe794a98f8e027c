// tile_plan.cpp — host planning of cache-blocked passes.
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
#include <array>
#include <cmath>

#include "qi_internal.h"
#include "exec_plan.h"

namespace qi {

int tile_low_qubits(int prec) { return prec == QI_C128 ? 3 : 4; }  // 128-byte rows

int default_tile_qubits(const State &s) { return s.prec == QI_C128 ? 11 : 12; }

// Tiles of up to 4096 complex64 / 2048 complex128 amplitudes run as one wave
// (no LDS, no barriers, 64 / 32 amplitudes per lane, the per-op scalar work
// paid once per tile); larger tiles use 256 threads, 512 for the 64 KiB tiles
// (twice the waves per CU at the same LDS).
int default_tile_threads_log2(int prec, int K) {
    if (K <= (prec == QI_C64 ? 12 : 11)) return 6;
    return K >= 12 + (prec == QI_C128 ? 0 : 1) ? 9 : 8;
}

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

// Number of gates (from `first`, list order) a pass on qubit set q can take:
// a gate joins unless it fails to commute with an earlier skipped gate or
// acts non-diagonally outside q; skipped gates block later ones.
size_t pass_members(const std::vector<GateSets> &gs, const std::vector<char> &streamed, const std::vector<char> &done,
                    size_t first, uint64_t q, uint64_t all, std::vector<size_t> *members) {
    Blocked blk;
    size_t count = 0;
    for (size_t g = first; g < gs.size(); g++) {
        if (done[g]) continue;
        if (streamed[g] || blk.blocks(gs[g]) || (gs[g].need & ~q)) blk.add(gs[g]);
        else {
            count++;
            if (members) members->push_back(g);
        }
        if ((blk.nd & all) == all) break;  // every later gate is blocked
    }
    return count;
}

// The next pass: its tile's qubit set and members. The set starts greedy (the
// qubits of the earliest gates that fit, then the lowest unused qubits: longer
// HBM rows); with `search`, single-qubit exchanges that let the pass take more
// gates are then applied while they help, and when none does, exchanges of two
// tile qubits for two qubits the first skipped gates need (up to 8 sweeps).
// On random 30-qubit circuits the single exchanges remove a quarter or more of
// the passes, the pair exchanges 0-2 more (about 0.15 s of planning).
constexpr int kSearchSweeps = 8;     // improvement sweeps of the tile qubit search
constexpr int kSearchSkipped = 24;   // skipped gates whose qubits a pair exchange may bring in
uint64_t select_pass(const std::vector<GateSets> &gs, const std::vector<char> &streamed, std::vector<char> &done,
                     size_t first, int K, int n, int low, bool search, std::vector<size_t> &members) {
    const uint64_t all = n >= 64 ? ~uint64_t(0) : ((uint64_t(1) << n) - 1);
    const uint64_t fixed = (uint64_t(1) << low) - 1;
    uint64_t active = fixed;
    Blocked blk;
    for (size_t g = first; g < gs.size(); g++) {
        if (done[g]) continue;
        if (streamed[g] || blk.blocks(gs[g])) blk.add(gs[g]);
        else if (__builtin_popcountll(active | gs[g].need) <= K) active |= gs[g].need;
        else blk.add(gs[g]);
        if ((blk.nd & all) == all) break;
    }
    uint64_t q = active;
    for (int b = 0; __builtin_popcountll(q) < K && b < n; b++) q |= uint64_t(1) << b;
    if (search) {
        size_t best = pass_members(gs, streamed, done, first, q, all, nullptr);
        const auto bit = [](int b) { return uint64_t(1) << b; };
        for (int sweep = 0; sweep < kSearchSweeps; sweep++) {
            bool improved = false;
            for (int out = 0; out < n; out++) {
                if (!((q >> out) & 1) || ((fixed >> out) & 1)) continue;
                for (int in = 0; in < n; in++) {
                    if ((q >> in) & 1) continue;
                    const uint64_t q2 = (q & ~bit(out)) | bit(in);
                    const size_t c = pass_members(gs, streamed, done, first, q2, all, nullptr);
                    if (c > best) { best = c; q = q2; improved = true; break; }
                }
            }
            if (improved) continue;
            // pair exchanges: two tile qubits out, two qubits that the first
            // skipped gates need in
            std::vector<size_t> mem;
            pass_members(gs, streamed, done, first, q, all, &mem);
            std::vector<char> taken(gs.size(), 0);
            for (size_t g : mem) taken[g] = 1;
            uint64_t want = 0;
            for (size_t g = first, seen = 0; g < gs.size() && seen < kSearchSkipped; g++)
                if (!done[g] && !taken[g] && !streamed[g]) { want |= gs[g].need & ~q; seen++; }
            const uint64_t outs = q & ~fixed;
            for (uint64_t o1 = outs; o1 && !improved; o1 &= o1 - 1)
                for (uint64_t o2 = o1 & (o1 - 1); o2 && !improved; o2 &= o2 - 1)
                    for (uint64_t i1 = want; i1 && !improved; i1 &= i1 - 1)
                        for (uint64_t i2 = i1 & (i1 - 1); i2 && !improved; i2 &= i2 - 1) {
                            const uint64_t q2 = (q & ~(bit(__builtin_ctzll(o1)) | bit(__builtin_ctzll(o2)))) |
                                                bit(__builtin_ctzll(i1)) | bit(__builtin_ctzll(i2));
                            const size_t c = pass_members(gs, streamed, done, first, q2, all, nullptr);
                            if (c > best) { best = c; q = q2; improved = true; }
                        }
            if (!improved) break;
        }
    }
    pass_members(gs, streamed, done, first, q, all, &members);
    for (size_t g : members) done[g] = 1;
    return q;
}

// States from this size on spend planning time on the tile qubit search (a
// pass over them costs milliseconds).
constexpr uint32_t kSearchQubits = 20;

bool is_relabel(const Lowered &L) { return L.kind == LK_SWAP && L.ctrl_mask == 0; }
bool is_pair(const Lowered &L) { return L.kind <= LK_U2; }

// ---------------------------------------------------------- register passes --
// Targets of 2x2 gates must sit on register bits or lane bits; the two wave
// bits hold local bits no gate of the current group targets. Register
// targets are the cheapest (pairs inside a thread), so the most used targets
// and controls of a group go to registers.
struct RegPlanner {
    int K = 0, RB = 0, L = 0, TB = 8;
    bool f32 = false;  // complex64 kernels: rotations may run as shears
    // Uncontrolled RZ = e^{-i theta/2} diag(1, e^{i theta}): emitted as a
    // one-sided phase (half the work on a register target) while the global
    // factor collects here, applied once by the plan's last pass.
    double *gphase = nullptr;  // (re, im)
    // Uncontrolled multiples of H run unscaled (RO_HS) and multiples of a
    // rotation as unit shears (RO_RRS, RO_RXS) on register bits; their
    // factors (1/sqrt(2) for H) multiply up here and go into the global
    // phase. Every kMaxUnscaled of them a scaling diagonal is emitted, so
    // complex64 amplitudes grow by at most 2 per such op between scalings,
    // 2^32 in all: far below overflow also for unnormalised states.
    int *unscaled = nullptr;
    double *hfactor = nullptr;
    static constexpr int kMaxUnscaled = 32;
    // The last uncontrolled diagonal op on each diagonal-op position (register
    // bit, local bit, or qubit outside the tile) that no later op has acted on
    // non-diagonally: a new uncontrolled diagonal there folds into it
    // (diagonals commute with everything diagonal on their target).
    int last_diag[3][64];
    void forget_diagonals() {
        for (auto &a : last_diag)
            for (int &x : a) x = -1;
    }
    int lbit[64];       // qubit -> local bit (-1: outside the tile)
    int qof[16];        // local bit -> qubit
    uint8_t cr[8], ct[12];  // layout: local bit of each register / thread bit
    std::vector<RegOp> out;

    uint32_t io_set() const {
        uint32_t m = 0;
        for (int k = 0; k < L; k++) m |= 1u << lbit[k];
        return m;
    }
    uint32_t wave_set() const {
        uint32_t m = 0;
        for (int i = kLaneBits; i < TB; i++) m |= 1u << ct[i];
        return m;
    }
    int reg_index(int lb) const {
        for (int j = 0; j < RB; j++)
            if (cr[j] == lb) return j;
        return -1;
    }
    int lane_index(int lb) const {
        for (int i = 0; i < kLaneBits; i++)
            if (ct[i] == lb) return i;
        return -1;
    }
    int thread_index(int lb) const {  // thread-index bit holding local bit lb
        for (int i = 0; i < TB; i++)
            if (ct[i] == lb) return i;
        return -1;
    }
    bool io_ok() const {
        for (int k = 0; k < L; k++)
            if (ct[k] != lbit[k]) return false;
        return true;
    }
    // Layout for a group: row qubits on lanes 0..L-1; wave bits from the bits
    // the group does not target (latest next use first); register bits by
    // score (pair targets and controls of the group); the rest on lanes.
    void make(uint32_t need, const int *score, const int *next_use, uint8_t *r, uint8_t *t) const {
        const uint32_t all = (1u << K) - 1, io = io_set();
        int i = 0;
        for (int k = 0; k < L; k++) t[i++] = (uint8_t)lbit[k];
        uint32_t used = io;
        // waves
        const int NW = TB - kLaneBits;
        int w[4], nw = 0;
        for (int pass = 0; pass < NW; pass++) {
            int best = -1;
            for (int b = 0; b < K; b++) {
                if (((used >> b) & 1) || ((need >> b) & 1)) continue;
                if (best < 0 || next_use[b] > next_use[best] || (next_use[b] == next_use[best] && b > best)) best = b;
            }
            if (best < 0)  // the group targets every non-row bit: take any (planner bounds prevent this)
                for (int b = K - 1; b >= 0 && best < 0; b--)
                    if (!((used >> b) & 1)) best = b;
            w[nw++] = best;
            used |= 1u << best;
        }
        // registers: highest score first
        int j = 0;
        while (j < RB) {
            int best = -1;
            for (int b = 0; b < K; b++) {
                if ((used >> b) & 1) continue;
                if (best < 0 || score[b] > score[best]) best = b;
            }
            r[j++] = (uint8_t)best;
            used |= 1u << best;
        }
        // Register r's amplitude is loaded and stored in register order, so the
        // register bits set the order in which a wave walks its tile in HBM:
        // highest qubit on register bit 0 measured fastest (12% over the
        // lowest-first and 20% over an arbitrary order, one-wave tiles at 30q).
        std::sort(r, r + RB, [&](uint8_t x, uint8_t y) { return qof[x] > qof[y]; });
        for (; j < 8; j++) r[j] = 0;
        // remaining lanes
        for (int b = 0; b < K; b++)
            if (!((used >> b) & 1)) { t[i++] = (uint8_t)b; used |= 1u << b; }
        for (int k = 0; k < NW; k++) t[kLaneBits + k] = (uint8_t)w[k];
        for (int k = TB; k < 12; k++) t[k] = 0;
        (void)all;
    }
    void set_layout(const uint8_t *r, const uint8_t *t, bool emit_op) {
        if (!memcmp(r, cr, 8) && !memcmp(t, ct, 12)) return;
        if (emit_op) {
            RegOp op{};
            op.kind = RO_TRANSPOSE;
            memcpy(op.lay_r, r, 8);
            memcpy(op.lay_t, t, 12);
            out.push_back(op);
            forget_diagonals();  // diagonal positions are layout-relative
        }
        memcpy(cr, r, 8);
        memcpy(ct, t, 12);
    }
    // One-wave tiles: before a 2x2 op on a lane-held local bit, exchange that
    // lane bit with a register bit (RO_XLR) when the pass's remaining ops on
    // it make that cheaper. Per-op costs in units of a lane-bit 2x2 op (the
    // partner of every amplitude through ds_bpermute, 128 per op): a
    // register-bit 2x2 op kRegCost, an exchange kXlrCost (half the amplitudes
    // cross lanes). The register given up holds the local bit with the fewest
    // remaining 2x2 ops, which then run on the lane; a row local bit (loaded
    // and stored on lanes 0..L-1) pays a second exchange to go back at the end.
    static constexpr double kRegCost = 0.18, kXlrCost = 0.5;
    void lane_to_register(const std::vector<int> &tgt_lb, size_t at) {
        const int b = tgt_lb[at];
        const int i = lane_index(b);
        if (i < 0) return;  // already on a register bit
        auto uses = [&](int lb) {
            int k = 0;
            for (size_t s = at; s < tgt_lb.size(); s++) k += tgt_lb[s] == lb;
            return k;
        };
        const int k = uses(b);
        bool row = false;
        for (int r = 0; r < L; r++) row = row || lbit[r] == b;
        int best = -1, best_m = 1 << 30;
        for (int j = 0; j < RB; j++) {
            const int m = uses(cr[j]);
            if (m < best_m) { best_m = m; best = j; }
        }
        if (best < 0) return;
        const double keep = k * 1.0;
        const double move = kXlrCost * (row ? 2 : 1) + k * kRegCost + best_m * (1.0 - kRegCost);
        if (move >= keep) return;
        RegOp op{};
        op.kind = RO_XLR;
        op.tgt = i;
        op.dbit = best;
        std::swap(ct[i], cr[best]);
        memcpy(op.lay_t, ct, sizeof op.lay_t);  // the kernel's lane-held local bits from here on
        forget_diagonals();  // diagonal positions are layout-relative
        out.push_back(op);
    }
    void controls(uint64_t cm, RegOp &op) const {
        for (; cm; cm &= cm - 1) {
            const int q = __builtin_ctzll(cm);
            const int lb = lbit[q];
            if (lb < 0) op.gcm |= uint64_t(1) << q;
            else if (reg_index(lb) >= 0) op.rcm |= 1u << reg_index(lb);
            else op.tcm |= 1u << thread_index(lb);
        }
    }
    // ---- Pauli frame: uncontrolled X, Y and Z gates cost no work --------
    // The tile's true amplitudes are F phi, phi what the registers hold and
    // F = i^fph X^fx Z^fz (masks over local bits, the X part left of the Z
    // part). X, Y and Z multiply into F; H and single-control CNOT and CZ
    // carry F through themselves (Clifford gates); any other gate U on a bit
    // with a frame runs conjugated, F_t^-1 U F_t, or the frame is applied
    // first (controls with an X part, controlled targets, parametric and LDS
    // gates). When the pass closes, the X part is an address flip at the
    // store, the Z part diagonal ops (folded into earlier diagonals where
    // possible) and i^fph joins the global phase.
    bool frames = false;
    uint32_t fx = 0, fz = 0;
    int fph = 0;
    Status flush_z(int b, uint32_t src) {  // registers get Z_b: F = ... X_b^x Z_b
        if (b < 0 || !((fz >> b) & 1)) return Status{};
        fz &= ~(1u << b);
        Lowered Z;
        Z.kind = LK_Z;
        Z.t0 = (uint32_t)qof[b];
        return emit_op(Z, src);
    }
    Status flush_x(int b, uint32_t src) {  // registers get X_b: X_b Z_b^z = (-1)^z Z_b^z X_b
        if (b < 0 || !((fx >> b) & 1)) return Status{};
        fx &= ~(1u << b);
        fph += 2 * ((fz >> b) & 1);
        Lowered X;
        X.kind = LK_X;
        X.t0 = (uint32_t)qof[b];
        return emit_op(X, src);
    }
    Status flush_bit(int b, uint32_t src) {
        Status e = flush_z(b, src);
        return e.ok() ? flush_x(b, src) : e;
    }
    // The 2x2 matrix of a one-qubit lowered gate (row-major, (re, im) pairs).
    static void matrix_of(const Lowered &G, double *m) {
        for (int k = 0; k < 8; k++) m[k] = 0;
        const double *c = G.c;
        switch (G.kind) {
        case LK_H: m[0] = m[2] = m[4] = c[0]; m[6] = -c[0]; break;
        case LK_X: m[2] = m[4] = 1; break;
        case LK_Y: m[2] = c[0]; m[3] = c[1]; m[4] = c[2]; m[5] = c[3]; break;
        case LK_RX: m[0] = m[6] = c[0]; m[2] = -c[1]; m[3] = -c[2]; m[4] = c[3]; m[5] = c[4]; break;
        case LK_RY: m[0] = m[6] = c[0]; m[2] = -c[1]; m[4] = c[1]; break;
        default: for (int k = 0; k < 8; k++) m[k] = c[k]; break;  // LK_U2
        }
    }
    Status emit(const Lowered &G, uint32_t src) {
        if (!frames || is_relabel(G)) return emit_op(G, src);
        const int t = lbit[G.t0];
        uint32_t cl = 0;  // controls inside the tile (local bits)
        bool ctl_out = false;
        for (uint64_t cm = G.ctrl_mask; cm; cm &= cm - 1) {
            const int q = __builtin_ctzll(cm);
            if (lbit[q] >= 0) cl |= 1u << lbit[q];
            else ctl_out = true;
        }
        const uint32_t tb = t >= 0 ? 1u << t : 0u;
        const int xt = t >= 0 ? (int)((fx >> t) & 1) : 0, zt = t >= 0 ? (int)((fz >> t) & 1) : 0;
        Status e;
        if (t < 0) {  // a diagonal on a qubit outside the tile: no frame there; X parts of controls out
            for (uint32_t c = cl; c; c &= c - 1)
                if (!(e = flush_x(__builtin_ctz(c), src)).ok()) return e;
            return emit_op(G, src);
        }
        if (G.kind == LK_SWAP || G.kind == LK_MATCH || G.slot >= 0) {
            for (uint32_t c = cl; c; c &= c - 1)
                if (!(e = flush_x(__builtin_ctz(c), src)).ok()) return e;
            if (!(e = flush_bit(t, src)).ok()) return e;
            if (G.kind == LK_SWAP || G.kind == LK_MATCH)
                if (!(e = flush_bit(lbit[G.t1], src)).ok()) return e;
            return emit_op(G, src);
        }
        if (G.ctrl_mask) {
            if (__builtin_popcountll(G.ctrl_mask) == 1 && !ctl_out && t >= 0 && (G.kind == LK_X || G.kind == LK_Z)) {
                const int c = __builtin_ctz(cl), xc = (int)((fx >> c) & 1);
                if (G.kind == LK_X) {  // CNOT: X_c -> X_c X_t, Z_t -> Z_c Z_t
                    if (xc) fx ^= tb;
                    if (zt) fz ^= 1u << c;
                } else {  // CZ: X_c -> X_c Z_t, X_t -> Z_c X_t, sign (-1)^(x_c x_t)
                    if (xc) fz ^= tb;
                    if (xt) fz ^= 1u << c;
                    fph += 2 * (xc & xt);
                }
                return emit_op(G, src);
            }
            // Z parts of controls commute with a controlled gate, X parts are
            // applied first. On the target, F_t C(U) F_t = C(F_t U F_t): the
            // part of F_t that commutes with U stays (X with X, Z with a
            // diagonal), the rest is applied first.
            for (uint32_t c = cl; c; c &= c - 1)
                if (!(e = flush_x(__builtin_ctz(c), src)).ok()) return e;
            const bool diag = G.kind == LK_Z || G.kind == LK_PHASE || G.kind == LK_RZ;
            if (G.kind == LK_X) e = flush_z(t, src);
            else if (diag) e = flush_x(t, src);
            else e = flush_bit(t, src);
            if (!e.ok()) return e;
            return emit_op(G, src);
        }
        switch (G.kind) {
        case LK_X: fx ^= tb; return Status{};
        case LK_Z: fph += 2 * xt; fz ^= tb; return Status{};  // Z_t X^x = (-1)^x X^x Z_t
        case LK_Y:  // Y = i X Z: Y F = i^(1 + 2 x_t) X^(x ^ t) Z^(z ^ t)
            if (G.c[0] == 0 && G.c[1] == -1 && G.c[2] == 0 && G.c[3] == 1) {
                fph += 1 + 2 * xt;
                fx ^= tb;
                fz ^= tb;
                return Status{};
            }
            break;
        case LK_H:  // H X^a Z^b H = (-1)^(ab) X^b Z^a
            fx = (fx & ~tb) | (zt ? tb : 0u);
            fz = (fz & ~tb) | (xt ? tb : 0u);
            fph += 2 * (xt & zt);
            return emit_op(G, src);
        default: break;
        }
        if (!xt && !zt) return emit_op(G, src);
        // conjugated: the registers get Z^z X^x U X^x Z^z
        Lowered C = G;
        if (G.kind == LK_Z || G.kind == LK_PHASE || G.kind == LK_RZ) {  // diagonal: X swaps the entries, Z keeps them
            if (xt) {
                double d[4];
                if (G.kind == LK_Z) { d[0] = 1; d[1] = 0; d[2] = -1; d[3] = 0; }
                else if (G.kind == LK_PHASE) { d[0] = 1; d[1] = 0; d[2] = G.c[0]; d[3] = G.c[1]; }
                else for (int k = 0; k < 4; k++) d[k] = G.c[k];
                C.kind = LK_RZ;
                C.c[0] = d[2]; C.c[1] = d[3]; C.c[2] = d[0]; C.c[3] = d[1];
            }
            return emit_op(C, src);
        }
        double m[8];
        matrix_of(G, m);
        if (xt) {  // X M X: rows and columns swapped
            double n[8] = {m[6], m[7], m[4], m[5], m[2], m[3], m[0], m[1]};
            memcpy(m, n, sizeof n);
        }
        if (zt) { m[2] = -m[2]; m[3] = -m[3]; m[4] = -m[4]; m[5] = -m[5]; }  // Z M Z: off-diagonals negated
        C.kind = LK_U2;
        for (int k = 0; k < 8; k++) C.c[k] = m[k];
        return emit_op(C, src);
    }
    // The frame out at the end of the pass: Z part as diagonals (folded into
    // the last foldable diagonal on the bit where there is one), the phase
    // into the global phase; returns the X part as a qubit mask (the store's
    // address flip).
    Status close_frame(uint32_t src, uint64_t &flip) {
        flip = 0;
        if (!frames) return Status{};
        for (uint32_t z = fz; z; z &= z - 1) {
            const int b = __builtin_ctz(z);
            const int j = reg_index(b), i = thread_index(b);
            const int loc = j >= 0 ? 0 : 1, pos = j >= 0 ? j : i;
            if (pos >= 0 && last_diag[loc][pos] >= 0) {
                RegOp zop{};
                zop.kind = RO_Z;
                fold_diagonal(out[(size_t)last_diag[loc][pos]], zop);
                fz &= ~(1u << b);
                continue;
            }
            Status e = flush_z(b, src);
            if (!e.ok()) return e;
        }
        for (uint32_t x = fx; x; x &= x - 1) flip |= uint64_t(1) << qof[__builtin_ctz(x)];
        fx = 0;
        const int k = ((fph % 4) + 4) % 4;
        static const double ipow[4][2] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
        if (k && gphase) {
            const double re = gphase[0] * ipow[k][0] - gphase[1] * ipow[k][1];
            const double im = gphase[0] * ipow[k][1] + gphase[1] * ipow[k][0];
            gphase[0] = re;
            gphase[1] = im;
        }
        fph = 0;
        return Status{};
    }
    Status emit_op(const Lowered &G, uint32_t src) {
        RegOp op{};
        op.src = src;
        const bool param = G.slot >= 0;  // no rewrite that depends on the angle's value
        op.param = param ? G.slot + 1 : 0;
        switch (G.kind) {
        case LK_Z: case LK_PHASE: case LK_RZ: {
            if (G.kind == LK_Z) op.kind = RO_Z;
            else if (G.kind == LK_PHASE) { op.kind = RO_P1; op.c[0] = 1; op.c[2] = G.c[0]; op.c[3] = G.c[1]; }
            else if (gphase && !param && G.ctrl_mask == 0 && (G.c[0] != 0 || G.c[1] != 0)) {
                // diag(p0, p1) = p0 diag(1, p1 / p0)
                const double d = G.c[0] * G.c[0] + G.c[1] * G.c[1];
                op.kind = RO_P1; op.c[0] = 1;
                op.c[2] = (G.c[2] * G.c[0] + G.c[3] * G.c[1]) / d;
                op.c[3] = (G.c[3] * G.c[0] - G.c[2] * G.c[1]) / d;
                const double re = gphase[0] * G.c[0] - gphase[1] * G.c[1];
                const double im = gphase[0] * G.c[1] + gphase[1] * G.c[0];
                gphase[0] = re; gphase[1] = im;
            } else { op.kind = RO_DG; for (int k = 0; k < 4; k++) op.c[k] = G.c[k]; }
            const int lb = lbit[G.t0];
            if (lb < 0) { op.dloc = 2; op.dbit = (int)G.t0; }
            else if (reg_index(lb) >= 0) { op.dloc = 0; op.dbit = reg_index(lb); }
            else { op.dloc = 1; op.dbit = thread_index(lb); }
            controls(G.ctrl_mask, op);
            if (op.kind == RO_Z && op.tcm == 0 && __builtin_popcount(op.rcm) == 1) {
                op.kind = RO_ZC1;
                op.lp0 = __builtin_ctz(op.rcm);
            }
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
            for (uint64_t cm = G.ctrl_mask; cm; cm &= cm - 1) {
                const int q = __builtin_ctzll(cm);
                if (lbit[q] >= 0) op.lcm |= 1u << lbit[q];
                else op.gcm |= uint64_t(1) << q;
            }
        } break;
        default: {
            const int lb = lbit[G.t0];
            const int j = reg_index(lb), i = lane_index(lb);
            if (j >= 0) op.tgt = j;
            else if (i >= 0) op.tgt = kLaneTarget + i;
            else return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "tile planner: target on a wave bit");
            switch (G.kind) {
            case LK_X: op.kind = RO_X; break;  // one register control: RO_XC1 below
            case LK_H:  // [[h, h], [h, -h]]
                op.kind = RO_RR; op.c[0] = G.c[0]; op.c[1] = G.c[0]; op.c[2] = G.c[0]; op.c[3] = -G.c[0]; break;
            case LK_RY:  // [[c, -s], [s, c]]
                op.kind = RO_RR; op.c[0] = G.c[0]; op.c[1] = -G.c[1]; op.c[2] = G.c[1]; op.c[3] = G.c[0]; break;
            case LK_Y:  // [[0, (c0,c1)], [(c2,c3), 0]]; as lowered: [[0, -i], [i, 0]]
                if (G.c[0] == 0 && G.c[1] == -1 && G.c[2] == 0 && G.c[3] == 1) { op.kind = RO_Y; break; }
                op.kind = RO_GEN;
                op.c[2] = G.c[0]; op.c[3] = G.c[1]; op.c[4] = G.c[2]; op.c[5] = G.c[3];
                break;
            case LK_RX:  // [[c0, -(c1,c2)], [(c3,c4), c0]]; as lowered [[cos, -i sin], [-i sin, cos]]
                if (G.c[1] == 0 && G.c[3] == 0) {
                    op.kind = RO_RXI;
                    op.c[0] = G.c[0]; op.c[1] = -G.c[2]; op.c[2] = G.c[4]; op.c[3] = G.c[0];
                    break;
                }
                op.kind = RO_GEN;
                op.c[0] = G.c[0]; op.c[2] = -G.c[1]; op.c[3] = -G.c[2]; op.c[4] = G.c[3]; op.c[5] = G.c[4];
                op.c[6] = G.c[0];
                break;
            default:  // U2, row-major (c0,c1) (c2,c3) / (c4,c5) (c6,c7)
                if (G.c[1] == 0 && G.c[3] == 0 && G.c[5] == 0 && G.c[7] == 0) {  // real
                    op.kind = RO_RR; op.c[0] = G.c[0]; op.c[1] = G.c[2]; op.c[2] = G.c[4]; op.c[3] = G.c[6];
                } else if (G.c[1] == 0 && G.c[2] == 0 && G.c[4] == 0 && G.c[7] == 0) {  // real diagonal, imaginary off-diagonal
                    op.kind = RO_RXI; op.c[0] = G.c[0]; op.c[1] = G.c[3]; op.c[2] = G.c[5]; op.c[3] = G.c[6];
                } else {
                    op.kind = RO_GEN;
                    for (int k = 0; k < 8; k++) op.c[k] = G.c[k];
                }
                break;
            }
            controls(G.ctrl_mask, op);
            if (op.kind == RO_X && op.tcm == 0 && __builtin_popcount(op.rcm) == 1) {
                op.kind = RO_XC1;
                op.lp0 = __builtin_ctz(op.rcm);
            }
            if (f32 && !param && op.kind == RO_RR && op.tgt < kLaneTarget && !op.rcm && !op.tcm && !op.gcm &&
                op.c[0] == op.c[1] && op.c[0] == op.c[2] && op.c[0] == -op.c[3] && op.c[0] != 0) {
                // a multiple of [[1, 1], [1, -1]] (H, and the X basis changes of Trotter steps)
                const double f = op.c[0];
                op.kind = RO_HS;
                for (int k = 0; k < 8; k++) op.c[k] = 0;
                out.push_back(op);
                defer_factor(f, op.src);
                last_diag[0][op.tgt] = -1;
                return Status{};
            }
            if (f32 && !param) shear_form(op);
        } break;
        }
        if (param && (op.kind == RO_P1 || op.kind == RO_DG)) {
            // neither folds into an earlier diagonal nor takes later ones
            if (!op.rcm && !op.tcm && !op.gcm) last_diag[op.dloc][op.dbit] = -1;
        } else if (op.kind == RO_Z || op.kind == RO_P1 || op.kind == RO_DG) {
            if (!op.rcm && !op.tcm && !op.gcm) {
                int &last = last_diag[op.dloc][op.dbit];
                if (last >= 0) {
                    fold_diagonal(out[(size_t)last], op);
                    return Status{};
                }
                last = (int)out.size();
            }
        } else if (op.kind <= RO_Y || op.kind == RO_XC1 || op.kind == RO_RRS || op.kind == RO_RXS) {  // non-diagonal
            if (op.tgt >= kLaneTarget) last_diag[1][op.tgt - kLaneTarget] = -1;  // thread-index key
            else last_diag[0][op.tgt] = -1;
        } else if (op.kind != RO_ZC1) {
            forget_diagonals();
        }
        out.push_back(op);
        return Status{};
    }
    // A rotation on a register bit with no control inside the tile: three
    // in-place shears per component pair (3 multiply-adds instead of 4). The
    // angle is brought into [-pi/2, pi/2] (tan(phi/2) bounded by 1) by
    // R(phi) = -R(phi - pi), the sign going to the global phase, which needs
    // the op uncontrolled; otherwise wide angles keep the plain form.
    void shear_form(RegOp &op) {
        if (op.tgt >= kLaneTarget || op.rcm || op.tcm) return;
        double c, s;
        if (op.kind == RO_RR && op.c[0] == op.c[3] && op.c[1] == -op.c[2]) { c = op.c[0]; s = op.c[2]; }
        else if (op.kind == RO_RXI && op.c[0] == op.c[3] && op.c[1] == op.c[2]) { c = op.c[0]; s = op.c[1]; }
        else return;
        // a multiple r of a rotation (uncontrolled only: r joins the deferred factors)
        const double r2 = c * c + s * s;
        double r = 1.0;
        if (std::fabs(r2 - 1.0) > 1e-12) {
            if (op.gcm || r2 == 0.0) return;
            r = std::sqrt(r2);
            c /= r;
            s /= r;
        }
        double phi = std::atan2(s, c);
        if (std::fabs(phi) > M_PI / 2) {
            if (op.gcm) return;
            phi -= std::copysign(M_PI, phi);
            gphase[0] = -gphase[0];
            gphase[1] = -gphase[1];
        }
        op.kind = op.kind == RO_RR ? RO_RRS : RO_RXS;
        op.c[0] = -std::tan(phi / 2);
        op.c[1] = std::sin(phi);
        op.c[2] = op.c[3] = 0;
        if (r != 1.0) defer_factor(r, op.src);
    }
    // A scalar factor of an uncontrolled op, applied later: multiplied into
    // the pending product; every kMaxUnscaled factors the product is applied
    // by a diagonal on every amplitude (emitted here).
    void defer_factor(double f, uint32_t src) {
        *hfactor *= f;
        if (++*unscaled < kMaxUnscaled) return;
        RegOp sc{};
        sc.kind = RO_DG; sc.dloc = 0; sc.dbit = 0;
        sc.c[0] = sc.c[2] = *hfactor;
        sc.src = src;
        *unscaled = 0;
        *hfactor = 1.0;
        forget_diagonals();  // not folded: it must act on every amplitude now
        out.push_back(sc);
    }
    // diag(p0, p1) of an uncontrolled diagonal op
    static void diag_of(const RegOp &o, double *p) {
        if (o.kind == RO_Z) { p[0] = 1; p[1] = 0; p[2] = -1; p[3] = 0; }
        else if (o.kind == RO_P1) { p[0] = 1; p[1] = 0; p[2] = o.c[2]; p[3] = o.c[3]; }
        else for (int k = 0; k < 4; k++) p[k] = o.c[k];
    }
    // prev <- the product of prev and op (both uncontrolled, same position);
    // an identity product becomes a no-op dropped when the pass is closed
    static void fold_diagonal(RegOp &prev, const RegOp &op) {
        double a[4], b[4];
        diag_of(prev, a);
        diag_of(op, b);
        const double p0r = a[0] * b[0] - a[1] * b[1], p0i = a[0] * b[1] + a[1] * b[0];
        const double p1r = a[2] * b[2] - a[3] * b[3], p1i = a[2] * b[3] + a[3] * b[2];
        prev.c[0] = p0r; prev.c[1] = p0i; prev.c[2] = p1r; prev.c[3] = p1i;
        if (p0r == 1 && p0i == 0) {
            if (p1r == 1 && p1i == 0) prev.kind = RO_NOP;
            else if (p1r == -1 && p1i == 0) prev.kind = RO_Z;
            else prev.kind = RO_P1;
        } else {
            prev.kind = RO_DG;
        }
    }
};

// One-wave tiles: a diagonal op whose factor depends on the lane alone
// (target on a lane bit or outside the tile, no register controls, not
// parametric) can multiply a per-lane pending phase (RO_LPH: a few
// instructions per wave) instead of every amplitude of the lane. The phase
// commutes with every op that moves no amplitude between lanes it depends on,
// so it is applied to the amplitudes (RO_LPF) just before the first op that
// does (a 2x2 op on such a lane bit, an exchange of one) or before the store.
// A run of deferrable ops between two such points is deferred when that costs
// less than applying each op (in practice: two or more ops).
static void defer_lane_phases(std::vector<RegOp> &ops) {
    auto deferrable = [](const RegOp &op) {
        return (op.kind == RO_Z || op.kind == RO_P1 || op.kind == RO_DG) && op.dloc != 0 && op.rcm == 0 && op.param == 0;
    };
    // cost per amplitude in VALU instructions: sign flips or a complex product, plus the selects of lane controls
    auto cost = [](const RegOp &op) { return (op.kind == RO_Z ? 2.0 : 4.0) + (op.tcm ? 2.0 : 0.0); };
    auto lanes_moved = [](const RegOp &op) -> uint32_t {
        if ((op.kind <= RO_Y || op.kind == RO_XC1) && op.tgt >= kLaneTarget) return 1u << (op.tgt - kLaneTarget);
        if (op.kind == RO_XLR) return 1u << op.tgt;
        if (op.kind == RO_XLL) return (1u << op.tgt) | (1u << op.dbit);
        return 0;
    };
    std::vector<RegOp> out;
    out.reserve(ops.size() + 8);
    std::vector<size_t> run;  // positions in `out` of the run's deferrable ops
    uint32_t dep = 0;         // lane bits the run's factors depend on
    bool sign = true;         // every factor of the run is +-1
    double saved = 0;
    auto close = [&](size_t at) {
        const double flush = sign ? 2.0 : 4.0, per_op = 0.25;
        if (run.size() >= 2 && saved > flush + per_op * (double)run.size()) {
            for (size_t k : run) {
                RegOp &o = out[k];
                double d[4];
                RegPlanner::diag_of(o, d);
                o.kind = RO_LPH;
                for (int q = 0; q < 4; q++) o.c[q] = d[q];
            }
            RegOp f{};
            f.kind = RO_LPF;
            f.lp0 = sign ? 1 : 0;
            f.src = out[run.back()].src;
            out.insert(out.begin() + (std::ptrdiff_t)at, f);
        }
        run.clear();
        dep = 0;
        sign = true;
        saved = 0;
    };
    for (const RegOp &op : ops) {
        if (dep & lanes_moved(op)) close(out.size());
        if (deferrable(op)) {
            run.push_back(out.size());
            saved += cost(op);
            if (op.dloc == 1) dep |= 1u << op.dbit;
            dep |= op.tcm;
            sign = sign && op.kind == RO_Z;
        }
        out.push_back(op);
    }
    close(out.size());
    ops.swap(out);
}

struct Phase {
    std::vector<size_t> ops;  // indices into the pass members, in execution order
    uint32_t need = 0;        // local bits the group's 2x2 gates target
};

}  // namespace

static Status plan_tiled_list(const State &s, const std::vector<Lowered> &gates, int K, int TB, ExecPlan &plan) {
    if (K <= 0) K = default_tile_qubits(s);
    if (TB <= 0) TB = default_tile_threads_log2(s.prec, K);
    if (TB != 6 && TB != 8 && TB != 9) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "tile threads out of range");
    const int kmax = s.prec == QI_C128 ? 12 : 13;
    if (K < 10 || K > kmax) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "tile width out of range");
    // one-wave tiles keep up to 128 amplitude components per lane in registers
    if (TB == 6 && K > (s.prec == QI_C64 ? 12 : 11))
        return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0,
                        "one-wave tiles hold at most 12 (complex64) or 11 (complex128) tile qubits");
    if ((int)s.n < K) {  // state smaller than a tile: stream gate by gate
        for (const Lowered &L : gates) plan.steps.push_back(ExecStep::gate(L));
        return Status{};
    }
    const int low = tile_low_qubits(s.prec), RB = K - TB;
    std::vector<GateSets> gs(gates.size());
    for (size_t g = 0; g < gates.size(); g++) gs[g] = gate_sets(gates[g]);
    std::vector<char> done(gates.size(), 0);
    // One-wave tiles have no LDS: controlled SWAPs and matchgates run as
    // streaming steps between passes (later gates pass them only if they commute).
    std::vector<char> streamed(gates.size(), 0);
    if (TB == 6)
        for (size_t g = 0; g < gates.size(); g++)
            streamed[g] = (gates[g].kind == LK_SWAP || gates[g].kind == LK_MATCH) && !is_relabel(gates[g]);
    double gphase[2] = {1.0, 0.0};
    int unscaled = 0;      // unscaled H ops since the last scaling
    double hfactor = 1.0;  // their factors, not applied yet
    size_t first = 0;
    while (first < gates.size()) {
        if (streamed[first]) {
            plan.steps.push_back(ExecStep::gate(gates[first]));
            done[first] = 1;
            while (first < gates.size() && done[first]) first++;
            continue;
        }
        std::vector<size_t> members;
        const uint64_t q = select_pass(gs, streamed, done, first, K, (int)s.n, low, s.n >= kSearchQubits, members);
        RegPlanner P;
        P.K = K; P.RB = RB; P.L = low; P.TB = TB;
        P.f32 = s.prec == QI_C64;
        P.gphase = gphase;
        P.unscaled = &unscaled;
        P.hfactor = &hfactor;
        P.forget_diagonals();
        // one-wave tiles only: every local bit is a lane or register bit there,
        // so a frame can be applied wherever it sits (not so on wave bits)
        P.frames = gphase != nullptr && TB == kLaneBits;
        for (int b = 0; b < 64; b++) P.lbit[b] = -1;
        RegPass rp{};
        rp.qmask = q;
        rp.K = K;
        rp.tb = TB;
        for (int b = 0, j = 0; b < (int)s.n; b++)
            if ((q >> b) & 1) { P.lbit[b] = j; P.qof[j] = b; j++; }
        // ---- phases: ops whose targets fit the register and lane bits (all but
        // the two wave bits), commuting-greedy; relabels tracked on the way
        const int CAP = K - (TB - kLaneBits);
        std::vector<Phase> phases;
        std::vector<std::vector<int>> phase_lb;  // labels (qubit -> local bit) at each phase start
        {
            int lb[64];
            memcpy(lb, P.lbit, sizeof lb);
            std::vector<char> pd(members.size(), 0);
            size_t left = members.size();
            while (left) {
                Phase ph;
                phase_lb.emplace_back(lb, lb + 64);
                Blocked blk;
                uint32_t io_u = 0;  // row-qubit bits over the phase (relabels may move them)
                for (size_t i = 0; i < members.size(); i++) {
                    if (pd[i]) continue;
                    const Lowered &G = gates[members[i]];
                    const GateSets &g = gs[members[i]];
                    if (blk.blocks(g)) { blk.add(g); continue; }
                    const uint32_t need = is_pair(G) ? 1u << lb[G.t0] : 0u;
                    uint32_t io = io_u;  // the row qubits stay on lanes: they count against the capacity
                    for (int k = 0; k < low; k++) io |= 1u << lb[k];
                    if (__builtin_popcount(ph.need | need | io) > CAP) { blk.add(g); continue; }
                    io_u = io;
                    ph.need |= need;
                    ph.ops.push_back(i);
                    pd[i] = 1;
                    left--;
                    if (is_relabel(G)) std::swap(lb[G.t0], lb[G.t1]);
                }
                phases.push_back(std::move(ph));
            }
        }
        // ---- per phase: scores for register placement (targets count most, then controls)
        std::vector<std::array<int, 16>> score(phases.size());
        for (size_t p = 0; p < phases.size(); p++) {
            score[p].fill(0);
            int lb[64];
            std::copy(phase_lb[p].begin(), phase_lb[p].end(), lb);
            for (size_t i : phases[p].ops) {
                const Lowered &G = gates[members[i]];
                if (is_pair(G)) score[p][lb[G.t0]] += 4;
                for (uint64_t cm = G.ctrl_mask; cm; cm &= cm - 1) {
                    const int q = __builtin_ctzll(cm);
                    if (lb[q] >= 0) score[p][lb[q]] += 1;
                }
                if (G.kind == LK_Z || G.kind == LK_PHASE || G.kind == LK_RZ)
                    if (lb[G.t0] >= 0) score[p][lb[G.t0]] += 1;
                if (is_relabel(G)) std::swap(lb[G.t0], lb[G.t1]);
            }
        }
        // next use of each local bit as a target, from each phase on
        auto next_use = [&](size_t from, int *nu) {
            for (int b = 0; b < 16; b++) nu[b] = 1 << 30;
            for (size_t p = phases.size(); p-- > from;)
                for (int b = 0; b < K; b++)
                    if ((phases[p].need >> b) & 1) nu[b] = (int)p;
        };
        // ---- layouts: a transpose only when a target sits on a wave bit
        int nu[16];
        next_use(0, nu);
        {
            uint8_t r[8], t[12];
            P.make(phases.empty() ? 0u : phases[0].need, phases.empty() ? nu : score[0].data(), nu, r, t);
            P.set_layout(r, t, false);
            memcpy(rp.lay_r, P.cr, 8);
            memcpy(rp.lay_t, P.ct, 12);
            for (int j = 0; j < RB; j++) rp.qr_in[j] = (uint8_t)P.qof[P.cr[j]];
            for (int i = 0; i < TB; i++) rp.qt_in[i] = (uint8_t)P.qof[P.ct[i]];
        }
        // The local bit each 2x2 op targets, in emission order (relabels move
        // qubits between local bits, not data): the lookahead of the register
        // allocation below.
        std::vector<int> tgt_lb;
        {
            int lb[64];
            memcpy(lb, P.lbit, sizeof lb);
            for (size_t p = 0; p < phases.size(); p++)
                for (size_t i : phases[p].ops) {
                    const Lowered &G = gates[members[i]];
                    // (uncontrolled X and Y only change the frame: no op)
                    const bool framed = P.frames && G.ctrl_mask == 0 && G.slot < 0 && (G.kind == LK_X || G.kind == LK_Y);
                    tgt_lb.push_back(is_pair(G) && !is_relabel(G) && !framed ? lb[G.t0] : -1);
                    if (is_relabel(G)) std::swap(lb[G.t0], lb[G.t1]);
                }
        }
        size_t seq = 0;
        for (size_t p = 0; p < phases.size(); p++) {
            if (phases[p].need & P.wave_set()) {
                next_use(p, nu);
                uint8_t r[8], t[12];
                P.make(phases[p].need, score[p].data(), nu, r, t);
                P.set_layout(r, t, true);
            }
            for (size_t i : phases[p].ops) {
                if (TB == kLaneBits && tgt_lb[seq] >= 0) P.lane_to_register(tgt_lb, seq);
                Status e = P.emit(gates[members[i]], (uint32_t)members[i]);
                if (!e.ok()) return e;
                seq++;
            }
        }
        {
            Status e = P.close_frame(members.empty() ? 0u : (uint32_t)members.back(), rp.flip);
            if (!e.ok()) return e;
        }
        // ---- store: row qubits back on lanes 0..L-1 (relabels may have moved them)
        if (TB == kLaneBits) {  // one-wave tiles: lane/register exchanges, lane k in turn
            for (int k = 0; k < low; k++) {
                const int lb = P.lbit[k];
                if (P.ct[k] == lb) continue;
                RegOp op{};
                const int i2 = P.lane_index(lb), j = P.reg_index(lb);
                op.tgt = k;
                if (i2 >= 0) { op.kind = RO_XLL; op.dbit = i2; std::swap(P.ct[k], P.ct[i2]); }
                else { op.kind = RO_XLR; op.dbit = j; std::swap(P.ct[k], P.cr[j]); }
                memcpy(op.lay_t, P.ct, sizeof op.lay_t);
                P.out.push_back(op);
            }
        }
        if (!P.io_ok()) {
            int zero[16] = {0}, far[16];
            for (int b = 0; b < 16; b++) far[b] = 1 << 30;
            uint8_t r[8], t[12];
            P.make(0, zero, far, r, t);
            P.set_layout(r, t, true);
            if (!P.io_ok()) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "tile planner: store layout");
        }
        for (int j = 0; j < RB; j++) rp.qr_out[j] = (uint8_t)P.qof[P.cr[j]];
        for (int i = 0; i < TB; i++) rp.qt_out[i] = (uint8_t)P.qof[P.ct[i]];
        // SWAP relabels or transposes may send an amplitude to an address another
        // thread loads: the kernel then waits for every load before storing
        rp.sync_store = memcmp(rp.qt_out, rp.qt_in, 12) != 0 || memcmp(rp.qr_out, rp.qr_in, 8) != 0 || rp.flip != 0;
        P.out.erase(std::remove_if(P.out.begin(), P.out.end(), [](const RegOp &op) { return op.kind == RO_NOP; }),
                    P.out.end());
        if (TB == kLaneBits) defer_lane_phases(P.out);
        rp.op0 = (int)plan.reg_ops.size();
        rp.nops = (int)P.out.size();
        for (RegOp &op : P.out) {
            // (c64) a one-sided phase on a register bit as a rotation of (re, im) by shears,
            // once no later diagonal can fold into it
            if (P.f32 && op.kind == RO_P1 && !op.param && op.dloc == 0 && !op.rcm && !op.tcm &&
                std::fabs(op.c[2] * op.c[2] + op.c[3] * op.c[3] - 1.0) <= 1e-12) {
                const double phi = std::atan2(op.c[3], op.c[2]);
                if (std::fabs(phi) <= M_PI / 2) {
                    op.kind = RO_P1S;
                    op.c[0] = -std::tan(phi / 2);
                    op.c[1] = std::sin(phi);
                    op.c[2] = op.c[3] = 0;
                }
            }
            for (int q = 0; q < 8; q++) op.cf[q] = (float)op.c[q];
        }
        plan.reg_ops.insert(plan.reg_ops.end(), P.out.begin(), P.out.end());
        RegOp end{};
        end.kind = RO_NOP;
        plan.reg_ops.push_back(end);  // the kernel reads one op ahead
        plan.steps.push_back(ExecStep::reg_step(rp));
        while (first < gates.size() && done[first]) first++;
    }
    gphase[0] *= hfactor;
    gphase[1] *= hfactor;
    if (gphase[0] != 1.0 || gphase[1] != 0.0) {
        // the collected global phase, as a last op of the last pass (whose ops
        // close the op table): it commutes with everything
        for (size_t k = plan.steps.size(); k-- > 0;) {
            if (plan.steps[k].type != ExecStep::REG) continue;
            RegOp op{};
            op.kind = RO_DG; op.dloc = 0; op.dbit = 0;
            op.c[0] = op.c[2] = gphase[0];
            op.c[1] = op.c[3] = gphase[1];
            for (int q = 0; q < 8; q++) op.cf[q] = (float)op.c[q];
            plan.reg_ops.insert(plan.reg_ops.end() - 1, op);
            plan.steps[k].reg.nops++;
            break;
        }
    }
    return Status{};
}

// Uncontrolled SWAPs as one relabeling of the whole circuit: a SWAP moves no
// data, it renames the two qubits for every later gate (map: logical qubit ->
// the physical bit holding it). The net renaming is undone at the end by
// SWAPs on physical bits, emitted as transpositions in ascending or
// descending order of the logical qubit put back in place. A pass then never
// needs both qubits of a SWAP in its tile; the price is the final
// permutation (a few passes).
static std::vector<Lowered> relabel_swaps(const std::vector<Lowered> &in, int n, bool descending,
                                          int *holds_out = nullptr) {
    int at[64], holds[64];  // logical -> physical, physical -> logical
    for (int q = 0; q < 64; q++) at[q] = holds[q] = q;
    std::vector<Lowered> out;
    out.reserve(in.size() + n);
    for (const Lowered &L : in) {
        if (is_relabel(L)) {
            std::swap(at[L.t0], at[L.t1]);
            holds[at[L.t0]] = (int)L.t0;
            holds[at[L.t1]] = (int)L.t1;
            continue;
        }
        Lowered M = L;
        M.t0 = (uint32_t)at[L.t0];
        M.t1 = (uint32_t)at[L.t1];
        M.ctrl_mask = 0;
        for (uint64_t c = L.ctrl_mask; c; c &= c - 1) M.ctrl_mask |= uint64_t(1) << at[__builtin_ctzll(c)];
        out.push_back(M);
    }
    if (holds_out) {  // the caller moves the data: physical bit p holds logical qubit holds_out[p]
        for (int p = 0; p < 64; p++) holds_out[p] = holds[p];
        return out;
    }
    for (int i = 0; i < n; i++) {
        const int l = descending ? n - 1 - i : i;
        if (at[l] == l) continue;
        const int p = at[l], m = holds[l];  // logical l sits at p; bit l holds logical m
        Lowered S;
        S.kind = LK_SWAP;
        S.t0 = (uint32_t)l;
        S.t1 = (uint32_t)p;
        out.push_back(S);
        at[l] = l; holds[l] = l;
        at[m] = p; holds[p] = m;
    }
    return out;
}

static void append_plan(ExecPlan &dst, const ExecPlan &src) {
    const int32_t base = (int32_t)dst.reg_ops.size();
    for (ExecStep st : src.steps) {
        if (st.type == ExecStep::REG) st.reg.op0 += base;
        dst.steps.push_back(st);
    }
    dst.reg_ops.insert(dst.reg_ops.end(), src.reg_ops.begin(), src.reg_ops.end());
}

// The gate list as given and with SWAPs relabeled (both emission orders of
// the final permutation): the plan with the fewest launches wins (the given
// list on ties).
// The last pass of an out-of-place plan: the state moves back to its own
// buffer with physical bit p stored at bit pi[p]. Its tile holds qubits 0..3
// (lanes 0..3 at the load: contiguous reads) and the four qubits that go to bits
// 0..3 (brought onto lanes 0..3 by lane/register exchanges: contiguous writes);
// the other tile qubits land where pi sends them through the store layout and
// the bits outside the tile through the store's base permutation.
static void permutation_pass(int n, int K, const int *pi, ExecPlan &plan) {
    int src_of[64];  // src_of[d]: the physical bit stored at bit d
    for (int p = 0; p < n; p++) src_of[pi[p]] = p;
    uint64_t q = 0xFull;
    for (int d = 0; d < 4; d++) q |= uint64_t(1) << src_of[d];
    for (int b = 0; b < n && __builtin_popcountll(q) < K; b++) q |= uint64_t(1) << b;
    int qof[16], lb[64];
    for (int b = 0, j = 0; b < n; b++) {
        lb[b] = -1;
        if ((q >> b) & 1) { lb[b] = j; qof[j] = b; j++; }
    }
    constexpr int TB = kLaneBits;
    const int RB = K - TB;
    // load layout: qubits 0..3 on lanes 0..3; the qubits headed for bits 0..3 next
    // (registers first, where an exchange reaches them), then the rest
    uint8_t ct[12] = {0}, cr[8] = {0};
    std::vector<int> rest;
    for (int d = 0; d < 4; d++)
        if (src_of[d] > 3) rest.push_back(lb[src_of[d]]);
    for (int j = 0; j < K; j++)
        if (qof[j] > 3 && std::find(rest.begin(), rest.end(), j) == rest.end()) rest.push_back(j);
    for (int i = 0; i < 4; i++) ct[i] = (uint8_t)lb[i];
    size_t r = 0;
    for (int j = 0; j < RB; j++) cr[j] = (uint8_t)rest[r++];
    for (int i = 4; i < TB; i++) ct[i] = (uint8_t)rest[r++];
    RegPass rp{};
    rp.qmask = q;
    rp.K = K;
    rp.tb = TB;
    memcpy(rp.lay_r, cr, 8);
    memcpy(rp.lay_t, ct, 12);
    for (int j = 0; j < RB; j++) rp.qr_in[j] = (uint8_t)qof[cr[j]];
    for (int i = 0; i < TB; i++) rp.qt_in[i] = (uint8_t)qof[ct[i]];
    std::vector<RegOp> ops;
    for (int i = 0; i < 4; i++) {  // lane i gets the qubit stored at bit i
        const int want = lb[src_of[i]];
        if (ct[i] == want) continue;
        RegOp op{};
        op.tgt = i;
        int k = -1, j = -1;
        for (int x = 0; x < TB; x++) if (ct[x] == want) k = x;
        for (int x = 0; x < RB; x++) if (cr[x] == want) j = x;
        if (k >= 0) { op.kind = RO_XLL; op.dbit = k; std::swap(ct[i], ct[k]); }
        else { op.kind = RO_XLR; op.dbit = j; std::swap(ct[i], cr[j]); }
        memcpy(op.lay_t, ct, sizeof op.lay_t);
        ops.push_back(op);
    }
    for (int j = 0; j < RB; j++) rp.qr_out[j] = (uint8_t)pi[qof[cr[j]]];
    for (int i = 0; i < TB; i++) rp.qt_out[i] = (uint8_t)pi[qof[ct[i]]];
    rp.oop = 1;
    rp.permute = 1;
    for (int p = 0; p < 64; p++) rp.perm[p] = (uint8_t)(p < n ? pi[p] : p);
    rp.op0 = (int)plan.reg_ops.size();
    rp.nops = (int)ops.size();
    plan.reg_ops.insert(plan.reg_ops.end(), ops.begin(), ops.end());
    RegOp end{};
    end.kind = RO_NOP;
    plan.reg_ops.push_back(end);
    plan.steps.push_back(ExecStep::reg_step(rp));
}

Status plan_tiled(const State &s, const std::vector<Lowered> &gates, int K, int TB, ExecPlan &plan, bool allow_oop) {
    ExecPlan best;
    Status st = plan_tiled_list(s, gates, K, TB, best);
    if (!st.ok()) return st;
    const bool any = std::any_of(gates.begin(), gates.end(), is_relabel);
    const int Kr = K <= 0 ? default_tile_qubits(s) : K;
    if (any && (int)s.n >= Kr) {
        for (int order = 0; order < 2; order++) {
            ExecPlan alt;
            if (!plan_tiled_list(s, relabel_swaps(gates, (int)s.n, order == 1), K, TB, alt).ok()) continue;
            if (alt.steps.size() < best.steps.size()) best = std::move(alt);
        }
        const int TBr = TB <= 0 ? default_tile_threads_log2(s.prec, Kr) : TB;
        if (allow_oop && TBr == kLaneBits && (int)s.n >= Kr + 4) {
            // relabelled, the data moved by the first pass (to the second
            // buffer) and the last (back, through the final permutation)
            int holds[64];
            ExecPlan alt;
            if (plan_tiled_list(s, relabel_swaps(gates, (int)s.n, false, holds), K, TB, alt).ok()) {
                bool identity = true;
                for (int p = 0; p < (int)s.n; p++) identity = identity && holds[p] == p;
                size_t first = alt.steps.size();
                for (size_t k = 0; k < alt.steps.size() && first == alt.steps.size(); k++)
                    if (alt.steps[k].type == ExecStep::REG) first = k;
                if (!identity && first < alt.steps.size()) {
                    alt.steps[first].reg.oop = 1;
                    alt.steps[first].reg.sync_store = 0;  // loads and stores in different buffers
                    permutation_pass((int)s.n, Kr, holds, alt);
                    if (alt.steps.size() < best.steps.size()) best = std::move(alt);
                }
            }
        }
    }
    append_plan(plan, best);
    return Status{};
}

void param_op_coeffs(RegOp &op, double theta) {
    static const int kind_of[] = {-1, QI_GATE_RY, -1, QI_GATE_RX, -1, -1, QI_GATE_RZ, QI_GATE_P};
    qi_gate g{};
    g.kind = op.kind < 8 ? kind_of[op.kind] : -1;
    g.ntargets = 1;
    g.params[0] = theta;
    Lowered G;
    (void)lower_gate(64, g, G);
    for (int k = 0; k < 8; k++) op.c[k] = 0;
    switch (op.kind) {  // the forms RegPlanner::emit gives these gates
    case RO_RR: op.c[0] = G.c[0]; op.c[1] = -G.c[1]; op.c[2] = G.c[1]; op.c[3] = G.c[0]; break;   // RY
    case RO_RXI: op.c[0] = G.c[0]; op.c[1] = -G.c[2]; op.c[2] = G.c[4]; op.c[3] = G.c[0]; break;  // RX
    case RO_DG: for (int k = 0; k < 4; k++) op.c[k] = G.c[k]; break;                              // RZ
    case RO_P1: op.c[0] = 1; op.c[2] = G.c[0]; op.c[3] = G.c[1]; break;                           // P
    default: break;
    }
    for (int k = 0; k < 8; k++) op.cf[k] = (float)op.c[k];
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
    const std::vector<char> none(gates.size(), 0);  // every gate can join an LDS pass
    size_t first = 0;
    while (first < gates.size()) {
        std::vector<size_t> members;
        const uint64_t q = select_pass(gs, none, done, first, K, (int)s.n, tile_low_qubits(s.prec), false, members);
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
