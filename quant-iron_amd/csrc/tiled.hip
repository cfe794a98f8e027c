// tiled.hip — cache-blocked circuit execution: a "pass" loads tiles of 2^K
// amplitudes spanned by a chosen qubit set Q into LDS, applies every gate of
// the pass inside the tile, and writes the tile back, so one HBM round trip
// carries many gates instead of one.
//
// A gate fits a pass when its non-diagonal qubits (pair targets, both SWAP /
// Matchgate targets) lie in Q. Controls and diagonal targets outside Q are
// constant across a tile (the tile fixes all other index bits), so they become
// a per-tile predicate or phase. Q always contains the lowest qubits, which
// makes every tile row a contiguous run of at least 128 bytes in HBM.
//
// Scheduling: gates are taken in list order; a gate may move ahead of
// earlier gates only when their qubit supports are disjoint (they commute),
// so the pass sequence applies exactly the circuit's unitary. Within a tile
// each gate uses the same per-element formulas as the streaming kernels.
#include <algorithm>

#include "qi_internal.h"
#include "exec_plan.h"

#pragma clang fp contract(off)

namespace qi {

namespace {

template <typename R> struct V2;
template <> struct V2<float> { using T = float2; };
template <> struct V2<double> { using T = double2; };

template <typename T> __device__ __forceinline__ T cmul(T a, T b) {
    T r; r.x = a.x * b.x - a.y * b.y; r.y = a.x * b.y + a.y * b.x; return r;
}
template <typename T> __device__ __forceinline__ T cadd(T a, T b) { T r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename T> __device__ __forceinline__ T csub(T a, T b) { T r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename T, typename R> __device__ __forceinline__ T cscale(R f, T a) { T r; r.x = f * a.x; r.y = f * a.y; return r; }
template <typename T> __device__ __forceinline__ T mk(decltype(T::x) re, decltype(T::x) im) { T r; r.x = re; r.y = im; return r; }

__device__ __forceinline__ uint32_t ins0(uint32_t k, unsigned p) { return ((k >> p) << (p + 1)) | (k & ((1u << p) - 1)); }

__device__ __forceinline__ uint64_t insert_zeros64(uint64_t k, uint64_t fixed) {
    while (fixed) {
        unsigned p = __builtin_ctzll(fixed);
        k = ((k >> p) << (p + 1)) | (k & ((uint64_t(1) << p) - 1));
        fixed &= fixed - 1;
    }
    return k;
}

// Deposit the low bits of e onto the qubit positions listed in pos[].
__device__ __forceinline__ uint64_t deposit_pos(uint32_t e, const uint8_t *pos, int k) {
    uint64_t r = 0;
#pragma unroll
    for (int j = 0; j < k; j++) r |= uint64_t((e >> j) & 1) << pos[j];
    return r;
}

constexpr int kTileThreads = 256;

template <typename R, int K>
__global__ __launch_bounds__(kTileThreads) void k_tile_pass(typename V2<R>::T *__restrict__ psi, TilePass pass,
                                                           const TileOp *__restrict__ ops) {
    using T = typename V2<R>::T;
    constexpr int N = 1 << K;
    constexpr int PER = N / kTileThreads;
    __shared__ T tile[N];
    const uint32_t tid = threadIdx.x;
    const uint64_t base = insert_zeros64(uint64_t(blockIdx.x), pass.qmask);
    // element e = tid + it * 256: its global offset splits into a per-thread
    // part (low 8 local bits) and a per-iteration part (uniform)
    const uint64_t off_t = deposit_pos(tid, pass.pos, 8);
#pragma unroll
    for (int it = 0; it < PER; it++) {
        uint64_t off_i = deposit_pos(uint32_t(it) << 8, pass.pos, K);
        tile[tid + it * kTileThreads] = psi[base | off_t | off_i];
    }
    __syncthreads();
    for (int o = 0; o < pass.nops; o++) {
        const TileOp op = ops[pass.op0 + o];
        if ((base & op.gcm) != op.gcm) continue;  // controls outside the tile
        const uint32_t lcm = op.lcm;
        switch (op.kind) {
        case LK_Z: case LK_PHASE: case LK_RZ: {
            // diagonal: phase of the target bit (local or fixed by the tile)
            T p0 = mk<T>(R(op.c[0]), R(op.c[1])), p1 = mk<T>(R(op.c[2]), R(op.c[3]));
            for (int it = 0; it < PER; it++) {
                uint32_t e = tid + it * kTileThreads;
                if ((e & lcm) != lcm) continue;
                bool one = op.outside ? ((base >> op.gt) & 1) : ((e >> op.lp0) & 1);
                T a = tile[e];
                if (op.kind == LK_Z) { if (one) { a.x = -a.x; a.y = -a.y; } }
                else if (op.kind == LK_PHASE) { if (one) a = cmul(a, p1); }
                else a = cmul(a, one ? p1 : p0);
                tile[e] = a;
            }
        } break;
        case LK_SWAP: case LK_MATCH: {
            const uint32_t b0 = 1u << op.lp0, b1 = 1u << op.lp1;
            const unsigned lo = op.lp0 < op.lp1 ? op.lp0 : op.lp1, hi = op.lp0 < op.lp1 ? op.lp1 : op.lp0;
            for (int it = 0; it < PER / 4; it++) {
                uint32_t q = tid + it * kTileThreads;
                uint32_t l = ins0(ins0(q, lo), hi);
                if (op.kind == LK_SWAP) {
                    uint32_t i = l | b0, j = l | b1;
                    if ((i & lcm) != lcm) continue;
                    T a = tile[i], b = tile[j];
                    tile[i] = b; tile[j] = a;
                } else {
                    uint32_t i01 = l | b0, i10 = l | b1, i11 = l | b0 | b1;
                    if ((i01 & lcm) == lcm) {
                        T a01 = tile[i01], a10 = tile[i10];
                        tile[i01] = csub(cscale<T>(R(op.c[0]), a01), cmul(mk<T>(R(op.c[2]), R(op.c[3])), a10));
                        tile[i10] = cadd(cscale<T>(R(op.c[1]), a01), cmul(mk<T>(R(op.c[4]), R(op.c[5])), a10));
                    }
                    if ((i11 & lcm) == lcm) tile[i11] = cmul(tile[i11], mk<T>(R(op.c[6]), R(op.c[7])));
                }
            }
            // Matchgate controls may sit on q2 (only overlap with q1 is rejected,
            // operator.rs:214-273); those are tested per element above.
        } break;
        default: {  // pair gates on a local target
            const uint32_t tb = 1u << op.lp0;
            for (int it = 0; it < PER / 2; it++) {
                uint32_t q = tid + it * kTileThreads;
                uint32_t i = ins0(q, op.lp0), j = i | tb;
                if ((i & lcm) != lcm) continue;
                T a = tile[i], b = tile[j], na, nb;
                switch (op.kind) {
                case LK_H: na = cscale<T>(R(op.c[0]), cadd(a, b)); nb = cscale<T>(R(op.c[0]), csub(a, b)); break;
                case LK_X: na = b; nb = a; break;
                case LK_Y: na = cmul(mk<T>(R(op.c[0]), R(op.c[1])), b); nb = cmul(mk<T>(R(op.c[2]), R(op.c[3])), a); break;
                case LK_RX:
                    na = csub(cscale<T>(R(op.c[0]), a), cmul(mk<T>(R(op.c[1]), R(op.c[2])), b));
                    nb = cadd(cmul(mk<T>(R(op.c[3]), R(op.c[4])), a), cscale<T>(R(op.c[0]), b));
                    break;
                case LK_RY:
                    na = csub(cscale<T>(R(op.c[0]), a), cscale<T>(R(op.c[1]), b));
                    nb = cadd(cscale<T>(R(op.c[1]), a), cscale<T>(R(op.c[0]), b));
                    break;
                default:  // LK_U2
                    na = cadd(cmul(mk<T>(R(op.c[0]), R(op.c[1])), a), cmul(mk<T>(R(op.c[2]), R(op.c[3])), b));
                    nb = cadd(cmul(mk<T>(R(op.c[4]), R(op.c[5])), a), cmul(mk<T>(R(op.c[6]), R(op.c[7])), b));
                    break;
                }
                tile[i] = na; tile[j] = nb;
            }
        } break;
        }
        __syncthreads();
    }
#pragma unroll
    for (int it = 0; it < PER; it++) {
        uint64_t off_i = deposit_pos(uint32_t(it) << 8, pass.pos, K);
        psi[base | off_t | off_i] = tile[tid + it * kTileThreads];
    }
}

// Qubits a gate needs inside the tile, and all qubits it touches.
void gate_sets(const Lowered &L, uint64_t &need, uint64_t &supp) {
    uint64_t t0 = uint64_t(1) << L.t0, t1 = uint64_t(1) << L.t1;
    switch (L.kind) {
    case LK_Z: case LK_PHASE: case LK_RZ: need = 0; supp = t0 | L.ctrl_mask; break;
    case LK_SWAP: case LK_MATCH: need = t0 | t1; supp = t0 | t1 | L.ctrl_mask; break;
    default: need = t0; supp = t0 | L.ctrl_mask; break;
    }
}

template <typename R, int K> void launch_pass(State &s, const TilePass &p, const TileOp *ops, hipStream_t st) {
    using T = typename V2<R>::T;
    unsigned tiles = (unsigned)(uint64_t(1) << (s.n - K));
    hipLaunchKernelGGL((k_tile_pass<R, K>), dim3(tiles), dim3(kTileThreads), 0, st, (T *)s.d, p, ops);
}

}  // namespace

int tile_low_qubits(int prec) { return prec == QI_C128 ? 3 : 4; }  // 128-byte rows

int default_tile_qubits(const State &s) { return s.prec == QI_C128 ? 12 : 13; }

Status plan_tiled(const State &s, const std::vector<Lowered> &gates, int K, ExecPlan &plan) {
    if (K <= 0) K = default_tile_qubits(s);
    const int kmax = s.prec == QI_C128 ? 12 : 13;
    if (K < 10 || K > kmax) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "tile width out of range");
    if ((int)s.n < K) {  // state smaller than a tile: stream gate by gate
        for (const Lowered &L : gates) plan.steps.push_back(ExecStep::gate(L));
        return Status{};
    }
    const uint64_t all = (s.n >= 64) ? ~uint64_t(0) : ((uint64_t(1) << s.n) - 1);
    const uint64_t low = (uint64_t(1) << tile_low_qubits(s.prec)) - 1;
    std::vector<char> done(gates.size(), 0);
    size_t first = 0;
    while (first < gates.size()) {
        uint64_t active = low, blocked = 0;
        std::vector<size_t> members;
        for (size_t g = first; g < gates.size(); g++) {
            if (done[g]) continue;
            uint64_t need, supp;
            gate_sets(gates[g], need, supp);
            if (supp & blocked) { blocked |= supp; }
            else if (__builtin_popcountll(active | need) <= K) { active |= need; members.push_back(g); done[g] = 1; }
            else blocked |= supp;
            if ((blocked & all) == all) break;
        }
        // fill the tile with the lowest unused qubits (longer contiguous rows)
        uint64_t q = active;
        for (int b = 0; __builtin_popcountll(q) < K && b < (int)s.n; b++) q |= uint64_t(1) << b;
        TilePass p{};
        p.qmask = q;
        p.K = K;
        int j = 0;
        uint8_t local[64];
        for (int b = 0; b < (int)s.n; b++)
            if ((q >> b) & 1) { local[b] = (uint8_t)j; p.pos[j++] = (uint8_t)b; }
        p.op0 = (int)plan.tile_ops.size();
        for (size_t g : members) {
            const Lowered &L = gates[g];
            TileOp op{};
            op.kind = L.kind;
            for (int c = 0; c < 8; c++) op.c[c] = L.c[c];
            if (L.kind == LK_Z) { op.c[2] = -1.0; op.c[3] = 0.0; }
            if (L.kind == LK_PHASE) { op.c[2] = L.c[0]; op.c[3] = L.c[1]; }
            uint64_t tb0 = uint64_t(1) << L.t0;
            if ((q & tb0) == 0) { op.outside = 1; op.gt = (uint8_t)L.t0; }
            else op.lp0 = local[L.t0];
            if (L.kind == LK_SWAP || L.kind == LK_MATCH) op.lp1 = local[L.t1];
            uint64_t lc = L.ctrl_mask & q, gc = L.ctrl_mask & ~q;
            op.gcm = gc;
            uint32_t lcm = 0;
            for (int b = 0; b < (int)s.n; b++)
                if ((lc >> b) & 1) lcm |= 1u << local[b];
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

Status launch_tile_pass(State &s, const TilePass &p, const TileOp *dops, hipStream_t st) {
    if (s.prec == QI_C128) {
        switch (p.K) {
        case 10: launch_pass<double, 10>(s, p, dops, st); break;
        case 11: launch_pass<double, 11>(s, p, dops, st); break;
        case 12: launch_pass<double, 12>(s, p, dops, st); break;
        default: return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "tile width");
        }
    } else {
        switch (p.K) {
        case 10: launch_pass<float, 10>(s, p, dops, st); break;
        case 11: launch_pass<float, 11>(s, p, dops, st); break;
        case 12: launch_pass<float, 12>(s, p, dops, st); break;
        case 13: launch_pass<float, 13>(s, p, dops, st); break;
        default: return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "tile width");
        }
    }
    QI_HIP_TRY(hipGetLastError());
    return Status{};
}

}  // namespace qi
