// gates.hip — streaming per-gate kernels (one HBM pass per gate) and the small
// elementwise kernels (init, scale, axpy, copy).
//
// Each gate kernel enumerates exactly the amplitudes the gate changes: the
// index space with the target and control bits removed is walked linearly and
// the fixed bits are re-inserted (controls = 1, target = 0 or 1), so a CNOT
// reads and writes half the state and a Toffoli a quarter. Memory moves in
// 16-byte units (two complex64 amplitudes, or one complex128) with
// non-temporal loads and stores; consecutive threads own consecutive units, so
// every wave's loads are contiguous runs. A complex64 gate with a control on
// qubit 0 rewrites the untouched even amplitudes of its units as read: HBM
// moves whole 64-byte lines, so this costs no extra traffic.
//
// Arithmetic: the per-element formulas keep the reference's operand order
// (operator.rs, e.g. Hadamard :396-412, RotateX :1756-1757, RotateY :1897-1898,
// Unitary2 :2253-2254, Matchgate :1000-1006) and this file is compiled
// without FMA contraction, so a complex128 state gets the reference's results
// bit for bit (up to the sign of zero). All constants are computed on the host.
#include "qi_internal.h"

#pragma clang fp contract(off)

namespace qi {

namespace {

template <typename R> struct V2;
template <> struct V2<float> { using T = float2; };
template <> struct V2<double> { using T = double2; };

template <typename T> __device__ __forceinline__ T cmul(T a, T b) {
    T r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename T> __device__ __forceinline__ T cadd(T a, T b) { T r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename T> __device__ __forceinline__ T csub(T a, T b) { T r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
template <typename T, typename R> __device__ __forceinline__ T cscale(R f, T a) { T r; r.x = f * a.x; r.y = f * a.y; return r; }
template <typename T> __device__ __forceinline__ T mk(decltype(T::x) re, decltype(T::x) im) { T r; r.x = re; r.y = im; return r; }

// Insert a zero bit at every set bit of `fixed` (ascending), turning a free
// index into a full index. `fixed` is uniform across the launch.
__device__ __forceinline__ uint64_t insert_zeros(uint64_t k, uint64_t fixed) {
    while (fixed) {
        unsigned p = __builtin_ctzll(fixed);
        uint64_t lo = k & ((uint64_t(1) << p) - 1);
        k = ((k >> p) << (p + 1)) | lo;
        fixed &= fixed - 1;
    }
    return k;
}

constexpr int kThreads = 256;
constexpr int kItems = 4;  // independent pairs per thread (memory-level parallelism)

template <typename R> struct GateK {
    uint64_t items;   // number of free indices
    uint64_t fixed;   // bits removed from the free index
    uint64_t set;     // bits forced to 1 in the base index
    uint64_t tb;      // target bit
    uint64_t pmask;   // pair partner = base ^ pmask
    uint64_t tb2;     // second target bit (swap / matchgate)
    uint64_t cmask;   // control mask tested per element (matchgate only)
    uint64_t base;    // first item of this launch (a gate over more items than one grid holds is split)
    int32_t chunk;    // workgroup order (xcd_block_order) of the pair and diagonal kernels
    int32_t lane;     // lane-exchange kernels: lane offset of the partner (1, 2, 4, 8 or 16)
    R c[10];
};

// Per-pair and per-element updates shared by every layout below.
template <typename R, int KIND, typename T>
__device__ __forceinline__ void pair_update(const GateK<R> &g, T a, T b, T &na, T &nb) {
    if constexpr (KIND == LK_H) {
        R s = g.c[0];
        na = cscale<T>(s, cadd(a, b));
        nb = cscale<T>(s, csub(a, b));
    } else if constexpr (KIND == LK_X) {
        na = b; nb = a;
    } else if constexpr (KIND == LK_Y) {
        na = cmul(mk<T>(g.c[0], g.c[1]), b);   // (-i) * b
        nb = cmul(mk<T>(g.c[2], g.c[3]), a);   // ( i) * a
    } else if constexpr (KIND == LK_RX) {
        na = csub(cscale<T>(g.c[0], a), cmul(mk<T>(g.c[1], g.c[2]), b));
        nb = cadd(cmul(mk<T>(g.c[3], g.c[4]), a), cscale<T>(g.c[0], b));
    } else if constexpr (KIND == LK_RY) {
        na = csub(cscale<T>(g.c[0], a), cscale<T>(g.c[1], b));
        nb = cadd(cscale<T>(g.c[1], a), cscale<T>(g.c[0], b));
    } else if constexpr (KIND == LK_U2) {
        na = cadd(cmul(mk<T>(g.c[0], g.c[1]), a), cmul(mk<T>(g.c[2], g.c[3]), b));
        nb = cadd(cmul(mk<T>(g.c[4], g.c[5]), a), cmul(mk<T>(g.c[6], g.c[7]), b));
    } else {  // LK_RZ
        na = cmul(a, mk<T>(g.c[0], g.c[1]));
        nb = cmul(b, mk<T>(g.c[2], g.c[3]));
    }
}

template <typename R, int KIND, typename T>
__device__ __forceinline__ T diag_update(const GateK<R> &g, T a) {
    if constexpr (KIND == LK_Z) { T r; r.x = -a.x; r.y = -a.y; return r; }
    else return cmul(a, mk<T>(g.c[0], g.c[1]));
}

// 16-byte units moved with non-temporal loads and stores: the state is
// streamed once per gate and never re-read from cache, so the accesses skip
// cache retention.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <typename V> __device__ __forceinline__ V ldnt(const V *p) { return __builtin_nontemporal_load(p); }
template <typename V> __device__ __forceinline__ void stnt(V v, V *p) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ float2 lo(f32x4 v) { return make_float2(v.x, v.y); }
__device__ __forceinline__ float2 hi(f32x4 v) { return make_float2(v.z, v.w); }
__device__ __forceinline__ f32x4 pack(float2 a, float2 b) { f32x4 r; r.x = a.x; r.y = a.y; r.z = b.x; r.w = b.y; return r; }
__device__ __forceinline__ double2 amp(f64x2 v) { return make_double2(v.x, v.y); }
__device__ __forceinline__ f64x2 unamp(double2 a) { f64x2 r; r.x = a.x; r.y = a.y; return r; }

// complex64 pair gates on units of two neighbouring amplitudes (index bit 0
// inside the unit; `g` is expressed in unit indices).
//   MODE 0: bit 0 is free: both halves pair with the partner unit's halves.
//   MODE 1: bit 0 is a control: only the odd halves change; the even halves
//           are written back as read (each unit has exactly one owner).
//   MODE 2: bit 0 is the target: every unit holds one pair; one unit per
//           thread (a plain stream: one 16-byte unit per thread measured
//           85% of HBM for a copy, two 78%).
template <int KIND, int MODE>
__global__ __launch_bounds__(kThreads) void k_pair_c64(f32x4 *__restrict__ psi, GateK<float> g) {
    if constexpr (MODE == 2) {
        const uint64_t k = g.base + xcd_block_order(blockIdx.x, g.chunk) * kThreads + threadIdx.x;
        if (k >= g.items) return;
        const uint64_t j = insert_zeros(k, g.fixed) | g.set;
        const f32x4 u = ldnt(psi + j);
        float2 na, nb;
        pair_update<float, KIND>(g, lo(u), hi(u), na, nb);
        stnt(pack(na, nb), psi + j);
    } else {
        const uint64_t k = g.base + xcd_block_order(blockIdx.x, g.chunk) * kThreads + threadIdx.x;
        if (k >= g.items) return;
        const uint64_t ja = insert_zeros(k, g.fixed) | g.set, jb = ja ^ g.pmask;
        const f32x4 x = ldnt(psi + ja), y = ldnt(psi + jb);
        float2 a0 = lo(x), b0 = lo(y), a1, b1;
        pair_update<float, KIND>(g, hi(x), hi(y), a1, b1);
        if constexpr (MODE == 0) pair_update<float, KIND>(g, lo(x), lo(y), a0, b0);
        stnt(pack(a0, a1), psi + ja);
        stnt(pack(b0, b1), psi + jb);
    }
}

// Low targets (partner 16-256 bytes away, unit bits 0..4 of the index without
// the control bits): a wave reads 64 consecutive units, one fully used 1 KiB
// row, one unit per lane, and takes each partner from the lane that holds it
// (in the pair layout above each load instruction would touch twice the cache
// lines it uses, or move two units per thread). `g`: fixed = controls only,
// items = every unit with the controls satisfied.
template <typename T, typename R, int KIND>
__device__ __forceinline__ T lane_pick(const GateK<R> &g, T own, T other, bool upper) {
    T na, nb;
    pair_update<R, KIND>(g, upper ? other : own, upper ? own : other, na, nb);
    return upper ? nb : na;
}

__device__ __forceinline__ f32x4 shfl_xor4(f32x4 v, int m) {
    f32x4 r;
    r.x = __shfl_xor(v.x, m); r.y = __shfl_xor(v.y, m); r.z = __shfl_xor(v.z, m); r.w = __shfl_xor(v.w, m);
    return r;
}

template <int KIND, int MODE>
__global__ __launch_bounds__(kThreads) void k_pair_c64_lane(f32x4 *__restrict__ psi, GateK<float> g) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t k = g.base + xcd_block_order(blockIdx.x, g.chunk) * kThreads + threadIdx.x;
    const bool h = k < g.items;  // a partner is in range iff its lane's unit is
    uint64_t j = 0;
    f32x4 u = {0, 0, 0, 0};
    if (h) { j = insert_zeros(k, g.fixed) | g.set; u = ldnt(psi + j); }
    const f32x4 p = shfl_xor4(u, g.lane);
    const bool upper = (lane & uint32_t(g.lane)) != 0;  // this lane holds the target-1 units
    // MODE 0: both halves of a unit pair up; MODE 1: qubit 0 is a control, only odd halves change
    const float2 a = MODE == 0 ? lane_pick<float2, float, KIND>(g, lo(u), lo(p), upper) : lo(u);
    if (h) stnt(pack(a, lane_pick<float2, float, KIND>(g, hi(u), hi(p), upper)), psi + j);
}

// complex64 diagonal gates (Z, phases) on two-amplitude units. MODE 0: bit 0
// free (both halves change); MODE 1: bit 0 is set by the gate (a control or
// the target), so only the odd halves change.
template <int KIND, int MODE>
__global__ __launch_bounds__(kThreads) void k_diag_c64(f32x4 *__restrict__ psi, GateK<float> g) {
    const uint64_t k = g.base + xcd_block_order(blockIdx.x, g.chunk) * kThreads + threadIdx.x;
    if (k >= g.items) return;
    const uint64_t j = insert_zeros(k, g.fixed) | g.set;
    const f32x4 u = ldnt(psi + j);
    stnt(pack(MODE == 0 ? diag_update<float, KIND>(g, lo(u)) : lo(u), diag_update<float, KIND>(g, hi(u))), psi + j);
}

// Pair gates whose fixed bits (target, controls) all sit at bit 5 of the unit
// index or above, one 16-byte unit per thread: lanes 0..31 of a wave load the
// a-units of 32 consecutive pairs (512 contiguous bytes), lanes 32..63 their
// partners, and the two halves trade through v_permlane32_swap (the four
// 32-bit components of a unit). With one pair (two units) per thread the same
// gate measured 79% of HBM at target 15 and 71% at the widest strides; with
// one unit per thread 83% and 80%. MODE as k_pair_c64 (complex128: 0).
template <typename U> __device__ __forceinline__ U lane_partner32(U v) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 w = __builtin_bit_cast(u32x4, v);
    u32x4 o;
    const bool upper = (threadIdx.x & 32) != 0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        // lanes < 32 keep their value in the first result and get lane + 32's
        // in the second; lanes >= 32 the other way round
        const auto r = __builtin_amdgcn_permlane32_swap(w[c], w[c], false, false);
        o[c] = upper ? r[0] : r[1];
    }
    return __builtin_bit_cast(U, o);
}
template <int KIND, int MODE>
__global__ __launch_bounds__(kThreads) void k_pair_c64_split(f32x4 *__restrict__ psi, GateK<float> g) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t k = g.base + (xcd_block_order(blockIdx.x, g.chunk) * (kThreads / 64) + (threadIdx.x >> 6)) * 32 + (lane & 31);
    if (k >= g.items) return;  // lanes l and l + 32 share k
    const bool upper = lane >= 32;
    const uint64_t ja = insert_zeros(k, g.fixed) | g.set, j = upper ? (ja ^ g.pmask) : ja;
    const f32x4 v = ldnt(psi + j), p = lane_partner32(v);
    const f32x4 x = upper ? p : v, y = upper ? v : p;  // the a- and b-units
    float2 a0 = lo(x), b0 = lo(y), a1, b1;
    pair_update<float, KIND>(g, hi(x), hi(y), a1, b1);
    if constexpr (MODE == 0) pair_update<float, KIND>(g, lo(x), lo(y), a0, b0);
    stnt(upper ? pack(b0, b1) : pack(a0, a1), psi + j);
}
template <int KIND>
__global__ __launch_bounds__(kThreads) void k_pair_c128_split(f64x2 *__restrict__ psi, GateK<double> g) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t k = g.base + (xcd_block_order(blockIdx.x, g.chunk) * (kThreads / 64) + (threadIdx.x >> 6)) * 32 + (lane & 31);
    if (k >= g.items) return;
    const bool upper = lane >= 32;
    const uint64_t ia = insert_zeros(k, g.fixed) | g.set, j = upper ? (ia ^ g.pmask) : ia;
    const f64x2 v = ldnt(psi + j), p = lane_partner32(v);
    double2 na, nb;
    pair_update<double, KIND>(g, amp(upper ? p : v), amp(upper ? v : p), na, nb);
    stnt(unamp(upper ? nb : na), psi + j);
}

// complex128: one amplitude is already a 16-byte unit; one pair per thread.
template <int KIND>
__global__ __launch_bounds__(kThreads) void k_pair_c128(f64x2 *__restrict__ psi, GateK<double> g) {
    const uint64_t k = g.base + xcd_block_order(blockIdx.x, g.chunk) * kThreads + threadIdx.x;
    if (k >= g.items) return;
    const uint64_t ia = insert_zeros(k, g.fixed) | g.set, ib = ia ^ g.pmask;
    const double2 a = amp(ldnt(psi + ia)), b = amp(ldnt(psi + ib));
    double2 na, nb;
    pair_update<double, KIND>(g, a, b, na, nb);
    stnt(unamp(na), psi + ia);
    stnt(unamp(nb), psi + ib);
}

// complex128 low targets (partner 16-64 bytes away), as k_pair_c64_lane.
template <int KIND>
__global__ __launch_bounds__(kThreads) void k_pair_c128_lane(f64x2 *__restrict__ psi, GateK<double> g) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t k = g.base + xcd_block_order(blockIdx.x, g.chunk) * kThreads + threadIdx.x;
    const bool h = k < g.items;
    uint64_t j = 0;
    f64x2 u = {0, 0};
    if (h) { j = insert_zeros(k, g.fixed) | g.set; u = ldnt(psi + j); }
    const double2 p = make_double2(__shfl_xor(u.x, g.lane), __shfl_xor(u.y, g.lane));
    const bool upper = (lane & uint32_t(g.lane)) != 0;
    if (h) stnt(unamp(lane_pick<double2, double, KIND>(g, amp(u), p, upper)), psi + j);
}

template <int KIND>
__global__ __launch_bounds__(kThreads) void k_diag_c128(f64x2 *__restrict__ psi, GateK<double> g) {
    const uint64_t k = g.base + xcd_block_order(blockIdx.x, g.chunk) * kThreads + threadIdx.x;
    if (k >= g.items) return;
    const uint64_t j = insert_zeros(k, g.fixed) | g.set;
    stnt(unamp(diag_update<double, KIND>(g, amp(ldnt(psi + j)))), psi + j);
}

// complex64 SWAP with qubit 0 as one of its targets: the exchanged
// amplitudes sit in different halves of their units, so this one walks
// single amplitudes (8-byte accesses, several pairs per thread in flight).
__global__ __launch_bounds__(kThreads) void k_swap_bit0(float2 *__restrict__ psi, GateK<float> g) {
    const uint64_t first = g.base + (uint64_t(blockIdx.x) * kItems) * kThreads + threadIdx.x;
    uint64_t i0[kItems];
    float2 a[kItems], b[kItems];
#pragma unroll
    for (int u = 0; u < kItems; u++) {
        uint64_t k = first + uint64_t(u) * kThreads;
        i0[u] = k < g.items ? (insert_zeros(k, g.fixed) | g.set) : ~uint64_t(0);
        if (k < g.items) { a[u] = psi[i0[u]]; b[u] = psi[i0[u] ^ g.pmask]; }
    }
#pragma unroll
    for (int u = 0; u < kItems; u++) {
        if (i0[u] == ~uint64_t(0)) continue;
        psi[i0[u]] = b[u];
        psi[i0[u] ^ g.pmask] = a[u];
    }
}

// Matchgate on (q1, q1+1); controls are tested on i01 for the rotation and on
// i11 for the phase, as the reference does (operator.rs:993-1004).
template <typename R>
__global__ __launch_bounds__(kThreads) void k_match(typename V2<R>::T *__restrict__ psi, GateK<R> g) {
    using T = typename V2<R>::T;
    const uint64_t k = g.base + uint64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (k >= g.items) return;
    uint64_t l = insert_zeros(k, g.fixed);
    uint64_t i01 = l | g.tb, i10 = l | g.tb2, i11 = l | g.tb | g.tb2;
    if ((i01 & g.cmask) == g.cmask) {
        T a01 = psi[i01], a10 = psi[i10];
        psi[i01] = csub(cscale<T>(g.c[0], a01), cmul(mk<T>(g.c[2], g.c[3]), a10));
        psi[i10] = cadd(cscale<T>(g.c[1], a01), cmul(mk<T>(g.c[4], g.c[5]), a10));
    }
    if ((i11 & g.cmask) == g.cmask) psi[i11] = cmul(psi[i11], mk<T>(g.c[6], g.c[7]));
}

template <typename R> GateK<R> make_args(const Lowered &L, uint32_t n) {
    GateK<R> g{};
    for (int q = 0; q < 8; q++) g.c[q] = R(L.c[q]);
    g.tb = uint64_t(1) << L.t0;
    g.tb2 = uint64_t(1) << L.t1;
    uint64_t fixed = L.ctrl_mask | g.tb;
    if (L.kind == LK_SWAP) fixed |= g.tb2;
    if (L.kind == LK_MATCH) { fixed = g.tb | g.tb2; g.cmask = L.ctrl_mask; }
    g.fixed = fixed;
    g.set = L.ctrl_mask;
    g.pmask = g.tb;
    if (L.kind == LK_Z || L.kind == LK_PHASE) g.set |= g.tb;
    if (L.kind == LK_SWAP) { g.set |= g.tb; g.pmask = g.tb | g.tb2; }
    if (L.kind == LK_MATCH) g.set = 0;
    g.items = uint64_t(1) << (n - __builtin_popcountll(fixed));
    return g;
}

inline uint64_t grid_for(uint64_t items, uint64_t per_block) { return (items + per_block - 1) / per_block; }

// A grid holds at most 2^32 - 1 work-items; a gate on a larger state (2^32
// unit pairs at 34 complex64 qubits) runs as several launches of at most 2^22
// workgroups, each with `base` advanced past the items of the earlier ones.
constexpr uint64_t kMaxGridBlocks = uint64_t(1) << 22;

// Runs of consecutive workgroups per XCD (log2) for the pair and diagonal
// kernels: 1-5% faster single gates at 30 qubits than the identity order.
constexpr int kChunkLog2 = 8;

#define QI_PAIR_KINDS(M) M(LK_H) M(LK_X) M(LK_Y) M(LK_RX) M(LK_RY) M(LK_U2) M(LK_RZ)
#define QI_LANE_KINDS(M) M(LK_H) M(LK_X) M(LK_Y) M(LK_RX) M(LK_RY) M(LK_U2)

// Lane-exchange form of a one-target gate whose (unit-space) target bit `tb`
// lands on bit 0..4 of the index with the controls `ctrl` removed: returns
// the launch arguments, or false when the pair layout is the better one.
template <typename R> bool lane_args(const GateK<R> &g, uint64_t tb, uint64_t ctrl, int kind, GateK<R> &out) {
    if (kind == LK_RZ || (tb & (tb - 1)) || !tb) return false;
    const int b = __builtin_ctzll(tb);
    const int kb = b - __builtin_popcountll(ctrl & (tb - 1));
    if (kb > 4) return false;
    out = g;
    out.fixed = ctrl;
    out.items = g.items * 2;
    out.lane = 1 << kb;
    out.chunk = xcd_chunk_log2(grid_for(out.items, kThreads), kChunkLog2);
    return true;
}

// Launches `launch(grid, args)` over `blocks` workgroups of `per_block` items:
// one launch, or pieces of kMaxGridBlocks workgroups with args.base advanced
// (and the XCD order of each piece's own grid).
template <typename R, typename F> void split_launch(uint64_t blocks, uint64_t per_block, const GateK<R> &g, F &&launch) {
    for (uint64_t b0 = 0; b0 < blocks; b0 += kMaxGridBlocks) {
        const uint64_t nb = std::min(kMaxGridBlocks, blocks - b0);
        GateK<R> h = g;
        h.base = b0 * per_block;
        h.chunk = xcd_chunk_log2(nb, kChunkLog2);
        launch(dim3((unsigned)nb), h);
    }
}

Status launch_c64(State &s, const Lowered &L, hipStream_t st) {
    GateK<float> g = make_args<float>(L, s.n);
    const dim3 block(kThreads);
    if (L.kind == LK_MATCH) {
        split_launch(grid_for(g.items, kThreads), kThreads, g, [&](dim3 gr, const GateK<float> &h) {
            hipLaunchKernelGGL((k_match<float>), gr, block, 0, st, (float2 *)s.d, h);
        });
        QI_HIP_TRY(hipGetLastError());
        return Status{};
    }
    if (L.kind == LK_SWAP && (L.t0 == 0 || L.t1 == 0)) {
        const uint64_t per = uint64_t(kThreads) * kItems;
        split_launch(grid_for(g.items, per), per, g, [&](dim3 gr, const GateK<float> &h) {
            hipLaunchKernelGGL(k_swap_bit0, gr, block, 0, st, (float2 *)s.d, h);
        });
        QI_HIP_TRY(hipGetLastError());
        return Status{};
    }
    // Re-express the index sets in two-amplitude units.
    const bool diag = L.kind == LK_Z || L.kind == LK_PHASE;
    const int mode = diag ? ((g.set & 1) ? 1 : 0) : (g.pmask & 1) ? 2 : (g.set & 1) ? 1 : 0;
    GateK<float> u = g;
    u.fixed = g.fixed >> 1;
    u.set = g.set >> 1;
    u.pmask = g.pmask >> 1;
    u.items = uint64_t(1) << (s.n - 1 - __builtin_popcountll(u.fixed));
    f32x4 *psi = reinterpret_cast<f32x4 *>(s.d);
    const uint64_t b1 = grid_for(u.items, kThreads);
    const int kind = L.kind == LK_SWAP ? LK_X : L.kind;  // SWAP: exchange base|tb with base|tb2
    GateK<float> ul;
    if (L.kind != LK_SWAP && mode != 2 && !diag && lane_args(u, u.pmask, u.fixed & ~u.pmask, kind, ul)) {
        bool ok = true;
        split_launch(grid_for(ul.items, kThreads), kThreads, ul, [&](dim3 gl, const GateK<float> &h) {
            switch (kind * 2 + mode) {
#define QI_LANE(K) \
            case K * 2 + 0: hipLaunchKernelGGL((k_pair_c64_lane<K, 0>), gl, block, 0, st, psi, h); break; \
            case K * 2 + 1: hipLaunchKernelGGL((k_pair_c64_lane<K, 1>), gl, block, 0, st, psi, h); break;
            QI_LANE_KINDS(QI_LANE)
#undef QI_LANE
            default: ok = false;
            }
        });
        if (!ok) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "gate kind has no streaming kernel");
        QI_HIP_TRY(hipGetLastError());
        return Status{};
    }
    bool ok = true;
    if (!diag && mode != 2 && (u.fixed & 31) == 0) {  // one unit per thread, partners in the other half-wave
        split_launch(grid_for(u.items, kThreads / 2), kThreads / 2, u, [&](dim3 gr, const GateK<float> &h) {
            switch (kind * 4 + mode) {
#define QI_SPLIT(K) \
            case K * 4 + 0: hipLaunchKernelGGL((k_pair_c64_split<K, 0>), gr, block, 0, st, psi, h); break; \
            case K * 4 + 1: hipLaunchKernelGGL((k_pair_c64_split<K, 1>), gr, block, 0, st, psi, h); break;
            QI_PAIR_KINDS(QI_SPLIT)
#undef QI_SPLIT
            default: ok = false;
            }
        });
        if (!ok) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "gate kind has no streaming kernel");
        QI_HIP_TRY(hipGetLastError());
        return Status{};
    }
    split_launch(b1, kThreads, u, [&](dim3 gr, const GateK<float> &h) {
        switch (kind * 4 + mode) {
#define QI_PAIR(K) \
        case K * 4 + 0: hipLaunchKernelGGL((k_pair_c64<K, 0>), gr, block, 0, st, psi, h); break; \
        case K * 4 + 1: hipLaunchKernelGGL((k_pair_c64<K, 1>), gr, block, 0, st, psi, h); break; \
        case K * 4 + 2: hipLaunchKernelGGL((k_pair_c64<K, 2>), gr, block, 0, st, psi, h); break;
        QI_PAIR_KINDS(QI_PAIR)
#undef QI_PAIR
        case LK_Z * 4 + 0: hipLaunchKernelGGL((k_diag_c64<LK_Z, 0>), gr, block, 0, st, psi, h); break;
        case LK_Z * 4 + 1: hipLaunchKernelGGL((k_diag_c64<LK_Z, 1>), gr, block, 0, st, psi, h); break;
        case LK_PHASE * 4 + 0: hipLaunchKernelGGL((k_diag_c64<LK_PHASE, 0>), gr, block, 0, st, psi, h); break;
        case LK_PHASE * 4 + 1: hipLaunchKernelGGL((k_diag_c64<LK_PHASE, 1>), gr, block, 0, st, psi, h); break;
        default: ok = false;
        }
    });
    if (!ok) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "gate kind has no streaming kernel");
    QI_HIP_TRY(hipGetLastError());
    return Status{};
}

Status launch_c128(State &s, const Lowered &L, hipStream_t st) {
    GateK<double> g = make_args<double>(L, s.n);
    const dim3 block(kThreads);
    const uint64_t b1 = grid_for(g.items, kThreads);
    f64x2 *psi = reinterpret_cast<f64x2 *>(s.d);
    GateK<double> gl;
    bool ok = true;
    if (L.kind != LK_SWAP && L.kind != LK_MATCH && L.kind != LK_Z && L.kind != LK_PHASE &&
        lane_args(g, g.pmask, g.fixed & ~g.pmask, L.kind, gl)) {
        split_launch(grid_for(gl.items, kThreads), kThreads, gl, [&](dim3 gg, const GateK<double> &h) {
            switch (L.kind) {
#define QI_LANE(K) case K: hipLaunchKernelGGL((k_pair_c128_lane<K>), gg, block, 0, st, psi, h); break;
            QI_LANE_KINDS(QI_LANE)
#undef QI_LANE
            default: ok = false;
            }
        });
        if (!ok) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "gate kind has no streaming kernel");
        QI_HIP_TRY(hipGetLastError());
        return Status{};
    }
    const int kind = L.kind == LK_SWAP ? LK_X : L.kind;
    if (kind != LK_Z && kind != LK_PHASE && kind != LK_MATCH && (g.fixed & 31) == 0) {  // one amplitude per thread
        split_launch(grid_for(g.items, kThreads / 2), kThreads / 2, g, [&](dim3 gr, const GateK<double> &h) {
            switch (kind) {
#define QI_SPLIT(K) case K: hipLaunchKernelGGL((k_pair_c128_split<K>), gr, block, 0, st, psi, h); break;
            QI_PAIR_KINDS(QI_SPLIT)
#undef QI_SPLIT
            default: ok = false;
            }
        });
        if (!ok) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "gate kind has no streaming kernel");
        QI_HIP_TRY(hipGetLastError());
        return Status{};
    }
    split_launch(b1, kThreads, g, [&](dim3 gr, const GateK<double> &h) {
        switch (kind) {
#define QI_PAIR(K) case K: hipLaunchKernelGGL((k_pair_c128<K>), gr, block, 0, st, psi, h); break;
        QI_PAIR_KINDS(QI_PAIR)
#undef QI_PAIR
        case LK_Z: hipLaunchKernelGGL((k_diag_c128<LK_Z>), gr, block, 0, st, psi, h); break;
        case LK_PHASE: hipLaunchKernelGGL((k_diag_c128<LK_PHASE>), gr, block, 0, st, psi, h); break;
        case LK_MATCH: hipLaunchKernelGGL((k_match<double>), gr, block, 0, st, (double2 *)s.d, h); break;
        default: ok = false;
        }
    });
    if (!ok) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "gate kind has no streaming kernel");
    QI_HIP_TRY(hipGetLastError());
    return Status{};
}

// ------------------------------------------------------ elementwise kernels --
template <typename R>
__global__ __launch_bounds__(256) void k_init(typename V2<R>::T *psi, uint64_t dim, int kind, uint64_t basis,
                                              double amp) {
    using T = typename V2<R>::T;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < dim;
         i += uint64_t(gridDim.x) * blockDim.x) {
        T v; v.x = 0; v.y = 0;
        switch (kind) {
        case QI_INIT_ZERO: if (i == 0) v.x = 1; break;
        case QI_INIT_BASIS: if (i == basis) v.x = 1; break;
        case QI_INIT_PLUS: v.x = R(amp); break;
        case QI_INIT_MINUS: v.x = (__popcll(i) & 1) ? -R(amp) : R(amp); break;
        case QI_INIT_GHZ: if (i == 0 || i == dim - 1) v.x = R(amp); break;
        }
        psi[i] = v;
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_scale(typename V2<R>::T *psi, uint64_t dim, R re, R im) {
    using T = typename V2<R>::T;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < dim;
         i += uint64_t(gridDim.x) * blockDim.x)
        psi[i] = cmul(psi[i], mk<T>(re, im));
}

template <typename R>
__global__ __launch_bounds__(256) void k_div(typename V2<R>::T *psi, uint64_t dim, R f) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < dim;
         i += uint64_t(gridDim.x) * blockDim.x) {
        auto v = psi[i];
        v.x = v.x / f; v.y = v.y / f;
        psi[i] = v;
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_axpy(typename V2<R>::T *y, const typename V2<R>::T *x, uint64_t dim,
                                              R re, R im) {
    using T = typename V2<R>::T;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < dim;
         i += uint64_t(gridDim.x) * blockDim.x)
        y[i] = cadd(y[i], cmul(x[i], mk<T>(re, im)));
}

// State Add / Sub (state.rs:2773-2800, 2836-2862): out = a + b or a - b per
// component, one rounding each, so the result is the reference's bit for bit.
template <typename R, bool SUB>
__global__ __launch_bounds__(256) void k_addsub(typename V2<R>::T *out, const typename V2<R>::T *a,
                                                const typename V2<R>::T *b, uint64_t dim) {
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < dim;
         i += uint64_t(gridDim.x) * blockDim.x)
        out[i] = SUB ? csub(a[i], b[i]) : cadd(a[i], b[i]);
}

// Whole-state copy in 16-byte units, streamed like the gate kernels.
__global__ __launch_bounds__(kThreads) void k_copy16(f32x4 *__restrict__ dst, const f32x4 *__restrict__ src,
                                                     uint64_t units, uint64_t base) {
    const uint64_t i = base + uint64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i < units) stnt(ldnt(src + i), dst + i);
}

}  // namespace

int pick_blocks(uint64_t items, int per_block) {
    uint64_t b = (items + per_block - 1) / per_block;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return (int)b;
}


Status launch_gate(State &s, const Lowered &L, hipStream_t st) {
    if (L.kind == LK_NOP) return Status{};
    return s.prec == QI_C128 ? launch_c128(s, L, st) : launch_c64(s, L, st);
}

Status launch_init(State &s, int kind, uint64_t basis, hipStream_t st) {
    uint64_t dim = uint64_t(1) << s.n;
    double amp = 0.0;
    if (kind == QI_INIT_PLUS || kind == QI_INIT_MINUS) amp = 1.0 / std::sqrt(double(dim));  // state.rs:233
    if (kind == QI_INIT_GHZ) amp = 0.70710678118654752440;                                    // FRAC_1_SQRT_2
    int blocks = pick_blocks(dim, 256);
    if (s.prec == QI_C128)
        hipLaunchKernelGGL(k_init<double>, dim3(blocks), dim3(256), 0, st, (double2 *)s.d, dim, kind, basis, amp);
    else
        hipLaunchKernelGGL(k_init<float>, dim3(blocks), dim3(256), 0, st, (float2 *)s.d, dim, kind, basis, amp);
    QI_HIP_TRY(hipGetLastError());
    return Status{};
}

Status launch_scale(State &s, double re, double im, hipStream_t st) {
    uint64_t dim = uint64_t(1) << s.n;
    int blocks = pick_blocks(dim, 256);
    if (s.prec == QI_C128)
        hipLaunchKernelGGL(k_scale<double>, dim3(blocks), dim3(256), 0, st, (double2 *)s.d, dim, re, im);
    else
        hipLaunchKernelGGL(k_scale<float>, dim3(blocks), dim3(256), 0, st, (float2 *)s.d, dim, (float)re, (float)im);
    QI_HIP_TRY(hipGetLastError());
    return Status{};
}

Status launch_div_real(State &s, double f, hipStream_t st) {
    uint64_t dim = uint64_t(1) << s.n;
    int blocks = pick_blocks(dim, 256);
    if (s.prec == QI_C128)
        hipLaunchKernelGGL(k_div<double>, dim3(blocks), dim3(256), 0, st, (double2 *)s.d, dim, f);
    else
        hipLaunchKernelGGL(k_div<float>, dim3(blocks), dim3(256), 0, st, (float2 *)s.d, dim, (float)f);
    QI_HIP_TRY(hipGetLastError());
    return Status{};
}

Status launch_axpy(State &y, double re, double im, const State &x, hipStream_t st) {
    uint64_t dim = uint64_t(1) << y.n;
    int blocks = pick_blocks(dim, 256);
    if (y.prec == QI_C128)
        hipLaunchKernelGGL(k_axpy<double>, dim3(blocks), dim3(256), 0, st, (double2 *)y.d, (const double2 *)x.d, dim, re, im);
    else
        hipLaunchKernelGGL(k_axpy<float>, dim3(blocks), dim3(256), 0, st, (float2 *)y.d, (const float2 *)x.d, dim,
                           (float)re, (float)im);
    QI_HIP_TRY(hipGetLastError());
    return Status{};
}

Status launch_addsub(State &out, const State &a, const State &b, bool sub, hipStream_t st) {
    uint64_t dim = uint64_t(1) << out.n;
    int blocks = pick_blocks(dim, 256);
    if (out.prec == QI_C128) {
        auto k = sub ? k_addsub<double, true> : k_addsub<double, false>;
        hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, st, (double2 *)out.d, (const double2 *)a.d, (const double2 *)b.d, dim);
    } else {
        auto k = sub ? k_addsub<float, true> : k_addsub<float, false>;
        hipLaunchKernelGGL(k, dim3(blocks), dim3(256), 0, st, (float2 *)out.d, (const float2 *)a.d, (const float2 *)b.d, dim);
    }
    QI_HIP_TRY(hipGetLastError());
    return Status{};
}

Status launch_copy(State &dst, const State &src, hipStream_t st) {
    const uint64_t units = src.bytes() / 16;  // a state is at least two complex64 amplitudes
    const uint64_t blocks = grid_for(units, kThreads);
    for (uint64_t b0 = 0; b0 < blocks; b0 += kMaxGridBlocks)
        hipLaunchKernelGGL(k_copy16, dim3((unsigned)std::min(kMaxGridBlocks, blocks - b0)), dim3(kThreads), 0, st,
                           reinterpret_cast<f32x4 *>(dst.d), reinterpret_cast<const f32x4 *>(src.d), units, b0 * kThreads);
    QI_HIP_TRY(hipGetLastError());
    return Status{};
}

}  // namespace qi
