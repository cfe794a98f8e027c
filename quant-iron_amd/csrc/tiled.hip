// tiled.hip — register-resident execution of a tile pass (planned in
// tile_plan.hip).
//
// 256 threads hold a tile of 2^K amplitudes, 2^(K-8) per thread in
// registers. A layout assigns the K local index bits to 8 thread bits and
// K-8 register bits. A 2x2 gate on a register-held target touches only the
// thread's own registers: no LDS traffic and no barrier. Diagonal gates and
// controls work on any bit. When the next gates need other register bits,
// the tile is transposed to a new layout through LDS. Controlled SWAPs and
// Matchgates run on the LDS copy of the tile.
//
// Arithmetic: complex128 follows the reference's operand order (no
// contraction); complex64 uses fused multiply-adds.
#include "qi_internal.h"
#include "exec_plan.h"

#pragma clang fp contract(off)

namespace qi {

namespace {

template <typename R> struct V2;
template <> struct V2<float> { using T = float2; };
template <> struct V2<double> { using T = double2; };

template <typename T> __device__ __forceinline__ T mk(decltype(T::x) re, decltype(T::x) im) { T r; r.x = re; r.y = im; return r; }

// a * b for complex values
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return mk<float2>(__builtin_fmaf(-a.y, b.y, a.x * b.x), __builtin_fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return mk<double2>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> __device__ __forceinline__ T cadd(T a, T b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ T csub(T a, T b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <typename T, typename R> __device__ __forceinline__ T cscale(R f, T a) { return mk<T>(f * a.x, f * a.y); }

__device__ __forceinline__ uint32_t ins0(uint32_t k, unsigned p) { return ((k >> p) << (p + 1)) | (k & ((1u << p) - 1)); }

__device__ __forceinline__ uint64_t insert_zeros64(uint64_t k, uint64_t fixed) {
    while (fixed) {
        unsigned p = __builtin_ctzll(fixed);
        k = ((k >> p) << (p + 1)) | (k & ((uint64_t(1) << p) - 1));
        fixed &= fixed - 1;
    }
    return k;
}

constexpr int kTT = 1 << kRegThreadBits;

// The thread's share of the tile: E complex amplitudes held as one vector
// value. Element writes are whole-vector updates of an SSA value, so no
// code path can turn them into stores through selected addresses (which
// would keep the tile in scratch memory).
template <typename R, int E> struct Regs {
    using T = typename V2<R>::T;
    R d[2 * E];
    __device__ __forceinline__ T get(int r) const { return mk<T>(d[2 * r], d[2 * r + 1]); }
    __device__ __forceinline__ void set(int r, T x) { d[2 * r] = x.x; d[2 * r + 1] = x.y; }
};

// LDS slot of local index e: the bits above the low F fold into the low F
// (XOR-linear, a bijection) so strided thread bits spread over the banks.
template <int F> __device__ __forceinline__ uint32_t swz(uint32_t e) {
    static_assert(F == 3 || F == 4, "fold width");
    if constexpr (F == 4) return e ^ (((e >> 4) ^ (e >> 8) ^ (e >> 12)) & 15u);
    else return e ^ (((e >> 3) ^ (e >> 6) ^ (e >> 9) ^ (e >> 12)) & 7u);
}

// --------------------------------------------------------------- 2x2 ops --
template <int KIND, typename R> __device__ __forceinline__ void pair(float2 &a, float2 &b, const R *c) {
    if constexpr (KIND == RO_X) {
        const float2 t = a; a = b; b = t;
    } else if constexpr (KIND == RO_H) {
        const float2 s = cadd(a, b), d = csub(a, b);
        a = cscale<float2>(c[0], s);
        b = cscale<float2>(c[0], d);
    } else if constexpr (KIND == RO_REAL) {  // [[c, -s], [s, c]]
        const float2 na = mk<float2>(__builtin_fmaf(-c[1], b.x, c[0] * a.x), __builtin_fmaf(-c[1], b.y, c[0] * a.y));
        const float2 nb = mk<float2>(__builtin_fmaf(c[1], a.x, c[0] * b.x), __builtin_fmaf(c[1], a.y, c[0] * b.y));
        a = na; b = nb;
    } else {  // general complex 2x2, row-major
        float2 na, nb;
        na.x = __builtin_fmaf(-c[3], b.y, __builtin_fmaf(c[2], b.x, __builtin_fmaf(-c[1], a.y, c[0] * a.x)));
        na.y = __builtin_fmaf(c[3], b.x, __builtin_fmaf(c[2], b.y, __builtin_fmaf(c[1], a.x, c[0] * a.y)));
        nb.x = __builtin_fmaf(-c[7], b.y, __builtin_fmaf(c[6], b.x, __builtin_fmaf(-c[5], a.y, c[4] * a.x)));
        nb.y = __builtin_fmaf(c[7], b.x, __builtin_fmaf(c[6], b.y, __builtin_fmaf(c[5], a.x, c[4] * a.y)));
        a = na; b = nb;
    }
}
template <int KIND, typename R> __device__ __forceinline__ void pair(double2 &a, double2 &b, const R *c) {
    if constexpr (KIND == RO_X) {
        const double2 t = a; a = b; b = t;
    } else if constexpr (KIND == RO_H) {
        const double2 s = cadd(a, b), d = csub(a, b);
        a = cscale<double2>(c[0], s);
        b = cscale<double2>(c[0], d);
    } else if constexpr (KIND == RO_REAL) {
        const double2 na = csub(cscale<double2>(c[0], a), cscale<double2>(c[1], b));
        const double2 nb = cadd(cscale<double2>(c[1], a), cscale<double2>(c[0], b));
        a = na; b = nb;
    } else {
        const double2 na = cadd(cmul(mk<double2>(c[0], c[1]), a), cmul(mk<double2>(c[2], c[3]), b));
        const double2 nb = cadd(cmul(mk<double2>(c[4], c[5]), a), cmul(mk<double2>(c[6], c[7]), b));
        a = na; b = nb;
    }
}

// Element updates carry no control flow: a controlled op computes every pair
// and keeps the old values by a select, so the tile stays in the same
// registers on every path through the op switch.
template <typename T> __device__ __forceinline__ T sel(bool on, T x, T y) { return mk<T>(on ? x.x : y.x, on ? x.y : y.y); }

template <int B, int KIND, bool CTRL, int E, typename T, typename R>
__device__ __forceinline__ void pairs(Regs<R, E> &v, const R *c, uint32_t rcm, bool tok) {
#pragma unroll
    for (int r = 0; r < E; r++) {
        if (r & (1 << B)) continue;
        const T a = v.get(r), b = v.get(r | (1 << B));
        T na = a, nb = b;
        pair<KIND, R>(na, nb, c);
        if constexpr (CTRL) {
            const bool on = tok && ((uint32_t(r) & rcm) == rcm);
            na = sel(on, na, a);
            nb = sel(on, nb, b);
        }
        v.set(r, na);
        v.set(r | (1 << B), nb);
    }
}

template <int KIND, bool CTRL, int RB, int E, typename T, typename R>
__device__ __forceinline__ void pairs_at(Regs<R, E> &v, int rb, const R *c, uint32_t rcm, bool tok) {
#define QI_RB(BB) case BB: if constexpr (BB < RB) pairs<BB, KIND, CTRL, E, T, R>(v, c, rcm, tok); break;
    switch (rb) { QI_RB(0) QI_RB(1) QI_RB(2) QI_RB(3) QI_RB(4) QI_RB(5) default: break; }
#undef QI_RB
}

// ---------------------------------------------------------- diagonal ops --
template <int KIND, typename T> __device__ __forceinline__ T dphase(T a, T p0, T p1, bool one) {
    if constexpr (KIND == RO_Z) return sel(one, mk<T>(-a.x, -a.y), a);
    else if constexpr (KIND == RO_PHASE) return sel(one, cmul(a, p1), a);
    else return cmul(a, sel(one, p1, p0));
}

// B >= 0: target on register bit B; B < 0: target fixed per thread (`tone`)
template <int B, int KIND, bool CTRL, int E, typename T, typename R>
__device__ __forceinline__ void diag(Regs<R, E> &v, T p0, T p1, uint32_t rcm, bool tok, bool tone) {
#pragma unroll
    for (int r = 0; r < E; r++) {
        const bool one = B >= 0 ? (((r >> (B >= 0 ? B : 0)) & 1) != 0) : tone;
        const T a = v.get(r);
        T na = dphase<KIND, T>(a, p0, p1, one);
        if constexpr (CTRL) na = sel(tok && ((uint32_t(r) & rcm) == rcm), na, a);
        v.set(r, na);
    }
}
template <int KIND, bool CTRL, int RB, int E, typename T, typename R>
__device__ __forceinline__ void diag_at(Regs<R, E> &v, int dloc, int dbit, T p0, T p1, uint32_t rcm, bool tok, bool tone) {
    if (dloc != 0) { diag<-1, KIND, CTRL, E, T, R>(v, p0, p1, rcm, tok, tone); return; }
#define QI_RB(BB) case BB: if constexpr (BB < RB) diag<BB, KIND, CTRL, E, T, R>(v, p0, p1, rcm, tok, tone); break;
    switch (dbit) { QI_RB(0) QI_RB(1) QI_RB(2) QI_RB(3) QI_RB(4) QI_RB(5) default: break; }
#undef QI_RB
}

// ------------------------------------------------------------- LDS ops --
// Controlled SWAP or Matchgate (CPU form, operator.rs:993-1004) on the LDS
// copy; local index e lives at slot swz(e).
template <typename R, int K, int F>
__device__ __forceinline__ void lds_gate(typename V2<R>::T *tile, const RegOp &op) {
    using T = typename V2<R>::T;
    constexpr int N = 1 << K;
    const uint32_t lcm = op.lcm, b0 = 1u << op.lp0, b1 = 1u << op.lp1;
    const unsigned lo = op.lp0 < op.lp1 ? op.lp0 : op.lp1, hi = op.lp0 < op.lp1 ? op.lp1 : op.lp0;
    const R c0 = R(op.c[0]), c1 = R(op.c[1]);
    const T m23 = mk<T>(R(op.c[2]), R(op.c[3])), m45 = mk<T>(R(op.c[4]), R(op.c[5])), m67 = mk<T>(R(op.c[6]), R(op.c[7]));
    for (uint32_t q = threadIdx.x; q < N / 4; q += kTT) {
        const uint32_t l = ins0(ins0(q, lo), hi);
        if (op.lkind == LK_SWAP) {
            const uint32_t i = l | b0, j = l | b1;
            if ((i & lcm) != lcm) continue;
            const T a = tile[swz<F>(i)], b = tile[swz<F>(j)];
            tile[swz<F>(i)] = b;
            tile[swz<F>(j)] = a;
        } else {
            const uint32_t i01 = l | b0, i10 = l | b1, i11 = l | b0 | b1;
            if ((i01 & lcm) == lcm) {
                const T a01 = tile[swz<F>(i01)], a10 = tile[swz<F>(i10)];
                tile[swz<F>(i01)] = csub(cscale<T>(c0, a01), cmul(m23, a10));
                tile[swz<F>(i10)] = cadd(cscale<T>(c1, a01), cmul(m45, a10));
            }
            if ((i11 & lcm) == lcm) tile[swz<F>(i11)] = cmul(tile[swz<F>(i11)], m67);
        }
    }
}

// ------------------------------------------------------------ the pass --
template <typename R, int K>
__global__ __launch_bounds__(kTT) void k_reg_pass(typename V2<R>::T *__restrict__ psi, const RegPass pass,
                                                  const RegOp *__restrict__ ops) {
    using T = typename V2<R>::T;
    constexpr int RB = K - kRegThreadBits, E = 1 << RB, F = sizeof(R) == 8 ? 3 : 4;
    __shared__ T lds[1 << K];
    const uint32_t tid = threadIdx.x;
    const uint64_t base = insert_zeros64(uint64_t(blockIdx.x), pass.qmask);

    Regs<R, E> v;
    uint32_t mr[RB];  // local-bit mask of each register bit (uniform)
    uint32_t tl = 0;  // local index bits held by this thread
    {
        uint64_t at = 0;
#pragma unroll
        for (int i = 0; i < kRegThreadBits; i++) {
            const uint32_t bit = (tid >> i) & 1u;
            tl |= bit << pass.lay_t[i];
            at |= uint64_t(bit) << pass.qt_in[i];
        }
        uint64_t ar[RB];
#pragma unroll
        for (int j = 0; j < RB; j++) { mr[j] = 1u << pass.lay_r[j]; ar[j] = uint64_t(1) << pass.qr_in[j]; }
        const T *src = psi + (base | at);
#pragma unroll
        for (int r = 0; r < E; r++) {
            uint64_t o = 0;
#pragma unroll
            for (int j = 0; j < RB; j++) if ((r >> j) & 1) o |= ar[j];
            v.set(r, src[o]);
        }
    }
    bool dirty = false;  // LDS written since the last barrier
    // the op table of a pass ends with a no-op, so op o + 1 can be fetched
    // while op o runs
    RegOp nxt = ops[pass.op0];
    for (int o = 0; o < pass.nops; o++) {
        const RegOp op = nxt;
        nxt = ops[pass.op0 + o + 1];
        const int kind = op.kind;
        if ((base & op.gcm) != op.gcm) continue;  // controls outside the tile (uniform)
        const bool ctrl = (op.rcm | op.tcm) != 0;  // controls held in the tile (uniform)
        const bool tok = (tl & op.tcm) == op.tcm;  // this thread's thread-held controls
        if (kind <= RO_GEN) {
            R c[8];
#pragma unroll
            for (int k = 0; k < 8; k++) c[k] = R(op.c[k]);
#define QI_PAIR(KD)                                                                     \
    case KD:                                                                            \
        if (ctrl) pairs_at<KD, true, RB, E, T, R>(v, op.rb, c, op.rcm, tok);            \
        else pairs_at<KD, false, RB, E, T, R>(v, op.rb, c, op.rcm, tok);                \
        break;
            switch (kind) { QI_PAIR(RO_X) QI_PAIR(RO_H) QI_PAIR(RO_REAL) QI_PAIR(RO_GEN) default: break; }
#undef QI_PAIR
        } else if (kind <= RO_RZ) {
            const T p0 = mk<T>(R(op.c[0]), R(op.c[1])), p1 = mk<T>(R(op.c[2]), R(op.c[3]));
            const bool tone = op.dloc == 1 ? ((tl >> op.dbit) & 1u) != 0 : ((base >> op.dbit) & 1u) != 0;
#define QI_DIAG(KD)                                                                             \
    case KD:                                                                                    \
        if (ctrl) diag_at<KD, true, RB, E, T, R>(v, op.dloc, op.dbit, p0, p1, op.rcm, tok, tone);  \
        else diag_at<KD, false, RB, E, T, R>(v, op.dloc, op.dbit, p0, p1, op.rcm, tok, tone);      \
        break;
            switch (kind) { QI_DIAG(RO_Z) QI_DIAG(RO_PHASE) QI_DIAG(RO_RZ) default: break; }
#undef QI_DIAG
        } else if (kind <= RO_LDS) {
            // through LDS: a transpose to a new layout, or a gate on the LDS copy
            if (dirty) __syncthreads();
            {
                const uint32_t st = swz<F>(tl);
#pragma unroll
                for (int r = 0; r < E; r++) {
                    uint32_t e = 0;
#pragma unroll
                    for (int j = 0; j < RB; j++) if ((r >> j) & 1) e |= mr[j];
                    lds[st ^ swz<F>(e)] = v.get(r);
                }
            }
            __syncthreads();
            if (kind == RO_TRANSPOSE) {
                tl = 0;
#pragma unroll
                for (int i = 0; i < kRegThreadBits; i++) tl |= ((tid >> i) & 1u) << op.lay_t[i];
#pragma unroll
                for (int j = 0; j < RB; j++) mr[j] = 1u << op.lay_r[j];
            } else {
                lds_gate<R, K, F>(lds, op);
                __syncthreads();
            }
            {
                const uint32_t st = swz<F>(tl);
#pragma unroll
                for (int r = 0; r < E; r++) {
                    uint32_t e = 0;
#pragma unroll
                    for (int j = 0; j < RB; j++) if ((r >> j) & 1) e |= mr[j];
                    v.set(r, lds[st ^ swz<F>(e)]);
                }
            }
            dirty = true;
        }
    }
    {
        // store: thread and register bits hold the qubits of the final layout
        uint64_t at = 0;
#pragma unroll
        for (int i = 0; i < kRegThreadBits; i++) at |= uint64_t((tid >> i) & 1u) << pass.qt_out[i];
        uint64_t ar[RB];
#pragma unroll
        for (int j = 0; j < RB; j++) ar[j] = uint64_t(1) << pass.qr_out[j];
        T *dst = psi + (base | at);
#pragma unroll
        for (int r = 0; r < E; r++) {
            uint64_t o = 0;
#pragma unroll
            for (int j = 0; j < RB; j++) if ((r >> j) & 1) o |= ar[j];
            dst[o] = v.get(r);
        }
    }
}

template <typename R, int K> void launch_pass(State &s, const RegPass &p, const RegOp *ops, hipStream_t st) {
    using T = typename V2<R>::T;
    const unsigned tiles = (unsigned)(uint64_t(1) << (s.n - K));
    hipLaunchKernelGGL((k_reg_pass<R, K>), dim3(tiles), dim3(kTT), 0, st, (T *)s.d, p, ops);
}

}  // namespace

Status launch_reg_pass(State &s, const RegPass &p, const RegOp *dops, hipStream_t st) {
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
