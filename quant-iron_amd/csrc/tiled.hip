// tiled.hip — register-resident execution of a tile pass (planned in
// tile_plan.hip).
//
// 256 threads hold a tile of 2^K amplitudes, 2^(K-8) per thread in
// registers. A layout assigns the K local index bits to 8 thread bits (6 lane
// bits, 2 wave bits) and K-8 register bits. A 2x2 gate acts on a lane-held
// target: every lane fetches its partner's amplitudes from lane ^ 2^i with
// ds_bpermute (the LDS crossbar, no LDS memory, no barrier) and computes its
// own half of the pair, so the code of an op is the same for every target
// and every register. Diagonal gates and controls test the local index
// (thread part | register part) of each amplitude. When the next gates need
// other target bits on lanes, the tile is transposed to a new layout through
// LDS. Controlled SWAPs and Matchgates run on the LDS copy of the tile.
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

template <typename T> __device__ __forceinline__ T sel(bool on, T x, T y) { return mk<T>(on ? x.x : y.x, on ? x.y : y.y); }

// The amplitude of lane `addr / 4` (ds_bpermute: the LDS crossbar, no LDS memory).
__device__ __forceinline__ float xlane(int addr, float x) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(x)));
}
__device__ __forceinline__ double xlane(int addr, double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_ds_bpermute(addr, (int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_ds_bpermute(addr, (int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <typename T> __device__ __forceinline__ T xlane2(int addr, T x) { return mk<T>(xlane(addr, x.x), xlane(addr, x.y)); }

// ------------------------------------------------------------- 2x2 ops --
// Lane half `bit` of the pair (a, b) on its target: bit 0 holds a and
// computes a', bit 1 holds b and computes b'; s = own, p = partner.
//   X:    a' = b,               b' = a
//   H:    a' = h (a + b),       b' = h (a - b)          -> h (sgn s + p)
//   REAL: a' = c a - s b,       b' = s a + c b          -> c s + (+-s) p
//   GEN:  a' = m00 a + m01 b,   b' = m10 a + m11 b      -> cs s + cp p
// The complex128 forms keep the reference's products and only reorder
// commutative sums, which IEEE arithmetic evaluates exactly alike.
template <int KIND, typename T, typename R>
__device__ __forceinline__ T pair_lane(T s, T p, const R *k) {
    if constexpr (KIND == RO_X) {
        return p;
    } else if constexpr (KIND == RO_H) {  // k0 = sgn, k1 = h
        if constexpr (sizeof(R) == 4) return mk<T>(k[1] * __builtin_fmaf(k[0], s.x, p.x), k[1] * __builtin_fmaf(k[0], s.y, p.y));
        else return mk<T>(k[1] * (k[0] * s.x + p.x), k[1] * (k[0] * s.y + p.y));
    } else if constexpr (KIND == RO_REAL) {  // k0 = c, k1 = +-s
        if constexpr (sizeof(R) == 4)
            return mk<T>(__builtin_fmaf(k[1], p.x, k[0] * s.x), __builtin_fmaf(k[1], p.y, k[0] * s.y));
        else return mk<T>(k[0] * s.x + k[1] * p.x, k[0] * s.y + k[1] * p.y);
    } else {  // k0,k1 = cs, k2,k3 = cp
        if constexpr (sizeof(R) == 4) {
            float x = __builtin_fmaf(-k[1], s.y, k[0] * s.x);
            x = __builtin_fmaf(k[2], p.x, x);
            x = __builtin_fmaf(-k[3], p.y, x);
            float y = __builtin_fmaf(k[1], s.x, k[0] * s.y);
            y = __builtin_fmaf(k[2], p.y, y);
            y = __builtin_fmaf(k[3], p.x, y);
            return mk<T>(x, y);
        } else {
            return mk<T>((k[0] * s.x - k[1] * s.y) + (k[2] * p.x - k[3] * p.y),
                         (k[0] * s.y + k[1] * s.x) + (k[2] * p.y + k[3] * p.x));
        }
    }
}

template <int KIND, bool CTRL, int E, typename T, typename R>
__device__ __forceinline__ void pairs(Regs<R, E> &v, int addr, const R *k, uint32_t tl, const uint32_t *rl, uint32_t cm) {
#pragma unroll
    for (int r = 0; r < E; r++) {
        const T s = v.get(r);
        const T p = xlane2<T>(addr, s);
        T n = pair_lane<KIND, T, R>(s, p, k);
        if constexpr (CTRL) n = sel(((tl | rl[r]) & cm) == cm, n, s);
        v.set(r, n);
    }
}

// ---------------------------------------------------------- diagonal ops --
template <int KIND, typename T> __device__ __forceinline__ T dphase(T a, T p0, T p1, bool one) {
    if constexpr (KIND == RO_Z) return sel(one, mk<T>(-a.x, -a.y), a);
    else if constexpr (KIND == RO_PHASE) return sel(one, cmul(a, p1), a);
    else return cmul(a, sel(one, p1, p0));
}

// one(r) = target bit of amplitude r: thread part `tb` | register part of rl[r]
template <int KIND, bool CTRL, int E, typename T, typename R>
__device__ __forceinline__ void diag(Regs<R, E> &v, T p0, T p1, uint32_t tb, uint32_t dmask, uint32_t tl, const uint32_t *rl,
                                     uint32_t cm) {
#pragma unroll
    for (int r = 0; r < E; r++) {
        const bool one = (tb | (rl[r] & dmask)) != 0;
        const T a = v.get(r);
        T n = dphase<KIND, T>(a, p0, p1, one);
        if constexpr (CTRL) n = sel(((tl | rl[r]) & cm) == cm, n, a);
        v.set(r, n);
    }
}

// ------------------------------------------------------------- LDS ops --
// Controlled SWAP or Matchgate (CPU form, operator.rs:993-1004) on the LDS
// copy; local index e lives at slot swz(e).
template <typename R, int K, int F>
__device__ __forceinline__ void lds_gate(typename V2<R>::T *tile, const RegOp &op) {
    using T = typename V2<R>::T;
    constexpr int N = 1 << K;
    const uint32_t lcm = op.cm, b0 = 1u << op.lp0, b1 = 1u << op.lp1;
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
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const uint64_t base = insert_zeros64(uint64_t(blockIdx.x), pass.qmask);

    Regs<R, E> v;
    uint32_t tl = 0;  // local index bits held by this thread
    uint32_t rl[E];   // local index bits of register r (uniform)
    {
        uint64_t at = 0;
#pragma unroll
        for (int i = 0; i < kRegThreadBits; i++) {
            const uint32_t bit = (tid >> i) & 1u;
            tl |= bit << pass.lay_t[i];
            at |= uint64_t(bit) << pass.qt_in[i];
        }
        uint64_t ar[RB];
        uint32_t mr[RB];
#pragma unroll
        for (int j = 0; j < RB; j++) { mr[j] = 1u << pass.lay_r[j]; ar[j] = uint64_t(1) << pass.qr_in[j]; }
        const T *src = psi + (base | at);
#pragma unroll
        for (int r = 0; r < E; r++) {
            uint64_t o = 0;
            uint32_t e = 0;
#pragma unroll
            for (int j = 0; j < RB; j++)
                if ((r >> j) & 1) { o |= ar[j]; e |= mr[j]; }
            rl[r] = e;
            v.set(r, src[o]);
        }
    }
    bool dirty = false;  // LDS written since the last barrier
    // The op table of a pass ends with a no-op, so op o + 1 can be fetched
    // while op o runs. In-register ops run in an inner loop; LDS ops
    // (transposes, LDS gates) sit outside it.
    RegOp nxt = ops[pass.op0];
    int o = 0;
    for (;;) {
        for (; o < pass.nops; o++) {
            const RegOp op = nxt;
            const int kind = op.kind;
            if (kind >= RO_TRANSPOSE) break;
            nxt = ops[pass.op0 + o + 1];
            if ((base & op.gcm) != op.gcm) continue;  // controls outside the tile (uniform)
            const bool ctrl = op.cm != 0;             // controls inside the tile (uniform)
            if (kind <= RO_GEN) {
                const int addr = int((lane ^ (1u << op.lane)) << 2);
                const bool bit = ((lane >> op.lane) & 1u) != 0;
                R k[4];
                if (kind == RO_H) {
                    k[0] = bit ? R(-1) : R(1);
                    k[1] = R(op.c[0]);
                } else if (kind == RO_REAL) {
                    k[0] = R(op.c[0]);
                    k[1] = bit ? R(op.c[1]) : -R(op.c[1]);
                } else {
                    k[0] = R(bit ? op.c[6] : op.c[0]);
                    k[1] = R(bit ? op.c[7] : op.c[1]);
                    k[2] = R(bit ? op.c[4] : op.c[2]);
                    k[3] = R(bit ? op.c[5] : op.c[3]);
                }
#define QI_PAIR(KD)                                                          \
    case KD:                                                                 \
        if (ctrl) pairs<KD, true, E, T, R>(v, addr, k, tl, rl, op.cm);       \
        else pairs<KD, false, E, T, R>(v, addr, k, tl, rl, op.cm);           \
        break;
                switch (kind) { QI_PAIR(RO_X) QI_PAIR(RO_H) QI_PAIR(RO_REAL) QI_PAIR(RO_GEN) default: break; }
#undef QI_PAIR
            } else {
                const T p0 = mk<T>(R(op.c[0]), R(op.c[1])), p1 = mk<T>(R(op.c[2]), R(op.c[3]));
                uint32_t tb, dmask;
                if (op.dloc == 0) { dmask = 1u << op.dbit; tb = tl & dmask; }
                else { dmask = 0; tb = uint32_t((base >> op.dbit) & 1u); }
#define QI_DIAG(KD)                                                                  \
    case KD:                                                                         \
        if (ctrl) diag<KD, true, E, T, R>(v, p0, p1, tb, dmask, tl, rl, op.cm);      \
        else diag<KD, false, E, T, R>(v, p0, p1, tb, dmask, tl, rl, op.cm);          \
        break;
                switch (kind) { QI_DIAG(RO_Z) QI_DIAG(RO_PHASE) QI_DIAG(RO_RZ) default: break; }
#undef QI_DIAG
            }
        }
        if (o >= pass.nops) break;
        // through LDS: a transpose to a new layout, or a gate on the LDS copy
        const RegOp op = nxt;
        nxt = ops[pass.op0 + o + 1];
        o++;
        if ((base & op.gcm) != op.gcm) continue;
        if (dirty) __syncthreads();
        char *ldsb = reinterpret_cast<char *>(lds);
        {
            const uint32_t stb = swz<F>(tl) * uint32_t(sizeof(T));  // XOR-linear: slot(tl | e) = slot(tl) ^ slot(e)
#pragma unroll
            for (int r = 0; r < E; r++)
                *reinterpret_cast<T *>(ldsb + (stb ^ (swz<F>(rl[r]) * uint32_t(sizeof(T))))) = v.get(r);
        }
        __syncthreads();
        if (op.kind == RO_TRANSPOSE) {
            tl = 0;
#pragma unroll
            for (int i = 0; i < kRegThreadBits; i++) tl |= ((tid >> i) & 1u) << op.lay_t[i];
            uint32_t mr[RB];
#pragma unroll
            for (int j = 0; j < RB; j++) mr[j] = 1u << op.lay_r[j];
#pragma unroll
            for (int r = 0; r < E; r++) {
                uint32_t e = 0;
#pragma unroll
                for (int j = 0; j < RB; j++) if ((r >> j) & 1) e |= mr[j];
                rl[r] = e;
            }
        } else {
            lds_gate<R, K, F>(lds, op);
            __syncthreads();
        }
        {
            const uint32_t stb = swz<F>(tl) * uint32_t(sizeof(T));
#pragma unroll
            for (int r = 0; r < E; r++)
                v.set(r, *reinterpret_cast<const T *>(ldsb + (stb ^ (swz<F>(rl[r]) * uint32_t(sizeof(T))))));
        }
        dirty = true;
    }
    // In-place permutation (SWAP relabels, a final layout that differs from the
    // load layout): an amplitude may be stored where another wave has not
    // loaded yet, so every wave's loads complete (s_waitcnt 0: the barrier alone
    // does not wait for loads in flight) before any wave stores.
    if (pass.sync_store) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
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
