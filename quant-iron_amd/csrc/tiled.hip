// tiled.hip — register-resident execution of a tile pass (planned in
// tile_plan.cpp).
//
// 2^TB threads hold a tile of 2^K amplitudes, 2^(K-TB) per thread in
// registers. A layout assigns the K local index bits to TB thread bits (6
// lane bits, TB-6 wave bits) and K-TB register bits. A 2x2 gate acts on a
// target held in registers (the pair lives in one thread) or on lanes (each
// lane fetches its partner with ds_bpermute, the LDS crossbar without LDS
// memory, and computes its half of the pair). Controls and diagonal targets
// on register bits are compile-time facts of each register; on thread bits
// they are a per-lane mask. With wave bits (TB = 8, 9), a target on a wave bit
// takes a transpose of the tile through LDS, and controlled SWAPs and
// Matchgates run on the LDS copy. A one-wave tile (TB = 6: complex64, up to
// 4096 amplitudes, 64 per lane) has no wave bits: no LDS and no barriers;
// its only layout changes are lane/register exchanges at the end of a pass.
//
// Arithmetic: both precisions use fused multiply-adds in the register ops
// (complex128 tiles agree with the reference's unfused products to about
// 1e-15 per gate); the LDS gates keep the reference's products.
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

// Tile loads and stores stream the state once per pass: non-temporal, so
// they do not displace cached data (the op table) for data never re-read.
template <typename T> __device__ __forceinline__ T ld_stream(const T *p) {
    typedef decltype(T::x) E2 __attribute__((ext_vector_type(2)));
    const E2 v = __builtin_nontemporal_load(reinterpret_cast<const E2 *>(p));
    T r; r.x = v.x; r.y = v.y;
    return r;
}
template <typename T> __device__ __forceinline__ void st_stream(T x, T *p) {
    typedef decltype(T::x) E2 __attribute__((ext_vector_type(2)));
    E2 v; v.x = x.x; v.y = x.y;
    __builtin_nontemporal_store(v, reinterpret_cast<E2 *>(p));
}

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

// ------------------------------------------------------ in-place updates --
// A tile lives in registers across a loop whose body dispatches on
// the op kind. Each op writes every amplitude component back into the
// register it was read from (tied asm operands), so all op paths leave the
// tile in the same registers and no copies are needed where they merge.
// Temporaries are ordinary values. Only VALU instructions appear here; the
// compiler places the waits for loads and LDS results as for its own code.
// Ops stay straight-line: with controls, new values are computed into
// temporaries and selected in place under a lane mask (no branches, whose
// joins would again need copies).
__device__ __forceinline__ void ip_fma(float &d, float k, float t) {  // d = k * d + t
    asm("v_fma_f32 %0, %1, %0, %2" : "+v"(d) : "v"(k), "v"(t));
}
__device__ __forceinline__ void ip_fma3(float &d, float k, float x, float t) {  // d = k * x + t (old d dead)
    asm("v_fma_f32 %0, %1, %2, %3" : "+v"(d) : "v"(k), "v"(x), "v"(t));
}
__device__ __forceinline__ void ip_fmac(float &d, float k, float x) {  // d += k * x
    asm("v_fmac_f32 %0, %1, %2" : "+v"(d) : "v"(k), "v"(x));
}
__device__ __forceinline__ void ip_set(float &d, float x) { asm("v_mov_b32 %0, %1" : "+v"(d) : "v"(x)); }
__device__ __forceinline__ void ip_swap(float &x, float &y) { asm("v_swap_b32 %0, %1" : "+v"(x), "+v"(y)); }
__device__ __forceinline__ void ip_neg(float &d) { asm("v_xor_b32 %0, 0x80000000, %0" : "+v"(d)); }
// d = lane in mask ? x : d
__device__ __forceinline__ void ip_sel(float &d, float x, uint64_t mask) {
    asm("v_cndmask_b32 %0, %0, %1, %2" : "+v"(d) : "v"(x), "s"(mask));
}
// lane in mask ? x : y, as a new value
__device__ __forceinline__ float selm(uint64_t mask, float x, float y) {
    float r;
    asm("v_cndmask_b32 %0, %1, %2, %3" : "=v"(r) : "v"(y), "v"(x), "s"(mask));
    return r;
}
// The same for complex128 components: two 32-bit selects on value operands
// (a plain ?: between two tile registers can become a load through a selected
// address, which keeps the whole tile in scratch memory).
__device__ __forceinline__ double selm(uint64_t mask, double x, double y) {
    const uint64_t bx = (uint64_t)__double_as_longlong(x), by = (uint64_t)__double_as_longlong(y);
    uint32_t lo, hi;
    asm("v_cndmask_b32 %0, %1, %2, %3" : "=v"(lo) : "v"((uint32_t)by), "v"((uint32_t)bx), "s"(mask));
    asm("v_cndmask_b32 %0, %1, %2, %3" : "=v"(hi) : "v"((uint32_t)(by >> 32)), "v"((uint32_t)(bx >> 32)), "s"(mask));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ void ip_xor(float &d, uint32_t m) { asm("v_xor_b32 %0, %1, %0" : "+v"(d) : "v"(m)); }
__device__ __forceinline__ void ip_xor3(float &d, float x, float m) {  // d = x ^ m (bit patterns; old d dead)
    asm("v_xor_b32 %0, %1, %2" : "+v"(d) : "v"(__float_as_uint(m)), "v"(x));
}
__device__ __forceinline__ float xor_bits(float x, float m) { return __uint_as_float(__float_as_uint(x) ^ __float_as_uint(m)); }
__device__ __forceinline__ float xor_sign(float x, uint32_t m) { return __uint_as_float(__float_as_uint(x) ^ m); }
__device__ __forceinline__ float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// The same in-place updates for complex128 tiles. FP64 arithmetic works on
// register pairs; moves, selects and sign flips work on the 32-bit halves
// (each half tied to its own register, so the pair stays in place).
__device__ __forceinline__ uint32_t lo32(double x) { return (uint32_t)(uint64_t)__double_as_longlong(x); }
__device__ __forceinline__ uint32_t hi32(double x) { return (uint32_t)((uint64_t)__double_as_longlong(x) >> 32); }
__device__ __forceinline__ double mk64(uint32_t lo, uint32_t hi) {
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ void ip_fma(double &d, double k, double t) {  // d = k * d + t
    asm("v_fma_f64 %0, %1, %0, %2" : "+v"(d) : "v"(k), "v"(t));
}
__device__ __forceinline__ void ip_fma3(double &d, double k, double x, double t) {  // d = k * x + t (old d dead)
    asm("v_fma_f64 %0, %1, %2, %3" : "+v"(d) : "v"(k), "v"(x), "v"(t));
}
__device__ __forceinline__ void ip_fmac(double &d, double k, double x) {  // d += k * x
    asm("v_fmac_f64 %0, %1, %2" : "+v"(d) : "v"(k), "v"(x));
}
__device__ __forceinline__ void ip_set(double &d, double x) { asm("v_mov_b64 %0, %1" : "+v"(d) : "v"(x)); }
__device__ __forceinline__ void ip_swap(double &x, double &y) {
    uint32_t xl = lo32(x), xh = hi32(x), yl = lo32(y), yh = hi32(y);
    asm("v_swap_b32 %0, %1" : "+v"(xl), "+v"(yl));
    asm("v_swap_b32 %0, %1" : "+v"(xh), "+v"(yh));
    x = mk64(xl, xh);
    y = mk64(yl, yh);
}
__device__ __forceinline__ void ip_xor(double &d, uint32_t m) {  // sign-word XOR
    uint32_t h = hi32(d);
    asm("v_xor_b32 %0, %1, %0" : "+v"(h) : "v"(m));
    d = mk64(lo32(d), h);
}
__device__ __forceinline__ void ip_neg(double &d) {
    uint32_t h = hi32(d);
    asm("v_xor_b32 %0, 0x80000000, %0" : "+v"(h));
    d = mk64(lo32(d), h);
}
__device__ __forceinline__ void ip_sel(double &d, double x, uint64_t mask) {  // d = lane in mask ? x : d
    uint32_t l = lo32(d), h = hi32(d);
    asm("v_cndmask_b32 %0, %0, %1, %2" : "+v"(l) : "v"(lo32(x)), "s"(mask));
    asm("v_cndmask_b32 %0, %0, %1, %2" : "+v"(h) : "v"(hi32(x)), "s"(mask));
    d = mk64(l, h);
}
__device__ __forceinline__ void ip_xor3(double &d, double x, double m) {  // d = x ^ m (bit patterns; old d dead)
    uint32_t l = lo32(d), h = hi32(d);
    asm("v_xor_b32 %0, %1, %2" : "+v"(l) : "v"(lo32(m)), "v"(lo32(x)));
    asm("v_xor_b32 %0, %1, %2" : "+v"(h) : "v"(hi32(m)), "v"(hi32(x)));
    d = mk64(l, h);
}
__device__ __forceinline__ double xor_bits(double x, double m) {
    return __longlong_as_double(__double_as_longlong(x) ^ __double_as_longlong(m));
}
__device__ __forceinline__ double xor_sign(double x, uint32_t m) {
    return __longlong_as_double(__double_as_longlong(x) ^ (long long)((uint64_t)m << 32));
}
__device__ __forceinline__ double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }
// the sign-flip pattern of a component as a value of its own type
template <typename R> __device__ __forceinline__ R sign_bits(bool neg) {
    if constexpr (sizeof(R) == 4) return __uint_as_float(neg ? 0x80000000u : 0u);
    else return __longlong_as_double(neg ? (long long)0x8000000000000000ull : 0ll);
}

// m0 a + m1 b for complex m0 = (k0, k1), m1 = (k2, k3), as a new value
template <typename T, typename R> __device__ __forceinline__ T cmul_add(const R *k, T a, T b) {
    R x = fma_r(-k[1], a.y, k[0] * a.x);
    x = fma_r(k[2], b.x, x);
    x = fma_r(-k[3], b.y, x);
    R y = fma_r(k[1], a.x, k[0] * a.y);
    y = fma_r(k[2], b.y, y);
    y = fma_r(k[3], b.x, y);
    return mk<T>(x, y);
}
// (a, b) <- (m0 a + m1 b, m2 a + m3 b), real m, one component
template <typename R> __device__ __forceinline__ void ip_rr(R &a, R &b, const R *k) {
    const R t1 = k[2] * a, t2 = k[1] * b;
    ip_fma(b, k[3], t1);
    ip_fma(a, k[0], t2);
}
// (a, b) <- (c0 a + c1 b, c2 a + c3 b), complex c_i = (k[2i], k[2i+1]); nk = -k
template <typename R> __device__ __forceinline__ void ip_gen(R &ax, R &ay, R &bx, R &by, const R *k, const R *nk) {
    const R p = fma_r(nk[1], ay, k[0] * ax), q = fma_r(k[0], ay, k[1] * ax);
    const R r = fma_r(nk[5], ay, k[4] * ax), s = fma_r(k[4], ay, k[5] * ax);
    ip_fma3(ax, k[2], bx, p);
    ip_fmac(ax, nk[3], by);
    ip_fma3(ay, k[2], by, q);
    ip_fmac(ay, k[3], bx);
    const R z2 = fma_r(nk[7], by, r), z = fma_r(k[6], by, s);
    ip_fma3(by, k[7], bx, z);
    ip_fma(bx, k[6], z2);
}
// (a, b) <- (c0 a + i c1 b, i c2 a + c3 b), real c; nk = -k
template <typename R> __device__ __forceinline__ void ip_rxi(R &ax, R &ay, R &bx, R &by, const R *k, const R *nk) {
    const R t1 = nk[1] * by, t2 = k[1] * bx, t3 = nk[2] * ay, t4 = k[2] * ax;
    ip_fma(ax, k[0], t1);
    ip_fma(ay, k[0], t2);
    ip_fma(bx, k[3], t3);
    ip_fma(by, k[3], t4);
}
// a <- (px, py) a
template <typename R> __device__ __forceinline__ void ip_cmul(R &ax, R &ay, R px, R py, R npy) {
    const R t1 = npy * ay, t2 = py * ax;
    ip_fma(ax, px, t1);
    ip_fma(ay, px, t2);
}

// Op coefficient q in the kernel's precision (complex64 kernels read the
// table's float copy: no conversions per op).
template <typename R> __device__ __forceinline__ R coef(const RegOp &op, int q) {
    if constexpr (sizeof(R) == 4) return op.cf[q];
    else return op.c[q];
}

// ------------------------------------------------------------- 2x2 ops --
// Pair (a, b) = registers (r, r | 2^B) of one thread. With controls in the
// tile (CT) every register computes and keeps the old value unless its
// register-held controls (uniform per register) and this thread's
// thread-held controls are all 1; ops without controls compute straight.
template <int B, int KIND, bool CT, int E, typename T, typename R>
__device__ __forceinline__ void pairs_reg(Regs<R, E> &v, const R *k, uint32_t rcm, bool tok) {
    R nk[8];
#pragma unroll
    for (int q = 0; q < 8; q++) nk[q] = -k[q];
    const uint64_t tokm = CT ? __builtin_amdgcn_ballot_w64(tok) : 0;
#pragma unroll
    for (int r = 0; r < E; r++) {
        if (r & (1 << B)) continue;
        R &ax = v.d[2 * r], &ay = v.d[2 * r + 1];
        R &bx = v.d[2 * (r | (1 << B))], &by = v.d[2 * (r | (1 << B)) + 1];
        if constexpr (!CT) {
            if constexpr (KIND == RO_X) { ip_swap(ax, bx); ip_swap(ay, by); }
            else if constexpr (KIND == RO_RR) { ip_rr(ax, bx, k); ip_rr(ay, by, k); }
            else if constexpr (KIND == RO_RXI) ip_rxi(ax, ay, bx, by, k, nk);
            else if constexpr (KIND == RO_Y) {  // (a, b) <- (-i b, i a)
                ip_swap(ax, by); ip_swap(ay, bx);
                ip_neg(ay); ip_neg(bx);
            } else ip_gen(ax, ay, bx, by, k, nk);
        } else {
            const uint64_t m = (uint32_t(r) & rcm) == rcm ? tokm : 0;  // uniform per register
            if constexpr (KIND == RO_X) {  // controlled swap: new b first, then a and b in place
                const R tx = selm(m, ax, bx), ty = selm(m, ay, by);
                ip_sel(ax, bx, m); ip_sel(ay, by, m);
                ip_set(bx, tx); ip_set(by, ty);
                continue;
            }
            if constexpr (KIND == RO_Y) {  // all four new values first, then in place
                const R n0 = selm(m, by, ax), n1 = selm(m, -bx, ay), n2 = selm(m, -ay, bx), n3 = selm(m, ax, by);
                ip_set(ax, n0); ip_set(ay, n1); ip_set(bx, n2); ip_set(by, n3);
                continue;
            }
            R nax, nay, nbx, nby;
            if constexpr (KIND == RO_RR) {
                nax = fma_r(k[1], bx, k[0] * ax); nbx = fma_r(k[3], bx, k[2] * ax);
                nay = fma_r(k[1], by, k[0] * ay); nby = fma_r(k[3], by, k[2] * ay);
            } else if constexpr (KIND == RO_RXI) {
                nax = fma_r(nk[1], by, k[0] * ax); nay = fma_r(k[1], bx, k[0] * ay);
                nbx = fma_r(nk[2], ay, k[3] * bx); nby = fma_r(k[2], ax, k[3] * by);
            } else {
                const T na = cmul_add<T, R>(k, mk<T>(ax, ay), mk<T>(bx, by));
                const T nb = cmul_add<T, R>(k + 4, mk<T>(ax, ay), mk<T>(bx, by));
                nax = na.x; nay = na.y; nbx = nb.x; nby = nb.y;
            }
            ip_sel(ax, nax, m); ip_sel(ay, nay, m); ip_sel(bx, nbx, m); ip_sel(by, nby, m);
        }
    }
}

// Lane pairs: s = own amplitude, p = partner's; k holds this lane's row of
// the matrix (own coefficient first).
template <int KIND, bool CT, int E, typename T, typename R>
__device__ __forceinline__ void pairs_lane(Regs<R, E> &v, int addr, const R *k, uint32_t rcm, bool tok) {
    const R nk1 = -k[1], nk3 = -k[3];
    const uint64_t tokm = CT ? __builtin_amdgcn_ballot_w64(tok) : 0;
    // partners fetched a group of registers ahead: the exchange latency
    // overlaps the arithmetic of the previous registers
    constexpr int G0 = sizeof(R) == 4 ? 4 : 2, G = E < G0 ? E : G0;  // 8 partner registers
#pragma unroll
    for (int r0 = 0; r0 < E; r0 += G) {
        R px[G], py[G];
#pragma unroll
        for (int g = 0; g < G; g++) { px[g] = xlane(addr, v.d[2 * (r0 + g)]); py[g] = xlane(addr, v.d[2 * (r0 + g) + 1]); }
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int r = r0 + g;
            R &sx = v.d[2 * r], &sy = v.d[2 * r + 1];
            if constexpr (!CT) {
                if constexpr (KIND == RO_X) { ip_set(sx, px[g]); ip_set(sy, py[g]); }
                else if constexpr (KIND == RO_RR) { ip_fma(sx, k[0], k[1] * px[g]); ip_fma(sy, k[0], k[1] * py[g]); }
                else if constexpr (KIND == RO_RXI) { ip_fma(sx, k[0], nk1 * py[g]); ip_fma(sy, k[0], k[1] * px[g]); }
                else if constexpr (KIND == RO_Y) {  // own = (p.y, -p.x) on target-0 lanes, (-p.y, p.x) on target-1
                    ip_xor3(sx, py[g], k[0]);
                    ip_xor3(sy, px[g], k[1]);
                } else {
                    const R ux = fma_r(nk3, py[g], k[2] * px[g]), uy = fma_r(k[3], px[g], k[2] * py[g]);
                    const R z = fma_r(nk1, sy, ux), w = fma_r(k[1], sx, uy);
                    ip_fma(sy, k[0], w);
                    ip_fma(sx, k[0], z);
                }
            } else {
                const uint64_t m = (uint32_t(r) & rcm) == rcm ? tokm : 0;
                R nx, ny;
                if constexpr (KIND == RO_X) { nx = px[g]; ny = py[g]; }
                else if constexpr (KIND == RO_RR) {
                    nx = fma_r(k[1], px[g], k[0] * sx); ny = fma_r(k[1], py[g], k[0] * sy);
                } else if constexpr (KIND == RO_RXI) {
                    nx = fma_r(nk1, py[g], k[0] * sx); ny = fma_r(k[1], px[g], k[0] * sy);
                } else if constexpr (KIND == RO_Y) {
                    nx = xor_bits(py[g], k[0]);
                    ny = xor_bits(px[g], k[1]);
                } else {
                    const T n = cmul_add<T, R>(k, mk<T>(sx, sy), mk<T>(px[g], py[g]));
                    nx = n.x; ny = n.y;
                }
                ip_sel(sx, nx, m); ip_sel(sy, ny, m);
            }
        }
    }
}

template <int KIND, bool CT, int RB, int E, typename T, typename R>
__device__ __forceinline__ void pair_op(Regs<R, E> &v, const RegOp &op, uint32_t lane, uint32_t rcm, bool tok) {
    if (op.tgt >= kLaneTarget) {
        const int i = op.tgt - kLaneTarget;
        const int addr = int((lane ^ (1u << i)) << 2);
        const bool bit = ((lane >> i) & 1u) != 0;
        R k[4] = {R(0), R(0), R(0), R(0)};
        if constexpr (KIND == RO_RR || KIND == RO_RXI) {  // row 0: (m00, m01); row 1: (m11, m10)
            k[0] = bit ? coef<R>(op, 3) : coef<R>(op, 0);
            k[1] = bit ? coef<R>(op, 2) : coef<R>(op, 1);
        } else if constexpr (KIND == RO_Y) {  // sign patterns for the two components
            k[0] = sign_bits<R>(bit);
            k[1] = sign_bits<R>(!bit);
        } else if constexpr (KIND == RO_GEN) {
            k[0] = bit ? coef<R>(op, 6) : coef<R>(op, 0);
            k[1] = bit ? coef<R>(op, 7) : coef<R>(op, 1);
            k[2] = bit ? coef<R>(op, 4) : coef<R>(op, 2);
            k[3] = bit ? coef<R>(op, 5) : coef<R>(op, 3);
        }
        pairs_lane<KIND, CT, E, T, R>(v, addr, k, rcm, tok);
        return;
    }
    R k[8];
#pragma unroll
    for (int q = 0; q < 8; q++) k[q] = coef<R>(op, q);
#define QI_RB(BB) case BB: if constexpr (BB < RB) pairs_reg<BB, KIND, CT, E, T, R>(v, k, rcm, tok); break;
    switch (op.tgt) { QI_RB(0) QI_RB(1) QI_RB(2) QI_RB(3) QI_RB(4) QI_RB(5) default: break; }
#undef QI_RB
}

// ---------------------------------------------------------- diagonal ops --
// Target on register bit B: each register's target value is known at
// compile time (RO_Z and RO_P1 leave the target-0 registers alone).
template <int B, int KIND, bool CT, int E, typename T, typename R>
__device__ __forceinline__ void diag_reg(Regs<R, E> &v, T p0, T p1, uint32_t rcm, bool tok) {
    const R n0 = -p0.y, n1 = -p1.y;
    const uint64_t tokm = CT ? __builtin_amdgcn_ballot_w64(tok) : 0;
#pragma unroll
    for (int r = 0; r < E; r++) {
        const bool one = ((r >> B) & 1) != 0;
        if (!one && KIND != RO_DG) continue;
        R &ax = v.d[2 * r], &ay = v.d[2 * r + 1];
        const T p = one ? p1 : p0;
        if constexpr (!CT) {
            if constexpr (KIND == RO_Z) { ip_neg(ax); ip_neg(ay); }
            else ip_cmul(ax, ay, p.x, p.y, one ? n1 : n0);
        } else {
            const uint64_t m = (uint32_t(r) & rcm) == rcm ? tokm : 0;
            if constexpr (KIND == RO_Z) { ip_sel(ax, -ax, m); ip_sel(ay, -ay, m); }
            else {
                const T n = cmul(mk<T>(ax, ay), p);
                ip_sel(ax, n.x, m); ip_sel(ay, n.y, m);
            }
        }
    }
}
// Target fixed for the whole thread (thread bit or outside the tile).
template <int KIND, bool CT, int E, typename T, typename R>
__device__ __forceinline__ void diag_thr(Regs<R, E> &v, T f, bool one, uint32_t rcm, bool tok) {
    const uint32_t sgn = one ? 0x80000000u : 0u;  // Z: sign flip on the target-1 lanes
    const R nfy = -f.y;
    const uint64_t tokm = CT ? __builtin_amdgcn_ballot_w64(tok) : 0;
#pragma unroll
    for (int r = 0; r < E; r++) {
        R &ax = v.d[2 * r], &ay = v.d[2 * r + 1];
        if constexpr (!CT) {
            if constexpr (KIND == RO_Z) { ip_xor(ax, sgn); ip_xor(ay, sgn); }
            else ip_cmul(ax, ay, f.x, f.y, nfy);
        } else {
            const uint64_t m = (uint32_t(r) & rcm) == rcm ? tokm : 0;
            if constexpr (KIND == RO_Z) {
                ip_sel(ax, xor_sign(ax, sgn), m);
                ip_sel(ay, xor_sign(ay, sgn), m);
            } else {
                const T n = cmul(mk<T>(ax, ay), f);
                ip_sel(ax, n.x, m); ip_sel(ay, n.y, m);
            }
        }
    }
}

template <int KIND, bool CT, int RB, int E, typename T, typename R>
__device__ __forceinline__ void diag_op(Regs<R, E> &v, const RegOp &op, uint32_t tid, uint64_t base, uint32_t rcm, bool tok) {
    const T p0 = mk<T>(coef<R>(op, 0), coef<R>(op, 1)), p1 = mk<T>(coef<R>(op, 2), coef<R>(op, 3));
    if (op.dloc == 0) {
#define QI_RB(BB) case BB: if constexpr (BB < RB) diag_reg<BB, KIND, CT, E, T, R>(v, p0, p1, rcm, tok); break;
        switch (op.dbit) { QI_RB(0) QI_RB(1) QI_RB(2) QI_RB(3) QI_RB(4) QI_RB(5) default: break; }
#undef QI_RB
        return;
    }
    const bool one = op.dloc == 1 ? ((tid >> op.dbit) & 1u) != 0 : ((base >> op.dbit) & 1u) != 0;
    diag_thr<KIND, CT, E, T, R>(v, one ? p1 : p0, one, rcm, tok);
}

// ------------------------------------ single-register-control X and Z --
// CNOT, Toffoli and CZ whose one in-tile control sits on register bit C (no
// lane controls): registers without bit C are skipped at compile time and
// the rest is updated uncontrolled, instead of computing every register and
// selecting under a mask.
template <int B, int C, int E, typename R>
__device__ __forceinline__ void x_reg_c1(Regs<R, E> &v) {
#pragma unroll
    for (int r = 0; r < E; r++) {
        if (((r >> B) & 1) || !((r >> C) & 1)) continue;
        const int s = r | (1 << B);
        ip_swap(v.d[2 * r], v.d[2 * s]);
        ip_swap(v.d[2 * r + 1], v.d[2 * s + 1]);
    }
}
template <int C, int E, typename R>
__device__ __forceinline__ void x_lane_c1(Regs<R, E> &v, int addr) {
    constexpr int G0 = sizeof(R) == 4 ? 4 : 2, H = E / 2, G = H < G0 ? H : G0;
    int rs[H];
    {
        int n = 0;
#pragma unroll
        for (int r = 0; r < E; r++) if ((r >> C) & 1) rs[n++] = r;
    }
#pragma unroll
    for (int g0 = 0; g0 < H; g0 += G) {
        R px[G], py[G];
#pragma unroll
        for (int g = 0; g < G; g++) { px[g] = xlane(addr, v.d[2 * rs[g0 + g]]); py[g] = xlane(addr, v.d[2 * rs[g0 + g] + 1]); }
#pragma unroll
        for (int g = 0; g < G; g++) { ip_set(v.d[2 * rs[g0 + g]], px[g]); ip_set(v.d[2 * rs[g0 + g] + 1], py[g]); }
    }
}
template <int B, int C, int E, typename R>
__device__ __forceinline__ void z_reg_c1(Regs<R, E> &v) {
#pragma unroll
    for (int r = 0; r < E; r++)
        if (((r >> B) & 1) && ((r >> C) & 1)) { ip_neg(v.d[2 * r]); ip_neg(v.d[2 * r + 1]); }
}
template <int C, int E, typename R>
__device__ __forceinline__ void z_thr_c1(Regs<R, E> &v, uint32_t sgn) {
#pragma unroll
    for (int r = 0; r < E; r++)
        if ((r >> C) & 1) { ip_xor(v.d[2 * r], sgn); ip_xor(v.d[2 * r + 1], sgn); }
}
template <int B, int E, typename R> __device__ __forceinline__ void x_reg_c1_any(Regs<R, E> &v, int c) {
    constexpr int RB = __builtin_ctz(E);
    switch (c) {
    case 0: if constexpr (0 < RB && B != 0) x_reg_c1<B, 0, E, R>(v); break;
    case 1: if constexpr (1 < RB && B != 1) x_reg_c1<B, 1, E, R>(v); break;
    case 2: if constexpr (2 < RB && B != 2) x_reg_c1<B, 2, E, R>(v); break;
    case 3: if constexpr (3 < RB && B != 3) x_reg_c1<B, 3, E, R>(v); break;
    case 4: if constexpr (4 < RB && B != 4) x_reg_c1<B, 4, E, R>(v); break;
    case 5: if constexpr (5 < RB && B != 5) x_reg_c1<B, 5, E, R>(v); break;
    default: break;
    }
}
template <int B, int E, typename R> __device__ __forceinline__ void z_reg_c1_any(Regs<R, E> &v, int c) {
    constexpr int RB = __builtin_ctz(E);
    switch (c) {
    case 0: if constexpr (0 < RB && B != 0) z_reg_c1<B, 0, E, R>(v); break;
    case 1: if constexpr (1 < RB && B != 1) z_reg_c1<B, 1, E, R>(v); break;
    case 2: if constexpr (2 < RB && B != 2) z_reg_c1<B, 2, E, R>(v); break;
    case 3: if constexpr (3 < RB && B != 3) z_reg_c1<B, 3, E, R>(v); break;
    case 4: if constexpr (4 < RB && B != 4) z_reg_c1<B, 4, E, R>(v); break;
    case 5: if constexpr (5 < RB && B != 5) z_reg_c1<B, 5, E, R>(v); break;
    default: break;
    }
}
template <int E, typename R> __device__ __forceinline__ void xc1_op(Regs<R, E> &v, const RegOp &op, uint32_t lane) {
    constexpr int RB = __builtin_ctz(E);
    const int c = op.lp0;
    if (op.tgt >= kLaneTarget) {
        const int addr = int((lane ^ (1u << (op.tgt - kLaneTarget))) << 2);
        switch (c) {
        case 0: if constexpr (0 < RB) x_lane_c1<0, E, R>(v, addr); break;
        case 1: if constexpr (1 < RB) x_lane_c1<1, E, R>(v, addr); break;
        case 2: if constexpr (2 < RB) x_lane_c1<2, E, R>(v, addr); break;
        case 3: if constexpr (3 < RB) x_lane_c1<3, E, R>(v, addr); break;
        case 4: if constexpr (4 < RB) x_lane_c1<4, E, R>(v, addr); break;
        case 5: if constexpr (5 < RB) x_lane_c1<5, E, R>(v, addr); break;
        default: break;
        }
        return;
    }
#define QI_RB(BB) case BB: if constexpr (BB < RB) x_reg_c1_any<BB, E, R>(v, c); break;
    switch (op.tgt) { QI_RB(0) QI_RB(1) QI_RB(2) QI_RB(3) QI_RB(4) QI_RB(5) default: break; }
#undef QI_RB
}
template <int E, typename R>
__device__ __forceinline__ void zc1_op(Regs<R, E> &v, const RegOp &op, uint32_t tid, uint64_t base) {
    constexpr int RB = __builtin_ctz(E);
    const int c = op.lp0;
    if (op.dloc == 0) {
#define QI_RB(BB) case BB: if constexpr (BB < RB) z_reg_c1_any<BB, E, R>(v, c); break;
        switch (op.dbit) { QI_RB(0) QI_RB(1) QI_RB(2) QI_RB(3) QI_RB(4) QI_RB(5) default: break; }
#undef QI_RB
        return;
    }
    const bool one = op.dloc == 1 ? ((tid >> op.dbit) & 1u) != 0 : ((base >> op.dbit) & 1u) != 0;
    const uint32_t sgn = one ? 0x80000000u : 0u;
    switch (c) {
    case 0: if constexpr (0 < RB) z_thr_c1<0, E, R>(v, sgn); break;
    case 1: if constexpr (1 < RB) z_thr_c1<1, E, R>(v, sgn); break;
    case 2: if constexpr (2 < RB) z_thr_c1<2, E, R>(v, sgn); break;
    case 3: if constexpr (3 < RB) z_thr_c1<3, E, R>(v, sgn); break;
    case 4: if constexpr (4 < RB) z_thr_c1<4, E, R>(v, sgn); break;
    case 5: if constexpr (5 < RB) z_thr_c1<5, E, R>(v, sgn); break;
    default: break;
    }
}

// ------------------------------------------------ rotations as shears (c64) --
// R(phi) = [[1, a], [0, 1]] [[1, 0], [b, 1]] [[1, a], [0, 1]] with
// a = -tan(phi/2), b = sin(phi): three in-place multiply-adds per component
// pair. RO_RRS rotates (re a, re b) and (im a, im b); RO_RXS rotates
// (re a, im b) by phi and (im a, re b) by -phi.
template <int B, bool XS, int E>
__device__ __forceinline__ void shear_reg(Regs<float, E> &v, float a, float b) {
    const float na = -a, nb = -b;
#pragma unroll
    for (int r = 0; r < E; r++) {
        if ((r >> B) & 1) continue;
        const int s = r | (1 << B);
        float &p0 = v.d[2 * r], &q0 = v.d[2 * s + (XS ? 1 : 0)];   // first pair, angle phi
        float &p1 = v.d[2 * r + 1], &q1 = v.d[2 * s + (XS ? 0 : 1)];  // second pair, angle +-phi
        const float a1 = XS ? na : a, b1 = XS ? nb : b;
        ip_fmac(p0, a, q0);
        ip_fmac(p1, a1, q1);
        ip_fmac(q0, b, p0);
        ip_fmac(q1, b1, p1);
        ip_fmac(p0, a, q0);
        ip_fmac(p1, a1, q1);
    }
}
template <bool XS, int E> __device__ __forceinline__ void shear_op(Regs<float, E> &v, const RegOp &op) {
    constexpr int RB = __builtin_ctz(E);
    const float a = op.cf[0], b = op.cf[1];
#define QI_RB(BB) case BB: if constexpr (BB < RB) shear_reg<BB, XS, E>(v, a, b); break;
    switch (op.tgt) { QI_RB(0) QI_RB(1) QI_RB(2) QI_RB(3) QI_RB(4) QI_RB(5) default: break; }
#undef QI_RB
}

// Phase e^{i phi} on the amplitudes whose register bit B is 1, as a rotation
// of (re, im) by three in-place shears (|phi| <= pi/2).
template <int B, int E> __device__ __forceinline__ void p1s_reg(Regs<float, E> &v, float a, float b) {
#pragma unroll
    for (int r = 0; r < E; r++) {
        if (!((r >> B) & 1)) continue;
        float &x = v.d[2 * r], &y = v.d[2 * r + 1];
        ip_fmac(x, a, y);
        ip_fmac(y, b, x);
        ip_fmac(x, a, y);
    }
}
template <int E> __device__ __forceinline__ void p1s_op(Regs<float, E> &v, const RegOp &op) {
    constexpr int RB = __builtin_ctz(E);
    const float a = op.cf[0], b = op.cf[1];
#define QI_RB(BB) case BB: if constexpr (BB < RB) p1s_reg<BB, E>(v, a, b); break;
    switch (op.dbit) { QI_RB(0) QI_RB(1) QI_RB(2) QI_RB(3) QI_RB(4) QI_RB(5) default: break; }
#undef QI_RB
}

// Unscaled H on register bit B: (a, b) <- (a + b, a - b) per component as
// a += b, then b = a - 2b (two in-place instructions instead of four).
template <int B, int E> __device__ __forceinline__ void hs_reg(Regs<float, E> &v) {
#pragma unroll
    for (int r = 0; r < E; r++) {
        if ((r >> B) & 1) continue;
        const int s = r | (1 << B);
#pragma unroll
        for (int c = 0; c < 2; c++) {
            float &a = v.d[2 * r + c], &b = v.d[2 * s + c];
            asm("v_add_f32 %0, %0, %1" : "+v"(a) : "v"(b));
            asm("v_fma_f32 %0, -2.0, %0, %1" : "+v"(b) : "v"(a));
        }
    }
}
template <int E> __device__ __forceinline__ void hs_op(Regs<float, E> &v, const RegOp &op) {
    constexpr int RB = __builtin_ctz(E);
#define QI_RB(BB) case BB: if constexpr (BB < RB) hs_reg<BB, E>(v); break;
    switch (op.tgt) { QI_RB(0) QI_RB(1) QI_RB(2) QI_RB(3) QI_RB(4) QI_RB(5) default: break; }
#undef QI_RB
}

// ------------------------------------------------------- exchange ops --
// a's lanes with lane bit 4 (5) set trade with b's lanes where it is clear:
// v_permlane16_swap / v_permlane32_swap, one VALU instruction per 32-bit
// component, both registers updated in place (tied operands, like the other
// in-place ops). A VALU write of a register that a permlane swap reads needs
// two wait states before the swap: each swap carries its own s_nop 1, so no
// instruction the compiler places in between can break that.
__device__ __forceinline__ void swap_rows32(uint32_t &a, uint32_t &b, bool b5) {
    if (b5) asm volatile("s_nop 1\n\tv_permlane32_swap_b32 %0, %1" : "+v"(a), "+v"(b));
    else asm volatile("s_nop 1\n\tv_permlane16_swap_b32 %0, %1" : "+v"(a), "+v"(b));
}
__device__ __forceinline__ void swap_rows(float &a, float &b, bool b5) {
    uint32_t x = __float_as_uint(a), y = __float_as_uint(b);
    swap_rows32(x, y, b5);
    a = __uint_as_float(x);
    b = __uint_as_float(y);
}
__device__ __forceinline__ void swap_rows(double &a, double &b, bool b5) {
    uint32_t al = lo32(a), ah = hi32(a), bl = lo32(b), bh = hi32(b);
    swap_rows32(al, bl, b5);
    swap_rows32(ah, bh, b5);
    a = mk64(al, ah);
    b = mk64(bl, bh);
}
template <int B, int E, typename R>
__device__ __forceinline__ void xlr_reg(Regs<R, E> &v, uint32_t lane, int i) {
    if (i >= 4) {  // lane bits 4 and 5: one cross-row swap per component pair, no lane addressing
        const bool b5 = i == 5;
#pragma unroll
        for (int r = 0; r < E; r++) {
            if (r & (1 << B)) continue;
#pragma unroll
            for (int c = 0; c < 2; c++) swap_rows(v.d[2 * r + c], v.d[2 * (r | (1 << B)) + c], b5);
        }
        return;
    }
    const int addr = int((lane ^ (1u << i)) << 2);
    const uint64_t hi = __builtin_amdgcn_ballot_w64(((lane >> i) & 1u) != 0), lo = ~hi;
    constexpr int G = sizeof(R) == 4 ? 4 : 2;  // exchanges issued a group ahead of their use
    int rs[E / 2];
    {
        int n = 0;
#pragma unroll
        for (int r = 0; r < E; r++) if (!(r & (1 << B))) rs[n++] = r;
    }
#pragma unroll
    for (int g0 = 0; g0 < E / 2; g0 += G) {
        R y[G][2];
#pragma unroll
        for (int g = 0; g < G && g0 + g < E / 2; g++)
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const int r = rs[g0 + g];
                y[g][c] = xlane(addr, selm(hi, v.d[2 * r + c], v.d[2 * (r | (1 << B)) + c]));
            }
#pragma unroll
        for (int g = 0; g < G && g0 + g < E / 2; g++)
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const int r = rs[g0 + g];
                ip_sel(v.d[2 * (r | (1 << B)) + c], y[g][c], lo);
                ip_sel(v.d[2 * r + c], y[g][c], hi);
            }
    }
}
template <int E, typename R>
__device__ __forceinline__ void xlr_op(Regs<R, E> &v, uint32_t lane, int i, int j) {
    constexpr int RB = __builtin_ctz(E);
#define QI_RB(BB) case BB: if constexpr (BB < RB) xlr_reg<BB, E, R>(v, lane, i); break;
    switch (j) { QI_RB(0) QI_RB(1) QI_RB(2) QI_RB(3) QI_RB(4) QI_RB(5) default: break; }
#undef QI_RB
}
// Lane bits i <-> i2: lanes whose two bits differ trade with lane ^ (2^i | 2^i2).
template <int E, typename R>
__device__ __forceinline__ void xll_op(Regs<R, E> &v, uint32_t lane, int i, int i2) {
    const int addr = int((lane ^ ((1u << i) | (1u << i2))) << 2);
    const uint64_t m = __builtin_amdgcn_ballot_w64(((lane >> i) & 1u) != ((lane >> i2) & 1u));
    constexpr int G = sizeof(R) == 4 ? 8 : 4;
#pragma unroll
    for (int k0 = 0; k0 < 2 * E; k0 += G) {
        R y[G];
#pragma unroll
        for (int g = 0; g < G; g++) y[g] = xlane(addr, v.d[k0 + g]);
#pragma unroll
        for (int g = 0; g < G; g++) ip_sel(v.d[k0 + g], y[g], m);
    }
}

// ------------------------------------------------------------- LDS ops --
// Controlled SWAP or Matchgate (CPU form, operator.rs:993-1004) on the LDS
// copy; local index e lives at slot swz(e).
template <typename R, int K, int F, int NT>
__device__ __forceinline__ void lds_gate(typename V2<R>::T *tile, const RegOp &op) {
    using T = typename V2<R>::T;
    constexpr int N = 1 << K;
    const uint32_t lcm = op.lcm, b0 = 1u << op.lp0, b1 = 1u << op.lp1;
    const unsigned lo = op.lp0 < op.lp1 ? op.lp0 : op.lp1, hi = op.lp0 < op.lp1 ? op.lp1 : op.lp0;
    const R c0 = coef<R>(op, 0), c1 = coef<R>(op, 1);
    const T m23 = mk<T>(coef<R>(op, 2), coef<R>(op, 3)), m45 = mk<T>(coef<R>(op, 4), coef<R>(op, 5)),
            m67 = mk<T>(coef<R>(op, 6), coef<R>(op, 7));
    for (uint32_t q = threadIdx.x; q < N / 4; q += NT) {
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

// ------------------------------------------------------ lane phases --
// One-wave tiles carry a per-lane pending phase (ph, two registers): diagonal
// gates whose factor depends on the lane alone multiply into it, and it is
// applied to the lane's amplitudes before anything moves them between lanes
// (the planner places RO_LPF).
template <typename R> __device__ __forceinline__ uint32_t sign_word(R x) {
    if constexpr (sizeof(R) == 4) return __float_as_uint(x) & 0x80000000u;
    else return hi32(x) & 0x80000000u;
}
template <typename R>
__device__ __forceinline__ void lph_op(R *ph, const RegOp &op, uint64_t base, uint32_t lane) {
    const bool one = op.dloc == 1 ? ((lane >> op.dbit) & 1u) != 0 : ((base >> op.dbit) & 1u) != 0;
    const bool on = (lane & op.tcm) == op.tcm;
    const R fx = on ? (one ? coef<R>(op, 2) : coef<R>(op, 0)) : R(1);
    const R fy = on ? (one ? coef<R>(op, 3) : coef<R>(op, 1)) : R(0);
    const R nx = fma_r(ph[0], fx, -(ph[1] * fy)), ny = fma_r(ph[0], fy, ph[1] * fx);
    ph[0] = nx;
    ph[1] = ny;
}
template <typename R, int E>
__device__ __forceinline__ void lpf_op(Regs<R, E> &v, R *ph, const RegOp &op) {
    if (op.lp0) {  // +-1: flip sign bits
        const uint32_t sgn = sign_word(ph[0]);
#pragma unroll
        for (int k = 0; k < 2 * E; k++) ip_xor(v.d[k], sgn);
    } else {
        const R npy = -ph[1];
#pragma unroll
        for (int r = 0; r < E; r++) ip_cmul(v.d[2 * r], v.d[2 * r + 1], ph[0], ph[1], npy);
    }
    ph[0] = R(1);
    ph[1] = R(0);
}

// ------------------------------------------------------------ the pass --
// One in-register op on the thread's share of the tile.
template <typename R, int K, int TB>
__device__ __forceinline__ void apply_op(Regs<R, (1 << (K - TB))> &v, const RegOp &op, uint64_t base, uint32_t tid,
                                         uint32_t lane, R *ph) {
    using T = typename V2<R>::T;
    constexpr int RB = K - TB, E = 1 << RB;
    const bool tc = (op.tcm | op.rcm) != 0;   // controls inside the tile (uniform)
    const bool tok = (tid & op.tcm) == op.tcm;  // this thread's thread-held controls
#define QI_OP(KD, FN)                                                                     \
    case KD:                                                                              \
        if (tc) FN<KD, true, RB, E, T, R>(v, op, ARGS, op.rcm, tok);                      \
        else FN<KD, false, RB, E, T, R>(v, op, ARGS, op.rcm, tok);                        \
        break;
#define ARGS lane
    switch (op.kind) {
        QI_OP(RO_X, pair_op) QI_OP(RO_RR, pair_op) QI_OP(RO_GEN, pair_op) QI_OP(RO_RXI, pair_op) QI_OP(RO_Y, pair_op)
#undef ARGS
#define ARGS tid, base
        QI_OP(RO_Z, diag_op) QI_OP(RO_DG, diag_op) QI_OP(RO_P1, diag_op)
#undef ARGS
    case RO_XC1: xc1_op<E, R>(v, op, lane); break;
    case RO_ZC1: zc1_op<E, R>(v, op, tid, base); break;
    case RO_RRS:
        if constexpr (sizeof(R) == 4) shear_op<false, E>(v, op);
        break;
    case RO_RXS:
        if constexpr (sizeof(R) == 4) shear_op<true, E>(v, op);
        break;
    case RO_HS:
        if constexpr (sizeof(R) == 4) hs_op<E>(v, op);
        break;
    case RO_P1S:
        if constexpr (sizeof(R) == 4) p1s_op<E>(v, op);
        break;
    case RO_XLR:
        if constexpr (TB == kLaneBits) xlr_op<E, R>(v, lane, op.tgt, op.dbit);
        break;
    case RO_XLL:
        if constexpr (TB == kLaneBits) xll_op<E, R>(v, lane, op.tgt, op.dbit);
        break;
    case RO_LPH:
        if constexpr (TB == kLaneBits) lph_op<R>(ph, op, base, lane);
        break;
    case RO_LPF:
        if constexpr (TB == kLaneBits) lpf_op<R, E>(v, ph, op);
        break;
    default: break;
    }
#undef QI_OP
}

template <typename R, int K, int TB>
__global__ __launch_bounds__(1 << TB) void k_reg_pass(typename V2<R>::T *__restrict__ psi, const RegPass pass,
                                                      const RegOp *__restrict__ ops) {
    using T = typename V2<R>::T;
    constexpr int RB = K - TB, E = 1 << RB, F = sizeof(R) == 8 ? 3 : 4;
    __shared__ T lds[1 << K];
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    const uint64_t base = insert_zeros64(xcd_block_order(blockIdx.x, pass.chunk_log2), pass.qmask);

    Regs<R, E> v;
    uint32_t tl = 0;  // local index bits held by this thread
    uint32_t rl[E];   // local index bits of register r (uniform)
    {
        uint64_t at = 0;
#pragma unroll
        for (int i = 0; i < TB; i++) {
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
            v.set(r, ld_stream(src + o));
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
            apply_op<R, K, TB>(v, op, base, tid, lane, nullptr);  // (no lane phases in LDS tiles)
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
            for (int i = 0; i < TB; i++) tl |= ((tid >> i) & 1u) << op.lay_t[i];
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
            lds_gate<R, K, F, (1 << TB)>(lds, op);
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
        // (the frame's address flip: lane part per thread, register part uniform)
        uint64_t at = 0, tmask = 0;
#pragma unroll
        for (int i = 0; i < TB; i++) {
            at |= uint64_t((tid >> i) & 1u) << pass.qt_out[i];
            tmask |= uint64_t(1) << pass.qt_out[i];
        }
        uint64_t ar[RB];
#pragma unroll
        for (int j = 0; j < RB; j++) ar[j] = uint64_t(1) << pass.qr_out[j];
        const uint64_t rflip = pass.flip & ~tmask;
        T *dst = psi + (base | (at ^ (pass.flip & tmask)));
#pragma unroll
        for (int r = 0; r < E; r++) {
            uint64_t o = 0;
#pragma unroll
            for (int j = 0; j < RB; j++) if ((r >> j) & 1) o |= ar[j];
            st_stream(v.get(r), dst + (o ^ rflip));
        }
    }
}

// One-wave tiles (TB = 6): a wave holds a whole tile (64 amplitudes per lane
// for 4096-amplitude complex64 tiles, 32 for 2048-amplitude complex128 ones)
// and needs no LDS and no barriers, and the per-op scalar work is paid once
// per tile. (A second register set to prefetch the next tile costs more in
// occupancy than it gains: measured slower at 2048 and 1024 amplitudes.)
// src / dst: the pass's load and store buffers (the same buffer, or the
// state's two buffers for a pass that moves the state: tiles are disjoint in
// both, so no wave stores where another loads).
template <typename R, int K>
__global__ __launch_bounds__(64) void k_wave_pass(const typename V2<R>::T *src_psi, typename V2<R>::T *dst_psi,
                                                  const RegPass pass, const RegOp *__restrict__ ops) {
    using T = typename V2<R>::T;
    constexpr int TB = kLaneBits, RB = K - TB, E = 1 << RB;
    const uint32_t lane = threadIdx.x;
    const uint64_t base = insert_zeros64(xcd_block_order(blockIdx.x, pass.chunk_log2), pass.qmask);
    uint64_t at_in = 0, at_out = 0, tmask = 0;
#pragma unroll
    for (int i = 0; i < TB; i++) {
        const uint32_t bit = (lane >> i) & 1u;
        at_in |= uint64_t(bit) << pass.qt_in[i];
        at_out |= uint64_t(bit) << pass.qt_out[i];
        tmask |= uint64_t(1) << pass.qt_out[i];
    }
    Regs<R, E> v;
    {
        const T *src = src_psi + (base | at_in);
#pragma unroll
        for (int r = 0; r < E; r++) {
            uint64_t o = 0;
#pragma unroll
            for (int j = 0; j < RB; j++) if ((r >> j) & 1) o |= uint64_t(1) << pass.qr_in[j];
            v.set(r, ld_stream(src + o));
        }
    }
    // The op table of a pass ends with a no-op, so op o + 1 can be fetched
    // while op o runs. One-wave passes hold no LDS ops (the planner streams them).
    R ph[2] = {R(1), R(0)};  // the lane's pending phase
    RegOp nxt = ops[pass.op0];
    for (int o = 0; o < pass.nops; o++) {
        const RegOp op = nxt;
        nxt = ops[pass.op0 + o + 1];
        // controls outside the tile are uniform per tile
        if (op.kind < RO_TRANSPOSE && (base & op.gcm) == op.gcm) apply_op<R, K, TB>(v, op, base, lane, lane, ph);
    }
    // In-place permutation: a store may go where this wave loaded another
    // amplitude, so every load must be complete first.
    if (pass.sync_store) __builtin_amdgcn_s_waitcnt(0);
    // the frame's address flip: lane part per lane, register part uniform; a
    // permuting pass also moves the tile's base bits (uniform per tile)
    uint64_t base_out = base;
    if (pass.permute) {
        base_out = 0;
        for (uint64_t m = base; m; m &= m - 1) base_out |= uint64_t(1) << pass.perm[__builtin_ctzll(m)];
    }
    const uint64_t rflip = pass.flip & ~tmask;
    T *dst = dst_psi + (base_out | (at_out ^ (pass.flip & tmask)));
#pragma unroll
    for (int r = 0; r < E; r++) {
        uint64_t o = 0;
#pragma unroll
        for (int j = 0; j < RB; j++) if ((r >> j) & 1) o |= uint64_t(1) << pass.qr_out[j];
        st_stream(v.get(r), dst + (o ^ rflip));
    }
}

// ------------------------------------------------ batched one-wave passes --
// Coefficients of a parametric op from its slot's row of the table
// (kParamRow values per angle slot, in the kernel's precision), in the forms
// param_op_coeffs writes: RY (c, -s, s, c), RX (c, -s, -s, c), RZ (c, -s, c, s)
// with c, s = cos, sin of half the angle, P (1, 0, cos, sin) of the angle; and
// RY as shears (RO_RRS, complex64: the row's shear pair).
template <typename R> __device__ __forceinline__ void set_coef(RegOp &op, int q, R x) {
    if constexpr (sizeof(R) == 4) op.cf[q] = x;
    else op.c[q] = x;
}
template <typename R> __device__ __forceinline__ void param_coeffs(RegOp &op, const R *__restrict__ t) {
    const R ch = t[0], nsh = t[1], sh = t[2];
    switch (op.kind) {
    case RO_RR: set_coef<R>(op, 0, ch); set_coef<R>(op, 1, nsh); set_coef<R>(op, 2, sh); set_coef<R>(op, 3, ch); break;
    case RO_RXI: set_coef<R>(op, 0, ch); set_coef<R>(op, 1, nsh); set_coef<R>(op, 2, nsh); set_coef<R>(op, 3, ch); break;
    case RO_DG: set_coef<R>(op, 0, ch); set_coef<R>(op, 1, nsh); set_coef<R>(op, 2, ch); set_coef<R>(op, 3, sh); break;
    case RO_RRS: set_coef<R>(op, 0, t[5]); set_coef<R>(op, 1, t[6]); break;
    default: set_coef<R>(op, 0, R(1)); set_coef<R>(op, 1, R(0)); set_coef<R>(op, 2, t[3]); set_coef<R>(op, 3, t[4]); break;
    }
}

__device__ __forceinline__ int find_wave_task(const WaveTask *__restrict__ tasks, int ntasks, uint32_t b) {
    int lo = 0, hi = ntasks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tasks[mid].block0 <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Pass p of many circuits in one launch (the batched energies): workgroup b
// is one wave holding one tile of the circuit whose task covers b (tasks in
// block order); the task names the circuit's state in the arena, its pass
// and its first angle slot. The tile work is k_wave_pass's; the first pass of
// every circuit (INIT, the first launch) builds |0..0> in registers instead of
// loading.
template <typename R, int K, bool INIT>
__global__ __launch_bounds__(64) void k_batch_wave(typename V2<R>::T *__restrict__ arena, const WaveTask *__restrict__ tasks,
                                                   int ntasks, const RegPass *__restrict__ passes,
                                                   const RegOp *__restrict__ ops, const R *__restrict__ ptab) {
    using T = typename V2<R>::T;
    constexpr int TB = kLaneBits, RB = K - TB, E = 1 << RB;
    const uint32_t lane = threadIdx.x;
    const WaveTask &tk = tasks[find_wave_task(tasks, ntasks, blockIdx.x)];
    const RegPass &pass = passes[tk.pass];
    const uint64_t base = insert_zeros64(blockIdx.x - tk.block0, pass.qmask);
    T *psi = arena + tk.off;
    uint64_t at_in = 0, at_out = 0, tmask = 0;
#pragma unroll
    for (int i = 0; i < TB; i++) {
        const uint32_t bit = (lane >> i) & 1u;
        at_in |= uint64_t(bit) << pass.qt_in[i];
        at_out |= uint64_t(bit) << pass.qt_out[i];
        tmask |= uint64_t(1) << pass.qt_out[i];
    }
    Regs<R, E> v;
    if constexpr (INIT) {  // |0..0>: register 0 of the lane whose address is 0
        const bool one = (base | at_in) == 0;
#pragma unroll
        for (int r = 0; r < E; r++) v.set(r, mk<T>(r == 0 && one ? R(1) : R(0), R(0)));
    } else {
        const T *src = psi + (base | at_in);
#pragma unroll
        for (int r = 0; r < E; r++) {
            uint64_t o = 0;
#pragma unroll
            for (int j = 0; j < RB; j++) if ((r >> j) & 1) o |= uint64_t(1) << pass.qr_in[j];
            v.set(r, ld_stream(src + o));
        }
    }
    const R *prow = ptab + size_t(kParamRow) * size_t(tk.slot_base);
    R ph[2] = {R(1), R(0)};
    RegOp nxt = ops[pass.op0];
    for (int o = 0; o < pass.nops; o++) {
        RegOp op = nxt;
        nxt = ops[pass.op0 + o + 1];
        if (op.param > 0) param_coeffs<R>(op, prow + size_t(kParamRow) * size_t(op.param - 1));
        if (op.kind < RO_TRANSPOSE && (base & op.gcm) == op.gcm) apply_op<R, K, TB>(v, op, base, lane, lane, ph);
    }
    if (pass.sync_store) __builtin_amdgcn_s_waitcnt(0);
    const uint64_t rflip = pass.flip & ~tmask;
    T *dst = psi + (base | (at_out ^ (pass.flip & tmask)));
#pragma unroll
    for (int r = 0; r < E; r++) {
        uint64_t o = 0;
#pragma unroll
        for (int j = 0; j < RB; j++) if ((r >> j) & 1) o |= uint64_t(1) << pass.qr_out[j];
        st_stream(v.get(r), dst + (o ^ rflip));
    }
}

// Runs of consecutive tiles per XCD (log2; shorter when the grid is small):
// 2-3% faster passes than the identity order on 30-qubit states.
constexpr int kTileChunkLog2 = 8;

template <typename R, int K, int TB>
void launch_pass(State &s, const RegPass &p, const RegOp *ops, hipStream_t st, void *src, void *dst) {
    using T = typename V2<R>::T;
    const int tile_bits = (int)s.n - K;
    const uint64_t tiles = uint64_t(1) << tile_bits;
    RegPass q = p;
    q.chunk_log2 = xcd_chunk_log2(tiles, kTileChunkLog2);
    if constexpr (TB == kLaneBits) {
        static_assert(K <= (sizeof(R) == 4 ? 12 : 11), "one-wave tiles: at most 64 amplitude components per lane");
        hipLaunchKernelGGL((k_wave_pass<R, K>), dim3((unsigned)tiles), dim3(64), 0, st, (const T *)src, (T *)dst, q, ops);
    } else {
        hipLaunchKernelGGL((k_reg_pass<R, K, TB>), dim3((unsigned)tiles), dim3(1 << TB), 0, st, (T *)dst, q, ops);
    }
}

template <typename R, int K> bool launch_k(State &s, const RegPass &p, const RegOp *ops, hipStream_t st, void *src, void *dst) {
    switch (p.tb) {
    case 6:
        if constexpr (K <= (sizeof(R) == 4 ? 12 : 11)) { launch_pass<R, K, 6>(s, p, ops, st, src, dst); return true; }
        return false;
    case 8: launch_pass<R, K, 8>(s, p, ops, st, src, dst); return true;
    case 9: launch_pass<R, K, 9>(s, p, ops, st, src, dst); return true;
    default: return false;
    }
}

}  // namespace

Status launch_reg_pass(State &s, const RegPass &p, const RegOp *dops, hipStream_t st, void *src, void *dst) {
    // one workgroup per tile: the grid's work-item count must stay below 2^32
    if (p.K < 10 || (int)s.n - p.K + p.tb >= 32)
        return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "tile width too small for this state (grid over 2^32 threads)");
    // moving or permuting passes exist for one-wave tiles only
    if ((p.oop || p.permute) && (p.tb != kLaneBits || src == dst || !src || !dst))
        return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "out-of-place pass: one-wave tiles and two buffers");
    if (!p.oop && src != dst) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "in-place pass with two buffers");
    bool ok = false;
    if (s.prec == QI_C128) {
        switch (p.K) {
        case 10: ok = launch_k<double, 10>(s, p, dops, st, src, dst); break;
        case 11: ok = launch_k<double, 11>(s, p, dops, st, src, dst); break;
        case 12: ok = launch_k<double, 12>(s, p, dops, st, src, dst); break;
        default: break;
        }
    } else {
        switch (p.K) {
        case 10: ok = launch_k<float, 10>(s, p, dops, st, src, dst); break;
        case 11: ok = launch_k<float, 11>(s, p, dops, st, src, dst); break;
        case 12: ok = launch_k<float, 12>(s, p, dops, st, src, dst); break;
        case 13: ok = launch_k<float, 13>(s, p, dops, st, src, dst); break;
        default: break;
        }
    }
    if (!ok) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "tile width / threads");
    QI_HIP_TRY(hipGetLastError());
    return Status{};
}

Status launch_batch_wave(int prec, bool init, uint32_t blocks, hipStream_t st, void *arena, const WaveTask *tasks,
                         int ntasks, const RegPass *passes, const RegOp *ops, const void *ptab) {
    if (!blocks || ntasks <= 0) return Status{};
    const dim3 g(blocks), b(64);
    if (prec == QI_C128) {
        if (init)
            hipLaunchKernelGGL((k_batch_wave<double, kBatchWaveK128, true>), g, b, 0, st, (double2 *)arena, tasks, ntasks,
                               passes, ops, (const double *)ptab);
        else
            hipLaunchKernelGGL((k_batch_wave<double, kBatchWaveK128, false>), g, b, 0, st, (double2 *)arena, tasks, ntasks,
                               passes, ops, (const double *)ptab);
    } else {
        if (init)
            hipLaunchKernelGGL((k_batch_wave<float, kBatchWaveK64, true>), g, b, 0, st, (float2 *)arena, tasks, ntasks,
                               passes, ops, (const float *)ptab);
        else
            hipLaunchKernelGGL((k_batch_wave<float, kBatchWaveK64, false>), g, b, 0, st, (float2 *)arena, tasks, ntasks,
                               passes, ops, (const float *)ptab);
    }
    QI_HIP_TRY(hipGetLastError());
    return Status{};
}

}  // namespace qi
