// batch.hip — thousands of parametrised circuits of uneven qubit counts, each
// with its own SumOp, evaluated to energies in one hipGraph launch.
//
// Circuits that fit a workgroup's LDS (2^13 complex64 / 2^12 complex128
// amplitudes) run LDS-resident: one workgroup builds |0..0>, applies every gate
// and reduces the energy without touching HBM. Larger circuits live in one
// device arena and run tile passes, pass p of every circuit in one launch
// ("level" p), so a level boundary is the only synchronisation a circuit
// needs: circuits of 12+ (complex64) / 11+ (complex128) qubits run the tiled
// executor's one-wave register passes (k_batch_wave in tiled.hip), smaller
// ones LDS-tiled passes (k_batch_pass). Energies are then accumulated by
// tiled expectation passes (FP64, two-sum compensated per block, folded in
// block order). Parametric gates carry a slot into a device angle table, so
// new parameters are one upload and the graph is replayed.
#include <algorithm>
#include <cmath>
#include <memory>
#include <unordered_map>

#include "qi_internal.h"
#include "exec_plan.h"
#include "reduce_util.h"

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

template <int KMAX>
__device__ __forceinline__ uint64_t deposit_pos(uint32_t e, const uint8_t *pos, int k) {
    uint64_t r = 0;
#pragma unroll
    for (int j = 0; j < KMAX; j++)
        if (j < k) r |= uint64_t((e >> j) & 1) << pos[j];
    return r;
}

constexpr int kBT = 256;
// Largest LDS-resident circuit (qubits) of the complex64 / complex128 kernels.
constexpr int kSmallK64 = 13, kSmallK128 = 12;
// Op kind of two parametric, uncontrolled RY on different qubits (lp0, lp1)
// in one sweep of the tile, each optionally followed by its folded phase
// (slot2, slot4); the first RY's angle in pad, the second's in slot3.
constexpr int kTwoRY = 100;
// Op kind of a run of in-tile CNOTs as one permutation of the local index:
// the new amplitude at e is the old one at G(e) = XOR of column j for each set
// bit j of e; the columns are uint32 masks packed into c[] (at most 16).
constexpr int kPerm = 101;
// Threads of a tiled pass (a 2^K-amplitude tile in LDS; 256 measured fastest
// against 64, 128 and 512 on the 4096-circuit ansatz batch).
constexpr int kPT = 256, kPTLog = __builtin_ctz(kPT);

// Coefficients of one op; parametric ops take their angle from the table
// (same formulas as lower_gate, evaluated on the device).
struct OpC {
    double c[8];
};

// Per-slot trigonometry, once per run (not per op and thread): slot i holds
// (cos th/2, sin th/2, cos th, sin th) of angle i.
__global__ __launch_bounds__(256) void k_batch_resolve(const double *__restrict__ angles, uint32_t n,
                                                       double *__restrict__ coef) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const double th = angles[i];
    double sh, ch, s1, c1;
    sincos(th / 2.0, &sh, &ch);
    sincos(th, &s1, &c1);
    coef[4 * i] = ch; coef[4 * i + 1] = sh; coef[4 * i + 2] = c1; coef[4 * i + 3] = s1;
}

// Per-slot coefficient rows of the one-wave passes (kParamRow values per
// slot in the pass's precision, the format k_batch_wave reads), from the
// slot trigonometry above.
template <typename R>
__global__ __launch_bounds__(256) void k_batch_ptab(const double *__restrict__ coef, uint32_t n, R *__restrict__ out) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const double *k = coef + 4 * (size_t)i;
    R *o = out + (size_t)kParamRow * i;
    o[0] = R(k[0]); o[1] = R(-k[1]); o[2] = R(k[1]); o[3] = R(k[2]);
    o[4] = R(k[3]);
    // the rotation by psi = t/2 as shears: psi moved by pi into [-pi/2, pi/2]
    // (cos psi >= 0), a = -tan(psi/2) = -sin psi / (1 + cos psi), b = sin psi
    const double cr = k[0] < 0.0 ? -k[0] : k[0], sr = k[0] < 0.0 ? -k[1] : k[1];
    o[5] = R(-sr / (1.0 + cr)); o[6] = R(sr); o[7] = R(0);
}

__device__ __forceinline__ void resolve(const TileOp &op, const double *__restrict__ coef, int slot_base, double *c) {
#pragma unroll
    for (int q = 0; q < 8; q++) c[q] = op.c[q];
    const int slot = (int)op.pad;
    if (slot < 0) return;
    const double *k = coef + 4 * (size_t)(slot_base + slot);
    const double ch = k[0], sh = k[1];
    switch (op.kind) {
    case LK_RY: case kTwoRY: c[0] = ch; c[1] = sh; break;
    case LK_RZ: c[0] = ch; c[1] = -sh; c[2] = ch; c[3] = sh; break;
    case LK_RX: c[0] = ch; c[1] = 0.0 * sh; c[2] = sh; c[3] = -0.0 * sh; c[4] = -sh; break;
    case LK_PHASE: c[2] = k[2]; c[3] = k[3]; break;
    default: break;
    }
}

// (cos theta, sin theta) of a folded RZ(theta), or nullptr.
__device__ __forceinline__ const double *phase_of(const TileOp &op, const double *__restrict__ coef, int slot_base) {
    return op.slot2 >= 0 ? coef + 4 * (size_t)(slot_base + op.slot2) + 2 : nullptr;
}

// Apply one op to a tile held in LDS (shared by the LDS-resident and the tiled kernels).
// ph: e^{i theta} applied to the target-1 output of a one-qubit pair op
// (a folded RZ, see fold_rz), or nullptr.
template <typename R, int N, int NT>
__device__ __forceinline__ void tile_apply(typename V2<R>::T *tile, const TileOp &op, const double *c, uint64_t base,
                                           int nloc, const double *ph, const double *__restrict__ coef, int slot_base) {
    using T = typename V2<R>::T;
    const uint32_t tid = threadIdx.x, lcm = op.lcm;
    const uint32_t nel = 1u << nloc;
    switch (op.kind) {
    case LK_Z: case LK_PHASE: case LK_RZ: {
        T p0 = mk<T>(R(c[0]), R(c[1])), p1 = mk<T>(R(c[2]), R(c[3]));
        bool half = op.kind != LK_RZ;
        if (op.kind == LK_RZ && lcm == 0 && op.gcm == 0) {
            // this engine returns energies only, which no global phase changes:
            // an uncontrolled diag(p0, p1) acts as diag(1, p1 conj(p0))
            p1 = mk<T>(R(c[2] * c[0] + c[3] * c[1]), R(c[3] * c[0] - c[2] * c[1]));
            half = true;
        }
        if (half) {  // only the target-1 amplitudes change
            if (op.outside) {
                if (!((base >> op.gt) & 1)) break;
                for (uint32_t e = tid; e < nel; e += NT) {
                    if ((e & lcm) != lcm) continue;
                    T a = tile[e];
                    if (op.kind == LK_Z) { a.x = -a.x; a.y = -a.y; } else a = cmul(a, p1);
                    tile[e] = a;
                }
            } else {
                const uint32_t tb = 1u << op.lp0;
                for (uint32_t q = tid; q < (nel >> 1); q += NT) {
                    const uint32_t e = ins0(q, op.lp0) | tb;
                    if ((e & lcm) != lcm) continue;
                    T a = tile[e];
                    if (op.kind == LK_Z) { a.x = -a.x; a.y = -a.y; } else a = cmul(a, p1);
                    tile[e] = a;
                }
            }
            break;
        }
        for (uint32_t e = tid; e < nel; e += NT) {
            if ((e & lcm) != lcm) continue;
            const bool one = op.outside ? ((base >> op.gt) & 1) : ((e >> op.lp0) & 1);
            tile[e] = cmul(tile[e], one ? p1 : p0);
        }
    } break;
    case kPerm: if constexpr (sizeof(R) == 4) {  // complex64 only (complex128 keeps one sweep per CNOT)
        uint32_t cols[16];
        memcpy(cols, op.c, sizeof cols);
        T v[N / NT];
#pragma unroll
        for (int k = 0; k < N / NT; k++) {
            const uint32_t e = tid + (uint32_t)k * NT;
            if (e < nel) {
                uint32_t src = 0;
#pragma unroll
                for (int j = 0; j < 16; j++)  // fixed trip count: cols stays in registers
                    if (j < nloc && ((e >> j) & 1u)) src ^= cols[j];
                v[k] = tile[src];
            }
        }
        __syncthreads();  // every gather done before any element is overwritten
#pragma unroll
        for (int k = 0; k < N / NT; k++) {
            const uint32_t e = tid + (uint32_t)k * NT;
            if (e < nel) tile[e] = v[k];
        }
    } break;
    case kTwoRY: {
        // (cos, sin) of each half angle from the slot table; c[] holds the first RY's
        const double c1 = c[0], s1 = c[1];
        const double *k2 = coef + 4 * (size_t)(slot_base + op.slot3);
        const double c2 = k2[0], s2 = k2[1];
        const double *q2 = op.slot4 >= 0 ? coef + 4 * (size_t)(slot_base + op.slot4) + 2 : nullptr;
        const uint32_t b1 = 1u << op.lp0, b2 = 1u << op.lp1;
        const unsigned lo = op.lp0 < op.lp1 ? op.lp0 : op.lp1, hi = op.lp0 < op.lp1 ? op.lp1 : op.lp0;
        for (uint32_t q = tid; q < (nel >> 2); q += NT) {
            const uint32_t l = ins0(ins0(q, lo), hi);
            T a00 = tile[l], a01 = tile[l | b1], a10 = tile[l | b2], a11 = tile[l | b1 | b2];
            // first RY on (x0, x1) pairs along b1, then its phase on the b1-set outputs
            T n00 = csub(cscale<T>(R(c1), a00), cscale<T>(R(s1), a01)), n01 = cadd(cscale<T>(R(s1), a00), cscale<T>(R(c1), a01));
            T n10 = csub(cscale<T>(R(c1), a10), cscale<T>(R(s1), a11)), n11 = cadd(cscale<T>(R(s1), a10), cscale<T>(R(c1), a11));
            if (ph) { n01 = cmul(n01, mk<T>(R(ph[0]), R(ph[1]))); n11 = cmul(n11, mk<T>(R(ph[0]), R(ph[1]))); }
            // second RY along b2, then its phase on the b2-set outputs
            T m00 = csub(cscale<T>(R(c2), n00), cscale<T>(R(s2), n10)), m10 = cadd(cscale<T>(R(s2), n00), cscale<T>(R(c2), n10));
            T m01 = csub(cscale<T>(R(c2), n01), cscale<T>(R(s2), n11)), m11 = cadd(cscale<T>(R(s2), n01), cscale<T>(R(c2), n11));
            if (q2) { m10 = cmul(m10, mk<T>(R(q2[0]), R(q2[1]))); m11 = cmul(m11, mk<T>(R(q2[0]), R(q2[1]))); }
            tile[l] = m00; tile[l | b1] = m01; tile[l | b2] = m10; tile[l | b1 | b2] = m11;
        }
    } break;
    case LK_SWAP: case LK_MATCH: {
        const uint32_t b0 = 1u << op.lp0, b1 = 1u << op.lp1;
        const unsigned lo = op.lp0 < op.lp1 ? op.lp0 : op.lp1, hi = op.lp0 < op.lp1 ? op.lp1 : op.lp0;
        for (uint32_t q = tid; q < (nel >> 2); q += NT) {
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
                    tile[i01] = csub(cscale<T>(R(c[0]), a01), cmul(mk<T>(R(c[2]), R(c[3])), a10));
                    tile[i10] = cadd(cscale<T>(R(c[1]), a01), cmul(mk<T>(R(c[4]), R(c[5])), a10));
                }
                if ((i11 & lcm) == lcm) tile[i11] = cmul(tile[i11], mk<T>(R(c[6]), R(c[7])));
            }
        }
    } break;
    default: {
        const uint32_t tb = 1u << op.lp0;
        for (uint32_t q = tid; q < (nel >> 1); q += NT) {
            uint32_t i = ins0(q, op.lp0), j = i | tb;
            if ((i & lcm) != lcm) continue;
            T a = tile[i], b = tile[j], na, nb;
            switch (op.kind) {
            case LK_H: na = cscale<T>(R(c[0]), cadd(a, b)); nb = cscale<T>(R(c[0]), csub(a, b)); break;
            case LK_X: na = b; nb = a; break;
            case LK_Y: na = cmul(mk<T>(R(c[0]), R(c[1])), b); nb = cmul(mk<T>(R(c[2]), R(c[3])), a); break;
            case LK_RX:
                na = csub(cscale<T>(R(c[0]), a), cmul(mk<T>(R(c[1]), R(c[2])), b));
                nb = cadd(cmul(mk<T>(R(c[3]), R(c[4])), a), cscale<T>(R(c[0]), b));
                break;
            case LK_RY:
                na = csub(cscale<T>(R(c[0]), a), cscale<T>(R(c[1]), b));
                nb = cadd(cscale<T>(R(c[1]), a), cscale<T>(R(c[0]), b));
                break;
            default:
                na = cadd(cmul(mk<T>(R(c[0]), R(c[1])), a), cmul(mk<T>(R(c[2]), R(c[3])), b));
                nb = cadd(cmul(mk<T>(R(c[4]), R(c[5])), a), cmul(mk<T>(R(c[6]), R(c[7])), b));
                break;
            }
            if (ph) nb = cmul(nb, mk<T>(R(ph[0]), R(ph[1])));
            tile[i] = na; tile[j] = nb;
        }
    } break;
    }
}

// An energy term: coefficient with (-i)^{nY} folded into quarter turns. A
// circuit's terms are stored in groups of one x mask; within a group the
// terms whose sign depends on the element slot (see tile_energy) come first,
// and the group's first term holds the counts.
struct ETerm {
    uint64_t x, z;
    uint32_t q0;
    uint32_t lx;         // the x mask over the tile's local bits
    uint32_t zl;         // the z mask over the tile's local bits (z bits outside the tile act through its base)
    uint16_t ng, nslot;  // first term of a group: terms in the group, and how many of them come first as slot terms
    double c[2];         // the coefficient times i^q0
};

// Elements per thread and chunk of the energy sums over a 2^K-element tile,
// and the local index bits that select the element slot within a thread.
constexpr int energy_per(int K) { return (1 << K) / kBT < 8 ? (1 << K) / kBT : 8; }
constexpr uint32_t energy_slot_bits(int per) { return uint32_t(per - 1) * kBT; }

// Sum over the tile's elements e < nel of conj(psi_e) * c * i^{q(e)} * psi_{e ^ x}.
// x lies inside the tile, so flipping its global bits flips the same local
// bits (ETerm::lx). Thread tid takes elements e0 + tid + it * kBT (it < PER)
// of each chunk of PER * kBT elements. The terms of one x mask share
// w = conj(a) b: their signed coefficients sum first (g) and meet w once. The
// sign (-1)^popcount(index & z) is the product of the thread's part (tile
// base, chunk and thread index: folded into the coefficient once per term and
// thread) and the element slot's part (uniform: a scalar +-1). Terms whose z
// mask misses the slot bits have one sign for the thread's PER elements and
// are summed once per thread (gc).
template <typename R, int PER>
__device__ __forceinline__ void tile_energy(const typename V2<R>::T *tile, const ETerm *__restrict__ terms, int nt,
                                            uint64_t base, uint32_t nel, DD &sr, DD &si) {
    const uint32_t tid = threadIdx.x;
    for (uint32_t e0 = 0; e0 < nel; e0 += PER * kBT) {
        double hr[PER], hi[PER];
#pragma unroll
        for (int it = 0; it < PER; it++) hr[it] = hi[it] = 0.0;
        double ar[PER], ai[PER];  // the thread's elements, read once per chunk
#pragma unroll
        for (int it = 0; it < PER; it++) {
            const uint32_t e = e0 + tid + uint32_t(it) * kBT;
            ar[it] = e < nel ? double(tile[e].x) : 0.0;
            ai[it] = e < nel ? double(tile[e].y) : 0.0;
        }
        // the thread's signed coefficient: c i^q0 with the sign bit of the
        // base, chunk and thread parts of (-1)^popcount(index & z)
        const auto thread_coeff = [&](const ETerm &tm, double &cr, double &ci) {
            const uint32_t p = uint32_t(__popcll(base & tm.z) + __popc((e0 | tid) & tm.zl)) << 31;
            cr = __hiloint2double(__double2hiint(tm.c[0]) ^ (int)p, __double2loint(tm.c[0]));
            ci = __hiloint2double(__double2hiint(tm.c[1]) ^ (int)p, __double2loint(tm.c[1]));
        };
        for (int t = 0; t < nt;) {
            const ETerm head = terms[t];
            const uint32_t lx = head.lx;
            const int ng = head.ng, ns = head.nslot;
            double gcr = 0.0, gci = 0.0;  // terms with one sign for all of the thread's elements
            for (int k = ns; k < ng; k++) {
                double cr, ci;
                thread_coeff(terms[t + k], cr, ci);
                gcr += cr;
                gci += ci;
            }
            double gr[PER], gi[PER];
#pragma unroll
            for (int it = 0; it < PER; it++) { gr[it] = gcr; gi[it] = gci; }
            for (int k = 0; k < ns; k++) {
                const ETerm tm = terms[t + k];
                double cr, ci;
                thread_coeff(tm, cr, ci);
#pragma unroll
                for (int it = 0; it < PER; it++) {
                    const double su = (__popc((uint32_t(it) * kBT) & tm.zl) & 1) ? -1.0 : 1.0;
                    gr[it] = fma(su, cr, gr[it]);
                    gi[it] = fma(su, ci, gi[it]);
                }
            }
            // the group meets w = conj(a) b once per element
#pragma unroll
            for (int it = 0; it < PER; it++) {
                const uint32_t e = e0 + tid + uint32_t(it) * kBT;
                if (e < nel) {
                    const double br = tile[e ^ lx].x, bi = tile[e ^ lx].y;
                    const double wr = fma(ar[it], br, ai[it] * bi), wi = fma(ar[it], bi, -(ai[it] * br));
                    hr[it] = fma(wr, gr[it], fma(-wi, gi[it], hr[it]));
                    hi[it] = fma(wr, gi[it], fma(wi, gr[it], hi[it]));
                }
            }
            t += ng;
        }
#pragma unroll
        for (int it = 0; it < PER; it++)
            if (e0 + tid + uint32_t(it) * kBT < nel) {
                dd_acc(sr, hr[it]);
                dd_acc(si, hi[it]);
            }
    }
}

// ------------------------------------------------------- LDS-resident path --
struct SmallTask {
    uint32_t n, op0, nops, t0, nt;
    uint32_t out;       // energy slot
    int32_t slot_base;  // first angle of this circuit in the table
    uint32_t pad;
};

template <typename R, int KMAX>
__global__ __launch_bounds__(kBT) void k_batch_small(const SmallTask *__restrict__ tasks, const TileOp *__restrict__ ops,
                                                     const ETerm *__restrict__ terms, const double *__restrict__ coef,
                                                     double *__restrict__ energies) {
    using T = typename V2<R>::T;
    __shared__ T tile[1 << KMAX];
    __shared__ double sh[4 * 4];
    const SmallTask tk = tasks[blockIdx.x];
    const uint32_t nel = 1u << tk.n;
    for (uint32_t e = threadIdx.x; e < nel; e += kBT) { T z; z.x = e == 0 ? R(1) : R(0); z.y = R(0); tile[e] = z; }
    __syncthreads();
    for (uint32_t o = 0; o < tk.nops; o++) {
        const TileOp op = ops[tk.op0 + o];
        double c[8];
        resolve(op, coef, tk.slot_base, c);
        tile_apply<R, (1 << KMAX), kBT>(tile, op, c, 0, (int)tk.n, phase_of(op, coef, tk.slot_base), coef, tk.slot_base);
        __syncthreads();
    }
    DD v[2] = {dd_zero(), dd_zero()};
    tile_energy<R, energy_per(KMAX)>(tile, terms + tk.t0, (int)tk.nt, 0, nel, v[0], v[1]);
    block_reduce_dd<2>(v, sh);
    if (threadIdx.x == 0) {
        energies[2 * tk.out] = v[0].s + v[0].c;
        energies[2 * tk.out + 1] = v[1].s + v[1].c;
    }
}

// ------------------------------------------------------------- tiled path --
struct BigTask {
    uint64_t off;      // element offset of the circuit's state in the arena
    uint64_t qmask;    // tile qubits
    uint32_t block0;   // first block of this task in the launch
    uint32_t op0, nops;
    uint32_t init;     // 1: the state is |0..0> (first pass reads nothing)
    uint32_t t0, nt;   // energy terms (expectation passes)
    uint32_t part0;    // first partial slot (expectation passes)
    int32_t slot_base; // first angle of this circuit in the table
    uint8_t pos[48];
};

__device__ __forceinline__ int find_task(const BigTask *__restrict__ tasks, int ntasks, uint32_t b) {
    int lo = 0, hi = ntasks - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (tasks[mid].block0 <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

template <typename R, int K>
__global__ __launch_bounds__(kPT) void k_batch_pass(typename V2<R>::T *__restrict__ arena, const BigTask *__restrict__ tasks,
                                                    int ntasks, const TileOp *__restrict__ ops,
                                                    const double *__restrict__ coef) {
    using T = typename V2<R>::T;
    constexpr int N = 1 << K, PER = N / kPT;
    __shared__ T tile[N];
    const int ti = find_task(tasks, ntasks, blockIdx.x);
    const BigTask &tk = tasks[ti];
    const uint64_t base = insert_zeros64(uint64_t(blockIdx.x - tk.block0), tk.qmask);
    T *psi = arena + tk.off;
    const uint64_t off_t = deposit_pos<kPTLog>(threadIdx.x, tk.pos, kPTLog);
    if (tk.init) {
#pragma unroll
        for (int it = 0; it < PER; it++) {
            const uint32_t e = threadIdx.x + it * kPT;
            T z; z.x = (base == 0 && e == 0) ? R(1) : R(0); z.y = R(0);
            tile[e] = z;
        }
    } else {
#pragma unroll
        for (int it = 0; it < PER; it++)
            tile[threadIdx.x + it * kPT] = psi[base | off_t | deposit_pos<K>(uint32_t(it) << kPTLog, tk.pos, K)];
    }
    __syncthreads();
    for (uint32_t o = 0; o < tk.nops; o++) {
        const TileOp op = ops[tk.op0 + o];
        if ((base & op.gcm) != op.gcm) continue;
        double c[8];
        resolve(op, coef, tk.slot_base, c);
        tile_apply<R, N, kPT>(tile, op, c, base, K, phase_of(op, coef, tk.slot_base), coef, tk.slot_base);
        __syncthreads();
    }
#pragma unroll
    for (int it = 0; it < PER; it++)
        psi[base | off_t | deposit_pos<K>(uint32_t(it) << kPTLog, tk.pos, K)] = tile[threadIdx.x + it * kPT];
}

template <typename R, int K>
__global__ __launch_bounds__(kBT) void k_batch_expect(const typename V2<R>::T *__restrict__ arena,
                                                      const BigTask *__restrict__ tasks, int ntasks,
                                                      const ETerm *__restrict__ terms, double *__restrict__ part) {
    using T = typename V2<R>::T;
    constexpr int N = 1 << K, PER = N / kBT;
    __shared__ T tile[N];
    __shared__ double sh[4 * 4];
    const int ti = find_task(tasks, ntasks, blockIdx.x);
    const BigTask &tk = tasks[ti];
    const uint32_t local = blockIdx.x - tk.block0;
    const uint64_t base = insert_zeros64(uint64_t(local), tk.qmask);
    const T *psi = arena + tk.off;
    const uint64_t off_t = deposit_pos<8>(threadIdx.x, tk.pos, 8);
#pragma unroll
    for (int it = 0; it < PER; it++)
        tile[threadIdx.x + it * kBT] = psi[base | off_t | deposit_pos<K>(uint32_t(it) << 8, tk.pos, K)];
    __syncthreads();
    DD v[2] = {dd_zero(), dd_zero()};
    tile_energy<R, energy_per(K)>(tile, terms + tk.t0, (int)tk.nt, base, uint32_t(N), v[0], v[1]);
    block_reduce_dd<2>(v, sh);
    if (threadIdx.x == 0) {
        part[(size_t(tk.part0) + local) * 4 + 0] = v[0].s; part[(size_t(tk.part0) + local) * 4 + 1] = v[0].c;
        part[(size_t(tk.part0) + local) * 4 + 2] = v[1].s; part[(size_t(tk.part0) + local) * 4 + 3] = v[1].c;
    }
}

// Fold each circuit's partials (all its expectation tiles, in order) into its energy.
struct FoldTask {
    uint32_t part0, nparts, out, pad;
};

__global__ __launch_bounds__(256) void k_batch_fold(const FoldTask *__restrict__ f, const double *__restrict__ part,
                                                    double *__restrict__ energies) {
    __shared__ double sh[4 * 4];
    const FoldTask t = f[blockIdx.x];
    DD v[2] = {dd_zero(), dd_zero()};
    for (uint32_t k = threadIdx.x; k < t.nparts; k += blockDim.x) {
        const double *p = part + (size_t(t.part0) + k) * 4;
        v[0] = dd_add(v[0], DD{p[0], p[1]});
        v[1] = dd_add(v[1], DD{p[2], p[3]});
    }
    block_reduce_dd<2>(v, sh);
    if (threadIdx.x == 0) {
        energies[2 * t.out] = v[0].s + v[0].c;
        energies[2 * t.out + 1] = v[1].s + v[1].c;
    }
}

}  // namespace

// ------------------------------------------------------------------ host --
struct Batch {
    Ctx *ctx = nullptr;
    int prec = QI_C64;
    // a circuit shape shared by many circuits: gates and angle slots relative
    // to each circuit's slot base
    struct Program {
        uint32_t n;
        std::vector<qi_gate> gates;
        std::vector<int32_t> rel_slot;
        // planned form (filled by batch_plan)
        bool small = false;
        uint32_t op0 = 0, nops = 0;           // small: op range
        std::vector<TilePass> passes;         // tiled: passes (op0 relative to ops_base)
        uint32_t ops_base = 0;
        // one-wave form (planned once: it does not depend on the schedule)
        int wave = -1;                        // -1 not planned, 0 no one-wave plan, 1 planned
        std::vector<RegPass> wpasses;         // op0 relative to the program's wops
        std::vector<RegOp> wops;
        uint32_t wpass_base = 0;              // first pass in the uploaded pass table
    };
    struct Circ {
        uint32_t prog;
        int32_t slot_base;
        std::vector<qi_pauli_term> terms;
        std::vector<uint32_t> eterm_idx;  // where each term landed in the device table
    };
    std::vector<Program> progs;
    std::vector<Circ> circs;
    std::vector<double> angles;  // host copy of the angle table
    std::vector<ETerm> eterms;   // host copy of the term table
    // device
    DevBuf d_ops, d_terms, d_small, d_big, d_fold, d_angles, d_coef, d_part, d_energy, d_arena;
    DevBuf d_wops, d_wpasses, d_wtasks, d_ptab;     // one-wave passes: ops, passes, tasks, slot coefficients
    int exec = QI_BATCH_WAVE;
    std::vector<std::pair<size_t, size_t>> levels;  // [task begin, task end) of each pass level
    std::vector<uint32_t> level_blocks;
    std::vector<std::pair<size_t, size_t>> wlevels;  // the same for one-wave passes
    std::vector<uint32_t> wlevel_blocks;
    std::vector<std::pair<size_t, size_t>> elevels;
    std::vector<uint32_t> elevel_blocks;
    size_t nsmall = 0, nfold = 0;
    int K = 0, KS = 0;        // tile width of tiled circuits, largest LDS-resident circuit (0: defaults)
    std::unordered_multimap<uint64_t, uint32_t> prog_index;  // shape hash -> program
    bool planned = false;
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
    hipGraphExec_t graph = nullptr;
    double last_ms = 0.0;
    ~Batch() {
        for (DevBuf *b : {&d_ops, &d_terms, &d_small, &d_big, &d_fold, &d_angles, &d_coef, &d_part, &d_energy, &d_arena,
                          &d_wops, &d_wpasses, &d_wtasks, &d_ptab})
            if (b->p) (void)hipFree(b->p);
        if (graph) (void)hipGraphExecDestroy(graph);
        if (side) (void)hipStreamDestroy(side);
        if (fork) (void)hipEventDestroy(fork);
        if (join) (void)hipEventDestroy(join);
    }
};

// c * i^q0 (the device adds it with the sign (-1)^popcount(index & z))
static void set_eterm_coeff(ETerm &t, double re, double im) {
    switch (t.q0) {
    case 0: t.c[0] = re; t.c[1] = im; break;
    case 1: t.c[0] = -im; t.c[1] = re; break;
    case 2: t.c[0] = -re; t.c[1] = -im; break;
    default: t.c[0] = im; t.c[1] = -re; break;
    }
}

static ETerm make_eterm(const qi_pauli_term &p) {
    ETerm t{};
    t.x = p.x_mask; t.z = p.z_mask;
    t.q0 = (3u * (unsigned)__builtin_popcountll(p.x_mask & p.z_mask)) & 3u;
    set_eterm_coeff(t, p.c_re, p.c_im);
    return t;
}

// Runs of consecutive CNOTs inside the tile (one in-tile control, no other
// control) become one kPerm op. A CNOT c -> t maps index e to
// F(e) = e ^ (bit c of e) << t, and the state to new[F(e)] = old[e]; for the
// run F_k ... F_1 the gather map is G = F_1 o ... o F_k (each F its own
// inverse), built column by column. In place; returns the new length.
// complex128 ops are returned unchanged: the kPerm case is compiled for
// complex64 kernels only (the complex128 gather buffer would not stay in
// registers), so no kPerm op may reach a complex128 kernel.
static size_t perm_cnots(TileOp *ops, size_t n, int nloc, bool c128) {
    if (c128) return n;
    const auto cnot = [](const TileOp &o) {
        return o.kind == LK_X && !o.outside && o.gcm == 0 && __builtin_popcount(o.lcm) == 1 && (int32_t)o.pad < 0 &&
               o.slot2 < 0;
    };
    size_t w = 0;
    for (size_t r = 0; r < n;) {
        size_t e = r;
        while (e < n && cnot(ops[e])) e++;
        if (e - r < 2 || nloc > 16) {
            if (e == r) e = r + 1;
            for (size_t k = r; k < e; k++) ops[w++] = ops[k];
            r = e;
            continue;
        }
        uint32_t cols[16];
        for (int j = 0; j < 16; j++) cols[j] = j < nloc ? 1u << j : 0u;
        for (size_t k = e; k-- > r;) {  // G = F_r o ... o F_(e-1): apply the later maps first
            const int c = __builtin_ctz(ops[k].lcm), t = ops[k].lp0;
            for (int j = 0; j < nloc; j++) cols[j] ^= ((cols[j] >> c) & 1u) << t;
        }
        TileOp p{};
        p.kind = kPerm;
        p.pad = (uint32_t)-1;
        memcpy(p.c, cols, sizeof cols);
        ops[w++] = p;
        r = e;
    }
    return w;
}

// Adjacent uncontrolled parametric RY ops on two positions become one kTwoRY
// op (both act on every amplitude; they commute). In place; returns the new
// length.
static size_t pair_ry(TileOp *ops, size_t n) {
    size_t w = 0;
    for (size_t r = 0; r < n; r++) {
        TileOp op = ops[r];
        const auto single = [](const TileOp &o) {
            return o.kind == LK_RY && o.lcm == 0 && o.gcm == 0 && (int32_t)o.pad >= 0;
        };
        if (r + 1 < n && single(op) && single(ops[r + 1]) && ops[r + 1].lp0 != op.lp0) {
            op.kind = kTwoRY;
            op.lp1 = ops[r + 1].lp0;
            op.slot3 = (int32_t)ops[r + 1].pad;
            op.slot4 = ops[r + 1].slot2;
            r++;
        }
        ops[w++] = op;
    }
    return w;
}

// This engine returns energies only, which no global phase changes, so an
// uncontrolled RZ(theta) = e^{-i theta/2} diag(1, e^{i theta}). A parametric
// one on position p folds into the last op before it on p when that is an
// uncontrolled one-qubit pair gate: its target-1 output takes e^{i theta}
// (ops on other positions in between commute with both). In place on a slice
// whose `pad` holds angle slots; returns the new length.
static size_t fold_rz(TileOp *ops, size_t n) {
    int last[64];
    for (int &x : last) x = -1;
    size_t w = 0;
    for (size_t r = 0; r < n; r++) {
        const TileOp op = ops[r];
        if (op.kind == LK_RZ && !op.outside && op.lcm == 0 && op.gcm == 0 && (int32_t)op.pad >= 0 && last[op.lp0] >= 0) {
            TileOp &pv = ops[last[op.lp0]];
            const bool one_q = pv.kind == LK_H || pv.kind == LK_X || pv.kind == LK_Y || pv.kind == LK_RX ||
                               pv.kind == LK_RY || pv.kind == LK_U2;
            if (one_q && pv.lp0 == op.lp0 && pv.lcm == 0 && pv.gcm == 0 && pv.slot2 < 0) {
                pv.slot2 = (int32_t)op.pad;
                continue;
            }
        }
        ops[w] = op;
        const bool diag_outside = op.outside && (op.kind == LK_Z || op.kind == LK_PHASE || op.kind == LK_RZ);
        if (!diag_outside) last[op.lp0] = (int)w;
        if (op.kind == LK_SWAP || op.kind == LK_MATCH) last[op.lp1] = (int)w;
        for (uint32_t m = op.lcm; m; m &= m - 1) last[__builtin_ctz(m)] = (int)w;
        w++;
    }
    return w;
}

// Appends terms idx of circuit C to the term table for a tile whose local bit
// j is qubit pos[j] (K bits) and an energy kernel with `per` elements per
// thread: groups of one x mask, each with its slot terms (z on the
// element-slot bits) first.
static void append_terms(Batch &B, Batch::Circ &C, const std::vector<size_t> &idx, const uint8_t *pos, int K, int per) {
    const uint32_t slot_bits = energy_slot_bits(per);
    std::vector<ETerm> ts(idx.size());
    for (size_t i = 0; i < idx.size(); i++) {
        const qi_pauli_term &p = C.terms[idx[i]];
        ETerm &t = ts[i];
        t = make_eterm(p);
        for (int j = 0; j < K; j++) {
            t.lx |= uint32_t((p.x_mask >> pos[j]) & 1) << j;
            t.zl |= uint32_t((p.z_mask >> pos[j]) & 1) << j;
        }
    }
    std::vector<size_t> order(idx.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        const bool sa = (ts[a].zl & slot_bits) != 0, sb = (ts[b].zl & slot_bits) != 0;
        return ts[a].lx != ts[b].lx ? ts[a].lx < ts[b].lx : sa > sb;
    });
    for (size_t g = 0; g < order.size();) {
        size_t e = g, ns = 0;
        while (e < order.size() && ts[order[e]].lx == ts[order[g]].lx) {
            if (ts[order[e]].zl & slot_bits) ns++;
            e++;
        }
        ts[order[g]].ng = (uint16_t)(e - g);
        ts[order[g]].nslot = (uint16_t)ns;
        g = e;
    }
    for (size_t i : order) {
        C.eterm_idx[idx[i]] = (uint32_t)B.eterms.size();
        B.eterms.push_back(ts[i]);
    }
}

template <typename T> static Status upload_vec(DevBuf &b, const std::vector<T> &v) {
    Status e = ensure(b, std::max<size_t>(v.size() * sizeof(T), 16));
    if (!e.ok()) return e;
    if (!v.empty()) QI_HIP_TRY(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return Status{};
}

// The program's one-wave plan: the tiled executor's planner (kBatchWaveK
// tile qubits, one wave per tile) with every slotted gate kept parametric.
// Programs whose plan needs anything but in-place one-wave passes (streamed
// LDS gates) keep LDS tiles. This engine returns energies only, which no
// global phase changes, so an uncontrolled parametric RZ(theta) =
// e^{-i theta/2} diag(1, e^{i theta}) runs as the one-sided phase of P(theta)
// (the target-1 half only), and in complex64 an uncontrolled parametric RY on
// a register bit as three shears (RY(theta) = -RY(theta -+ 2 pi) keeps the
// shear coefficients bounded).
static void plan_wave(Batch::Program &P, int prec) {
    P.wave = 0;
    P.wpasses.clear();
    P.wops.clear();
    std::vector<Lowered> lw;
    for (size_t g = 0; g < P.gates.size(); g++) {
        Lowered Lg;
        if (!lower_gate(P.n, P.gates[g], Lg).ok()) return;
        if (Lg.kind == LK_NOP) continue;
        Lg.slot = P.rel_slot[g];
        lw.push_back(Lg);
    }
    if (lw.empty()) return;
    const int K = prec == QI_C128 ? kBatchWaveK128 : kBatchWaveK64;
    State shape;
    shape.n = P.n;
    shape.prec = prec;
    ExecPlan plan;
    if (!plan_tiled(shape, lw, K, kLaneBits, plan, false).ok() || plan.steps.empty()) return;
    for (const ExecStep &st : plan.steps)
        if (st.type != ExecStep::REG || st.reg.oop || st.reg.permute || st.reg.tb != kLaneBits || st.reg.K != K) return;
    for (RegOp &op : plan.reg_ops) {
        const bool free = !op.rcm && !op.tcm && !op.gcm;  // uncontrolled: a sign or phase of it is global
        if (op.param > 0 && op.kind == RO_DG && free) op.kind = RO_P1;
        // complex64: an uncontrolled RY on a register bit as three shears (the
        // rotation by pi moved out of its angle is a global sign)
        if (op.param > 0 && op.kind == RO_RR && op.tgt < kLaneTarget && free && prec == QI_C64) op.kind = RO_RRS;
    }
    for (const ExecStep &st : plan.steps) P.wpasses.push_back(st.reg);
    P.wops = std::move(plan.reg_ops);
    P.wave = 1;
}

// Plans every program once (LDS-resident op list or tiled passes), then lays
// out per-circuit tasks, expectation passes and the energy fold.
Status batch_plan(Batch &B) {
    const bool c128 = B.prec == QI_C128;
    if (!B.K) B.K = c128 ? 12 : 13;
    if (!B.KS) B.KS = B.K;
    if (B.KS < B.K) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "LDS-resident limit below the tile width");
    const int L = c128 ? 3 : 4;
    const int kw = c128 ? kBatchWaveK128 : kBatchWaveK64;
    const auto on_waves = [&](const Batch::Program &P) {
        return !P.small && B.exec == QI_BATCH_WAVE && (int)P.n >= kw && P.wave == 1;
    };
    std::vector<TileOp> ops;
    for (auto &P : B.progs) {
        std::vector<Lowered> lw;
        std::vector<int32_t> lslot;
        for (size_t g = 0; g < P.gates.size(); g++) {
            Lowered Lg;
            Status e = lower_gate(P.n, P.gates[g], Lg);
            if (!e.ok()) return e;
            if (Lg.kind == LK_NOP) continue;
            lw.push_back(Lg);
            lslot.push_back(P.rel_slot[g]);
        }
        P.small = (int)P.n <= B.KS;
        P.passes.clear();
        if (!P.small && B.exec == QI_BATCH_WAVE && (int)P.n >= kw) {
            if (P.wave < 0) plan_wave(P, B.prec);
            if (P.wave == 1) continue;
        }
        if (P.small) {
            P.op0 = (uint32_t)ops.size();
            for (size_t g = 0; g < lw.size(); g++) {
                TileOp op{};
                op.kind = lw[g].kind;
                for (int q = 0; q < 8; q++) op.c[q] = lw[g].c[q];
                if (op.kind == LK_Z) { op.c[2] = -1.0; op.c[3] = 0.0; }
                if (op.kind == LK_PHASE) { op.c[2] = lw[g].c[0]; op.c[3] = lw[g].c[1]; }
                op.lp0 = (uint8_t)lw[g].t0; op.lp1 = (uint8_t)lw[g].t1;
                op.lcm = (uint32_t)lw[g].ctrl_mask;
                op.pad = (uint32_t)lslot[g];
                ops.push_back(op);
            }
            const size_t nf = pair_ry(ops.data() + P.op0, fold_rz(ops.data() + P.op0, lw.size()));
            P.nops = (uint32_t)perm_cnots(ops.data() + P.op0, nf, (int)P.n, c128);
            ops.resize(P.op0 + P.nops);
            continue;
        }
        State fake;
        fake.n = P.n; fake.prec = B.prec;
        ExecPlan plan;
        Status e = plan_lds_passes(fake, lw, B.K, plan);
        if (!e.ok()) return e;
        P.ops_base = (uint32_t)ops.size();
        for (TileOp &op : plan.tile_ops) op.pad = (uint32_t)lslot[op.pad];
        std::vector<TileOp> packed;
        for (const ExecStep &s : plan.steps) {
            if (s.type != ExecStep::TILE) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "batched circuit not tileable");
            TilePass tp = s.tile;
            TileOp *slice = plan.tile_ops.data() + tp.op0;
            const size_t nf = pair_ry(slice, fold_rz(slice, (size_t)tp.nops));
            const size_t n = perm_cnots(slice, nf, B.K, c128);
            tp.op0 = (int32_t)packed.size();
            tp.nops = (int32_t)n;
            packed.insert(packed.end(), plan.tile_ops.begin() + s.tile.op0, plan.tile_ops.begin() + s.tile.op0 + n);
            P.passes.push_back(tp);
        }
        ops.insert(ops.end(), packed.begin(), packed.end());
    }
    std::vector<RegOp> wops;
    std::vector<RegPass> wpasses;
    for (auto &P : B.progs) {
        if (!on_waves(P)) continue;
        P.wpass_base = (uint32_t)wpasses.size();
        for (RegPass rp : P.wpasses) {
            rp.op0 += (int32_t)wops.size();
            wpasses.push_back(rp);
        }
        wops.insert(wops.end(), P.wops.begin(), P.wops.end());
    }
    std::vector<SmallTask> small;
    std::vector<std::vector<BigTask>> per_level, per_elevel;
    std::vector<std::vector<WaveTask>> per_wlevel;
    std::vector<FoldTask> folds;
    B.eterms.clear();
    uint64_t arena_elems = 0;
    uint32_t nparts = 0;
    for (size_t ci = 0; ci < B.circs.size(); ci++) {
        auto &C = B.circs[ci];
        const auto &P = B.progs[C.prog];
        C.eterm_idx.assign(C.terms.size(), 0);
        if (P.small) {
            SmallTask st{};
            st.n = P.n; st.op0 = P.op0; st.nops = P.nops; st.out = (uint32_t)ci; st.slot_base = C.slot_base;
            st.t0 = (uint32_t)B.eterms.size(); st.nt = (uint32_t)C.terms.size();
            std::vector<size_t> all(C.terms.size());
            for (size_t k = 0; k < all.size(); k++) all[k] = k;
            uint8_t ident[64];  // a whole-state tile: local bit = qubit
            for (int j = 0; j < 64; j++) ident[j] = (uint8_t)j;
            append_terms(B, C, all, ident, (int)P.n, energy_per(c128 ? kSmallK128 : kSmallK64));
            small.push_back(st);
            continue;
        }
        const uint64_t off = arena_elems;
        arena_elems += uint64_t(1) << P.n;
        const uint32_t tiles = uint32_t(1) << (P.n - B.K);
        if (on_waves(P)) {
            for (size_t p = 0; p < P.wpasses.size(); p++) {
                WaveTask wt{};
                wt.off = off; wt.pass = P.wpass_base + (uint32_t)p; wt.slot_base = C.slot_base;
                wt.block0 = uint32_t(1) << (P.n - kw);  // tile count; turned into offsets per level below
                if (per_wlevel.size() <= p) per_wlevel.resize(p + 1);
                per_wlevel[p].push_back(wt);
            }
        }
        size_t level = 0;
        for (const TilePass &tp : P.passes) {
            BigTask bt{};
            bt.off = off; bt.qmask = tp.qmask; bt.init = level == 0 ? 1 : 0; bt.slot_base = C.slot_base;
            memcpy(bt.pos, tp.pos, sizeof bt.pos);
            bt.op0 = P.ops_base + (uint32_t)tp.op0; bt.nops = (uint32_t)tp.nops;
            bt.block0 = tiles;  // tile count; turned into offsets per level below
            if (per_level.size() <= level) per_level.resize(level + 1);
            per_level[level].push_back(bt);
            level++;
        }
        if (P.passes.empty() && !on_waves(P)) {  // no gates: one pass that only writes |0..0>
            BigTask bt{};
            bt.off = off; bt.init = 1;
            uint64_t q = 0;
            for (int j = 0; j < B.K; j++) { q |= uint64_t(1) << j; bt.pos[j] = (uint8_t)j; }
            bt.qmask = q; bt.block0 = tiles;
            if (per_level.empty()) per_level.resize(1);
            per_level[0].push_back(bt);
        }
        // expectation passes: each term goes to the first pass whose tile holds its X/Y support
        std::vector<size_t> rest(C.terms.size());
        for (size_t k = 0; k < rest.size(); k++) rest[k] = k;
        uint32_t cparts = 0;
        const uint32_t part_first = nparts;
        size_t elevel = 0;
        while (!rest.empty()) {
            uint64_t q = (uint64_t(1) << L) - 1;
            std::vector<size_t> take, left;
            for (size_t k : rest) {
                uint64_t x = C.terms[k].x_mask;
                if (__builtin_popcountll(q | x) <= B.K) { q |= x; take.push_back(k); }
                else left.push_back(k);
            }
            if (take.empty()) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "Pauli term wider than the tile");
            for (int b = 0; __builtin_popcountll(q) < B.K; b++) q |= uint64_t(1) << b;
            BigTask bt{};
            bt.off = off; bt.qmask = q;
            int j = 0;
            for (int b = 0; b < (int)P.n; b++)
                if ((q >> b) & 1) bt.pos[j++] = (uint8_t)b;
            bt.t0 = (uint32_t)B.eterms.size(); bt.nt = (uint32_t)take.size();
            append_terms(B, C, take, bt.pos, B.K, energy_per(B.K));
            bt.part0 = nparts;
            nparts += tiles;
            cparts += tiles;
            bt.block0 = tiles;
            if (per_elevel.size() <= elevel) per_elevel.resize(elevel + 1);
            per_elevel[elevel].push_back(bt);
            elevel++;
            rest.swap(left);
        }
        FoldTask f{};
        f.part0 = part_first; f.nparts = cparts; f.out = (uint32_t)ci;
        folds.push_back(f);
    }
    // flatten levels: block0 becomes the running block offset inside the level
    std::vector<BigTask> big;
    B.levels.clear(); B.level_blocks.clear(); B.elevels.clear(); B.elevel_blocks.clear();
    auto flatten = [&](std::vector<std::vector<BigTask>> &lv, std::vector<std::pair<size_t, size_t>> &rng,
                       std::vector<uint32_t> &blocks) {
        for (auto &v : lv) {
            size_t b0 = big.size();
            uint32_t acc = 0;
            for (auto &t : v) { uint32_t tiles = t.block0; t.block0 = acc; acc += tiles; big.push_back(t); }
            rng.push_back({b0, big.size()});
            blocks.push_back(acc);
        }
    };
    flatten(per_level, B.levels, B.level_blocks);
    flatten(per_elevel, B.elevels, B.elevel_blocks);
    std::vector<WaveTask> wtasks;
    B.wlevels.clear(); B.wlevel_blocks.clear();
    for (auto &v : per_wlevel) {
        const size_t b0 = wtasks.size();
        uint32_t acc = 0;
        for (auto &t : v) { const uint32_t tiles = t.block0; t.block0 = acc; acc += tiles; wtasks.push_back(t); }
        B.wlevels.push_back({b0, wtasks.size()});
        B.wlevel_blocks.push_back(acc);
    }
    B.nsmall = small.size();
    B.nfold = folds.size();
    Status e;
    if (!(e = upload_vec(B.d_ops, ops)).ok()) return e;
    if (!(e = upload_vec(B.d_terms, B.eterms)).ok()) return e;
    if (!(e = upload_vec(B.d_small, small)).ok()) return e;
    if (!(e = upload_vec(B.d_big, big)).ok()) return e;
    if (!(e = upload_vec(B.d_fold, folds)).ok()) return e;
    if (!(e = upload_vec(B.d_angles, B.angles)).ok()) return e;
    if (!(e = ensure(B.d_coef, std::max<size_t>(B.angles.size(), 1) * 4 * sizeof(double))).ok()) return e;
    if (!(e = upload_vec(B.d_wops, wops)).ok()) return e;
    if (!(e = upload_vec(B.d_wpasses, wpasses)).ok()) return e;
    if (!(e = upload_vec(B.d_wtasks, wtasks)).ok()) return e;
    if (!(e = ensure(B.d_ptab, std::max<size_t>(B.angles.size(), 1) * kParamRow * (c128 ? 8 : 4))).ok()) return e;
    if (!(e = ensure(B.d_part, std::max<size_t>(nparts, 1) * 4 * sizeof(double))).ok()) return e;
    if (!(e = ensure(B.d_energy, std::max<size_t>(B.circs.size(), 1) * 2 * sizeof(double))).ok()) return e;
    const size_t esz = c128 ? 16 : 8;
    if (!(e = ensure(B.d_arena, std::max<uint64_t>(arena_elems, 1) * esz)).ok()) return e;
    if (!B.side) {
        QI_HIP_TRY(hipStreamCreateWithFlags(&B.side, hipStreamNonBlocking));
        QI_HIP_TRY(hipEventCreateWithFlags(&B.fork, hipEventDisableTiming));
        QI_HIP_TRY(hipEventCreateWithFlags(&B.join, hipEventDisableTiming));
    }
    if (B.graph) { (void)hipGraphExecDestroy(B.graph); B.graph = nullptr; }
    B.planned = true;
    return Status{};
}

// Pass level l of the LDS-tiled circuits and of the one-wave ones (disjoint
// circuits), then the expectation passes and the fold.
template <typename R, int K> static Status launch_levels(Batch &B, hipStream_t st) {
    using T = typename V2<R>::T;
    const BigTask *big = (const BigTask *)B.d_big.p;
    const WaveTask *wt = (const WaveTask *)B.d_wtasks.p;
    for (size_t l = 0; l < std::max(B.levels.size(), B.wlevels.size()); l++) {
        if (l < B.levels.size()) {
            auto r = B.levels[l];
            hipLaunchKernelGGL((k_batch_pass<R, K>), dim3(B.level_blocks[l]), dim3(kPT), 0, st, (T *)B.d_arena.p,
                               big + r.first, (int)(r.second - r.first), (const TileOp *)B.d_ops.p,
                               (const double *)B.d_coef.p);
        }
        if (l < B.wlevels.size()) {
            auto r = B.wlevels[l];
            Status e = launch_batch_wave(B.prec, l == 0, B.wlevel_blocks[l], st, B.d_arena.p, wt + r.first,
                                         (int)(r.second - r.first), (const RegPass *)B.d_wpasses.p,
                                         (const RegOp *)B.d_wops.p, B.d_ptab.p);
            if (!e.ok()) return e;
        }
    }
    for (size_t l = 0; l < B.elevels.size(); l++) {
        auto r = B.elevels[l];
        hipLaunchKernelGGL((k_batch_expect<R, K>), dim3(B.elevel_blocks[l]), dim3(kBT), 0, st, (const T *)B.d_arena.p,
                           big + r.first, (int)(r.second - r.first), (const ETerm *)B.d_terms.p, (double *)B.d_part.p);
    }
    if (B.nfold)
        hipLaunchKernelGGL(k_batch_fold, dim3((unsigned)B.nfold), dim3(256), 0, st, (const FoldTask *)B.d_fold.p,
                           (const double *)B.d_part.p, (double *)B.d_energy.p);
    return Status{};
}

static Status batch_launch(Batch &B, hipStream_t st) {
    if (!B.angles.empty()) {
        const unsigned g = (unsigned)((B.angles.size() + 255) / 256);
        const uint32_t na = (uint32_t)B.angles.size();
        hipLaunchKernelGGL(k_batch_resolve, dim3(g), dim3(256), 0, st, (const double *)B.d_angles.p, na,
                           (double *)B.d_coef.p);
        if (!B.wlevels.empty()) {
            if (B.prec == QI_C128)
                hipLaunchKernelGGL(k_batch_ptab<double>, dim3(g), dim3(256), 0, st, (const double *)B.d_coef.p, na,
                                   (double *)B.d_ptab.p);
            else
                hipLaunchKernelGGL(k_batch_ptab<float>, dim3(g), dim3(256), 0, st, (const double *)B.d_coef.p, na,
                                   (float *)B.d_ptab.p);
        }
    }
    // fork: LDS-resident circuits on the side stream, tiled levels on `st`
    QI_HIP_TRY(hipEventRecord(B.fork, st));
    QI_HIP_TRY(hipStreamWaitEvent(B.side, B.fork, 0));
    if (B.nsmall) {
        if (B.prec == QI_C128)
            hipLaunchKernelGGL((k_batch_small<double, kSmallK128>), dim3((unsigned)B.nsmall), dim3(kBT), 0, B.side,
                               (const SmallTask *)B.d_small.p, (const TileOp *)B.d_ops.p, (const ETerm *)B.d_terms.p,
                               (const double *)B.d_coef.p, (double *)B.d_energy.p);
        else
            hipLaunchKernelGGL((k_batch_small<float, kSmallK64>), dim3((unsigned)B.nsmall), dim3(kBT), 0, B.side,
                               (const SmallTask *)B.d_small.p, (const TileOp *)B.d_ops.p, (const ETerm *)B.d_terms.p,
                               (const double *)B.d_coef.p, (double *)B.d_energy.p);
    }
    Status le;
    if (B.prec == QI_C128) {
        switch (B.K) {
        case 10: le = launch_levels<double, 10>(B, st); break;
        case 11: le = launch_levels<double, 11>(B, st); break;
        default: le = launch_levels<double, 12>(B, st); break;
        }
    } else {
        switch (B.K) {
        case 10: le = launch_levels<float, 10>(B, st); break;
        case 11: le = launch_levels<float, 11>(B, st); break;
        case 12: le = launch_levels<float, 12>(B, st); break;
        default: le = launch_levels<float, 13>(B, st); break;
        }
    }
    if (!le.ok()) return le;
    QI_HIP_TRY(hipEventRecord(B.join, B.side));
    QI_HIP_TRY(hipStreamWaitEvent(st, B.join, 0));
    QI_HIP_TRY(hipGetLastError());
    return Status{};
}

Status batch_run(Batch &B, double *energies, bool use_graph) {
    Ctx &C = *B.ctx;
    hipStream_t st = C.stream;
    if (!B.planned) { Status e = batch_plan(B); if (!e.ok()) return e; }
    QI_HIP_TRY(hipEventRecord(C.ev0, st));
    if (use_graph) {
        if (!B.graph) {
            hipGraph_t g = nullptr;
            QI_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            Status e = batch_launch(B, st);
            hipError_t ce = hipStreamEndCapture(st, &g);
            if (!e.ok()) { if (g) (void)hipGraphDestroy(g); return e; }
            if (ce != hipSuccess) return make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(ce));
            hipError_t ie = hipGraphInstantiate(&B.graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) { B.graph = nullptr; return make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(ie)); }
        }
        QI_HIP_TRY(hipGraphLaunch(B.graph, st));
    } else {
        Status e = batch_launch(B, st);
        if (!e.ok()) return e;
    }
    QI_HIP_TRY(hipEventRecord(C.ev1, st));
    QI_HIP_TRY(hipMemcpyAsync(energies, B.d_energy.p, B.circs.size() * 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    QI_HIP_TRY(hipStreamSynchronize(st));
    float ms = 0.f;
    QI_HIP_TRY(hipEventElapsedTime(&ms, C.ev0, C.ev1));
    B.last_ms = ms;
    C.last_exec_ms = ms;
    return Status{};
}

}  // namespace qi

using namespace qi;

struct qi_batch_s {
    qi::Batch b;
};

#define QB_RET(st)                 \
    do {                           \
        Status s_ = (st);          \
        fill_err(err, s_);         \
        return (qi_status)s_.code; \
    } while (0)

extern "C" {

qi_status qi_batch_create(qi_ctx_t ctx, qi_prec prec, qi_batch_t *out, qi_error *err) {
    if (!ctx || !out) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    if (prec != QI_C64 && prec != QI_C128) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "precision"));
    auto *b = new qi_batch_s();
    b->b.ctx = &ctx->c;
    b->b.prec = prec;
    *out = b;
    QB_RET(Status{});
}

qi_status qi_batch_destroy(qi_batch_t b) {
    if (!b) return QI_OK;
    {
        std::lock_guard<std::mutex> lk(b->b.ctx->mu);
        (void)hipStreamSynchronize(b->b.ctx->stream);
        delete b;
    }
    return QI_OK;
}

qi_status qi_batch_add(qi_batch_t b, uint32_t n, const qi_gate *gates, size_t ngates, const int32_t *angle_slots,
                       const qi_pauli_term *terms, size_t nterms, qi_error *err) {
    if (!b || (ngates && !gates) || (nterms && !terms)) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    if (n == 0 || n > 34) QB_RET(make_err(QI_ERR_INVALID_NUMBER_OF_QUBITS, n));
    Batch &B = b->b;
    const int K = B.prec == QI_C128 ? 12 : 13, L = B.prec == QI_C128 ? 3 : 4;
    // slots relative to this circuit's first angle
    int32_t base = INT32_MAX;
    for (size_t g = 0; g < ngates; g++)
        if (angle_slots && angle_slots[g] >= 0) base = std::min(base, angle_slots[g]);
    if (base == INT32_MAX) base = 0;
    std::vector<int32_t> rel(ngates, -1);
    for (size_t g = 0; g < ngates; g++)
        if (angle_slots && angle_slots[g] >= 0) rel[g] = angle_slots[g] - base;
    for (size_t t = 0; t < nterms; t++) {
        uint64_t bad = (terms[t].x_mask | terms[t].z_mask) >> n;
        if (bad) QB_RET(make_err(QI_ERR_INVALID_QUBIT_INDEX, n + (uint64_t)__builtin_ctzll(bad), n));
        // a term's X/Y support must fit one tile next to the low qubits
        if ((int)n > K && __builtin_popcountll(terms[t].x_mask | ((uint64_t(1) << L) - 1)) > K)
            QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "Pauli term too wide for batched evaluation"));
    }
    std::lock_guard<std::mutex> lk(B.ctx->mu);
    // circuits of the same shape (gates and relative angle slots) share one program
    uint64_t h = 1469598103934665603ull ^ n;
    auto mix = [&h](const void *p, size_t len) {
        const unsigned char *c = (const unsigned char *)p;
        for (size_t i = 0; i < len; i++) { h ^= c[i]; h *= 1099511628211ull; }
    };
    if (ngates) mix(gates, ngates * sizeof(qi_gate));
    if (ngates) mix(rel.data(), ngates * sizeof(int32_t));
    uint32_t prog = (uint32_t)B.progs.size();
    for (auto it = B.prog_index.equal_range(h); it.first != it.second; ++it.first) {
        const auto &P = B.progs[it.first->second];
        if (P.n == n && P.gates.size() == ngates && P.rel_slot == rel &&
            (ngates == 0 || memcmp(P.gates.data(), gates, ngates * sizeof(qi_gate)) == 0)) {
            prog = it.first->second;
            break;
        }
    }
    if (prog == B.progs.size()) {
        Batch::Program P;
        P.n = n;
        for (size_t g = 0; g < ngates; g++) {
            if (gates[g].kind < QI_GATE_H || gates[g].kind > QI_GATE_TOFFOLI)
                QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "batched circuits take operator gates only"));
            Status v = validate_gate(n, gates[g]);
            if (!v.ok()) QB_RET(v);
            if (rel[g] >= 0 && gates[g].kind != QI_GATE_RX && gates[g].kind != QI_GATE_RY && gates[g].kind != QI_GATE_RZ &&
                gates[g].kind != QI_GATE_P)
                QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "angle slots apply to RX/RY/RZ/P only"));
        }
        P.gates.assign(gates, gates + ngates);
        P.rel_slot = rel;
        B.progs.push_back(std::move(P));
        B.prog_index.emplace(h, prog);
    }
    Batch::Circ C;
    C.prog = prog;
    C.slot_base = base;
    C.terms.assign(terms, terms + nterms);
    B.circs.push_back(std::move(C));
    B.planned = false;
    QB_RET(Status{});
}

qi_status qi_batch_set_angles(qi_batch_t b, const double *angles, size_t nangles, qi_error *err) {
    if (!b || (nangles && !angles)) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    Batch &B = b->b;
    std::lock_guard<std::mutex> lk(B.ctx->mu);
    const size_t before = B.angles.size();
    B.angles.assign(angles, angles + nangles);
    if (B.planned) {
        // the graph's slot kernel covers the planned table size
        if (nangles != before || B.d_angles.bytes < nangles * sizeof(double)) { B.planned = false; QB_RET(Status{}); }
        hipError_t e = hipMemcpyAsync(B.d_angles.p, B.angles.data(), nangles * sizeof(double), hipMemcpyHostToDevice,
                                      B.ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(B.ctx->stream);
        if (e != hipSuccess) QB_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    }
    QB_RET(Status{});
}

// Rewrites the coefficients of every circuit's terms (re, im pairs, in the
// order the terms were added) without replanning.
qi_status qi_batch_set_term_coeffs(qi_batch_t b, const double *coeffs, size_t nterms_total, qi_error *err) {
    if (!b || (nterms_total && !coeffs)) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    Batch &B = b->b;
    std::lock_guard<std::mutex> lk(B.ctx->mu);
    size_t total = 0;
    for (auto &C : B.circs) total += C.terms.size();
    if (total != nterms_total) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "term count differs from the batch"));
    size_t k = 0;
    for (auto &C : B.circs)
        for (size_t t = 0; t < C.terms.size(); t++, k++) {
            C.terms[t].c_re = coeffs[2 * k];
            C.terms[t].c_im = coeffs[2 * k + 1];
            if (B.planned) {
                ETerm &e = B.eterms[C.eterm_idx[t]];
                set_eterm_coeff(e, coeffs[2 * k], coeffs[2 * k + 1]);
            }
        }
    if (B.planned) {
        hipError_t e = hipMemcpy(B.d_terms.p, B.eterms.data(), B.eterms.size() * sizeof(ETerm), hipMemcpyHostToDevice);
        if (e != hipSuccess) QB_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    }
    QB_RET(Status{});
}

qi_status qi_batch_run(qi_batch_t b, double *energies, int use_graph, qi_error *err) {
    if (!b || !energies) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    Batch &B = b->b;
    std::lock_guard<std::mutex> lk(B.ctx->mu);
    // every angle slot must exist in the table
    for (auto &C : B.circs) {
        int32_t top = -1;
        for (int32_t r : B.progs[C.prog].rel_slot) top = std::max(top, r);
        if (top >= 0 && size_t(C.slot_base) + size_t(top) >= B.angles.size())
            QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "angle slot beyond the angle table"));
    }
    QB_RET(batch_run(B, energies, use_graph != 0));
}

// Sets the schedule: tile width K of the tiled circuits (10..13 for complex64,
// 10..12 for complex128) and the largest circuit run LDS-resident (K <= ks <= 13 / 12).
qi_status qi_batch_set_schedule(qi_batch_t b, uint32_t k, uint32_t ks, qi_error *err) {
    if (!b) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    Batch &B = b->b;
    const uint32_t kmax = B.prec == QI_C128 ? 12 : 13;
    if (k < 10 || k > kmax || ks < k || ks > kmax) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "schedule out of range"));
    std::lock_guard<std::mutex> lk(B.ctx->mu);
    B.K = (int)k;
    B.KS = (int)ks;
    B.planned = false;
    QB_RET(Status{});
}

qi_status qi_batch_get_schedule(qi_batch_t b, uint32_t *k, uint32_t *ks, qi_error *err) {
    if (!b) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    Batch &B = b->b;
    const int kd = B.prec == QI_C128 ? 12 : 13;
    if (k) *k = (uint32_t)(B.K ? B.K : kd);
    if (ks) *ks = (uint32_t)(B.KS ? B.KS : (B.K ? B.K : kd));
    QB_RET(Status{});
}

qi_status qi_batch_set_executor(qi_batch_t b, qi_batch_exec exec, qi_error *err) {
    if (!b) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    if (exec != QI_BATCH_LDS && exec != QI_BATCH_WAVE) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "executor"));
    Batch &B = b->b;
    std::lock_guard<std::mutex> lk(B.ctx->mu);
    B.exec = (int)exec;
    B.planned = false;
    QB_RET(Status{});
}

qi_status qi_batch_get_executor(qi_batch_t b, qi_batch_exec *exec, qi_error *err) {
    if (!b || !exec) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    *exec = (qi_batch_exec)b->b.exec;
    QB_RET(Status{});
}

// Autotuner: plans, captures and times every schedule (executor, K, ks) on this batch
// (best of `trials` graph replays after one warm-up run each, device time
// from HIP events) and keeps the fastest; schedules the batch cannot use
// (a Pauli term wider than the tile) are skipped. The batch is left planned
// with the winner.
qi_status qi_batch_autotune(qi_batch_t b, uint32_t trials, uint32_t *k_out, uint32_t *ks_out, double *best_ms,
                            qi_error *err) {
    if (!b) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    Batch &B = b->b;
    std::lock_guard<std::mutex> lk(B.ctx->mu);
    if (B.circs.empty()) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "empty batch"));
    for (auto &C : B.circs) {
        int32_t top = -1;
        for (int32_t r : B.progs[C.prog].rel_slot) top = std::max(top, r);
        if (top >= 0 && size_t(C.slot_base) + size_t(top) >= B.angles.size())
            QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "angle slot beyond the angle table"));
    }
    const int kmax = B.prec == QI_C128 ? 12 : 13;
    std::vector<double> en(2 * B.circs.size());
    double best = 1e300;
    int bk = 0, bks = 0, bex = QI_BATCH_WAVE;
    for (int ex : {QI_BATCH_WAVE, QI_BATCH_LDS})
        for (int k = 10; k <= kmax; k++)
            for (int ks = k; ks <= kmax; ks++) {
                B.exec = ex;
                B.K = k;
                B.KS = ks;
                B.planned = false;
                Status e = batch_run(B, en.data(), true);  // plan, capture, warm up
                if (e.code == QI_ERR_INVALID_ARGUMENT) continue;
                if (!e.ok()) QB_RET(e);
                double t = 1e300;
                for (uint32_t r = 0; r < std::max<uint32_t>(trials, 1); r++) {
                    if (!(e = batch_run(B, en.data(), true)).ok()) QB_RET(e);
                    t = std::min(t, B.last_ms);
                }
                if (t < best) { best = t; bk = k; bks = ks; bex = ex; }
            }
    if (!bk) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "no schedule fits the batch"));
    B.exec = bex;
    B.K = bk;
    B.KS = bks;
    B.planned = false;
    Status e = batch_run(B, en.data(), true);
    if (!e.ok()) QB_RET(e);
    if (k_out) *k_out = (uint32_t)bk;
    if (ks_out) *ks_out = (uint32_t)bks;
    if (best_ms) *best_ms = best;
    QB_RET(Status{});
}

qi_status qi_batch_info(qi_batch_t b, uint64_t *ncircuits, uint64_t *nsmall, uint64_t *nlevels, double *last_ms,
                        qi_error *err) {
    if (!b) QB_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    Batch &B = b->b;
    if (ncircuits) *ncircuits = B.circs.size();
    if (nsmall) *nsmall = B.nsmall;
    if (nlevels) *nlevels = std::max(B.levels.size(), B.wlevels.size()) + B.elevels.size();
    if (last_ms) *last_ms = B.last_ms;
    QB_RET(Status{});
}

}  // extern "C"
