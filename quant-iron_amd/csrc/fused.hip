// fused.hip — gate fusion into dense 2^k x 2^k blocks (k = 5 or 6) applied as
// complex GEMMs on the matrix cores.
//
// Consecutive gates whose qubits fit in a k-qubit set B are multiplied on the
// host (FP64) into one unitary U_B. Applying U_B to the state is the GEMM
// W = U_B V, where column c of V holds the 2^k amplitudes that share all bits
// outside B. A workgroup takes a tile of 2^(k + kc) amplitudes (B plus the kc
// lowest qubits outside B, so HBM rows stay contiguous), stores it in LDS as
// V[row][col], and multiplies with the real embedding
//     [Re U  -Im U] [Re V]   [Re W]
//     [Im U   Re U] [Im V] = [Im W]
// using v_mfma_f32_16x16x4f32 (complex64 states, FP32 accumulation) or
// v_mfma_f64_16x16x4f64 (complex128). Each wave keeps its rows of the
// embedded matrix in registers for the whole (persistent) launch.
#include <complex>

#include "qi_internal.h"
#include "exec_plan.h"

namespace qi {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename R> struct MF;
template <> struct MF<float> {
    using T = float2;
    using V4 = floatx4;
    static __device__ __forceinline__ V4 mfma(float a, float b, V4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <> struct MF<double> {
    using T = double2;
    using V4 = doublex4;
    static __device__ __forceinline__ V4 mfma(double a, double b, V4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

__device__ __forceinline__ uint64_t insert_zeros64(uint64_t k, uint64_t fixed) {
    while (fixed) {
        unsigned p = __builtin_ctzll(fixed);
        k = ((k >> p) << (p + 1)) | (k & ((uint64_t(1) << p) - 1));
        fixed &= fixed - 1;
    }
    return k;
}

// Split enumeration index bits [lo, hi) of e into its HBM offset, block row and column.
template <int KT>
__device__ __forceinline__ void split_bits(uint32_t e, int lo, int hi, const FusedBlock &b, uint64_t &off, uint32_t &row,
                                           uint32_t &col) {
    off = 0; row = 0; col = 0;
    int jr = 0, cr = 0;
#pragma unroll
    for (int t = 0; t < KT; t++) {
        uint32_t bit = (e >> t) & 1;
        bool inb = b.eisb[t] != 0;
        if (t >= lo && t < hi) {
            off |= uint64_t(bit) << b.epos[t];
            if (inb) row |= bit << jr; else col |= bit << cr;
        }
        if (inb) jr++; else cr++;
    }
}

// One wave per 16 rows of the embedded matrix: 4 waves for k = 5, 8 for k = 6.
template <int KB> constexpr int fused_threads() { return 64 * (2 << KB) / 16; }

template <typename R, int KB, int KC>
__global__ __launch_bounds__(fused_threads<KB>()) void k_fused_block(typename MF<R>::T *__restrict__ psi, FusedBlock blk,
                                                                    const R *__restrict__ A, uint64_t ntiles) {
    using T = typename MF<R>::T;
    using V4 = typename MF<R>::V4;
    constexpr int D = 1 << KB, M2 = 2 * D, NC = 1 << KC, KT = KB + KC, NE = 1 << KT;
    constexpr int NT = fused_threads<KB>();
    constexpr int TB = KB == 5 ? 8 : 9;  // log2(NT)
    constexpr int CB = NC / 16;          // column blocks
    constexpr int KS = M2 / 4;           // k steps
    constexpr int PER = NE / NT;
    static_assert((1 << TB) == NT, "thread count");
    __shared__ T v[NE];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // this wave's 16 rows of the embedded matrix, resident for the launch
    R areg[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ks++) areg[ks] = A[(wave * 16 + (lane & 15)) * M2 + ks * 4 + (lane >> 4)];
    uint64_t off_t; uint32_t row_t, col_t;
    split_bits<KT>(tid, 0, TB, blk, off_t, row_t, col_t);
    const uint64_t fixed = blk.bmask | blk.cmask;
    for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const uint64_t base = insert_zeros64(tile, fixed);
#pragma unroll
        for (int it = 0; it < PER; it++) {
            uint64_t off_i; uint32_t row_i, col_i;
            split_bits<KT>(uint32_t(it) << TB, TB, KT, blk, off_i, row_i, col_i);
            v[(row_t | row_i) * NC + (col_t | col_i)] = psi[base | off_t | off_i];
        }
        __syncthreads();
        V4 acc[CB];
#pragma unroll
        for (int cb = 0; cb < CB; cb++) acc[cb] = V4{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const int s = ks * 4 + (lane >> 4);  // real row of V
            const int j = s & (D - 1);
            const bool im = s >= D;
#pragma unroll
            for (int cb = 0; cb < CB; cb++) {
                T x = v[j * NC + cb * 16 + (lane & 15)];
                acc[cb] = MF<R>::mfma(areg[ks], im ? x.y : x.x, acc[cb]);
            }
        }
        __syncthreads();
        // accumulator layout: lane holds column lane % 16; its 4 registers are
        // rows 4*(lane/16) + r for FP32, but rows (lane/16) + 4*r for FP64
#pragma unroll
        for (int cb = 0; cb < CB; cb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int row = wave * 16 + (sizeof(R) == 8 ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r);
                const int col = cb * 16 + (lane & 15);
                R *w = reinterpret_cast<R *>(&v[(row & (D - 1)) * NC + col]);
                w[row >= D ? 1 : 0] = acc[cb][r];
            }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < PER; it++) {
            uint64_t off_i; uint32_t row_i, col_i;
            split_bits<KT>(uint32_t(it) << TB, TB, KT, blk, off_i, row_i, col_i);
            psi[base | off_t | off_i] = v[(row_t | row_i) * NC + (col_t | col_i)];
        }
        __syncthreads();
    }
}

using cd = std::complex<double>;

// Apply one lowered gate to a 2^k vector whose bit r is qubit bq[r].
void apply_local(cd *v, int k, const Lowered &L, const int *loc) {
    const uint32_t dim = 1u << k;
    uint32_t cm = 0;
    for (int q = 0; q < 64; q++)
        if ((L.ctrl_mask >> q) & 1) cm |= 1u << loc[q];
    const uint32_t tb = 1u << loc[L.t0];
    auto C = [&](int i) { return cd(L.c[2 * i], L.c[2 * i + 1]); };
    switch (L.kind) {
    case LK_Z: case LK_PHASE: case LK_RZ:
        for (uint32_t i = 0; i < dim; i++) {
            if ((i & cm) != cm) continue;
            bool one = (i & tb) != 0;
            if (L.kind == LK_Z) { if (one) v[i] = -v[i]; }
            else if (L.kind == LK_PHASE) { if (one) v[i] *= cd(L.c[0], L.c[1]); }
            else v[i] *= one ? cd(L.c[2], L.c[3]) : cd(L.c[0], L.c[1]);
        }
        break;
    case LK_SWAP: {
        const uint32_t tb2 = 1u << loc[L.t1];
        for (uint32_t i = 0; i < dim; i++)
            if ((i & tb) && !(i & tb2) && (i & cm) == cm) std::swap(v[i], v[i ^ tb ^ tb2]);
    } break;
    case LK_MATCH: {
        const uint32_t tb2 = 1u << loc[L.t1];
        for (uint32_t i = 0; i < dim; i++) {
            if (i & (tb | tb2)) continue;
            uint32_t i01 = i | tb, i10 = i | tb2, i11 = i | tb | tb2;
            if ((i01 & cm) == cm) {
                cd a = v[i01], b = v[i10];
                v[i01] = L.c[0] * a - C(1) * b;
                v[i10] = L.c[1] * a + C(2) * b;
            }
            if ((i11 & cm) == cm) v[i11] *= C(3);
        }
    } break;
    default:
        for (uint32_t i = 0; i < dim; i++) {
            if ((i & tb) || (i & cm) != cm) continue;
            cd a = v[i], b = v[i | tb], na, nb;
            switch (L.kind) {
            case LK_H: na = L.c[0] * (a + b); nb = L.c[0] * (a - b); break;
            case LK_X: na = b; nb = a; break;
            case LK_Y: na = C(0) * b; nb = C(1) * a; break;
            case LK_RX: na = L.c[0] * a - cd(L.c[1], L.c[2]) * b; nb = cd(L.c[3], L.c[4]) * a + L.c[0] * b; break;
            case LK_RY: na = L.c[0] * a - L.c[1] * b; nb = L.c[1] * a + L.c[0] * b; break;
            default: na = C(0) * a + C(1) * b; nb = C(2) * a + C(3) * b; break;
            }
            v[i] = na; v[i | tb] = nb;
        }
        break;
    }
}

template <typename R, int KB, int KC> Status launch_fb(State &s, const FusedBlock &b, const void *dmat, hipStream_t st) {
    using T = typename MF<R>::T;
    uint64_t ntiles = uint64_t(1) << (s.n - KB - KC);
    int cus = s.ctx->cu_count > 0 ? s.ctx->cu_count : 256;
    uint64_t grid = std::min<uint64_t>(ntiles, uint64_t(cus) * 4);
    hipLaunchKernelGGL((k_fused_block<R, KB, KC>), dim3((unsigned)grid), dim3(fused_threads<KB>()), 0, st, (T *)s.d, b,
                       (const R *)dmat, ntiles);
    QI_HIP_TRY(hipGetLastError());
    return Status{};
}

}  // namespace

constexpr int kFusedCols = 6;

Status plan_fused(const State &s, const std::vector<Lowered> &gates, int kb, ExecPlan &plan) {
    if (kb <= 0) kb = 6;
    if (kb != 5 && kb != 6) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "dense block width must be 5 or 6");
    if ((int)s.n < kb + kFusedCols) {  // state smaller than one tile: stream gate by gate
        for (const Lowered &L : gates) plan.steps.push_back(ExecStep::gate(L));
        return Status{};
    }
    const uint64_t all = (uint64_t(1) << s.n) - 1;
    std::vector<char> done(gates.size(), 0);
    size_t first = 0;
    const size_t esz = s.prec == QI_C128 ? sizeof(double) : sizeof(float);
    while (first < gates.size()) {
        uint64_t active = 0, blocked = 0;
        std::vector<size_t> members;
        for (size_t g = first; g < gates.size(); g++) {
            if (done[g]) continue;
            const Lowered &L = gates[g];
            uint64_t supp = (uint64_t(1) << L.t0) | L.ctrl_mask;
            if (L.kind == LK_SWAP || L.kind == LK_MATCH) supp |= uint64_t(1) << L.t1;
            if (supp & blocked) blocked |= supp;
            else if (__builtin_popcountll(active | supp) <= kb) { active |= supp; members.push_back(g); done[g] = 1; }
            else blocked |= supp;
            if ((blocked & all) == all) break;
        }
        uint64_t bm = active;
        for (int q = 0; __builtin_popcountll(bm) < kb && q < (int)s.n; q++) bm |= uint64_t(1) << q;
        uint64_t cmk = 0;
        for (int q = 0, c = 0; c < kFusedCols && q < (int)s.n; q++)
            if (!((bm >> q) & 1)) { cmk |= uint64_t(1) << q; c++; }
        FusedBlock fb{};
        fb.bmask = bm; fb.cmask = cmk; fb.kb = kb; fb.kc = kFusedCols; fb.ngates = (int)members.size();
        int t = 0;
        for (int q = 0; q < (int)s.n; q++)
            if (((bm | cmk) >> q) & 1) { fb.epos[t] = (uint8_t)q; fb.eisb[t] = (bm >> q) & 1; t++; }
        fb.order = (bm & 1) ? 1 : 0;
        // U = G_m ... G_1 over the block qubits (local bit r = r-th lowest block qubit)
        int loc[64];
        for (int q = 0, r = 0; q < 64; q++) loc[q] = ((bm >> q) & 1) ? r++ : 0;
        const int D = 1 << kb, M2 = 2 * D;
        std::vector<cd> U(size_t(D) * D);  // column-major: column c is U e_c
        for (int c = 0; c < D; c++) {
            cd *col = &U[size_t(c) * D];
            for (int r = 0; r < D; r++) col[r] = (r == c) ? 1.0 : 0.0;
            for (size_t g : members) apply_local(col, kb, gates[g], loc);
        }
        // real embedding, row-major M2 x M2
        size_t off = plan.mats.size();
        fb.mat_off = (int64_t)off;
        plan.mats.resize(off + size_t(M2) * M2 * esz);
        for (int r = 0; r < M2; r++)
            for (int c = 0; c < M2; c++) {
                cd u = U[size_t(c % D) * D + (r % D)];
                double val = (r < D) == (c < D) ? u.real() : (r < D ? -u.imag() : u.imag());
                unsigned char *dst = &plan.mats[off + (size_t(r) * M2 + c) * esz];
                if (esz == sizeof(double)) memcpy(dst, &val, 8);
                else { float f = (float)val; memcpy(dst, &f, 4); }
            }
        plan.steps.push_back(ExecStep::fused_step(fb));
        while (first < gates.size() && done[first]) first++;
    }
    return Status{};
}

Status launch_fused_block(State &s, const FusedBlock &b, const void *dmat, hipStream_t st) {
    if (b.kc != kFusedCols) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "fused column width");
    if (s.prec == QI_C128) {
        if (b.kb == 5) return launch_fb<double, 5, kFusedCols>(s, b, dmat, st);
        return launch_fb<double, 6, kFusedCols>(s, b, dmat, st);
    }
    if (b.kb == 5) return launch_fb<float, 5, kFusedCols>(s, b, dmat, st);
    return launch_fb<float, 6, kFusedCols>(s, b, dmat, st);
}

}  // namespace qi
