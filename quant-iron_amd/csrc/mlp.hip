// mlp.hip — the energy surrogate: an MLP (n_in -> H -> H -> H -> 1, GELU)
// trained and evaluated with hand-written kernels. Weights are BF16 for the
// matrix cores (gfx950 v_mfma_f32_16x16x32_bf16, FP32 accumulation) with FP32
// master copies for Adam; activations are stored BF16 (GEMM inputs) and FP32
// (pre-activations, needed by the GELU derivative).
//
// GEMM: a 128 x 128 output tile per workgroup, four waves of 64 x 64 (4 x 4
// MFMA tiles of 16 x 16 each), K in stages of 64 held in LDS with K
// contiguous per row (padded rows), double-buffered: the next stage's global
// loads are in registers while the current stage feeds the matrix cores. The
// instruction's operands: A row = lane % 16, k = 8 (lane / 16) + v (8 BF16 per
// lane, one 16-byte LDS read); B col = lane % 16, same k; D row = 4 (lane / 16)
// + r, col = lane % 16 (the GEMM feeds B first, so D is the tile's transpose).
#include <algorithm>
#include <cmath>
#include <memory>
#include <random>
#include <vector>

#include "qi_internal.h"

namespace qi {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint16_t bf16;

__device__ __host__ __forceinline__ bf16 f2bf(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (bf16)(u >> 16);
}
// Four FP32 values as BF16 (round to nearest even: the same values as f2bf
// for finite inputs) in one 8-byte word, by the gfx950 conversion instruction
// (v_cvt_pk_bf16_f32, two values per instruction).
__device__ __forceinline__ uint2 pack_bf16x4(const float *v) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    bf16x2 a, b;
    a.x = (__bf16)v[0]; a.y = (__bf16)v[1];
    b.x = (__bf16)v[2]; b.y = (__bf16)v[3];
    return make_uint2(__builtin_bit_cast(uint32_t, a), __builtin_bit_cast(uint32_t, b));
}
__device__ __host__ __forceinline__ float bf2f(bf16 h) {
    uint32_t u = uint32_t(h) << 16;
    float x;
    memcpy(&x, &u, 4);
    return x;
}

// erf(x) by Abramowitz & Stegun 7.1.26 (absolute error below 1.5e-7), for
// the GELU derivative of the backward pass: 1 - t (a1 + t (a2 + ...)) e^(-x^2),
// t = 1 / (1 + p |x|), odd in x. Returns erf(x) and e^(-x^2) (the derivative
// reuses the exponential).
__device__ __forceinline__ float erf_as(float x, float &ex2) {
    const float ax = fabsf(x);
    const float t = __frcp_rn(1.0f + 0.3275911f * ax);
    float p = 1.061405429f;
    p = fmaf(p, t, -1.453152027f);
    p = fmaf(p, t, 1.421413741f);
    p = fmaf(p, t, -0.284496736f);
    p = fmaf(p, t, 0.254829592f);
    ex2 = __expf(-ax * ax);
    const float e = 1.0f - p * t * ex2;
    return copysignf(e, x);
}
// GELU(z) = 0.5 z (1 + erf(z / sqrt 2)); gelu'(z) = 0.5 (1 + erf(z / sqrt 2)) + z phi(z).
// GELU in the forward epilogues (activations stored in BF16, whose spacing
// near 1 is 8e-3): erf(z / sqrt 2) as an odd polynomial of degree 15 in z on
// |z| <= 3 sqrt 2 (least-squares fit, 1 at the ends) and +-1 outside; GELU's
// absolute error is below 2e-4 for every z, with no transcendental
// instructions (the A&S form above needs a reciprocal and an exponential).
__device__ __forceinline__ float gelu_fwd(float z) {
    const float zc = fminf(fmaxf(z, -4.2426405f), 4.2426405f), z2 = zc * zc;
    float p = -2.155357003985614e-09f;
    p = fmaf(p, z2, 1.8428821135785256e-07f);
    p = fmaf(p, z2, -6.820219823566731e-06f);
    p = fmaf(p, z2, 0.00014502427075058222f);
    p = fmaf(p, z2, -0.0019942785147577524f);
    p = fmaf(p, z2, 0.019005531445145607f);
    p = fmaf(p, z2, -0.13190430402755737f);
    p = fmaf(p, z2, 0.7975210547447205f);
    const float e = fminf(fmaxf(zc * p, -1.0f), 1.0f), h = 0.5f * z;
    return fmaf(h, e, h);
}
__device__ __forceinline__ float gelu_grad(float z) {
    float ex2;  // e^(-z^2 / 2)
    const float e = erf_as(z * 0.70710678118654752f, ex2);
    return 0.5f * (1.0f + e) + z * 0.39894228040143268f * ex2;
}

// EPI_GELU_HEAD (prediction only): the last hidden layer's GELU output is not
// stored; its dot product with the head weights (passed as Zin) over this
// wave's columns goes to slot WN bx + wc of the row's partials (Cf, WN *
// gridDim.x per row), summed with the head bias by k_head_fold.
enum Epi { EPI_GELU = 0, EPI_STORE = 1, EPI_GELU_BWD = 2, EPI_GELU_HEAD = 3 };

constexpr int TM = 128, TN = 128, TK = 64, LDK = TK + 8;  // LDS row: 72 BF16 = 144 bytes (9 x 16)
constexpr int kMaxSplits = 16;  // K slices of a weight-gradient GEMM

// One operand tile (ROWS rows of the output dimension x 64 of K) as 16-byte
// chunks in registers, ROWS * 8 / NT per thread of NT. KC (K contiguous in memory: X row-major M x K, or W^T):
// chunk c = row c / 8, k = 8 (c % 8). Otherwise (K strided: the operand is
// stored K x rows): chunk c = k c / 16, rows 8 (c % 16) .. + 7, transposed
// into LDS on store. Out-of-range rows and k read as zero (K % 8 == 0).
template <bool KC, int ROWS, int NT>
__device__ __forceinline__ void load_tile(const bf16 *__restrict__ P, int ld, int rows, int K, int r0, int k0,
                                          uint4 (&reg)[ROWS * TK / 8 / NT]) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int i = 0; i < ROWS * TK / 8 / NT; i++) {
        const int c = tid + NT * i;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (KC) {
            const int r = r0 + c / 8, k = k0 + 8 * (c % 8);
            if (r < rows && k < K) v = *reinterpret_cast<const uint4 *>(P + size_t(r) * ld + k);
        } else {
            const int k = k0 + c / 16, r = r0 + 8 * (c % 16);
            if (k < K) {
                if (r + 7 < rows) {
                    v = *reinterpret_cast<const uint4 *>(P + size_t(k) * ld + r);
                } else {
                    bf16 t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                    for (int q = 0; q < 8; q++)
                        if (r + q < rows) t[q] = P[size_t(k) * ld + r + q];
                    memcpy(&v, t, 16);
                }
            }
        }
        reg[i] = v;
    }
}

template <bool KC, int ROWS, int NT>
__device__ __forceinline__ void store_tile(bf16 (*S)[LDK], const uint4 (&reg)[ROWS * TK / 8 / NT]) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int i = 0; i < ROWS * TK / 8 / NT; i++) {
        const int c = tid + NT * i;
        if (KC) {
            *reinterpret_cast<uint4 *>(&S[c / 8][8 * (c % 8)]) = reg[i];
        } else {
            bf16 t[8];
            memcpy(t, &reg[i], 16);
            const int k = c / 16, r = 8 * (c % 16);
#pragma unroll
            for (int q = 0; q < 8; q++) S[r + q][k] = t[q];
        }
    }
}

// C[M x N] = op(A)[M x K] op(B)[K x N]; op(A) = A (row-major M x K, lda) or
// A^T (A row-major K x M); op(B) = B (row-major K x N) or B^T (B row-major N x K).
// Workgroup tile BM x BN on WM x WN waves (each BM / WM x BN / WN). DEPTH 2:
// two register sets take the global loads of stages s + 2 and s + 3 (a
// 128 x 128 tile fits twice per CU); DEPTH 1: one set, stage s + 1 (the
// 256 x 256 tile of 8 waves, once per CU: 147 KiB of LDS).
template <bool TA, bool TB, int EPI, int BM = TM, int BN = TN, int WM = 2, int WN = 2, int DEPTH = 2>
__global__ __launch_bounds__(64 * WM * WN) void k_gemm(const bf16 *__restrict__ A, int lda, const bf16 *__restrict__ B,
                                                       int ldb, int M, int N, int K, float *__restrict__ Cf,
                                                       bf16 *__restrict__ Cb, int ldc, const float *__restrict__ bias,
                                                       const float *__restrict__ Zin, int kper) {
    constexpr int NT = 64 * WM * WN, FI = BM / WM / 16, FJ = BN / WN / 16, CA = BM * TK / 8 / NT, CB = BN * TK / 8 / NT;
    __shared__ __attribute__((aligned(16))) bf16 As[2][BM][LDK];
    __shared__ __attribute__((aligned(16))) bf16 Bs[2][BN][LDK];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // Workgroup w runs on XCD w % 8. The N tiles of one M tile read the same
    // A rows: they go to one XCD (consecutive r below), so A crosses HBM once
    // and the other N tiles find it in that XCD's L2.
    int bx = blockIdx.x, by = blockIdx.y;
    if ((gridDim.y & 7) == 0) {
        const uint32_t w = blockIdx.y * gridDim.x + blockIdx.x, r = w >> 3;
        bx = int(r % gridDim.x);
        by = int((r / gridDim.x) * 8 + (w & 7));
    }
    const int m0 = by * BM, n0 = bx * BN;
    // split K: slice z takes K range [kb, ke) and writes its partial product
    // to slice z of the output (EPI_STORE into an FP32 workspace)
    const int kb = blockIdx.z * kper, ke = min(K, kb + kper);
    if (gridDim.z > 1) Cf += size_t(blockIdx.z) * M * ldc;
    const int wr = wave / WN, wc = wave % WN;
    floatx4 acc[FI][FJ];
#pragma unroll
    for (int i = 0; i < FI; i++)
#pragma unroll
        for (int j = 0; j < FJ; j++) acc[i][j] = floatx4{0, 0, 0, 0};
    auto compute = [&](int cur) {
#pragma unroll
        for (int kk = 0; kk < TK; kk += 32) {
            bf16x8 a[FI], b[FJ];
#pragma unroll
            for (int i = 0; i < FI; i++)
                a[i] = *reinterpret_cast<const bf16x8 *>(&As[cur][wr * (BM / WM) + i * 16 + (lane & 15)][kk + 8 * (lane >> 4)]);
#pragma unroll
            for (int j = 0; j < FJ; j++)
                b[j] = *reinterpret_cast<const bf16x8 *>(&Bs[cur][wc * (BN / WN) + j * 16 + (lane & 15)][kk + 8 * (lane >> 4)]);
#pragma unroll
            for (int i = 0; i < FI; i++)
#pragma unroll
                for (int j = 0; j < FJ; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j], a[i], acc[i][j], 0, 0, 0);
        }
    };
    if constexpr (DEPTH == 1) {
        // stages past K load zeros
        const int nst = (ke - kb + TK - 1) / TK;
        uint4 ra[CA], rb[CB];
        load_tile<!TA, BM, NT>(A, lda, M, ke, m0, kb, ra);
        load_tile<TB, BN, NT>(B, ldb, N, ke, n0, kb, rb);
        store_tile<!TA, BM, NT>(As[0], ra);
        store_tile<TB, BN, NT>(Bs[0], rb);
        __syncthreads();
        for (int s = 0; s < nst; s++) {
            if (s + 1 < nst) {
                load_tile<!TA, BM, NT>(A, lda, M, ke, m0, kb + (s + 1) * TK, ra);
                load_tile<TB, BN, NT>(B, ldb, N, ke, n0, kb + (s + 1) * TK, rb);
            }
            compute(s & 1);
            if (s + 1 < nst) {
                store_tile<!TA, BM, NT>(As[(s + 1) & 1], ra);
                store_tile<TB, BN, NT>(Bs[(s + 1) & 1], rb);
            }
            __syncthreads();
        }
    } else {
        // Stage s's global loads are issued two stages ahead (register sets 0
        // and 1 alternate, the loop runs two stages per trip so the set is
        // static), then stored to LDS buffer s % 2 one stage ahead of its use.
        // The stage count is rounded up to even: stages past K load zeros.
        uint4 ra0[CA], rb0[CB], ra1[CA], rb1[CB];
        const int nst = ((ke - kb + 2 * TK - 1) / (2 * TK)) * 2;
        load_tile<!TA, BM, NT>(A, lda, M, ke, m0, kb, ra0);
        load_tile<TB, BN, NT>(B, ldb, N, ke, n0, kb, rb0);
        load_tile<!TA, BM, NT>(A, lda, M, ke, m0, kb + TK, ra1);
        load_tile<TB, BN, NT>(B, ldb, N, ke, n0, kb + TK, rb1);
        store_tile<!TA, BM, NT>(As[0], ra0);
        store_tile<TB, BN, NT>(Bs[0], rb0);
        __syncthreads();
        for (int s = 0; s < nst; s += 2) {
            // even stage s: LDS buffer 0; set 0 (stage s, stored) takes stage s + 2
            if (s + 2 < nst) {
                load_tile<!TA, BM, NT>(A, lda, M, ke, m0, kb + (s + 2) * TK, ra0);
                load_tile<TB, BN, NT>(B, ldb, N, ke, n0, kb + (s + 2) * TK, rb0);
            }
            compute(0);
            store_tile<!TA, BM, NT>(As[1], ra1);
            store_tile<TB, BN, NT>(Bs[1], rb1);
            __syncthreads();
            // odd stage s + 1: LDS buffer 1; set 1 takes stage s + 3
            if (s + 3 < nst) {
                load_tile<!TA, BM, NT>(A, lda, M, ke, m0, kb + (s + 3) * TK, ra1);
                load_tile<TB, BN, NT>(B, ldb, N, ke, n0, kb + (s + 3) * TK, rb1);
            }
            compute(1);
            if (s + 2 < nst) {
                store_tile<!TA, BM, NT>(As[0], ra0);
                store_tile<TB, BN, NT>(Bs[0], rb0);
            }
            __syncthreads();
        }
    }
    // acc[i][j] holds C^T of its 16 x 16 tile (B fed as the first operand):
    // this lane has row 16 i + lane % 16 and columns 16 j + 4 (lane / 16) + r,
    // four consecutive columns stored as one 16-byte FP32 / 8-byte BF16 access
    const bool vec = (ldc & 3) == 0;
    if constexpr (EPI == EPI_GELU_HEAD) {
        const int slots = WN * gridDim.x;
#pragma unroll
        for (int i = 0; i < FI; i++) {
            const int row = m0 + wr * (BM / WM) + i * 16 + (lane & 15);
            float hs = 0.f;
#pragma unroll
            for (int j = 0; j < FJ; j++) {
                const int col = n0 + wc * (BN / WN) + j * 16 + 4 * (lane >> 4);
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (col + r < N) hs += gelu_fwd(acc[i][j][r] + bias[col + r]) * Zin[col + r];
            }
            hs += __shfl_xor(hs, 16);
            hs += __shfl_xor(hs, 32);
            if (lane < 16 && row < M) Cf[size_t(row) * slots + WN * bx + wc] = hs;
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < FI; i++)
#pragma unroll
        for (int j = 0; j < FJ; j++) {
            const int row = m0 + wr * (BM / WM) + i * 16 + (lane & 15);
            const int col = n0 + wc * (BN / WN) + j * 16 + 4 * (lane >> 4);
            if (row >= M) continue;
            const size_t o = size_t(row) * ldc + col;
            float v[4] = {acc[i][j][0], acc[i][j][1], acc[i][j][2], acc[i][j][3]};
            if (vec && col + 3 < N) {
                if (EPI == EPI_GELU) {
                    const float4 bb = *reinterpret_cast<const float4 *>(bias + col);
                    v[0] += bb.x; v[1] += bb.y; v[2] += bb.z; v[3] += bb.w;
                    if (Cf) *reinterpret_cast<float4 *>(Cf + o) = make_float4(v[0], v[1], v[2], v[3]);
#pragma unroll
                    for (int r = 0; r < 4; r++) v[r] = gelu_fwd(v[r]);
                } else if (EPI == EPI_GELU_BWD) {
                    const float4 z = *reinterpret_cast<const float4 *>(Zin + o);
                    v[0] *= gelu_grad(z.x); v[1] *= gelu_grad(z.y); v[2] *= gelu_grad(z.z); v[3] *= gelu_grad(z.w);
                    *reinterpret_cast<float4 *>(Cf + o) = make_float4(v[0], v[1], v[2], v[3]);
                } else if (Cf) {
                    *reinterpret_cast<float4 *>(Cf + o) = make_float4(v[0], v[1], v[2], v[3]);
                }
                if (EPI != EPI_STORE || Cb) *reinterpret_cast<uint2 *>(Cb + o) = pack_bf16x4(v);
                continue;
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                if (col + r >= N) continue;
                float x = v[r];
                if (EPI == EPI_GELU) {
                    x += bias[col + r];
                    if (Cf) Cf[o + r] = x;
                    Cb[o + r] = f2bf(gelu_fwd(x));
                } else if (EPI == EPI_GELU_BWD) {
                    x *= gelu_grad(Zin[o + r]);
                    Cf[o + r] = x;
                    Cb[o + r] = f2bf(x);
                } else {
                    if (Cf) Cf[o + r] = x;
                    if (Cb) Cb[o + r] = f2bf(x);
                }
            }
        }
}

// y[r] = H[r,:] . w + b ; with targets: dy = 2 (y - t) / B, loss partials, and
// dZ[r,k] = dy w[k] gelu'(Z[r,k]) (FP32 and BF16).
__global__ __launch_bounds__(256) void k_head(const bf16 *__restrict__ H, const float *__restrict__ w, const float *__restrict__ bp, int rows,
                                              int hid, float *__restrict__ y, const float *__restrict__ t, float inv_b,
                                              const float *__restrict__ Z, float *__restrict__ dZf, bf16 *__restrict__ dZb,
                                              float *__restrict__ dy, float *__restrict__ loss_part) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = blockIdx.x * 4 + wave;
    __shared__ float lsum[4];
    float l2 = 0.f;
    if (r < rows) {
        float s = 0.f;
        for (int k = lane; k < hid; k += 64) s += bf2f(H[size_t(r) * hid + k]) * w[k];
        for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
        s += *bp;
        if (y) y[r] = s;
        if (t) {
            const float d = 2.f * (s - t[r]) * inv_b;
            if (lane == 0) dy[r] = d;
            l2 = (s - t[r]) * (s - t[r]);
            for (int k = lane; k < hid; k += 64) {
                const size_t o = size_t(r) * hid + k;
                const float g = d * w[k] * gelu_grad(Z[o]);
                dZf[o] = g;
                dZb[o] = f2bf(g);
            }
        }
    }
    if (loss_part) {
        if (lane == 0) lsum[wave] = l2;
        __syncthreads();
        if (threadIdx.x == 0) loss_part[blockIdx.x] = lsum[0] + lsum[1] + lsum[2] + lsum[3];
    }
}

// Column sums of X (rows x cols, FP32) into out[cols]: grid (ceil(cols/64), row chunks of 256).
__global__ __launch_bounds__(256) void k_colsum_part(const float *__restrict__ X, int rows, int cols, float *__restrict__ part) {
    __shared__ float sh[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63), g = threadIdx.x >> 6;
    const int r0 = blockIdx.y * 256;
    float s = 0.f;
    // (unrolled: the loads of eight rows are in flight together; the sum
    // keeps its row order)
    if (c < cols)
#pragma unroll 8
        for (int r = r0 + g; r < r0 + 256 && r < rows; r += 4) s += X[size_t(r) * cols + c];
    sh[g][threadIdx.x & 63] = s;
    __syncthreads();
    if (g == 0 && c < cols) part[size_t(blockIdx.y) * cols + c] = sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
}
__global__ __launch_bounds__(256) void k_colsum_fold(const float *__restrict__ part, int chunks, int cols, float *__restrict__ out) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= cols) return;
    float s = 0.f;
    for (int k = 0; k < chunks; k++) s += part[size_t(k) * cols + c];
    out[c] = s;
}

// y[r] = b + the row's head partials in slot order (EPI_GELU_HEAD).
__global__ __launch_bounds__(256) void k_head_fold(const float *__restrict__ part, int slots, int rows,
                                                   const float *__restrict__ bp, float *__restrict__ y) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= rows) return;
    float s = 0.f;
    for (int k = 0; k < slots; k++) s += part[size_t(r) * slots + k];
    y[r] = s + *bp;
}

// Partial sums of dw[k] = sum_r H[r,k] dy[r] (head weights) over chunks of 256
// rows, as k_colsum_part (folded by k_colsum_fold).
__global__ __launch_bounds__(256) void k_head_grad_part(const bf16 *__restrict__ H, const float *__restrict__ dy, int rows,
                                                        int hid, float *__restrict__ part) {
    __shared__ float sh[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63), g = threadIdx.x >> 6;
    const int r0 = blockIdx.y * 256;
    float s = 0.f;
    if (c < hid)
#pragma unroll 8
        for (int r = r0 + g; r < r0 + 256 && r < rows; r += 4) s += bf2f(H[size_t(r) * hid + c]) * dy[r];
    sh[g][threadIdx.x & 63] = s;
    __syncthreads();
    if (g == 0 && c < hid)
        part[size_t(blockIdx.y) * hid + c] = sh[0][threadIdx.x] + sh[1][threadIdx.x] + sh[2][threadIdx.x] + sh[3][threadIdx.x];
}

// Adam on FP32 masters; refreshes the BF16 copy (nullable) of each parameter.
__global__ __launch_bounds__(256) void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                                              const float *__restrict__ g, bf16 *__restrict__ wb, size_t n, float lr,
                                              float b1, float b2, float eps, float c1, float c2) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (wb) wb[i] = f2bf(wi);
    }
}

__global__ __launch_bounds__(256) void k_to_bf16(const float *__restrict__ x, bf16 *__restrict__ y, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) y[i] = f2bf(x[i]);
}

// y = BF16(x^T) for a rows x cols FP32 matrix (y is cols x rows): the weights
// in the K-contiguous layout the forward GEMMs stream.
__global__ __launch_bounds__(256) void k_transpose_bf16(const float *__restrict__ x, int rows, int cols, bf16 *__restrict__ y) {
    __shared__ float t[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8)
        if (r0 + k < rows && c0 + tx < cols) t[k][tx] = x[size_t(r0 + k) * cols + c0 + tx];
    __syncthreads();
    for (int k = ty; k < 32; k += 8)
        if (c0 + k < cols && r0 + tx < rows) y[size_t(c0 + k) * rows + r0 + tx] = f2bf(t[tx][k]);
}

// y = x^T for a BF16 matrix x (rows x cols), through a 32 x 32 LDS tile.
__global__ __launch_bounds__(256) void k_transpose_b16(const bf16 *__restrict__ x, int rows, int cols, bf16 *__restrict__ y) {
    __shared__ bf16 t[32][34];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32, tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int k = ty; k < 32; k += 8)
        if (r0 + k < rows && c0 + tx < cols) t[k][tx] = x[size_t(r0 + k) * cols + c0 + tx];
    __syncthreads();
    for (int k = ty; k < 32; k += 8)
        if (c0 + k < cols && r0 + tx < rows) y[size_t(c0 + k) * rows + r0 + tx] = t[tx][k];
}

// out[i] = sum over z of ws[z count + i], z in order (split-K partials).
__global__ __launch_bounds__(256) void k_sum_splits(const float *__restrict__ ws, int splits, size_t count, float *__restrict__ out) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += size_t(gridDim.x) * blockDim.x) {
        float s = 0.f;
        for (int z = 0; z < splits; z++) s += ws[size_t(z) * count + i];
        out[i] = s;
    }
}

// The epilogue of a layer GEMM whose K was split (few output tiles: small
// batches): the slices' FP32 partials (ws, slice z at z * M * ldc) summed in
// slice order, then EPI_GELU (bias, FP32 pre-activation if Cf, BF16 GELU) or
// EPI_GELU_BWD (times gelu'(Zin), FP32 and BF16), four columns per thread.
template <int EPI>
__global__ __launch_bounds__(256) void k_split_epilogue(const float *__restrict__ ws, int splits, int M, int N, int ldc,
                                                        float *__restrict__ Cf, bf16 *__restrict__ Cb,
                                                        const float *__restrict__ bias, const float *__restrict__ Zin) {
    const int q = N / 4;  // N % 4 == 0 (the caller checks)
    const size_t slice = size_t(M) * ldc;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < size_t(M) * q; i += size_t(gridDim.x) * blockDim.x) {
        const int row = int(i / q), col = 4 * int(i % q);
        const size_t o = size_t(row) * ldc + col;
        float4 a = *reinterpret_cast<const float4 *>(ws + o);
        for (int z = 1; z < splits; z++) {
            const float4 b = *reinterpret_cast<const float4 *>(ws + size_t(z) * slice + o);
            a.x += b.x; a.y += b.y; a.z += b.z; a.w += b.w;
        }
        float v[4] = {a.x, a.y, a.z, a.w};
        if (EPI == EPI_GELU) {
            const float4 bb = *reinterpret_cast<const float4 *>(bias + col);
            v[0] += bb.x; v[1] += bb.y; v[2] += bb.z; v[3] += bb.w;
            if (Cf) *reinterpret_cast<float4 *>(Cf + o) = make_float4(v[0], v[1], v[2], v[3]);
#pragma unroll
            for (int r = 0; r < 4; r++) v[r] = gelu_fwd(v[r]);
        } else {
            const float4 z = *reinterpret_cast<const float4 *>(Zin + o);
            v[0] *= gelu_grad(z.x); v[1] *= gelu_grad(z.y); v[2] *= gelu_grad(z.z); v[3] *= gelu_grad(z.w);
            *reinterpret_cast<float4 *>(Cf + o) = make_float4(v[0], v[1], v[2], v[3]);
        }
        *reinterpret_cast<uint2 *>(Cb + o) = pack_bf16x4(v);
    }
}

// gather rows idx[0..rows) of a BF16 matrix (cols wide) into a contiguous block
__global__ __launch_bounds__(256) void k_gather_rows(const bf16 *__restrict__ X, const uint32_t *__restrict__ idx, int rows,
                                                     int cols, bf16 *__restrict__ out, const float *__restrict__ t,
                                                     float *__restrict__ tout) {
    const int r = blockIdx.x;
    if (r >= rows) return;
    const size_t src = size_t(idx[r]) * cols;
    for (int c = threadIdx.x; c < cols; c += blockDim.x) out[size_t(r) * cols + c] = X[src + c];
    if (threadIdx.x == 0 && t) tout[r] = t[idx[r]];
}

}  // namespace

// acc += sum of the loss partials of one batch (one workgroup, fixed order)
__global__ __launch_bounds__(256) void k_loss_accum(const float *__restrict__ part, int n, double *__restrict__ acc) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *acc += sh[0];
}

// SplitMix64 draw k (1-based) of the stream seeded with `seed` (Steele, Lea &
// Flood 2014), as a uniform in [0, 1) with 53 bits.
__device__ __forceinline__ double sm64_uniform(uint64_t seed, uint64_t k) {
    uint64_t z = seed + k * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * 0x1.0p-53;
}

// cos and sin of 2 pi x (x in revolutions, reduced to [-1/2, 1/2], inside
// the range of the hardware v_cos_f32 / v_sin_f32; absolute error about
// 1e-6, far below the BF16 rounding of the network inputs).
__device__ __forceinline__ void sincos_rev(float x, float &s, float &c) {
    x -= rintf(x);
    s = __builtin_amdgcn_sinf(x);
    c = __builtin_amdgcn_cosf(x);
}
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    bf16x2 v;
    v.x = (__bf16)a;
    v.y = (__bf16)b;
    return __builtin_bit_cast(uint32_t, v);
}
// A row's features of angles j and j + 1 (cos at j, sin at nang + j; one
// 4-byte store each when both exist and nang is even, else 2-byte stores).
__device__ __forceinline__ void store_angle_pair(bf16 *row, int nang, int j, float ca, float sa, float cb, float sb,
                                                 bool two) {
    if (two && !(nang & 1)) {
        *reinterpret_cast<uint32_t *>(row + j) = pack_bf16x2(ca, cb);
        *reinterpret_cast<uint32_t *>(row + nang + j) = pack_bf16x2(sa, sb);
        return;
    }
    row[j] = f2bf(ca);
    row[nang + j] = f2bf(sa);
    if (two) {
        row[j + 1] = f2bf(cb);
        row[nang + j + 1] = f2bf(sb);
    }
}

// Network inputs of ansatz candidates c0 .. c0 + rows drawn from a SplitMix64
// stream: candidate c's angles are draws draw0 + c nang + j + 1 (theta = 2 pi u),
// its couplings draws draw0 + count nang + 3 c + i + 1 (J = 0.5 + u). Row
// layout: cos theta_j, sin theta_j, J_i - 1, zero padding to k_in. One item
// per (row, pair of angles), the last item of a row writes its couplings and
// padding. The inputs are rounded to BF16, so the trigonometry runs in FP32
// (the host regenerates the kept candidates' angles in FP64 for the exact
// evaluation).
__global__ __launch_bounds__(256) void k_ansatz_features(uint64_t seed, uint64_t draw0, uint64_t count, uint64_t c0, int rows,
                                                         int nang, int k_in, bf16 *__restrict__ X) {
    const uint32_t pairs = (uint32_t(nang) + 1) / 2, per = pairs + 1, items = uint32_t(rows) * per;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < items; i += gridDim.x * blockDim.x) {
        const uint32_t r = i / per, p = i - r * per;
        const uint64_t c = c0 + r;
        bf16 *row = X + size_t(r) * k_in;
        if (p < pairs) {
            const int j = 2 * int(p);
            const bool two = j + 1 < nang;
            float sa, ca, sb = 0.f, cb = 0.f;
            sincos_rev((float)sm64_uniform(seed, draw0 + c * nang + j + 1), sa, ca);
            if (two) sincos_rev((float)sm64_uniform(seed, draw0 + c * nang + j + 2), sb, cb);
            store_angle_pair(row, nang, j, ca, sa, cb, sb, two);
        } else {
            for (int q = 0; q < 3; q++)
                row[2 * nang + q] = f2bf((float)(0.5 + sm64_uniform(seed, draw0 + count * nang + 3 * c + q + 1) - 1.0));
            for (int k = 2 * nang + 3; k < k_in; k++) row[k] = f2bf(0.f);
        }
    }
}

// Candidates of an optimisation round around a centre theta0 (a local
// proposal): theta_j = theta0_j + sigma z_j with standard normals from the
// Box-Muller transform, both of its outputs used: for the pair p = j / 2 of
// candidate c, u1 and u2 are draws d + 1 and d + 2 of the stream `seed`,
// d = draw0 + 2 (c P + p), P = ceil(nang / 2), rho = sqrt(-2 ln(1 - u1)),
// z_2p = rho cos(2 pi u2), z_2p+1 = rho sin(2 pi u2); the couplings are the
// fixed J of the round. Same row layout and items as k_ansatz_features.
__global__ __launch_bounds__(256) void k_ansatz_features_local(uint64_t seed, uint64_t draw0, uint64_t c0, int rows, int nang,
                                                               int k_in, const double *__restrict__ theta0, double sigma,
                                                               double j0, double j1, double j2, bf16 *__restrict__ X) {
    const uint32_t pairs = (uint32_t(nang) + 1) / 2, per = pairs + 1, items = uint32_t(rows) * per;
    const float inv2pi = 0.15915494309189535f, sg = (float)sigma;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < items; i += gridDim.x * blockDim.x) {
        const uint32_t r = i / per, p = i - r * per;
        const uint64_t c = c0 + r;
        bf16 *row = X + size_t(r) * k_in;
        if (p < pairs) {
            const int j = 2 * int(p);
            const bool two = j + 1 < nang;
            const uint64_t d = draw0 + 2 * (c * pairs + p);
            const double u1 = sm64_uniform(seed, d + 1), u2 = sm64_uniform(seed, d + 2);
            const float rho = sqrtf(-2.f * __logf((float)(1.0 - u1)));
            float zs, zc;
            sincos_rev((float)u2, zs, zc);
            float sa, ca, sb = 0.f, cb = 0.f;
            sincos_rev(((float)theta0[j] + sg * rho * zc) * inv2pi, sa, ca);
            if (two) sincos_rev(((float)theta0[j + 1] + sg * rho * zs) * inv2pi, sb, cb);
            store_angle_pair(row, nang, j, ca, sa, cb, sb, two);
        } else {
            row[2 * nang] = f2bf((float)(j0 - 1.0));
            row[2 * nang + 1] = f2bf((float)(j1 - 1.0));
            row[2 * nang + 2] = f2bf((float)(j2 - 1.0));
            for (int k = 2 * nang + 3; k < k_in; k++) row[k] = f2bf(0.f);
        }
    }
}

// Rows per launch group of the candidate screens (activations of this many
// rows stay allocated: about 1.1 GB at 512 hidden units).
constexpr size_t kScreenChunk = 262144;

// ------------------------------------------------------------------ host --
struct Mlp {
    Ctx *ctx = nullptr;
    int n_in = 0, k_in = 0, hid = 0;  // k_in: n_in padded to a multiple of 16
    // parameters: W1 [k_in x hid], b1, W2 [hid x hid], b2, W3 [hid x hid], b3, w4 [hid], b4
    size_t off[8] = {0}, len[8] = {0}, nparam = 0;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // FP32 masters, Adam moments, grads
    bf16 *wb = nullptr;                                            // BF16 copies (W1, W2, W3)
    bf16 *wt = nullptr;                                            // BF16 W1^T, W2^T, W3^T (same offsets)
    int step = 0;
    // activations for a batch of up to cap rows
    size_t cap = 0;
    bf16 *X = nullptr, *H[3] = {nullptr, nullptr, nullptr}, *dZb[3] = {nullptr, nullptr, nullptr};
    float *Z[3] = {nullptr, nullptr, nullptr}, *dZf[3] = {nullptr, nullptr, nullptr};
    float *y = nullptr, *t = nullptr, *dy = nullptr, *part = nullptr, *lossp = nullptr;
    bf16 *T1 = nullptr, *T2 = nullptr;                             // transposed GEMM operands (weight gradients)
    float *ws = nullptr;                                           // split-K partials
    size_t ws_floats() const { return ws ? size_t(kMaxSplits) * size_t(std::max(hid, k_in)) * size_t(hid) : 0; }
    double *lacc = nullptr;                                        // epoch loss accumulator
    float *yall = nullptr;                                         // predictions of a candidate screen
    size_t ycap = 0;
    double *theta0 = nullptr;                                      // centre of a local candidate round
    size_t th0cap = 0;
    // resident training set
    bf16 *dataX = nullptr;
    float *dataT = nullptr;
    uint32_t *perm = nullptr;
    size_t ndata = 0;
    ~Mlp() {
        void *ps[] = {w, m, v, g, wb, wt, X, H[0], H[1], H[2], dZb[0], dZb[1], dZb[2], Z[0], Z[1], Z[2], dZf[0], dZf[1], dZf[2],
                      y, t, dy, part, lossp, T1, T2, ws, dataX, dataT, perm, lacc, yall, theta0};
        for (void *p : ps)
            if (p) (void)hipFree(p);
    }
};

static Status mlp_alloc(Mlp &N, size_t rows) {
    if (rows <= N.cap) return Status{};
    void **ps[] = {(void **)&N.X, (void **)&N.H[0], (void **)&N.H[1], (void **)&N.H[2], (void **)&N.dZb[0], (void **)&N.dZb[1],
                   (void **)&N.dZb[2], (void **)&N.Z[0], (void **)&N.Z[1], (void **)&N.Z[2], (void **)&N.dZf[0],
                   (void **)&N.dZf[1], (void **)&N.dZf[2], (void **)&N.y, (void **)&N.t, (void **)&N.dy, (void **)&N.part,
                   (void **)&N.lossp, (void **)&N.T1, (void **)&N.T2, (void **)&N.ws};
    QI_HIP_TRY(hipDeviceSynchronize());
    for (void **p : ps)
        if (*p) { (void)hipFree(*p); *p = nullptr; }
    const size_t H = N.hid;
    QI_HIP_TRY(hipMalloc(&N.X, rows * N.k_in * sizeof(bf16)));
    for (int l = 0; l < 3; l++) {
        QI_HIP_TRY(hipMalloc(&N.H[l], rows * H * sizeof(bf16)));
        QI_HIP_TRY(hipMalloc(&N.dZb[l], rows * H * sizeof(bf16)));
        QI_HIP_TRY(hipMalloc(&N.Z[l], rows * H * sizeof(float)));
        QI_HIP_TRY(hipMalloc(&N.dZf[l], rows * H * sizeof(float)));
    }
    QI_HIP_TRY(hipMalloc(&N.y, rows * sizeof(float)));
    QI_HIP_TRY(hipMalloc(&N.t, rows * sizeof(float)));
    QI_HIP_TRY(hipMalloc(&N.dy, rows * sizeof(float)));
    QI_HIP_TRY(hipMalloc(&N.part, ((rows + 255) / 256) * std::max<size_t>(H, N.k_in) * sizeof(float)));
    QI_HIP_TRY(hipMalloc(&N.lossp, ((rows + 3) / 4) * sizeof(float)));
    const size_t wide = std::max<size_t>(H, N.k_in);
    QI_HIP_TRY(hipMalloc(&N.T1, rows * wide * sizeof(bf16)));
    QI_HIP_TRY(hipMalloc(&N.T2, rows * H * sizeof(bf16)));
    QI_HIP_TRY(hipMalloc(&N.ws, size_t(kMaxSplits) * wide * H * sizeof(float)));
    if (!N.lacc) QI_HIP_TRY(hipMalloc(&N.lacc, sizeof(double)));
    N.cap = rows;
    return Status{};
}

// Few output tiles (small batches): each tile's K loop is a chain of
// dependent global-load round trips, so K is split over more workgroups
// (their FP32 partials in ws, ws_floats long) and k_split_epilogue applies
// the layer's epilogue.
constexpr int kSplitBelowTiles = 128;

// Many rows (screening chunks): 256 x 256 workgroup tiles of 8 waves (each
// 64 x 128), one per CU, loads one stage ahead; their fragment reads per
// matrix-core instruction are 3/4 of the 128 x 128 tile's.
constexpr int kBigBM = 256, kBigBN = 256, kBigWM = 4, kBigWN = 2;
static bool big_tiles(int M, int Nn) { return size_t(M) * size_t(Nn) >= size_t(512) * kBigBM * kBigBN; }
// partial sums per row of the EPI_GELU_HEAD layer
static int head_slots(int M, int Nn) {
    return big_tiles(M, Nn) ? kBigWN * ((Nn + kBigBN - 1) / kBigBN) : 2 * ((Nn + TN - 1) / TN);
}

template <bool TA, bool TB, int EPI>
static void gemm(hipStream_t st, const bf16 *A, int lda, const bf16 *B, int ldb, int M, int Nn, int K, float *Cf, bf16 *Cb,
                 int ldc, const float *bias = nullptr, const float *Zin = nullptr, float *ws = nullptr,
                 size_t ws_floats = 0) {
    if (big_tiles(M, Nn)) {
        const dim3 big((Nn + kBigBN - 1) / kBigBN, (M + kBigBM - 1) / kBigBM);
        hipLaunchKernelGGL((k_gemm<TA, TB, EPI, kBigBM, kBigBN, kBigWM, kBigWN, 1>), big, dim3(64 * kBigWM * kBigWN), 0, st,
                           A, lda, B, ldb, M, Nn, K, Cf, Cb, ldc, bias, Zin, K);
        return;
    }
    dim3 grid((Nn + TN - 1) / TN, (M + TM - 1) / TM);
    const int tiles = int(grid.x * grid.y);
    if constexpr (EPI == EPI_GELU || EPI == EPI_GELU_BWD) {
        if (ws && tiles < kSplitBelowTiles && Nn % 4 == 0 && ldc % 4 == 0 && Cb) {
            const size_t slice = size_t(M) * ldc;
            int splits = std::min({kMaxSplits, 256 / tiles, K / (2 * TK), int(ws_floats / slice)});
            if (splits >= 2) {
                const int kper = ((K + splits - 1) / splits + 2 * TK - 1) / (2 * TK) * (2 * TK);
                splits = (K + kper - 1) / kper;
                grid.z = splits;
                hipLaunchKernelGGL((k_gemm<TA, TB, EPI_STORE>), grid, dim3(256), 0, st, A, lda, B, ldb, M, Nn, K, ws,
                                   (bf16 *)nullptr, ldc, (const float *)nullptr, (const float *)nullptr, kper);
                const size_t items = size_t(M) * (Nn / 4);
                const unsigned blocks = (unsigned)std::min<size_t>(1024, (items + 255) / 256);
                hipLaunchKernelGGL((k_split_epilogue<EPI>), dim3(blocks), dim3(256), 0, st, (const float *)ws, splits, M,
                                   Nn, ldc, Cf, Cb, bias, Zin);
                return;
            }
        }
    }
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), grid, dim3(256), 0, st, A, lda, B, ldb, M, Nn, K, Cf, Cb, ldc, bias, Zin, K);
}

// C = A^T B for A (K x M) and B (K x N) row-major with a long K (weight
// gradients: K = rows of the batch): both operands transposed first (BF16
// copies in N.T1, N.T2), then K split over `splits` slices whose FP32
// partials (N.ws) are summed in slice order into C (FP32, ldc = N).
static void gemm_at_b(Mlp &N, hipStream_t st, const bf16 *A, int M, const bf16 *B, int Nn, int K, float *C) {
    hipLaunchKernelGGL(k_transpose_b16, dim3((M + 31) / 32, (K + 31) / 32), dim3(256), 0, st, A, K, M, N.T1);
    hipLaunchKernelGGL(k_transpose_b16, dim3((Nn + 31) / 32, (K + 31) / 32), dim3(256), 0, st, B, K, Nn, N.T2);
    const int tiles = ((M + TM - 1) / TM) * ((Nn + TN - 1) / TN);
    int splits = std::max(1, std::min({kMaxSplits, 256 / std::max(tiles, 1), K / (2 * TK)}));
    const int kper = ((K + splits - 1) / splits + 2 * TK - 1) / (2 * TK) * (2 * TK);
    splits = (K + kper - 1) / kper;
    dim3 grid((Nn + TN - 1) / TN, (M + TM - 1) / TM, splits);
    hipLaunchKernelGGL((k_gemm<false, true, EPI_STORE>), grid, dim3(256), 0, st, (const bf16 *)N.T1, K, (const bf16 *)N.T2, K, M,
                       Nn, K, splits > 1 ? N.ws : C, (bf16 *)nullptr, Nn, (const float *)nullptr, (const float *)nullptr, kper);
    if (splits > 1)
        hipLaunchKernelGGL(k_sum_splits, dim3(1024), dim3(256), 0, st, (const float *)N.ws, splits, size_t(M) * Nn, C);
}

static void colsum(Mlp &N, hipStream_t st, const float *X, int rows, int cols, float *out) {
    const int chunks = (rows + 255) / 256;
    hipLaunchKernelGGL(k_colsum_part, dim3((cols + 63) / 64, chunks), dim3(256), 0, st, X, rows, cols, N.part);
    hipLaunchKernelGGL(k_colsum_fold, dim3((cols + 255) / 256), dim3(256), 0, st, (const float *)N.part, chunks, cols, out);
}

// The BF16 copies of the weights (W and W^T) from the FP32 masters.
static void refresh_bf16(Mlp &N, hipStream_t st) {
    hipLaunchKernelGGL(k_to_bf16, dim3(1024), dim3(256), 0, st, (const float *)N.w, N.wb, N.nparam);
    const int rows[3] = {N.k_in, N.hid, N.hid};
    for (int l = 0; l < 3; l++) {
        const size_t o = N.off[2 * l];
        hipLaunchKernelGGL(k_transpose_bf16, dim3((N.hid + 31) / 32, (rows[l] + 31) / 32), dim3(256), 0, st,
                           (const float *)(N.w + o), rows[l], N.hid, N.wt + o);
    }
}

// Forward over rows already staged in N.X; keeps Z (pre-activations) when training.
// C = X W with B^T = W^T stored K-contiguous (the fast operand layout).
// Predictions y for `rows` staged rows: layers 1 and 2, then layer 3 with the
// head in its epilogue (the last hidden activations are never stored).
static void predict_rows(Mlp &N, hipStream_t st, int rows, float *y) {
    const bf16 *W1t = N.wt + N.off[0], *W2t = N.wt + N.off[2], *W3t = N.wt + N.off[4];
    const float *b1 = N.w + N.off[1], *b2 = N.w + N.off[3], *b3 = N.w + N.off[5];
    gemm<false, true, EPI_GELU>(st, N.X, N.k_in, W1t, N.k_in, rows, N.hid, N.k_in, nullptr, N.H[0], N.hid, b1);
    gemm<false, true, EPI_GELU>(st, N.H[0], N.hid, W2t, N.hid, rows, N.hid, N.hid, nullptr, N.H[1], N.hid, b2);
    const int slots = head_slots(rows, N.hid);
    gemm<false, true, EPI_GELU_HEAD>(st, N.H[1], N.hid, W3t, N.hid, rows, N.hid, N.hid, N.Z[2], nullptr, slots, b3,
                                     N.w + N.off[6]);
    hipLaunchKernelGGL(k_head_fold, dim3((rows + 255) / 256), dim3(256), 0, st, (const float *)N.Z[2], slots, rows,
                       (const float *)(N.w + N.off[7]), y);
}

static void forward(Mlp &N, hipStream_t st, int rows, bool keepZ) {
    const bf16 *W1t = N.wt + N.off[0], *W2t = N.wt + N.off[2], *W3t = N.wt + N.off[4];
    const float *b1 = N.w + N.off[1], *b2 = N.w + N.off[3], *b3 = N.w + N.off[5];
    const size_t wsf = N.ws_floats();
    gemm<false, true, EPI_GELU>(st, N.X, N.k_in, W1t, N.k_in, rows, N.hid, N.k_in, keepZ ? N.Z[0] : nullptr, N.H[0], N.hid, b1,
                                nullptr, N.ws, wsf);
    gemm<false, true, EPI_GELU>(st, N.H[0], N.hid, W2t, N.hid, rows, N.hid, N.hid, keepZ ? N.Z[1] : nullptr, N.H[1], N.hid, b2,
                                nullptr, N.ws, wsf);
    gemm<false, true, EPI_GELU>(st, N.H[1], N.hid, W3t, N.hid, rows, N.hid, N.hid, keepZ ? N.Z[2] : nullptr, N.H[2], N.hid, b3,
                                nullptr, N.ws, wsf);
}

// Gradients of the mean squared error over `rows` staged rows (targets in N.t) into N.g.
static void backward(Mlp &N, hipStream_t st, int rows) {
    const int H = N.hid;
    const float *w4 = N.w + N.off[6];
    const int hb = (rows + 3) / 4;
    hipLaunchKernelGGL(k_head, dim3(hb), dim3(256), 0, st, (const bf16 *)N.H[2], w4, (const float *)(N.w + N.off[7]), rows, H,
                       N.y, (const float *)N.t,
                       1.0f / rows, (const float *)N.Z[2], N.dZf[2], N.dZb[2], N.dy, N.lossp);
    {  // head: dw4 = H3^T dy, db4 = sum dy
        const int chunks = (rows + 255) / 256;
        hipLaunchKernelGGL(k_head_grad_part, dim3((H + 63) / 64, chunks), dim3(256), 0, st, (const bf16 *)N.H[2],
                           (const float *)N.dy, rows, H, N.part);
        hipLaunchKernelGGL(k_colsum_fold, dim3((H + 255) / 256), dim3(256), 0, st, (const float *)N.part, chunks, H,
                           N.g + N.off[6]);
        colsum(N, st, N.dy, rows, 1, N.g + N.off[7]);
    }
    // layer 3: dW3 = H2^T dZ3, db3 = colsum dZ3, dZ2 = (dZ3 W3^T) * gelu'(Z2)
    const bf16 *W2 = N.wb + N.off[2], *W3 = N.wb + N.off[4];
    gemm_at_b(N, st, N.H[1], H, N.dZb[2], H, rows, N.g + N.off[4]);
    colsum(N, st, N.dZf[2], rows, H, N.g + N.off[5]);
    gemm<false, true, EPI_GELU_BWD>(st, N.dZb[2], H, W3, H, rows, H, H, N.dZf[1], N.dZb[1], H, nullptr, N.Z[1], N.ws,
                                    N.ws_floats());
    // layer 2
    gemm_at_b(N, st, N.H[0], H, N.dZb[1], H, rows, N.g + N.off[2]);
    colsum(N, st, N.dZf[1], rows, H, N.g + N.off[3]);
    gemm<false, true, EPI_GELU_BWD>(st, N.dZb[1], H, W2, H, rows, H, H, N.dZf[0], N.dZb[0], H, nullptr, N.Z[0], N.ws,
                                    N.ws_floats());
    // layer 1: dW1 = X^T dZ1, db1
    gemm_at_b(N, st, N.X, N.k_in, N.dZb[0], H, rows, N.g + N.off[0]);
    colsum(N, st, N.dZf[0], rows, H, N.g + N.off[1]);
}

static Status stage_rows(Mlp &N, hipStream_t st, const float *x, const float *t, size_t rows) {
    // host FP32 rows (n_in wide) -> device BF16 rows (k_in wide, zero padded)
    std::vector<bf16> hx(rows * N.k_in, 0);
    for (size_t r = 0; r < rows; r++)
        for (int c = 0; c < N.n_in; c++) hx[r * N.k_in + c] = f2bf(x[r * N.n_in + c]);
    QI_HIP_TRY(hipMemcpyAsync(N.X, hx.data(), hx.size() * sizeof(bf16), hipMemcpyHostToDevice, st));
    if (t) QI_HIP_TRY(hipMemcpyAsync(N.t, t, rows * sizeof(float), hipMemcpyHostToDevice, st));
    QI_HIP_TRY(hipStreamSynchronize(st));  // hx is a host temporary
    return Status{};
}


}  // namespace qi

using namespace qi;

struct qi_mlp_s {
    qi::Mlp m;
};

#define QM_RET(st)                 \
    do {                           \
        Status s_ = (st);          \
        fill_err(err, s_);         \
        return (qi_status)s_.code; \
    } while (0)
#define QM_TRY(expr)                            \
    do {                                        \
        Status s2_ = (expr);                    \
        if (!s2_.ok()) QM_RET(s2_);             \
    } while (0)

extern "C" {

qi_status qi_mlp_create(qi_ctx_t ctx, uint32_t n_in, uint32_t hidden, uint64_t seed, qi_mlp_t *out, qi_error *err) {
    if (!ctx || !out) QM_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    if (n_in == 0 || hidden == 0 || hidden % 64) QM_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "hidden width must be a multiple of 64"));
    std::unique_ptr<qi_mlp_s> h(new qi_mlp_s());
    Mlp &N = h->m;
    N.ctx = &ctx->c;
    N.n_in = (int)n_in;
    N.k_in = (int)((n_in + 15) / 16 * 16);
    N.hid = (int)hidden;
    const size_t H = hidden;
    N.len[0] = size_t(N.k_in) * H; N.len[1] = H; N.len[2] = H * H; N.len[3] = H; N.len[4] = H * H; N.len[5] = H;
    N.len[6] = H; N.len[7] = 1;
    for (int i = 0; i < 8; i++) { N.off[i] = N.nparam; N.nparam += N.len[i]; }
    // He-style init (normal, std sqrt(2 / fan_in)), zero biases; padded input rows zero
    std::vector<float> w(N.nparam, 0.f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.f, 1.f);
    auto fill = [&](int p, int fan_in, int rows_real, int cols) {
        const float s = std::sqrt(2.0f / fan_in);
        for (int r = 0; r < rows_real; r++)
            for (int c = 0; c < cols; c++) w[N.off[p] + size_t(r) * cols + c] = s * nd(rng);
    };
    fill(0, N.n_in, N.n_in, N.hid);
    fill(2, N.hid, N.hid, N.hid);
    fill(4, N.hid, N.hid, N.hid);
    fill(6, N.hid, N.hid, 1);
    std::lock_guard<std::mutex> lk(N.ctx->mu);
    hipStream_t st = N.ctx->stream;
    hipError_t e;
    if ((e = hipMalloc(&N.w, N.nparam * 4)) || (e = hipMalloc(&N.m, N.nparam * 4)) || (e = hipMalloc(&N.v, N.nparam * 4)) ||
        (e = hipMalloc(&N.g, N.nparam * 4)) || (e = hipMalloc(&N.wb, N.nparam * 2)) || (e = hipMalloc(&N.wt, N.nparam * 2)))
        QM_RET(oom_err(e));
    if ((e = hipMemcpy(N.w, w.data(), N.nparam * 4, hipMemcpyHostToDevice)) || (e = hipMemset(N.m, 0, N.nparam * 4)) ||
        (e = hipMemset(N.v, 0, N.nparam * 4)) || (e = hipMemset(N.g, 0, N.nparam * 4)))
        QM_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    refresh_bf16(N, st);
    if ((e = hipStreamSynchronize(st))) QM_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    *out = h.release();
    QM_RET(Status{});
}

qi_status qi_mlp_destroy(qi_mlp_t m) {
    if (!m) return QI_OK;
    {
        std::lock_guard<std::mutex> lk(m->m.ctx->mu);
        (void)hipStreamSynchronize(m->m.ctx->stream);
        delete m;
    }
    return QI_OK;
}

qi_status qi_mlp_num_params(qi_mlp_t m, uint64_t *n, qi_error *err) {
    if (!m || !n) QM_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    *n = m->m.nparam;
    QM_RET(Status{});
}

// Parameters in the layout W1 [k_in x H] (rows >= n_in are zero), b1, W2, b2, W3, b3, w4, b4.
qi_status qi_mlp_get_params(qi_mlp_t m, float *out, qi_error *err) {
    if (!m || !out) QM_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    Mlp &N = m->m;
    std::lock_guard<std::mutex> lk(N.ctx->mu);
    hipError_t e = hipMemcpy(out, N.w, N.nparam * 4, hipMemcpyDeviceToHost);
    if (e) QM_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    QM_RET(Status{});
}

qi_status qi_mlp_set_params(qi_mlp_t m, const float *in, qi_error *err) {
    if (!m || !in) QM_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    Mlp &N = m->m;
    std::lock_guard<std::mutex> lk(N.ctx->mu);
    hipError_t e = hipMemcpy(N.w, in, N.nparam * 4, hipMemcpyHostToDevice);
    if (e) QM_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    refresh_bf16(N, N.ctx->stream);
    if ((e = hipStreamSynchronize(N.ctx->stream))) QM_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    QM_RET(Status{});
}

// Mean-squared-error gradient over `rows` samples (the check against the CPU oracle).
qi_status qi_mlp_gradients(qi_mlp_t m, const float *x, const float *t, size_t rows, float *grads, float *loss, qi_error *err) {
    if (!m || !x || !t || !grads || rows == 0 || rows % 16)
        QM_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "rows must be a positive multiple of 16"));
    Mlp &N = m->m;
    std::lock_guard<std::mutex> lk(N.ctx->mu);
    hipStream_t st = N.ctx->stream;
    QM_TRY(mlp_alloc(N, rows));
    QM_TRY(stage_rows(N, st, x, t, rows));
    forward(N, st, (int)rows, true);
    backward(N, st, (int)rows);
    std::vector<float> lp((rows + 3) / 4);
    hipError_t e = hipMemcpyAsync(grads, N.g, N.nparam * 4, hipMemcpyDeviceToHost, st);
    if (!e) e = hipMemcpyAsync(lp.data(), N.lossp, lp.size() * 4, hipMemcpyDeviceToHost, st);
    if (!e) e = hipStreamSynchronize(st);
    if (e) QM_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    if (loss) { double s = 0; for (float v : lp) s += v; *loss = (float)(s / rows); }
    QM_RET(Status{});
}

// Uploads a training set (rows x n_in inputs, rows targets) to the device once.
qi_status qi_mlp_set_data(qi_mlp_t m, const float *x, const float *t, size_t rows, qi_error *err) {
    if (!m || !x || !t || rows == 0) QM_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    Mlp &N = m->m;
    std::lock_guard<std::mutex> lk(N.ctx->mu);
    if (N.dataX) { (void)hipFree(N.dataX); (void)hipFree(N.dataT); (void)hipFree(N.perm); N.dataX = nullptr; }
    hipError_t e;
    if ((e = hipMalloc(&N.dataX, rows * N.k_in * sizeof(bf16))) || (e = hipMalloc(&N.dataT, rows * 4)) ||
        (e = hipMalloc(&N.perm, rows * 4)))
        QM_RET(oom_err(e));
    const size_t chunk = 65536;
    std::vector<bf16> hx(chunk * N.k_in);
    for (size_t r0 = 0; r0 < rows; r0 += chunk) {
        size_t nr = std::min(chunk, rows - r0);
        std::fill(hx.begin(), hx.end(), 0);
        for (size_t r = 0; r < nr; r++)
            for (int c = 0; c < N.n_in; c++) hx[r * N.k_in + c] = f2bf(x[(r0 + r) * N.n_in + c]);
        if ((e = hipMemcpy(N.dataX + r0 * N.k_in, hx.data(), nr * N.k_in * sizeof(bf16), hipMemcpyHostToDevice)))
            QM_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    }
    if ((e = hipMemcpy(N.dataT, t, rows * 4, hipMemcpyHostToDevice))) QM_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    N.ndata = rows;
    QM_RET(Status{});
}

// One epoch of Adam over the resident training set in shuffled mini-batches.
qi_status qi_mlp_train_epoch(qi_mlp_t m, uint32_t batch, float lr, uint64_t shuffle_seed, float *mean_loss, qi_error *err) {
    if (!m || batch == 0 || batch % 16) QM_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "batch must be a positive multiple of 16"));
    Mlp &N = m->m;
    if (!N.ndata) QM_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "no training data"));
    std::lock_guard<std::mutex> lk(N.ctx->mu);
    hipStream_t st = N.ctx->stream;
    QM_TRY(mlp_alloc(N, batch));
    std::vector<uint32_t> perm(N.ndata);
    for (size_t i = 0; i < N.ndata; i++) perm[i] = (uint32_t)i;
    std::mt19937_64 rng(shuffle_seed);
    std::shuffle(perm.begin(), perm.end(), rng);
    hipError_t e = hipMemcpy(N.perm, perm.data(), perm.size() * 4, hipMemcpyHostToDevice);
    if (e) QM_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    const size_t nb = N.ndata / batch;  // full batches only
    if ((e = hipMemsetAsync(N.lacc, 0, sizeof(double), st))) QM_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    (void)hipEventRecord(N.ctx->ev0, st);
    // no host round trip inside the epoch: the head bias is read on the
    // device and the loss accumulates on the device
    for (size_t b = 0; b < nb; b++) {
        hipLaunchKernelGGL(k_gather_rows, dim3(batch), dim3(256), 0, st, (const bf16 *)N.dataX, (const uint32_t *)N.perm + b * batch,
                           (int)batch, N.k_in, N.X, (const float *)N.dataT, N.t);
        forward(N, st, (int)batch, true);
        backward(N, st, (int)batch);
        hipLaunchKernelGGL(k_loss_accum, dim3(1), dim3(256), 0, st, (const float *)N.lossp, (int)((batch + 3) / 4), N.lacc);
        N.step++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)N.step), c2 = 1.f - std::pow(b2, (float)N.step);
        hipLaunchKernelGGL(k_adam, dim3(1024), dim3(256), 0, st, N.w, N.m, N.v, (const float *)N.g, N.wb, N.nparam, lr, b1, b2,
                           1e-8f, c1, c2);
        const int rows[3] = {N.k_in, N.hid, N.hid};
        for (int l = 0; l < 3; l++)
            hipLaunchKernelGGL(k_transpose_bf16, dim3((N.hid + 31) / 32, (rows[l] + 31) / 32), dim3(256), 0, st,
                               (const float *)(N.w + N.off[2 * l]), rows[l], N.hid, N.wt + N.off[2 * l]);
    }
    (void)hipEventRecord(N.ctx->ev1, st);
    double lsum = 0.0;
    if ((e = hipMemcpyAsync(&lsum, N.lacc, sizeof(double), hipMemcpyDeviceToHost, st))) QM_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    if ((e = hipStreamSynchronize(st))) QM_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, N.ctx->ev0, N.ctx->ev1);
    N.ctx->last_exec_ms = ms;
    if (mean_loss) *mean_loss = nb ? (float)(lsum / (double(nb) * batch)) : 0.f;
    QM_RET(Status{});
}

// Inference: y[r] for rows x n_in inputs (FP32 host arrays), in chunks.
qi_status qi_mlp_predict(qi_mlp_t m, const float *x, size_t rows, float *y, qi_error *err) {
    if (!m || !x || !y) QM_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    Mlp &N = m->m;
    std::lock_guard<std::mutex> lk(N.ctx->mu);
    hipStream_t st = N.ctx->stream;
    const size_t chunk = 65536;
    QM_TRY(mlp_alloc(N, std::min(chunk, std::max<size_t>(rows, 1))));
    hipError_t e = hipSuccess;
    double dev_ms = 0.0;
    for (size_t r0 = 0; r0 < rows; r0 += chunk) {
        const size_t nr = std::min(chunk, rows - r0);
        QM_TRY(stage_rows(N, st, x + r0 * N.n_in, nullptr, nr));
        hipEvent_t a = N.ctx->ev0, b = N.ctx->ev1;
        (void)hipEventRecord(a, st);
        predict_rows(N, st, (int)nr, N.y);
        (void)hipEventRecord(b, st);
        if ((e = hipMemcpyAsync(y + r0, N.y, nr * 4, hipMemcpyDeviceToHost, st))) break;
        if ((e = hipStreamSynchronize(st))) break;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, a, b);
        dev_ms += ms;
    }
    if (e) QM_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    N.ctx->last_exec_ms = dev_ms;
    QM_RET(Status{});
}

// Predictions for `count` ansatz candidates whose parameters are drawn on the
// device from a SplitMix64 stream (features as k_ansatz_features documents);
// the host can regenerate any candidate's parameters from its index.
qi_status qi_mlp_predict_ansatz(qi_mlp_t m, uint64_t seed, uint64_t draw0, uint64_t count, uint32_t nang, float *y,
                                qi_error *err) {
    if (!m || !y || count == 0) QM_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    Mlp &N = m->m;
    if ((uint64_t)N.n_in != 2ull * nang + 3) QM_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "network inputs != 2 nang + 3"));
    std::lock_guard<std::mutex> lk(N.ctx->mu);
    hipStream_t st = N.ctx->stream;
    const size_t chunk = std::min<size_t>(kScreenChunk, (size_t(1) << 31) / (size_t(nang) + 1));  // 32-bit item indices
    QM_TRY(mlp_alloc(N, std::min<size_t>(chunk, count)));
    hipError_t e;
    if (count > N.ycap) {
        if (N.yall) (void)hipFree(N.yall);
        N.yall = nullptr;
        N.ycap = 0;
        if ((e = hipMalloc(&N.yall, count * sizeof(float)))) QM_RET(oom_err(e));
        N.ycap = count;
    }
    (void)hipEventRecord(N.ctx->ev0, st);
    for (uint64_t c0 = 0; c0 < count; c0 += chunk) {
        const int nr = (int)std::min<uint64_t>(chunk, count - c0);
        const size_t items = size_t(nr) * ((nang + 1) / 2 + 1);
        hipLaunchKernelGGL(k_ansatz_features, dim3((unsigned)std::min<size_t>((items + 255) / 256, 16384)), dim3(256), 0, st, seed,
                           draw0, count, c0, nr, (int)nang, N.k_in, N.X);
        predict_rows(N, st, nr, N.yall + c0);
    }
    (void)hipEventRecord(N.ctx->ev1, st);
    if ((e = hipMemcpyAsync(y, N.yall, count * sizeof(float), hipMemcpyDeviceToHost, st)) || (e = hipStreamSynchronize(st)))
        QM_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, N.ctx->ev0, N.ctx->ev1);
    N.ctx->last_exec_ms = ms;
    QM_RET(Status{});
}

// Predictions for `count` candidates of a local optimisation round around
// theta0 (k_ansatz_features_local), couplings J[3] fixed; the host regenerates
// the kept candidates' angles from their indices with the same formula.
qi_status qi_mlp_predict_ansatz_local(qi_mlp_t m, uint64_t seed, uint64_t draw0, uint64_t count, uint32_t nang,
                                      const double *theta0, double sigma, const double *J, float *y, qi_error *err) {
    if (!m || !y || !theta0 || !J || count == 0) QM_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "NULL argument"));
    Mlp &N = m->m;
    if ((uint64_t)N.n_in != 2ull * nang + 3) QM_RET(make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "network inputs != 2 nang + 3"));
    std::lock_guard<std::mutex> lk(N.ctx->mu);
    hipStream_t st = N.ctx->stream;
    const size_t chunk = std::min<size_t>(kScreenChunk, (size_t(1) << 31) / (size_t(nang) + 1));  // 32-bit item indices
    QM_TRY(mlp_alloc(N, std::min<size_t>(chunk, count)));
    hipError_t e;
    if (count > N.ycap) {
        if (N.yall) (void)hipFree(N.yall);
        N.yall = nullptr;
        N.ycap = 0;
        if ((e = hipMalloc(&N.yall, count * sizeof(float)))) QM_RET(oom_err(e));
        N.ycap = count;
    }
    if (nang > N.th0cap) {
        if (N.theta0) (void)hipFree(N.theta0);
        N.theta0 = nullptr;
        N.th0cap = 0;
        if ((e = hipMalloc(&N.theta0, nang * sizeof(double)))) QM_RET(oom_err(e));
        N.th0cap = nang;
    }
    if ((e = hipMemcpyAsync(N.theta0, theta0, nang * sizeof(double), hipMemcpyHostToDevice, st)))
        QM_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    (void)hipEventRecord(N.ctx->ev0, st);
    for (uint64_t c0 = 0; c0 < count; c0 += chunk) {
        const int nr = (int)std::min<uint64_t>(chunk, count - c0);
        const size_t items = size_t(nr) * ((nang + 1) / 2 + 1);
        hipLaunchKernelGGL(k_ansatz_features_local, dim3((unsigned)std::min<size_t>((items + 255) / 256, 16384)), dim3(256), 0, st,
                           seed, draw0, c0, nr, (int)nang, N.k_in, (const double *)N.theta0, sigma, J[0], J[1], J[2], N.X);
        predict_rows(N, st, nr, N.yall + c0);
    }
    (void)hipEventRecord(N.ctx->ev1, st);
    if ((e = hipMemcpyAsync(y, N.yall, count * sizeof(float), hipMemcpyDeviceToHost, st)) || (e = hipStreamSynchronize(st)))
        QM_RET(make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e)));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, N.ctx->ev0, N.ctx->ev1);
    N.ctx->last_exec_ms = ms;
    QM_RET(Status{});
}

}  // extern "C"
