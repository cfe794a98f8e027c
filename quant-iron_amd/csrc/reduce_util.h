// reduce_util.h — FP64 compensated (double-double) accumulation and the
// deterministic two-stage block reduction shared by the reduction kernels.
// Each thread keeps a (sum, error) pair per real component (Knuth's two-sum),
// blocks combine pairs in a fixed tree order, and a single-block second stage
// folds the per-block partials in index order, so a reduction over the same
// state always returns the same bits.
#pragma once

#include <hip/hip_runtime.h>

namespace qi {

struct DD {
    double s, c;
};

__device__ __forceinline__ DD dd_zero() { return DD{0.0, 0.0}; }

// s + x with the rounding error of the addition carried in c.
__device__ __forceinline__ void dd_acc(DD &a, double x) {
    double s = a.s + x;
    double bb = s - a.s;
    double e = (a.s - (s - bb)) + (x - bb);
    a.s = s;
    a.c += e;
}

__device__ __forceinline__ DD dd_add(DD a, DD b) {
    double s = a.s + b.s;
    double bb = s - a.s;
    double e = (a.s - (s - bb)) + (b.s - bb);
    e += a.c + b.c;
    double r = s + e;
    double c = e - (r - s);
    return DD{r, c};
}

__device__ __forceinline__ DD dd_shfl_xor(DD a, int m) {
    return DD{__shfl_xor(a.s, m), __shfl_xor(a.c, m)};
}

// Reduce NV double-double values across a block of `threads` (multiple of 64).
// Result valid in thread 0. `sh` must hold NV * 2 * (threads / 64) doubles.
template <int NV>
__device__ __forceinline__ void block_reduce_dd(DD (&v)[NV], double *sh) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1)
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] = dd_add(v[k], dd_shfl_xor(v[k], m));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) { sh[(wave * NV + k) * 2] = v[k].s; sh[(wave * NV + k) * 2 + 1] = v[k].c; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < nw; w++)
#pragma unroll
            for (int k = 0; k < NV; k++) v[k] = dd_add(v[k], DD{sh[(w * NV + k) * 2], sh[(w * NV + k) * 2 + 1]});
    }
    __syncthreads();
}

// Second stage: fold `nblocks` partials of NV double-doubles (layout
// [block][k][s,c]) in block order; writes NV doubles (s + c) to out.
template <int NV>
__global__ __launch_bounds__(256) void k_reduce_partials(const double *__restrict__ part, int nblocks,
                                                         double *__restrict__ out) {
    __shared__ double sh[NV * 2 * 4];
    DD v[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = dd_zero();
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < NV; k++) v[k] = dd_add(v[k], DD{part[(b * NV + k) * 2], part[(b * NV + k) * 2 + 1]});
    block_reduce_dd<NV>(v, sh);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) out[k] = v[k].s + v[k].c;
}

}  // namespace qi
