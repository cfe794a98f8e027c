// reduce.hip — FP64 compensated reductions over a state: sum |a_i|^2 and the
// inner product sum conj(a_i) b_i (State::inner_product, state.rs:890-917).
// Complex64 amplitudes are widened to FP64 before they are multiplied.
#include "qi_internal.h"
#include "reduce_util.h"

namespace qi {

namespace {

constexpr int kRedThreads = 256;
constexpr int kRedBlocksMax = 2048;

template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_norm_part(const T *__restrict__ a, uint64_t dim,
                                                           double *__restrict__ part) {
    __shared__ double sh[2 * 4];
    DD v[1] = {dd_zero()};
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < dim; i += uint64_t(gridDim.x) * blockDim.x) {
        double re = a[i].x, im = a[i].y;
        dd_acc(v[0], re * re + im * im);
    }
    block_reduce_dd<1>(v, sh);
    if (threadIdx.x == 0) { part[blockIdx.x * 2] = v[0].s; part[blockIdx.x * 2 + 1] = v[0].c; }
}

template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_inner_part(const T *__restrict__ a, const T *__restrict__ b,
                                                            uint64_t dim, double *__restrict__ part) {
    __shared__ double sh[4 * 4];
    DD v[2] = {dd_zero(), dd_zero()};
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < dim; i += uint64_t(gridDim.x) * blockDim.x) {
        double ar = a[i].x, ai = a[i].y, br = b[i].x, bi = b[i].y;
        // conj(a) * b
        dd_acc(v[0], ar * br + ai * bi);
        dd_acc(v[1], ar * bi - ai * br);
    }
    block_reduce_dd<2>(v, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x * 4 + 0] = v[0].s; part[blockIdx.x * 4 + 1] = v[0].c;
        part[blockIdx.x * 4 + 2] = v[1].s; part[blockIdx.x * 4 + 3] = v[1].c;
    }
}

}  // namespace

Status reduce_state(State &a, const State *b, int kind, double *out2, hipStream_t st) {
    Ctx &C = *a.ctx;
    uint64_t dim = uint64_t(1) << a.n;
    int blocks = (int)std::min<uint64_t>(kRedBlocksMax, (dim + kRedThreads * 8 - 1) / (kRedThreads * 8));
    if (blocks < 1) blocks = 1;
    Status s = ensure(C.red, size_t(blocks) * 4 * sizeof(double));
    if (!s.ok()) return s;
    s = ensure(C.red_out, 64 * sizeof(double));
    if (!s.ok()) return s;
    double *part = (double *)C.red.p, *out = (double *)C.red_out.p;
    if (kind == RK_NORM) {
        if (a.prec == QI_C128)
            hipLaunchKernelGGL(k_norm_part<double2>, dim3(blocks), dim3(kRedThreads), 0, st, (const double2 *)a.d, dim, part);
        else
            hipLaunchKernelGGL(k_norm_part<float2>, dim3(blocks), dim3(kRedThreads), 0, st, (const float2 *)a.d, dim, part);
        hipLaunchKernelGGL(k_reduce_partials<1>, dim3(1), dim3(256), 0, st, part, blocks, out);
    } else {
        if (a.prec == QI_C128)
            hipLaunchKernelGGL(k_inner_part<double2>, dim3(blocks), dim3(kRedThreads), 0, st, (const double2 *)a.d,
                               (const double2 *)b->d, dim, part);
        else
            hipLaunchKernelGGL(k_inner_part<float2>, dim3(blocks), dim3(kRedThreads), 0, st, (const float2 *)a.d,
                               (const float2 *)b->d, dim, part);
        hipLaunchKernelGGL(k_reduce_partials<2>, dim3(1), dim3(256), 0, st, part, blocks, out);
    }
    QI_HIP_TRY(hipGetLastError());
    QI_HIP_TRY(hipMemcpyAsync(C.host_pinned, out, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    QI_HIP_TRY(hipStreamSynchronize(st));
    out2[0] = C.host_pinned[0];
    out2[1] = kind == RK_NORM ? 0.0 : C.host_pinned[1];
    return Status{};
}

}  // namespace qi
