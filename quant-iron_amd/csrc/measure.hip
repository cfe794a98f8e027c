// measure.hip — State::measure / measure_n on the device (state.rs:525-784):
// FP64 outcome histogram over the measured qubits, FP64 cumulative scan in
// the reference's sequential order, per-shot search with injected uniforms,
// and collapse + renormalisation into a destination state.
#include "qi_internal.h"
#include "reduce_util.h"

namespace qi {

namespace {

__device__ __forceinline__ uint64_t insert_zeros(uint64_t k, uint64_t fixed) {
    while (fixed) {
        unsigned p = __builtin_ctzll(fixed);
        k = ((k >> p) << (p + 1)) | (k & ((uint64_t(1) << p) - 1));
        fixed &= fixed - 1;
    }
    return k;
}

// Spread the low bits of `o` onto the set bits of `mask` (ascending).
__device__ __host__ __forceinline__ uint64_t deposit(uint64_t o, uint64_t mask) {
    uint64_t r = 0;
    for (int b = 0; mask; b++) {
        uint64_t low = mask & (~mask + 1);
        if ((o >> b) & 1) r |= low;
        mask &= mask - 1;
    }
    return r;
}

template <typename T> __device__ __forceinline__ double nsq(T a) {
    double re = a.x, im = a.y;
    return re * re + im * im;
}

// One thread per bin: sum over the 2^rest indices of the bin, in order.
template <typename T>
__global__ __launch_bounds__(256) void k_hist_thread(const T *__restrict__ psi, uint64_t dmask, uint64_t bins,
                                                     uint64_t rest_items, double *__restrict__ probs) {
    uint64_t o = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (o >= bins) return;
    uint64_t base = deposit(o, dmask);
    DD v = dd_zero();
    for (uint64_t r = 0; r < rest_items; r++) dd_acc(v, nsq(psi[insert_zeros(r, dmask) | base]));
    probs[o] = v.s + v.c;
}

// Chunked: block (o, chunk) sums `per_chunk` indices of bin o; partials are
// folded per bin by k_fold_bins.
template <typename T>
__global__ __launch_bounds__(256) void k_hist_chunk(const T *__restrict__ psi, uint64_t dmask, uint64_t rest_items,
                                                    uint64_t per_chunk, int chunks, double *__restrict__ part) {
    __shared__ double sh[2 * 4];
    const uint64_t o = blockIdx.y;
    const uint64_t base = deposit(o, dmask);
    const uint64_t lo = uint64_t(blockIdx.x) * per_chunk;
    uint64_t hi = lo + per_chunk;
    if (hi > rest_items) hi = rest_items;
    DD v[1] = {dd_zero()};
    for (uint64_t r = lo + threadIdx.x; r < hi; r += blockDim.x) dd_acc(v[0], nsq(psi[insert_zeros(r, dmask) | base]));
    block_reduce_dd<1>(v, sh);
    if (threadIdx.x == 0) {
        size_t w = (size_t(o) * chunks + blockIdx.x) * 2;
        part[w] = v[0].s; part[w + 1] = v[0].c;
    }
}

__global__ __launch_bounds__(64) void k_fold_bins(const double *__restrict__ part, int chunks, double *__restrict__ probs) {
    if (threadIdx.x != 0) return;
    const uint64_t o = blockIdx.x;
    DD v = dd_zero();
    for (int c = 0; c < chunks; c++) v = dd_add(v, DD{part[(o * chunks + c) * 2], part[(o * chunks + c) * 2 + 1]});
    probs[o] = v.s + v.c;
}

// Expand a histogram over the distinct measured qubits to the 2^m outcomes of
// the listed qubits (duplicates allowed); inconsistent outcomes get 0.
__global__ __launch_bounds__(256) void k_expand(const double *__restrict__ hd, double *__restrict__ hm, uint64_t bins_m,
                                                const int *__restrict__ slot, int m) {
    uint64_t o = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (o >= bins_m) return;
    uint64_t d = 0, seen = 0;
    bool ok = true;
    for (int b = 0; b < m; b++) {
        int s = slot[b];
        uint64_t bit = (o >> b) & 1;
        if ((seen >> s) & 1) { if (((d >> s) & 1) != bit) ok = false; }
        else { seen |= uint64_t(1) << s; d |= bit << s; }
    }
    hm[o] = ok ? hd[d] : 0.0;
}

// Sequential scan (state.rs:590-606): total = sum p_i in index order, then
// C_i = C_{i-1} + p_i / total in index order, with the reference's rounding.
// One wave: lanes load 64 bins at a time (coalesced) and every lane runs the
// same sequential sum over the broadcast values; lane j keeps C_{base+j}.
__global__ __launch_bounds__(64) void k_scan_seq(const double *__restrict__ p, uint64_t bins, double *__restrict__ cdf,
                                                 double *__restrict__ total_out) {
    const int lane = threadIdx.x;
    double total = 0.0;
    for (uint64_t base = 0; base < bins; base += 64) {
        double x = base + lane < bins ? p[base + lane] : 0.0;
        int cnt = bins - base < 64 ? int(bins - base) : 64;
        for (int j = 0; j < cnt; j++) total += __shfl(x, j);
    }
    double cum = 0.0;
    for (uint64_t base = 0; base < bins; base += 64) {
        double x = base + lane < bins ? p[base + lane] : 0.0;
        int cnt = bins - base < 64 ? int(bins - base) : 64;
        double mine = 0.0;
        for (int j = 0; j < cnt; j++) {
            cum += __shfl(x, j) / total;
            if (lane == j) mine = cum;
        }
        if (base + lane < bins) cdf[base + lane] = mine;
    }
    if (lane == 0) *total_out = total;
}

// Large histograms: chunk totals, their sequential prefix, then each chunk is
// scanned sequentially from its prefix. Agrees with the sequential rule up to
// rounding of the chunk prefixes.
constexpr uint64_t kScanChunk = 1024;
__global__ __launch_bounds__(256) void k_scan_chunk_tot(const double *__restrict__ p, uint64_t bins, double *__restrict__ tot) {
    uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t lo = c * kScanChunk;
    if (lo >= bins) return;
    uint64_t hi = lo + kScanChunk < bins ? lo + kScanChunk : bins;
    DD v = dd_zero();
    for (uint64_t i = lo; i < hi; i++) dd_acc(v, p[i]);
    tot[2 * c] = v.s; tot[2 * c + 1] = v.c;
}
__global__ void k_scan_chunk_prefix(double *__restrict__ tot, uint64_t nchunks, double *__restrict__ total_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    DD run = dd_zero();
    for (uint64_t c = 0; c < nchunks; c++) {
        DD t{tot[2 * c], tot[2 * c + 1]};
        tot[2 * c] = run.s + run.c;  // exclusive prefix
        run = dd_add(run, t);
    }
    *total_out = run.s + run.c;
}
__global__ __launch_bounds__(256) void k_scan_chunk_fill(const double *__restrict__ p, uint64_t bins, const double *__restrict__ tot,
                                                         const double *__restrict__ total_p, double *__restrict__ cdf) {
    uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t lo = c * kScanChunk;
    if (lo >= bins) return;
    uint64_t hi = lo + kScanChunk < bins ? lo + kScanChunk : bins;
    double total = *total_p;
    double cum = tot[2 * c] / total;
    for (uint64_t i = lo; i < hi; i++) { cum += p[i] / total; cdf[i] = cum; }
}

// Per shot: first i with u < C_i, else the last bin (state.rs:604-619).
__global__ __launch_bounds__(256) void k_sample(const double *__restrict__ cdf, uint64_t bins, const double *__restrict__ u,
                                                uint64_t shots, uint64_t *__restrict__ out) {
    uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= shots) return;
    double x = u[s];
    uint64_t lo = 0, hi = bins;  // first index with x < cdf[i]
    while (lo < hi) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (x < cdf[mid]) hi = mid; else lo = mid + 1;
    }
    out[s] = lo < bins ? lo : bins - 1;
}

// Keep the amplitudes of the sampled outcome, zero the rest; partial sums of
// the kept |a|^2.
template <typename T>
__global__ __launch_bounds__(256) void k_collapse(const T *src, T *dst, uint64_t dim,
                                                  uint64_t dmask, uint64_t want, double *__restrict__ part) {
    __shared__ double sh[2 * 4];
    DD v[1] = {dd_zero()};
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < dim; i += uint64_t(gridDim.x) * blockDim.x) {
        T a = src[i];
        if ((i & dmask) == want) dd_acc(v[0], nsq(a));
        else { a.x = 0; a.y = 0; }
        dst[i] = a;
    }
    block_reduce_dd<1>(v, sh);
    if (threadIdx.x == 0) { part[blockIdx.x * 2] = v[0].s; part[blockIdx.x * 2 + 1] = v[0].c; }
}

}  // namespace

// Histogram over `qubits` (bit b = qubits[b]); returns a device pointer to
// 2^m FP64 probabilities inside ctx.probs.
Status measure_probs_device(State &s, const uint64_t *qubits, uint32_t m, double **dprobs, hipStream_t st) {
    Ctx &C = *s.ctx;
    // distinct qubits, in ascending order for the bin index
    uint64_t dmask = 0;
    for (uint32_t b = 0; b < m; b++) dmask |= uint64_t(1) << qubits[b];
    const int d = __builtin_popcountll(dmask);
    const uint64_t bins_d = uint64_t(1) << d, bins_m = uint64_t(1) << m;
    const uint64_t rest = uint64_t(1) << (s.n - d);
    // layout in ctx.probs: [hist_d | hist_m | cdf | partials]
    size_t need = (bins_d + 2 * bins_m + 64) * sizeof(double);
    Status e = ensure(C.probs, need);
    if (!e.ok()) return e;
    double *hd = (double *)C.probs.p;
    double *hm = hd + bins_d;
    if (rest <= 1024 || bins_d >= 65536) {
        int blocks = (int)((bins_d + 255) / 256);
        if (s.prec == QI_C128)
            hipLaunchKernelGGL(k_hist_thread<double2>, dim3(blocks), dim3(256), 0, st, (const double2 *)s.d, dmask, bins_d, rest, hd);
        else
            hipLaunchKernelGGL(k_hist_thread<float2>, dim3(blocks), dim3(256), 0, st, (const float2 *)s.d, dmask, bins_d, rest, hd);
    } else {
        int chunks = (int)std::max<uint64_t>(1, std::min<uint64_t>(rest / 4096, std::max<uint64_t>(1, 4096 / bins_d)));
        uint64_t per_chunk = (rest + chunks - 1) / chunks;
        e = ensure(C.red, bins_d * chunks * 2 * sizeof(double));
        if (!e.ok()) return e;
        double *part = (double *)C.red.p;
        dim3 grid(chunks, (unsigned)bins_d);
        if (s.prec == QI_C128)
            hipLaunchKernelGGL(k_hist_chunk<double2>, grid, dim3(256), 0, st, (const double2 *)s.d, dmask, rest, per_chunk, chunks, part);
        else
            hipLaunchKernelGGL(k_hist_chunk<float2>, grid, dim3(256), 0, st, (const float2 *)s.d, dmask, rest, per_chunk, chunks, part);
        hipLaunchKernelGGL(k_fold_bins, dim3((unsigned)bins_d), dim3(64), 0, st, part, chunks, hd);
    }
    // slot of each listed qubit among the distinct ones (ascending order)
    std::vector<int> slot(m);
    for (uint32_t b = 0; b < m; b++) slot[b] = __builtin_popcountll(dmask & ((uint64_t(1) << qubits[b]) - 1));
    e = ensure(C.aux, 64 * sizeof(int) + 64);
    if (!e.ok()) return e;
    QI_HIP_TRY(hipMemcpyAsync(C.aux.p, slot.data(), m * sizeof(int), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_expand, dim3((unsigned)((bins_m + 255) / 256)), dim3(256), 0, st, hd, hm, bins_m, (const int *)C.aux.p, (int)m);
    QI_HIP_TRY(hipGetLastError());
    QI_HIP_TRY(hipStreamSynchronize(st));  // slot table is host memory owned by this frame
    *dprobs = hm;
    return Status{};
}

// Keep the amplitudes whose measured bits spell `outcome` (bit b of the
// outcome = qubit qubits[b]), renormalise by sqrt(sum |a|^2) when that sum
// exceeds f64::EPSILON (state.rs:622-654). `dst` may be `s`.
Status collapse_outcome(State &s, State &dst, const uint64_t *qubits, uint32_t m, uint64_t outcome, hipStream_t st) {
    Ctx &C = *s.ctx;
    uint64_t dmask = 0, want = 0;
    for (uint32_t b = 0; b < m; b++) {
        dmask |= uint64_t(1) << qubits[b];
        if ((outcome >> b) & 1) want |= uint64_t(1) << qubits[b];
    }
    uint64_t dim = uint64_t(1) << s.n;
    int blocks = (int)std::min<uint64_t>(2048, (dim + 2047) / 2048);
    Status e = ensure(C.red, size_t(blocks) * 2 * sizeof(double));
    if (!e.ok()) return e;
    e = ensure(C.red_out, 64 * sizeof(double));
    if (!e.ok()) return e;
    if (s.prec == QI_C128)
        hipLaunchKernelGGL(k_collapse<double2>, dim3(blocks), dim3(256), 0, st, (const double2 *)s.d, (double2 *)dst.d, dim,
                           dmask, want, (double *)C.red.p);
    else
        hipLaunchKernelGGL(k_collapse<float2>, dim3(blocks), dim3(256), 0, st, (const float2 *)s.d, (float2 *)dst.d, dim,
                           dmask, want, (double *)C.red.p);
    hipLaunchKernelGGL(k_reduce_partials<1>, dim3(1), dim3(256), 0, st, (const double *)C.red.p, blocks, (double *)C.red_out.p);
    QI_HIP_TRY(hipGetLastError());
    QI_HIP_TRY(hipMemcpyAsync(C.host_pinned, C.red_out.p, sizeof(double), hipMemcpyDeviceToHost, st));
    QI_HIP_TRY(hipStreamSynchronize(st));
    double ns = C.host_pinned[0];
    if (ns > 2.220446049250313e-16) {  // state.rs:649-654
        e = launch_div_real(dst, std::sqrt(ns), st);
        if (!e.ok()) return e;
    }
    return Status{};
}

Status measure_sample_collapse(State &s, State &dst, const uint64_t *qubits, uint32_t m, const double *uniforms,
                               size_t shots, uint64_t *outcomes_host, bool collapse, hipStream_t st) {
    Ctx &C = *s.ctx;
    double *hm = nullptr;
    Status e = measure_probs_device(s, qubits, m, &hm, st);
    if (!e.ok()) return e;
    const uint64_t bins = uint64_t(1) << m;
    double *cdf = hm + bins;
    double *tot = cdf + bins;  // 64 doubles reserved; chunk totals go to ctx.red
    // total and cdf
    if (bins <= (uint64_t(1) << 16)) {
        hipLaunchKernelGGL(k_scan_seq, dim3(1), dim3(64), 0, st, (const double *)hm, bins, cdf, tot);
    } else {
        uint64_t nch = (bins + kScanChunk - 1) / kScanChunk;
        e = ensure(C.red, nch * 2 * sizeof(double));
        if (!e.ok()) return e;
        double *ct = (double *)C.red.p;
        hipLaunchKernelGGL(k_scan_chunk_tot, dim3((unsigned)((nch + 255) / 256)), dim3(256), 0, st, hm, bins, ct);
        hipLaunchKernelGGL(k_scan_chunk_prefix, dim3(1), dim3(64), 0, st, ct, nch, tot);
        hipLaunchKernelGGL(k_scan_chunk_fill, dim3((unsigned)((nch + 255) / 256)), dim3(256), 0, st, hm, bins, ct, tot, cdf);
    }
    // uniforms and outcomes through ctx.aux: [outcomes | uniforms]
    e = ensure(C.aux, shots * (sizeof(uint64_t) + sizeof(double)) + 64);
    if (!e.ok()) return e;
    uint64_t *dout = (uint64_t *)C.aux.p;
    double *du = (double *)(dout + shots);
    QI_HIP_TRY(hipMemcpyAsync(du, uniforms, shots * sizeof(double), hipMemcpyHostToDevice, st));
    // total < EPSILON -> UnknownError (state.rs:593-595)
    QI_HIP_TRY(hipMemcpyAsync(C.host_pinned, tot, sizeof(double), hipMemcpyDeviceToHost, st));
    hipLaunchKernelGGL(k_sample, dim3((unsigned)((shots + 255) / 256)), dim3(256), 0, st, cdf, bins, du, (uint64_t)shots, dout);
    QI_HIP_TRY(hipGetLastError());
    QI_HIP_TRY(hipMemcpyAsync(outcomes_host, dout, shots * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    QI_HIP_TRY(hipStreamSynchronize(st));
    if (C.host_pinned[0] < 2.220446049250313e-16) return make_err(QI_ERR_UNKNOWN);
    if (!collapse) return Status{};
    return collapse_outcome(s, dst, qubits, m, outcomes_host[0], st);
}

}  // namespace qi
