// pauli.hip — PauliString / SumOp kernels (pauli_string.rs:139-287, 453-507).
//
// A Pauli string with masks (x, z) (Y where both are set) acts as
//     (P psi)_i = (-i)^{nY} (-1)^{popcount(i & z)} psi_{i ^ x},
// the product of the single-qubit X/Y/Z actions of operator.rs:474-606 in any
// order. The phase is a whole number of quarter turns, applied exactly by
// swapping and negating parts, so the result equals the reference's
// op-by-op application. The coefficient multiply, cosh/sinh mix and term sums
// then follow the reference's complex operations one for one.
#include "qi_internal.h"
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

// Multiply by i^q exactly.
template <typename T> __device__ __forceinline__ T rot(T a, unsigned q) {
    T r;
    switch (q & 3) {
    case 0: r = a; break;
    case 1: r.x = -a.y; r.y = a.x; break;
    case 2: r.x = -a.x; r.y = -a.y; break;
    default: r.x = a.y; r.y = -a.x; break;
    }
    return r;
}

// Quarter turns of (-i)^{nY} (-1)^{popcount(i & z)}.
__device__ __forceinline__ unsigned quarter(uint64_t i, uint64_t z, unsigned q0) {
    return q0 + 2u * (unsigned)__popcll(i & z);
}

__device__ __forceinline__ uint64_t insert_zero(uint64_t k, unsigned p) {
    return ((k >> p) << (p + 1)) | (k & ((uint64_t(1) << p) - 1));
}

struct TermK {
    uint64_t x, z;
    unsigned q0;     // 3 * nY mod 4
    unsigned pivot;  // lowest bit of x (pair kernels)
    double c[2];     // coefficient
    double ch[2], sh[2];
};

// out = (P in) * c. Pairs (i, i ^ x) with the pivot bit of i clear, so the
// same kernel is correct in place.
template <typename R>
__global__ __launch_bounds__(256) void k_pauli_apply(const typename V2<R>::T *in, typename V2<R>::T *out,
                                                     uint64_t items, TermK t) {
    using T = typename V2<R>::T;
    T c; c.x = R(t.c[0]); c.y = R(t.c[1]);
    for (uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < items; k += uint64_t(gridDim.x) * blockDim.x) {
        if (t.x == 0) {
            T a = in[k];
            out[k] = cmul(rot(a, quarter(k, t.z, t.q0)), c);
        } else {
            uint64_t i = insert_zero(k, t.pivot), j = i ^ t.x;
            T a = in[i], b = in[j];
            out[i] = cmul(rot(b, quarter(i, t.z, t.q0)), c);
            out[j] = cmul(rot(a, quarter(j, t.z, t.q0)), c);
        }
    }
}

// psi <- psi * cosh(alpha) + (P psi) * sinh(alpha)   (pauli_string.rs:250-261)
template <typename R>
__global__ __launch_bounds__(256) void k_pauli_exp(typename V2<R>::T *psi, uint64_t items, TermK t) {
    using T = typename V2<R>::T;
    T ch, sh;
    ch.x = R(t.ch[0]); ch.y = R(t.ch[1]); sh.x = R(t.sh[0]); sh.y = R(t.sh[1]);
    for (uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < items; k += uint64_t(gridDim.x) * blockDim.x) {
        if (t.x == 0) {
            T a = psi[k];
            psi[k] = cadd(cmul(a, ch), cmul(rot(a, quarter(k, t.z, t.q0)), sh));
        } else {
            uint64_t i = insert_zero(k, t.pivot), j = i ^ t.x;
            T a = psi[i], b = psi[j];
            psi[i] = cadd(cmul(a, ch), cmul(rot(b, quarter(i, t.z, t.q0)), sh));
            psi[j] = cadd(cmul(b, ch), cmul(rot(a, quarter(j, t.z, t.q0)), sh));
        }
    }
}

// out_i = sum_t (P_t in)_i * c_t, folded in term order (SumOp::apply).
template <typename R>
__global__ __launch_bounds__(256) void k_sumop_apply(const typename V2<R>::T *__restrict__ in,
                                                     typename V2<R>::T *__restrict__ out, uint64_t dim,
                                                     const TermK *__restrict__ terms, int nt) {
    using T = typename V2<R>::T;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < dim; i += uint64_t(gridDim.x) * blockDim.x) {
        T acc;
        for (int k = 0; k < nt; k++) {
            T c; c.x = R(terms[k].c[0]); c.y = R(terms[k].c[1]);
            T v = cmul(rot(in[i ^ terms[k].x], quarter(i, terms[k].z, terms[k].q0)), c);
            acc = k == 0 ? v : cadd(acc, v);
        }
        out[i] = acc;
    }
}

// Expectation of a group of terms sharing the same x mask:
//   sum_i conj(psi_i) psi_{i^x} h(i),  h(i) = sum_t c_t i^{q_t(i)}.
// FP64 accumulation with two-sum compensation; one partial per block.
template <typename T>
__global__ __launch_bounds__(256) void k_expect_group(const T *__restrict__ psi, uint64_t items, uint64_t x,
                                                      unsigned pivot, const TermK *__restrict__ terms, int nt,
                                                      double *__restrict__ part) {
    __shared__ double sh[4 * 4];
    DD v[2] = {dd_zero(), dd_zero()};
    for (uint64_t k = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < items; k += uint64_t(gridDim.x) * blockDim.x) {
        uint64_t i = x == 0 ? k : insert_zero(k, pivot);
        uint64_t j = i ^ x;
        double ar = psi[i].x, ai = psi[i].y, br = psi[j].x, bi = psi[j].y;
        // w = conj(a) * b for index i; for the partner j it is conj(w)
        double wr = ar * br + ai * bi, wi = ar * bi - ai * br;
        double hr = 0, hi = 0, gr = 0, gi = 0;
        for (int t = 0; t < nt; t++) {
            unsigned qi_ = quarter(i, terms[t].z, terms[t].q0) & 3;
            double cr = terms[t].c[0], ci = terms[t].c[1];
            // c * i^q
            double pr = qi_ == 0 ? cr : qi_ == 1 ? -ci : qi_ == 2 ? -cr : ci;
            double pi = qi_ == 0 ? ci : qi_ == 1 ? cr : qi_ == 2 ? -ci : -cr;
            hr += pr; hi += pi;
            if (x != 0) {
                unsigned qj = quarter(j, terms[t].z, terms[t].q0) & 3;
                double qr = qj == 0 ? cr : qj == 1 ? -ci : qj == 2 ? -cr : ci;
                double qi2 = qj == 0 ? ci : qj == 1 ? cr : qj == 2 ? -ci : -cr;
                gr += qr; gi += qi2;
            }
        }
        // conj(psi_i) psi_j h(i)
        dd_acc(v[0], wr * hr - wi * hi);
        dd_acc(v[1], wr * hi + wi * hr);
        if (x != 0) {  // conj(psi_j) psi_i h(j) = conj(w) h(j)
            dd_acc(v[0], wr * gr + wi * gi);
            dd_acc(v[1], wr * gi - wi * gr);
        }
    }
    block_reduce_dd<2>(v, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x * 4 + 0] = v[0].s; part[blockIdx.x * 4 + 1] = v[0].c;
        part[blockIdx.x * 4 + 2] = v[1].s; part[blockIdx.x * 4 + 3] = v[1].c;
    }
}

// A run of diagonal Pauli terms (x = 0) applied in one pass: amplitude i is
// multiplied by T_b[bits of i on the qubits of table b] for each table b, a
// product over the run's terms of cosh(a_t) + (-1)^{popcount(i & z_t)} sinh(a_t)
// computed on the host in FP64 (apply_exp_factor of a diagonal string,
// pauli_string.rs:237-262). Tables live in LDS for the whole launch.
template <typename R>
__global__ __launch_bounds__(256) void k_diag_tables(typename V2<R>::T *psi, uint64_t dim, DiagPass p,
                                                     const double2 *__restrict__ tables) {
    using T = typename V2<R>::T;
    extern __shared__ double2 lut[];  // p.entries of them
    for (int e = threadIdx.x; e < p.entries; e += blockDim.x) lut[e] = tables[e];
    __syncthreads();
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < dim; i += uint64_t(gridDim.x) * blockDim.x) {
        const T a = psi[i];
        double re = a.x, im = a.y;
        for (int b = 0; b < p.ntab; b++) {
            uint32_t idx = 0;
            for (int k = 0; k < p.tab[b].bits; k++) idx |= uint32_t((i >> p.tab[b].pos[k]) & 1u) << k;
            const double2 f = lut[p.tab[b].off + idx];
            const double nr = re * f.x - im * f.y, ni = re * f.y + im * f.x;
            re = nr; im = ni;
        }
        T r; r.x = R(re); r.y = R(im);
        psi[i] = r;
    }
}

// Sequential sum of the group results in group order (the order the host
// sums them in sumop_expectation), so both entry points return the same bits.
__global__ void k_fold_groups(const double *__restrict__ g, int ng, double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double sr = 0.0, si = 0.0;
    for (int k = 0; k < ng; k++) { sr += g[2 * k]; si += g[2 * k + 1]; }
    out[0] = sr;
    out[1] = si;
}

TermK make_term(const qi_pauli_term &p) {
    TermK t{};
    t.x = p.x_mask; t.z = p.z_mask;
    unsigned ny = (unsigned)__builtin_popcountll(p.x_mask & p.z_mask);
    t.q0 = (3u * ny) & 3u;
    t.pivot = p.x_mask ? (unsigned)__builtin_ctzll(p.x_mask) : 0;
    t.c[0] = p.c_re; t.c[1] = p.c_im;
    return t;
}

}  // namespace

Status launch_pauli_apply(const State &in, State &out, const qi_pauli_term &p, hipStream_t st) {
    TermK t = make_term(p);
    uint64_t dim = uint64_t(1) << in.n;
    uint64_t items = t.x ? dim / 2 : dim;
    int blocks = pick_blocks(items, 256);
    if (in.prec == QI_C128)
        hipLaunchKernelGGL(k_pauli_apply<double>, dim3(blocks), dim3(256), 0, st, (const double2 *)in.d, (double2 *)out.d, items, t);
    else
        hipLaunchKernelGGL(k_pauli_apply<float>, dim3(blocks), dim3(256), 0, st, (const float2 *)in.d, (float2 *)out.d, items, t);
    QI_HIP_TRY(hipGetLastError());
    return Status{};
}

Status launch_pauli_exp(State &s, const qi_pauli_term &p, double ch_re, double ch_im, double sh_re, double sh_im,
                        hipStream_t st) {
    TermK t = make_term(p);
    t.ch[0] = ch_re; t.ch[1] = ch_im; t.sh[0] = sh_re; t.sh[1] = sh_im;
    uint64_t dim = uint64_t(1) << s.n;
    uint64_t items = t.x ? dim / 2 : dim;
    int blocks = pick_blocks(items, 256);
    if (s.prec == QI_C128)
        hipLaunchKernelGGL(k_pauli_exp<double>, dim3(blocks), dim3(256), 0, st, (double2 *)s.d, items, t);
    else
        hipLaunchKernelGGL(k_pauli_exp<float>, dim3(blocks), dim3(256), 0, st, (float2 *)s.d, items, t);
    QI_HIP_TRY(hipGetLastError());
    return Status{};
}

Status launch_diag_tables(State &s, const DiagPass &p, const double2 *dtables, hipStream_t st) {
    uint64_t dim = uint64_t(1) << s.n;
    int blocks = pick_blocks(dim, 256 * 4);
    const size_t lds = size_t(p.entries) * sizeof(double2);
    if (p.entries < 1 || p.entries > (kDiagTabMax << kDiagTabBits) || p.ntab < 1 || p.ntab > kDiagTabMax)
        return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "diagonal table pass");
    if (s.prec == QI_C128)
        hipLaunchKernelGGL(k_diag_tables<double>, dim3(blocks), dim3(256), lds, st, (double2 *)s.d, dim, p, dtables);
    else
        hipLaunchKernelGGL(k_diag_tables<float>, dim3(blocks), dim3(256), lds, st, (float2 *)s.d, dim, p, dtables);
    QI_HIP_TRY(hipGetLastError());
    return Status{};
}

Status launch_sumop_apply(const State &in, State &out, const qi_pauli_term *terms, size_t nt, hipStream_t st) {
    Ctx &C = *in.ctx;
    std::vector<TermK> h(nt);
    for (size_t k = 0; k < nt; k++) h[k] = make_term(terms[k]);
    Status s = ensure(C.ops, nt * sizeof(TermK));
    if (!s.ok()) return s;
    QI_HIP_TRY(hipMemcpyAsync(C.ops.p, h.data(), nt * sizeof(TermK), hipMemcpyHostToDevice, st));
    uint64_t dim = uint64_t(1) << in.n;
    int blocks = pick_blocks(dim, 256);
    if (in.prec == QI_C128)
        hipLaunchKernelGGL(k_sumop_apply<double>, dim3(blocks), dim3(256), 0, st, (const double2 *)in.d, (double2 *)out.d,
                           dim, (const TermK *)C.ops.p, (int)nt);
    else
        hipLaunchKernelGGL(k_sumop_apply<float>, dim3(blocks), dim3(256), 0, st, (const float2 *)in.d, (float2 *)out.d,
                           dim, (const TermK *)C.ops.p, (int)nt);
    QI_HIP_TRY(hipGetLastError());
    QI_HIP_TRY(hipStreamSynchronize(st));  // the term table is reused by the next call
    return Status{};
}

Status sumop_expectation(State &s, const qi_pauli_term *terms, size_t nt, double *re, double *im, hipStream_t st) {
    Ctx &C = *s.ctx;
    // group terms by x mask, keeping first-appearance order of the groups
    std::vector<uint64_t> xs;
    std::vector<std::vector<TermK>> groups;
    for (size_t k = 0; k < nt; k++) {
        size_t g = 0;
        while (g < xs.size() && xs[g] != terms[k].x_mask) g++;
        if (g == xs.size()) { xs.push_back(terms[k].x_mask); groups.emplace_back(); }
        groups[g].push_back(make_term(terms[k]));
    }
    size_t ng = xs.size();
    std::vector<TermK> flat;
    std::vector<size_t> off(ng + 1, 0);
    for (size_t g = 0; g < ng; g++) { off[g] = flat.size(); flat.insert(flat.end(), groups[g].begin(), groups[g].end()); }
    off[ng] = flat.size();
    uint64_t dim = uint64_t(1) << s.n;
    const int blocks = (int)std::min<uint64_t>(1024, (dim + 2047) / 2048);
    Status e = ensure(C.ops, flat.size() * sizeof(TermK));
    if (!e.ok()) return e;
    e = ensure(C.red, ng * size_t(blocks) * 4 * sizeof(double));
    if (!e.ok()) return e;
    e = ensure(C.red_out, ng * 2 * sizeof(double) + 64);
    if (!e.ok()) return e;
    if (C.host_pinned_words < ng * 2) {
        if (C.host_pinned) (void)hipHostFree(C.host_pinned);
        C.host_pinned = nullptr;
        QI_HIP_TRY(hipHostMalloc((void **)&C.host_pinned, ng * 2 * sizeof(double) + 64));
        C.host_pinned_words = ng * 2 + 8;
    }
    QI_HIP_TRY(hipMemcpyAsync(C.ops.p, flat.data(), flat.size() * sizeof(TermK), hipMemcpyHostToDevice, st));
    const TermK *dterms = (const TermK *)C.ops.p;
    double *part = (double *)C.red.p, *out = (double *)C.red_out.p;
    for (size_t g = 0; g < ng; g++) {
        uint64_t x = xs[g];
        uint64_t items = x ? dim / 2 : dim;
        unsigned pivot = x ? (unsigned)__builtin_ctzll(x) : 0;
        double *pg = part + g * size_t(blocks) * 4;
        if (s.prec == QI_C128)
            hipLaunchKernelGGL(k_expect_group<double2>, dim3(blocks), dim3(256), 0, st, (const double2 *)s.d, items, x, pivot,
                               dterms + off[g], (int)(off[g + 1] - off[g]), pg);
        else
            hipLaunchKernelGGL(k_expect_group<float2>, dim3(blocks), dim3(256), 0, st, (const float2 *)s.d, items, x, pivot,
                               dterms + off[g], (int)(off[g + 1] - off[g]), pg);
        hipLaunchKernelGGL(k_reduce_partials<2>, dim3(1), dim3(256), 0, st, pg, blocks, out + 2 * g);
    }
    QI_HIP_TRY(hipGetLastError());
    QI_HIP_TRY(hipMemcpyAsync(C.host_pinned, out, ng * 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    QI_HIP_TRY(hipStreamSynchronize(st));
    double sr = 0.0, si = 0.0;
    for (size_t g = 0; g < ng; g++) { sr += C.host_pinned[2 * g]; si += C.host_pinned[2 * g + 1]; }
    *re = sr; *im = si;
    return Status{};
}

// ---------------------------------------------------- no-allocation variant --
// qi_sumop_expectation_noalloc keeps its term table, block partials and group
// results in this fixed device array instead of the context's growable
// buffers: no device allocation on any call. Each device that the library
// runs on holds its own copy until the process ends. Every location a call
// reads is written earlier in the same call, so nothing needs clearing; calls
// that overlap on one device (different contexts, or threads without the
// context's lock) race on it.
namespace {
constexpr int kNoAllocBlocks = 1024;
struct NoAllocScratch {
    TermK terms[kNoAllocMaxTerms];
    double part[kNoAllocBlocks * 4];
    double group[kNoAllocMaxTerms * 2];
    double total[2];
};
}  // namespace

static __device__ __align__(256) unsigned char scratch[sizeof(NoAllocScratch)];

Status noalloc_scratch(int device, void **out) {
    static std::mutex mu;
    static void *ptrs[64] = {nullptr};
    if (device < 0 || device >= 64) return make_err(QI_ERR_INVALID_ARGUMENT, 0, 0, 0, 0, "device index");
    std::lock_guard<std::mutex> lk(mu);
    if (!ptrs[device]) {
        // the host-side name of `scratch` is not a device address: take the
        // address of this device's copy once, with the device current
        void *p = nullptr;
        QI_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptrs[device] = p;
    }
    *out = ptrs[device];
    return Status{};
}

Status sumop_expectation_noalloc(State &s, const qi_pauli_term *terms, size_t nt, double *re, double *im, hipStream_t st) {
    void *base = s.ctx->noalloc;
    if (!base) return make_err(QI_ERR_DEVICE, 0, 0, 0, 0, "no scratch address for this device");
    if (nt > (size_t)kNoAllocMaxTerms)
        return make_err(QI_ERR_INVALID_ARGUMENT, nt, kNoAllocMaxTerms, 0, 0, "more terms than the no-allocation maximum (512)");
    *re = 0.0; *im = 0.0;
    if (nt == 0) return Status{};
    NoAllocScratch *sc = reinterpret_cast<NoAllocScratch *>(base);
    // group by x mask in first-appearance order, as sumop_expectation does (host stack only)
    uint64_t xs[kNoAllocMaxTerms];
    uint16_t gid[kNoAllocMaxTerms];
    uint32_t cnt[kNoAllocMaxTerms + 1];
    TermK flat[kNoAllocMaxTerms];
    size_t ng = 0;
    for (size_t k = 0; k < nt; k++) {
        size_t g = 0;
        while (g < ng && xs[g] != terms[k].x_mask) g++;
        if (g == ng) { xs[ng] = terms[k].x_mask; cnt[ng] = 0; ng++; }
        gid[k] = (uint16_t)g;
        cnt[g]++;
    }
    uint32_t off[kNoAllocMaxTerms + 1];
    off[0] = 0;
    for (size_t g = 0; g < ng; g++) off[g + 1] = off[g] + cnt[g];
    uint32_t fill[kNoAllocMaxTerms];
    for (size_t g = 0; g < ng; g++) fill[g] = off[g];
    for (size_t k = 0; k < nt; k++) flat[fill[gid[k]]++] = make_term(terms[k]);
    QI_HIP_TRY(hipMemcpyAsync(sc->terms, flat, nt * sizeof(TermK), hipMemcpyHostToDevice, st));
    const uint64_t dim = uint64_t(1) << s.n;
    const int blocks = (int)std::min<uint64_t>(kNoAllocBlocks, (dim + 2047) / 2048);
    for (size_t g = 0; g < ng; g++) {
        const uint64_t x = xs[g];
        const uint64_t items = x ? dim / 2 : dim;
        const unsigned pivot = x ? (unsigned)__builtin_ctzll(x) : 0;
        if (s.prec == QI_C128)
            hipLaunchKernelGGL(k_expect_group<double2>, dim3(blocks), dim3(256), 0, st, (const double2 *)s.d, items, x, pivot,
                               (const TermK *)(sc->terms + off[g]), (int)(off[g + 1] - off[g]), sc->part);
        else
            hipLaunchKernelGGL(k_expect_group<float2>, dim3(blocks), dim3(256), 0, st, (const float2 *)s.d, items, x, pivot,
                               (const TermK *)(sc->terms + off[g]), (int)(off[g + 1] - off[g]), sc->part);
        // the partials are reused by the next group: this stream orders it after the fold
        hipLaunchKernelGGL(k_reduce_partials<2>, dim3(1), dim3(256), 0, st, (const double *)sc->part, blocks, sc->group + 2 * g);
    }
    hipLaunchKernelGGL(k_fold_groups, dim3(1), dim3(64), 0, st, (const double *)sc->group, (int)ng, sc->total);
    QI_HIP_TRY(hipGetLastError());
    double res[2];
    QI_HIP_TRY(hipMemcpyAsync(res, sc->total, sizeof res, hipMemcpyDeviceToHost, st));
    QI_HIP_TRY(hipStreamSynchronize(st));  // also retires the copy from `flat` on this stack
    *re = res[0];
    *im = res[1];
    return Status{};
}

}  // namespace qi
