// qi_internal.h — shared declarations of the gfx950 state-vector engine.
// Host orchestration (api.hip) talks to the kernel files through the launcher
// functions declared here; nothing in this header crosses the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/qiron.h"

namespace qi {

// ---------------------------------------------------------------- errors --
struct Status {
    int code = QI_OK;
    uint64_t a = 0, b = 0;
    double re = 0, im = 0;
    std::string msg;
    bool ok() const { return code == QI_OK; }
};
Status make_err(int code, uint64_t a = 0, uint64_t b = 0, double re = 0, double im = 0,
                const char *detail = nullptr);
void fill_err(qi_error *err, const Status &st);
// A failed device allocation: HIP keeps the error as its last error, which the
// next launch check would report, so it is read (cleared) here.
inline Status oom_err(hipError_t e) {
    (void)hipGetLastError();
    return make_err(QI_ERR_OUT_OF_MEMORY, 0, 0, 0, 0, hipGetErrorString(e));
}

#define QI_HIP_TRY(expr)                                                          \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess)                                                     \
            return ::qi::make_err(QI_ERR_DEVICE, 0, 0, 0, 0, hipGetErrorString(e_)); \
    } while (0)

// --------------------------------------------------------- device buffers --
struct Ctx;
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// Scratch that grows on demand (reductions, histograms, schedules).
Status ensure(DevBuf &b, size_t bytes);

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    DevBuf red;      // reduction partials
    DevBuf red_out;  // reduction results (device)
    DevBuf ops;      // per-call op tables (tiles, Pauli terms)
    DevBuf probs;    // measurement histograms / CDF
    DevBuf aux;      // misc (sample outcomes, uniforms)
    DevBuf plan;     // op tables and fused matrices of the last gate-list plan
    DevBuf trot;     // phase tables of the diagonal runs of a Trotter step
    DevBuf alt;      // second state buffer of out-of-place tiled plans (grows with the largest such state)
    size_t last_plan_steps = 0;
    double *host_pinned = nullptr;  // pinned host words for results
    size_t host_pinned_words = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_exec_ms = 0.0;
    int cu_count = 0;
    // SplitMix64 stream for measurements whose uniform draw is not injected
    // (the reference draws from rand::rng(), state.rs:601-602)
    uint64_t rng = 0;
    void *noalloc = nullptr;  // this device's copy of the no-allocation scratch array (pauli.hip)
    // cached graph of the last captured gate list
    hipGraphExec_t graph_exec = nullptr;
    std::vector<unsigned char> graph_key;
    uint64_t graph_state = 0;
    // the last gate-list plan (an ExecPlan), whose op tables are in `plan`, and its key
    std::shared_ptr<void> cached_plan;
    std::vector<unsigned char> plan_key;
};

struct State {
    Ctx *ctx = nullptr;
    uint32_t n = 0;
    int prec = QI_C128;
    void *d = nullptr;  // 2^n complex amplitudes (float2 or double2)
    size_t bytes() const { return (size_t(1) << n) * (prec == QI_C128 ? 16 : 8); }
};

// ------------------------------------------------------------ gate kernels --
// A gate lowered for the streaming kernels. The host resolves all constants
// (cos/sin of angles, phases) in FP64 with the host libm, so device results in
// FP64 reproduce the reference's arithmetic.
enum LoweredKind : int {
    LK_H = 0, LK_X, LK_Y, LK_RX, LK_RY, LK_U2,      // pair kernels
    LK_Z, LK_PHASE,                                  // target-1 diagonal kernels
    LK_RZ,                                           // two-phase diagonal
    LK_SWAP,                                         // permutation of two targets
    LK_MATCH,                                        // matchgate
    LK_NOP
};

struct Lowered {
    int kind = LK_NOP;
    uint32_t t0 = 0, t1 = 0;        // targets
    uint64_t ctrl_mask = 0;         // controls (must be 1)
    double c[8] = {0};              // coefficients (kind-specific, see gates.hip)
    int32_t slot = -1;              // parametric gate (RX, RY, RZ, P): its angle slot; the planner keeps
                                    // such gates in forms whose coefficients can be rewritten later
};

// Argument checks of Operator::apply / Gate::apply (validate_qubits and the
// per-operator checks, operator.rs:214-273 and siblings).
Status validate_gate(uint32_t n, const qi_gate &g);
Status lower_gate(uint32_t n, const qi_gate &g, Lowered &out);
Status launch_gate(State &s, const Lowered &L, hipStream_t st);

// ------------------------------------------------------- misc kernels ------
Status launch_init(State &s, int kind, uint64_t basis, hipStream_t st);
Status launch_scale(State &s, double re, double im, hipStream_t st);
Status launch_axpy(State &y, double re, double im, const State &x, hipStream_t st);
// out <- a + b (sub = false) or a - b (sub = true), elementwise; out may alias a or b.
Status launch_addsub(State &out, const State &a, const State &b, bool sub, hipStream_t st);
Status launch_copy(State &dst, const State &src, hipStream_t st);
Status launch_div_real(State &s, double f, hipStream_t st);

// FP64 compensated reductions: partials in ctx.red, the result in ctx.red_out,
// copied to pinned host memory before returning.
enum RedKind : int { RK_NORM = 0, RK_INNER = 1 };
Status reduce_state(State &a, const State *b, int kind, double *out2, hipStream_t st);

// Pauli kernels (pauli.hip)
Status launch_pauli_apply(const State &in, State &out, const qi_pauli_term &t, hipStream_t st);
Status launch_pauli_exp(State &s, const qi_pauli_term &t, double ch_re, double ch_im, double sh_re,
                        double sh_im, hipStream_t st);
Status launch_sumop_apply(const State &in, State &out, const qi_pauli_term *terms, size_t nt,
                          hipStream_t st);
// A run of diagonal Pauli-string exponentials in one pass: per-amplitude
// factors looked up in tables over at most kDiagTabBits qubits each.
constexpr int kDiagTabBits = 10;
constexpr int kDiagTabMax = 4;
struct DiagTab {
    int32_t bits, off;            // qubits in the table, first entry
    uint8_t pos[kDiagTabBits];    // qubit of table-index bit k
    uint8_t pad[6];
};
struct DiagPass {
    int32_t ntab, entries;
    DiagTab tab[kDiagTabMax];
};
Status launch_diag_tables(State &s, const DiagPass &p, const double2 *dtables, hipStream_t st);
// The no-allocation expectation: at most kNoAllocMaxTerms terms, all scratch in
// a static device array whose per-device address noalloc_scratch returns.
constexpr int kNoAllocMaxTerms = 512;
Status noalloc_scratch(int device, void **out);
Status sumop_expectation_noalloc(State &s, const qi_pauli_term *terms, size_t nt, double *re, double *im,
                                 hipStream_t st);
// Expectation of a list of terms; result (re, im) summed in FP64 compensated.
Status sumop_expectation(State &s, const qi_pauli_term *terms, size_t nt, double *re, double *im,
                         hipStream_t st);

// Measurement kernels (measure.hip)
Status measure_probs_device(State &s, const uint64_t *qubits, uint32_t m, double **dprobs,
                            hipStream_t st);
Status collapse_outcome(State &s, State &dst, const uint64_t *qubits, uint32_t m, uint64_t outcome, hipStream_t st);
Status measure_sample_collapse(State &s, State &dst, const uint64_t *qubits, uint32_t m,
                               const double *uniforms, size_t shots, uint64_t *outcomes_host,
                               bool collapse, hipStream_t st);

int pick_blocks(uint64_t items, int per_block);

// Next uniform in [0, 1) of a SplitMix64 stream (Steele, Lea & Flood 2014):
// the top 53 bits of the output word times 2^-53.
inline double splitmix_uniform(uint64_t &state) {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * (1.0 / 9007199254740992.0);
}

// ------------------------------------------------------- workgroup order --
// Workgroups go to the 8 XCDs in turn (workgroup b to XCD b % 8), so with the
// identity order every XCD walks the whole state with stride 8. With c > 0,
// XCD x's r-th workgroup takes work item (r >> c) * 8 * 2^c + x * 2^c +
// (r mod 2^c): each XCD works through runs of 2^c consecutive items. A
// bijection on [0, grid) when 8 * 2^c divides the grid.
__device__ __forceinline__ uint64_t xcd_block_order(uint32_t b, int c) {
    if (c == 0) return b;
    const uint32_t x = b & 7u, r = b >> 3;
    return (uint64_t(r >> c) << (c + 3)) | (uint64_t(x) << c) | (r & ((1u << c) - 1u));
}
// Largest c <= cmax for which xcd_block_order is a bijection on `grid`.
inline int xcd_chunk_log2(uint64_t grid, int cmax) {
    if (grid == 0 || (grid & (grid - 1))) return 0;
    const int lg = __builtin_ctzll(grid) - 3;
    return lg < 0 ? 0 : (lg < cmax ? lg : cmax);
}

}  // namespace qi

struct qi_ctx_s {
    qi::Ctx c;
};
struct qi_state_s {
    qi::State s;
};
