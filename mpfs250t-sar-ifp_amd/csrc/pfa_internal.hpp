// pfa_internal.hpp -- context, plan and launcher declarations shared by the
// host API (pfa_api.cpp) and the kernel translation units (*.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "mi355x_pfa.h"

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

struct TimerMark {
    int kind;
    hipEvent_t start, stop;
};

struct pfa_hostio;                       // staged host copies (pfa_hostio.cpp)

struct pfa_ctx {
    int device = 0;
    int num_cu = 256;                     // compute units of the device
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    std::map<std::string, DevBuf> ws;     // grow-only named device workspace
    bool timing = false;
    std::vector<TimerMark> marks;
    std::vector<hipEvent_t> event_pool;
    int *d_counters = nullptr;            // [0]=emax pass A, [1]=emax pass B, [2..] scratch
    unsigned *d_maxbits = nullptr;        // quantiser max |x| as float bits
    pfa_hostio *io = nullptr;             // created on the first large host copy
    bool io_failed = false;               // staging unavailable: copies go direct
    std::set<pfa_plan *> plans;           // live plans (orphaned by pfa_destroy)
    bool no_timers = false;               // set while an allocation-free call runs
    void *scratch = nullptr;              // this device's copy of the static scratch array
};

enum PlanKind { PLAN_FIXED = 1, PLAN_FLOAT = 2 };

// Internal flag of a fixed plan (beside the PFA_FIXED_* flags): the sinc
// table's half sums allow the one-dot-product-per-pair gather.
constexpr int FIXED_SINC_SPLIT = 0x100;
// Words per half of the packed float-path tap table (20 rows of 256 phases and
// a 32-word bank offset); the device code stages it, the host builds the global
// copy the 16384-point azimuth lines read.
constexpr int PFA_TAP2_HALF = 20 * 256 + 32;


struct pfa_plan {
    PlanKind kind;
    pfa_ctx *ctx;
    int m = 0, n = 0, grid = 0, flags = 0;
    // fixed-mode device tables
    float *a32 = nullptr, *f0 = nullptr, *df = nullptr, *tan_s = nullptr;
    float *krp = nullptr, *kcp = nullptr;
    int32_t *order = nullptr, *hamr = nullptr, *hamc = nullptr;
    int16_t *sinc = nullptr;
    const void *tw16 = nullptr;           // grid-point Q15 twiddles (short2[grid / 2])
    // float-mode device tables
    int sgn = -1, m2 = 0, n2 = 0;
    double *pr64 = nullptr, *a64 = nullptr, *tan_s64 = nullptr, *KR = nullptr, *KC = nullptr;
    double *f0d = nullptr, *dfd = nullptr, *hm = nullptr, *hn = nullptr;
    // KR / KC as np.linspace parameters {start, step, stop} when they are
    // exactly such grids (the kernels then form them in registers)
    double kr_lin_v[3] = {0, 0, 0}, kc_lin_v[3] = {0, 0, 0};
    const double *kr_lin = nullptr, *kc_lin = nullptr;
    // bracket guesses of the azimuth pass: tan_bin[b] = #{p : tan_s[p] < tb_lo + b / tb_inv_w},
    // b = 0..tb_bins (a starting point only; the kernels fix brackets up exactly)
    int32_t *tan_bin = nullptr;
    unsigned *taps2 = nullptr;            // packed Q15 tap pairs in global memory (16384-point lines)
    bool af_ready = false;                // autofocus checkpoint (range pass) computed
    void *af_g1t = nullptr;               // that checkpoint: the plan's own corner-turned range pass
    int tb_bins = 0;
    double tb_lo = 0.0, tb_inv_w = 0.0;
    std::vector<void *> owned;            // device allocations of the plan
};

// ---------------------------------------------------------------- helpers
// Exceptions never cross the C ABI: every entry point runs its body between
// PFA_TRY and PFA_CATCH(ctx).  Host memory exhaustion (std::bad_alloc) and a
// host resource failure such as a thread that cannot start (std::system_error)
// return PFA_ERR_NOMEM; any other exception returns PFA_ERR_STATE.
int pfa_exception(pfa_ctx *ctx) noexcept;
#define PFA_TRY try {
#define PFA_CATCH(ctx)                     \
    }                                      \
    catch (...) {                          \
        return pfa_exception(ctx);         \
    }
int pfa_fail(pfa_ctx *ctx, int code, const std::string &msg);
int pfa_fail(pfa_ctx *ctx, int code, const char *msg) noexcept;
int pfa_hip_fail(pfa_ctx *ctx, hipError_t e, const char *what);
void *pfa_ws(pfa_ctx *ctx, const char *slot, size_t bytes, int *status);
void pfa_timer_begin(pfa_ctx *ctx, int kind);
void pfa_timer_end(pfa_ctx *ctx, int kind);
// Host <-> device copies of the host-buffer entry points: large ones staged
// through pinned buffers on a copy stream (pfa_hostio.cpp).  h2d leaves the
// context stream waiting for the data; d2h returns when the host has it.
int pfa_io_h2d(pfa_ctx *ctx, void *dst, const void *src, size_t bytes);
// Address of this device's static scratch array (pfa_noalloc.hip); called by
// pfa_create with the context's device current.
int pfa_scratch_init(pfa_ctx *ctx);
int pfa_io_d2h(pfa_ctx *ctx, void *dst, const void *src, size_t bytes);
void pfa_hostio_destroy(pfa_ctx *ctx);

#define CK(x)                        \
    do {                             \
        int s_ = (x);                \
        if (s_ != PFA_OK) return s_; \
    } while (0)

#define PFA_HIP(ctx, call)                                        \
    do {                                                          \
        hipError_t e_ = (call);                                   \
        if (e_ != hipSuccess) return pfa_hip_fail((ctx), e_, #call); \
    } while (0)

// ---------------------------------------------------------------- fixed launchers
struct FixedFFTArgs {
    const void *in;          // short2 rows (row-major, ld = in_ld)
    int in_ld;
    void *out;               // short2 rows (ld = n)
    int32_t *e_row;
    int *emax;               // may be null
    int rows, n;
};
int launch_fft_bfp_rows(pfa_ctx *ctx, const FixedFFTArgs &a);
int launch_renorm_rows(pfa_ctx *ctx, void *y, const int32_t *e_row, const int *emax,
                       int rows, int n);
int launch_window_fixed(pfa_ctx *ctx, const void *g2, int rows, int cols,
                        const int32_t *hamr, const int32_t *hamc, void *out);
int launch_detect_fixed(pfa_ctx *ctx, const void *z, int64_t count, uint16_t *out);
int launch_quantize(pfa_ctx *ctx, const float *sig, int64_t count, int16_t *out, int *exp_out);
int launch_zero_i32(pfa_ctx *ctx, int *p, int count);

int launch_fixed_range(pfa_plan *p, const void *sig, void *g1t, int ldp);
int launch_fixed_azimuth_fft(pfa_plan *p, const void *g1t, int ldp, void *ya,
                             int32_t *ea, int *emax_a, bool do_fft, bool do_window);
int launch_fixed_fft_b(pfa_plan *p, const void *ya, const int32_t *ea, const int *emax_a,
                       void *yb, int32_t *eb, int *emax_b);
int launch_fixed_detect_t(pfa_plan *p, const void *yb, const int32_t *eb, const int *emax_b,
                          uint16_t *out);

// ---------------------------------------------------------------- float launchers
int launch_float_range(pfa_plan *p, const void *sig, const double *freq, void *g1t, int ldp);
// rot: optional per sorted pulse e^{j phi} applied to the pass-1 rows (autofocus)
int launch_float_azimuth(pfa_plan *p, int mode, const void *g1t, int ldp, const void *g2in,
                         void *out, const void *rot = nullptr);
// gradient autofocus (pfa_af.hip)
int launch_af_rot(pfa_plan *p, const float *phi, void *rot);
int launch_af_stats(pfa_plan *p, const void *img, double *partial, int nparts, double *stats);
int launch_af_fftb_adj(pfa_plan *p, const void *img, const double *stats, void *B);
int launch_af_az_adj(pfa_plan *p, const void *g1t, int ldp, const void *rot, const void *B,
                     float *partial, int *nparts);
int launch_af_grad_reduce(pfa_plan *p, const float *partial, int nparts, const double *stats,
                          float *grad);
int launch_float_fft_b(pfa_plan *p, const void *A, void *out, int out_kind);
int launch_kr_extrema(pfa_ctx *ctx, const double *freq, const double *pr, int m, int n,
                      double *row_max, double *row_min);
// matrix-core FFT form; fk: 1 BF16, 2 FP16, 3 BF16 with hi/lo DFT matrices
int launch_float_azimuth_mfma(pfa_plan *p, const void *g1t, int ldp, unsigned *A16, int *expA,
                              int fk);
int launch_float_fft_b_mfma(pfa_plan *p, const unsigned *A16, const int *expA, void *out,
                            int out_kind, int fk);
