// pfa_noalloc.hip -- the allocation-free variant of the fixed-point focus
// (pfa_form_image_fixed_dev_noalloc) and the static scratch array it runs in.
//
// The fixed focus needs device memory beyond its input and output: the
// corner-turned range pass (reused for the range-axis FFT's output), the
// cross-axis FFT's output, the per-row exponents when the caller does not ask
// for them, and the two pass maxima.  The allocating entry point takes them
// from the context's grow-only workspace (hipMalloc on first use); this one
// takes them from `scratch`, a file-scope device array every GPU that loads
// the library holds until the process ends, sized for grids up to
// PFA_NOALLOC_MAX_GRID.  Calls on one device share the array: each call
// clears what must start at zero (the pass maxima, with a fill kernel) and
// overwrites the rest, and calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "pfa_internal.hpp"

namespace {

constexpr int MAX_GRID = PFA_NOALLOC_MAX_GRID;
constexpr size_t PLANE = (size_t)MAX_GRID * MAX_GRID * 4;          // short2 [grid][grid]
constexpr size_t OFF_G1T = 0;                                      // range pass, then FFT-B output
constexpr size_t OFF_YA = PLANE;                                   // FFT-A output
constexpr size_t OFF_EA = 2 * PLANE;                               // int32 [grid] x 2
constexpr size_t OFF_EMAX = OFF_EA + 2 * (size_t)MAX_GRID * 4;     // int32 [2] pass maxima
constexpr size_t SCRATCH_BYTES = OFF_EMAX + 256;

__device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void *g_addr[MAX_DEVICES] = {};      // each device's own address of `scratch`

}  // namespace

int pfa_scratch_init(pfa_ctx *ctx) {
    if (ctx->device < 0 || ctx->device >= MAX_DEVICES) return PFA_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_addr[ctx->device]) {                      // the context's device is current
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) return pfa_hip_fail(ctx, e, "scratch symbol address");
        g_addr[ctx->device] = ptr;
    }
    ctx->scratch = g_addr[ctx->device];
    return PFA_OK;
}

extern "C" int pfa_form_image_fixed_dev_noalloc(pfa_plan *p, const int16_t *sig, uint16_t *out,
                                                int32_t *exp_a, int32_t *exp_b) {
    PFA_TRY
    if (!p || !p->ctx || p->kind != PLAN_FIXED) return PFA_ERR_STATE;
    pfa_ctx *ctx = p->ctx;
    if (!sig || !out) return pfa_fail(ctx, PFA_ERR_ARG, "form_image_noalloc: null pointer");
    if (p->grid > MAX_GRID)
        return pfa_fail(ctx, PFA_ERR_DIMS, "form_image_noalloc: grid above PFA_NOALLOC_MAX_GRID "
                                           "(use pfa_form_image_fixed_dev)");
    if (!ctx->scratch || !p->tw16) return pfa_fail(ctx, PFA_ERR_STATE, "form_image_noalloc: no scratch");
    unsigned char *base = (unsigned char *)ctx->scratch;   // never the array's own name here
    const int G = p->grid, ldp = (p->m + 63) / 64 * 64;
    void *g1t = base + OFF_G1T;
    void *ya = base + OFF_YA;
    int32_t *ea = exp_a ? exp_a : (int32_t *)(base + OFF_EA);
    int32_t *eb = exp_b ? exp_b : (int32_t *)(base + OFF_EA) + G;
    int *emax = (int *)(base + OFF_EMAX);
    struct NoTimers {                             // no timer events are created here
        pfa_ctx *c;
        explicit NoTimers(pfa_ctx *c_) : c(c_) { c->no_timers = true; }
        ~NoTimers() { c->no_timers = false; }
    } guard(ctx);
    CK(launch_zero_i32(ctx, emax, 2));
    CK(launch_fixed_range(p, sig, g1t, ldp));
    CK(launch_fixed_azimuth_fft(p, g1t, ldp, ya, ea, emax, true, true));
    CK(launch_fixed_fft_b(p, ya, ea, emax, g1t, eb, emax + 1));
    CK(launch_fixed_detect_t(p, g1t, eb, emax + 1, out));
    return PFA_OK;
    PFA_CATCH((p ? p->ctx : nullptr))
}
