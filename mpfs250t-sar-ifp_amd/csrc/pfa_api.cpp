// pfa_api.cpp -- host side of the C ABI (include/mi355x_pfa.h): contexts,
// workspace, plans, argument validation and the host-pointer wrappers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <system_error>
#include <vector>

#include "pfa_internal.hpp"

// ------------------------------------------------------------------ errors / workspace
int pfa_fail(pfa_ctx *ctx, int code, const char *msg) noexcept {
    if (ctx) {
        try {
            ctx->err = msg;
        } catch (...) {                   // no memory for the message: keep the code
            ctx->err.clear();
        }
    }
    return code;
}

int pfa_fail(pfa_ctx *ctx, int code, const std::string &msg) { return pfa_fail(ctx, code, msg.c_str()); }

int pfa_exception(pfa_ctx *ctx) noexcept {
    try {
        throw;
    } catch (const std::bad_alloc &) {
        return pfa_fail(ctx, PFA_ERR_NOMEM, "host memory exhausted");
    } catch (const std::system_error &e) {
        return pfa_fail(ctx, PFA_ERR_NOMEM, e.what());
    } catch (const std::exception &e) {
        return pfa_fail(ctx, PFA_ERR_STATE, e.what());
    } catch (...) {
        return pfa_fail(ctx, PFA_ERR_STATE, "unexpected exception");
    }
}

int pfa_hip_fail(pfa_ctx *ctx, hipError_t e, const char *what) {
    return pfa_fail(ctx, e == hipErrorOutOfMemory ? PFA_ERR_NOMEM : PFA_ERR_HIP,
                    std::string(what) + ": " + hipGetErrorString(e));
}

void *pfa_ws(pfa_ctx *ctx, const char *slot, size_t bytes, int *status) {
    DevBuf &b = ctx->ws[slot];
    if (b.bytes >= bytes && b.ptr) return b.ptr;
    if (b.ptr) {
        hipStreamSynchronize(ctx->stream);
        hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        *status = pfa_hip_fail(ctx, e, (std::string("workspace ") + slot).c_str());
        return nullptr;
    }
    b.bytes = want;
    return b.ptr;
}

static hipEvent_t take_event(pfa_ctx *ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
}

void pfa_timer_begin(pfa_ctx *ctx, int kind) {
    if (!ctx->timing || ctx->no_timers) return;
    TimerMark mk{kind, take_event(ctx), take_event(ctx)};
    hipEventRecord(mk.start, ctx->stream);
    ctx->marks.push_back(mk);
}

void pfa_timer_end(pfa_ctx *ctx, int kind) {
    if (!ctx->timing || ctx->no_timers || ctx->marks.empty()) return;
    TimerMark &mk = ctx->marks.back();
    if (mk.kind == kind) hipEventRecord(mk.stop, ctx->stream);
}

// ------------------------------------------------------------------ context
extern "C" int pfa_abi_version(void) { return PFA_ABI_VERSION; }

extern "C" int pfa_create(pfa_ctx **out, int device) {
    PFA_TRY
    if (!out) return PFA_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PFA_ERR_HIP;
    if (device < 0 || device >= ndev) return PFA_ERR_ARG;
    pfa_ctx *ctx = new (std::nothrow) pfa_ctx();
    if (!ctx) return PFA_ERR_NOMEM;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->own_stream) != hipSuccess) {
        delete ctx;
        return PFA_ERR_HIP;
    }
    ctx->stream = ctx->own_stream;
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cu > 0)
        ctx->num_cu = cu;
    if (hipMalloc(&ctx->d_counters, 64 * sizeof(int)) != hipSuccess ||
        hipMalloc(&ctx->d_maxbits, 64 * sizeof(unsigned)) != hipSuccess ||
        hipMemset(ctx->d_counters, 0, 64 * sizeof(int)) != hipSuccess) {
        hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return PFA_ERR_NOMEM;
    }
    if (pfa_scratch_init(ctx) != PFA_OK) {
        hipFree(ctx->d_counters);
        hipFree(ctx->d_maxbits);
        hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return PFA_ERR_HIP;
    }
    *out = ctx;
    return PFA_OK;
    PFA_CATCH(nullptr)
}

static int plan_register(pfa_plan *p, pfa_plan **out);

// Releases a plan's device tables (the plan object itself stays until
// pfa_plan_destroy).
static void plan_release(pfa_plan *p) noexcept {
    for (void *d : p->owned) hipFree(d);
    p->owned.clear();
    p->af_g1t = nullptr;
    p->af_ready = false;
}

extern "C" void pfa_destroy(pfa_ctx *ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    // plans still alive are orphaned: their device tables go with the context,
    // and every later call on them returns PFA_ERR_STATE until pfa_plan_destroy
    for (pfa_plan *p : ctx->plans) {
        plan_release(p);
        p->ctx = nullptr;
    }
    ctx->plans.clear();
    pfa_hostio_destroy(ctx);
    for (auto &kv : ctx->ws)
        if (kv.second.ptr) hipFree(kv.second.ptr);
    for (auto &mk : ctx->marks) {
        hipEventDestroy(mk.start);
        hipEventDestroy(mk.stop);
    }
    for (auto e : ctx->event_pool) hipEventDestroy(e);
    hipFree(ctx->d_counters);
    hipFree(ctx->d_maxbits);
    hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

extern "C" const char *pfa_last_error(const pfa_ctx *ctx) {
    return ctx ? ctx->err.c_str() : "null context";
}

extern "C" int pfa_set_stream(pfa_ctx *ctx, void *s) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    ctx->stream = s ? (hipStream_t)s : ctx->own_stream;
    return PFA_OK;
    PFA_CATCH(ctx)
}

extern "C" int pfa_synchronize(pfa_ctx *ctx) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    PFA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return PFA_OK;
    PFA_CATCH(ctx)
}

extern "C" int pfa_timer_enable(pfa_ctx *ctx, int on) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    ctx->timing = on != 0;
    return PFA_OK;
    PFA_CATCH(ctx)
}

extern "C" int pfa_timer_reset(pfa_ctx *ctx) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    PFA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (auto &mk : ctx->marks) {
        ctx->event_pool.push_back(mk.start);
        ctx->event_pool.push_back(mk.stop);
    }
    ctx->marks.clear();
    return PFA_OK;
    PFA_CATCH(ctx)
}

extern "C" int pfa_timer_read(pfa_ctx *ctx, int kind, float *ms_total, int *launches) {
    PFA_TRY
    if (!ctx || !ms_total || !launches || kind < 0 || kind >= PFA_TIMER_KINDS) return PFA_ERR_ARG;
    PFA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float tot = 0.f;
    int cnt = 0;
    for (auto &mk : ctx->marks) {
        if (mk.kind != kind) continue;
        float ms = 0.f;
        PFA_HIP(ctx, hipEventElapsedTime(&ms, mk.start, mk.stop));
        tot += ms;
        ++cnt;
    }
    *ms_total = tot;
    *launches = cnt;
    return PFA_OK;
    PFA_CATCH(ctx)
}

extern "C" int pfa_fallback_counts(pfa_ctx *ctx, uint32_t out[2], int reset) {
    PFA_TRY
    if (!ctx || !out) return PFA_ERR_ARG;
    PFA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    PFA_HIP(ctx, hipMemcpy(out, ctx->d_counters + 32, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (reset) PFA_HIP(ctx, hipMemset(ctx->d_counters + 32, 0, 2 * sizeof(uint32_t)));
    return PFA_OK;
    PFA_CATCH(ctx)
}

// ------------------------------------------------------------------ tables
extern "C" int pfa_twiddle_codes(int n, int bits, int32_t *out) {
    PFA_TRY
    if (n < 2 || (n & (n - 1)) || bits < 2 || bits > 30 || !out) return PFA_ERR_ARG;
    const double scale = std::ldexp(1.0, bits - 1);
    const double full = scale - 1.0;
    for (int k = 0; k < n / 2; ++k) {
        // the angle exactly as numpy forms -2j*pi*k/n: fl(2*pi*k) / n
        const double th = (2.0 * M_PI) * (double)k / (double)n;
        const double re = std::floor(std::cos(th) * scale);
        const double im = std::floor(-std::sin(th) * scale);
        out[2 * k] = (int32_t)std::min(std::max(re, -full - 1.0), full);
        out[2 * k + 1] = (int32_t)std::min(std::max(im, -full - 1.0), full);
    }
    return PFA_OK;
    PFA_CATCH(nullptr)
}

// ------------------------------------------------------------------ helpers
static bool pow2_in(int n, int lo, int hi) { return n >= lo && n <= hi && (n & (n - 1)) == 0; }

template <typename T>
static T *upload(pfa_plan *p, const T *h, size_t count, int *st) {
    void *d = nullptr;
    hipError_t e = hipMalloc(&d, std::max<size_t>(count * sizeof(T), 64));
    if (e != hipSuccess) {
        *st = pfa_hip_fail(p->ctx, e, "plan table");
        return nullptr;
    }
    p->owned.push_back(d);
    if (count) {
        e = hipMemcpy(d, h, count * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            *st = pfa_hip_fail(p->ctx, e, "plan table upload");
            return nullptr;
        }
    }
    return (T *)d;
}

// ------------------------------------------------------------------ fixed: stage functions
extern "C" int pfa_quantize_iq16_dev(pfa_ctx *ctx, const float *sig, int64_t count, int16_t *out,
                                     int *exp_out) {
    PFA_TRY
    if (!ctx || (count && (!sig || !out)) || !exp_out || count < 0) return pfa_fail(ctx, PFA_ERR_ARG, "quantize: bad argument");
    return launch_quantize(ctx, sig, count * 2, out, exp_out);
    PFA_CATCH(ctx)
}

extern "C" int pfa_quantize_iq16(pfa_ctx *ctx, const float *sig, int64_t count, int16_t *out,
                                 int *exp_out) {
    PFA_TRY
    if (!ctx || (count && (!sig || !out)) || !exp_out || count < 0) return pfa_fail(ctx, PFA_ERR_ARG, "quantize: bad argument");
    int st = PFA_OK;
    const size_t in_b = (size_t)count * 8, out_b = (size_t)count * 4;
    float *d_in = (float *)pfa_ws(ctx, "io_in", in_b, &st);
    if (!d_in) return st;
    int16_t *d_out = (int16_t *)pfa_ws(ctx, "io_out", out_b, &st);
    if (!d_out) return st;
    CK(pfa_io_h2d(ctx, d_in, sig, in_b));
    CK(launch_quantize(ctx, d_in, count * 2, d_out, exp_out));
    CK(pfa_io_d2h(ctx, out, d_out, out_b));
    return PFA_OK;
    PFA_CATCH(ctx)
}

static int fft_rows_dev(pfa_ctx *ctx, const int16_t *x, int rows, int n, int16_t *y, int32_t *e,
                        bool renorm) {
    if (!ctx) return PFA_ERR_ARG;
    if (!pow2_in(n, 64, 8192)) return pfa_fail(ctx, PFA_ERR_DIMS, "FFT length must be a power of two in 64..8192");
    if (rows < 0 || (rows && (!x || !y || !e))) return pfa_fail(ctx, PFA_ERR_ARG, "fft: bad argument");
    if (rows == 0) return PFA_OK;
    int *emax = ctx->d_counters + 8;
    CK(launch_zero_i32(ctx, emax, 1));
    FixedFFTArgs a{x, n, y, e, emax, rows, n};
    CK(launch_fft_bfp_rows(ctx, a));
    if (renorm) CK(launch_renorm_rows(ctx, y, e, emax, rows, n));
    return PFA_OK;
}

extern "C" int pfa_fft_pass_bfp_dev(pfa_ctx *ctx, const int16_t *x, int rows, int n, int16_t *y,
                                    int32_t *e) {
    PFA_TRY
    return fft_rows_dev(ctx, x, rows, n, y, e, true);
    PFA_CATCH(ctx)
}

static int fft_rows_host(pfa_ctx *ctx, const int16_t *x, int rows, int n, int16_t *y, int32_t *e,
                         bool renorm) {
    if (!ctx) return PFA_ERR_ARG;
    if (rows < 0 || (rows && (!x || !y || !e))) return pfa_fail(ctx, PFA_ERR_ARG, "fft: bad argument");
    if (!pow2_in(n, 64, 8192)) return pfa_fail(ctx, PFA_ERR_DIMS, "FFT length must be a power of two in 64..8192");
    if (rows == 0) return PFA_OK;
    int st = PFA_OK;
    const size_t b = (size_t)rows * n * 4;
    void *dx = pfa_ws(ctx, "io_in", b, &st);
    if (!dx) return st;
    void *dy = pfa_ws(ctx, "io_out", b, &st);
    if (!dy) return st;
    int32_t *de = (int32_t *)pfa_ws(ctx, "io_e", (size_t)rows * 4, &st);
    if (!de) return st;
    CK(pfa_io_h2d(ctx, dx, x, b));
    CK(fft_rows_dev(ctx, (const int16_t *)dx, rows, n, (int16_t *)dy, de, renorm));
    PFA_HIP(ctx, hipMemcpyAsync(e, de, (size_t)rows * 4, hipMemcpyDeviceToHost, ctx->stream));
    CK(pfa_io_d2h(ctx, y, dy, b));
    return PFA_OK;
}

extern "C" int pfa_fft_pass_bfp(pfa_ctx *ctx, const int16_t *x, int rows, int n, int16_t *y,
                                int32_t *e) {
    PFA_TRY
    return fft_rows_host(ctx, x, rows, n, y, e, true);
    PFA_CATCH(ctx)
}

extern "C" int pfa_fft_bfp_perrow(pfa_ctx *ctx, const int16_t *x, int rows, int n, int16_t *y,
                                  int32_t *e) {
    PFA_TRY
    return fft_rows_host(ctx, x, rows, n, y, e, false);
    PFA_CATCH(ctx)
}

extern "C" int pfa_window_fixed(pfa_ctx *ctx, const int16_t *g2, int rows, int cols,
                                const int32_t *hamr, const int32_t *hamc, int16_t *out) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    if (rows < 0 || cols < 0 || ((rows && cols) && (!g2 || !hamr || !hamc || !out)))
        return pfa_fail(ctx, PFA_ERR_ARG, "window: bad argument");
    for (int i = 0; i < rows; ++i)
        if (hamr[i] < 0 || hamr[i] > 32768) return pfa_fail(ctx, PFA_ERR_GEOMETRY, "window: hamr outside [0, 32768]");
    for (int i = 0; i < cols; ++i)
        if (hamc[i] < 0 || hamc[i] > 32768) return pfa_fail(ctx, PFA_ERR_GEOMETRY, "window: hamc outside [0, 32768]");
    const size_t b = (size_t)rows * cols * 4;
    if (!b) return PFA_OK;
    int st = PFA_OK;
    void *dx = pfa_ws(ctx, "io_in", b, &st);
    if (!dx) return st;
    void *dy = pfa_ws(ctx, "io_out", b, &st);
    if (!dy) return st;
    int32_t *dh = (int32_t *)pfa_ws(ctx, "io_ham", (size_t)(rows + cols) * 4, &st);
    if (!dh) return st;
    CK(pfa_io_h2d(ctx, dx, g2, b));
    CK(pfa_io_h2d(ctx, dh, hamr, (size_t)rows * 4));
    CK(pfa_io_h2d(ctx, dh + rows, hamc, (size_t)cols * 4));
    CK(launch_window_fixed(ctx, dx, rows, cols, dh, dh + rows, dy));
    CK(pfa_io_d2h(ctx, out, dy, b));
    return PFA_OK;
    PFA_CATCH(ctx)
}

extern "C" int pfa_detect_fixed_dev(pfa_ctx *ctx, const int16_t *z, int64_t count, uint16_t *out) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    if (count < 0 || (count && (!z || !out))) return pfa_fail(ctx, PFA_ERR_ARG, "detect: bad argument");
    if (!count) return PFA_OK;
    return launch_detect_fixed(ctx, z, count, out);
    PFA_CATCH(ctx)
}

extern "C" int pfa_detect_fixed(pfa_ctx *ctx, const int16_t *z, int64_t count, uint16_t *out) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    if (count < 0 || (count && (!z || !out))) return pfa_fail(ctx, PFA_ERR_ARG, "detect: bad argument");
    if (!count) return PFA_OK;
    int st = PFA_OK;
    void *dz = pfa_ws(ctx, "io_in", (size_t)count * 4, &st);
    if (!dz) return st;
    void *dy = pfa_ws(ctx, "io_out", (size_t)count * 2, &st);
    if (!dy) return st;
    CK(pfa_io_h2d(ctx, dz, z, (size_t)count * 4));
    CK(launch_detect_fixed(ctx, dz, count, (uint16_t *)dy));
    CK(pfa_io_d2h(ctx, out, dy, (size_t)count * 2));
    return PFA_OK;
    PFA_CATCH(ctx)
}

// ------------------------------------------------------------------ fixed: plan + pipeline
extern "C" int pfa_plan_fixed(pfa_ctx *ctx, pfa_plan **out, int m, int n, int grid,
                              const float *f0, const float *df, const float *pr,
                              const float *tan_s, const int32_t *inv_order, const float *krp,
                              const float *kcp, const int32_t *hamr, const int32_t *hamc,
                              const int16_t *sinc_tab, int flags) {
    PFA_TRY
    if (!ctx || !out) return PFA_ERR_ARG;
    *out = nullptr;
    if (!f0 || !df || !pr || !tan_s || !inv_order || !krp || !kcp || !hamr || !hamc)
        return pfa_fail(ctx, PFA_ERR_ARG, "plan_fixed: null table");
    if (flags & ~(PFA_FIXED_RANGE_SINC | PFA_FIXED_AZ_SINC | PFA_FIXED_LERP_EXACT))
        return pfa_fail(ctx, PFA_ERR_ARG, "plan_fixed: unknown flag");
    if (!pow2_in(grid, 64, 8192)) return pfa_fail(ctx, PFA_ERR_DIMS, "grid must be a power of two in 64..8192");
    if (m < 2 || n < 2 || m > grid || n > grid)
        return pfa_fail(ctx, PFA_ERR_DIMS, "plan_fixed: need 2 <= m, n <= grid");
    if ((flags & (PFA_FIXED_RANGE_SINC | PFA_FIXED_AZ_SINC)) && !sinc_tab)
        return pfa_fail(ctx, PFA_ERR_ARG, "plan_fixed: sinc flags need the 256x32 tap table");
    bool split = sinc_tab != nullptr;                // each half of every phase: sum |c| <= 65535
    if (sinc_tab) {
        for (int ph = 0; ph < 256; ++ph) {
            int sum = 0, a0 = 0, a1 = 0;
            for (int t = 0; t < 32; ++t) {
                sum += sinc_tab[ph * 32 + t];
                (t < 16 ? a0 : a1) += std::abs((int)sinc_tab[ph * 32 + t]);
            }
            if (sum != 32768) return pfa_fail(ctx, PFA_ERR_GEOMETRY, "sinc table phase does not sum to 2^15");
            split = split && a0 <= 65535 && a1 <= 65535;
        }
    }
    for (int i = 0; i < m; ++i)
        if (!std::isfinite(f0[i]) || !std::isfinite(df[i]) || !std::isfinite(pr[i]) || !std::isfinite(tan_s[i]))
            return pfa_fail(ctx, PFA_ERR_GEOMETRY, "plan_fixed: non-finite per-pulse table");
    for (int i = 0; i < grid; ++i)
        if (!std::isfinite(krp[i]) || !std::isfinite(kcp[i]))
            return pfa_fail(ctx, PFA_ERR_GEOMETRY, "plan_fixed: non-finite KR/KC grid");
    for (int i = 0; i < grid; ++i)
        if (hamr[i] < 0 || hamr[i] > 32768 || hamc[i] < 0 || hamc[i] > 32768)
            return pfa_fail(ctx, PFA_ERR_GEOMETRY, "plan_fixed: taper outside [0, 32768]");
    std::vector<int32_t> order(m, -1);
    for (int i = 0; i < m; ++i) {
        const int p = inv_order[i];
        if (p < 0 || p >= m || order[p] != -1) return pfa_fail(ctx, PFA_ERR_GEOMETRY, "inv_order is not a permutation");
        order[p] = i;
    }
    // a = f32(f32(2*pr)/f32(C)) (silicon_emulator.py:144, NumPy 2 float32 scalar math)
    std::vector<float> a32(m);
    const volatile float c32 = 299792458.0f;
    for (int i = 0; i < m; ++i) {
        volatile float two = 2.0f * pr[i];
        a32[i] = two / c32;
        if (!std::isfinite(a32[i]) || a32[i] == 0.0f || df[i] == 0.0f)
            return pfa_fail(ctx, PFA_ERR_GEOMETRY, "plan_fixed: degenerate range geometry (pr or df zero)");
    }
    pfa_plan *p = new (std::nothrow) pfa_plan();
    if (!p) return PFA_ERR_NOMEM;
    p->kind = PLAN_FIXED;
    p->ctx = ctx;
    p->m = m; p->n = n; p->grid = grid; p->flags = flags | (split ? FIXED_SINC_SPLIT : 0);
    int st = PFA_OK;
    bool ok = (p->a32 = upload(p, a32.data(), m, &st)) && (p->f0 = upload(p, f0, m, &st)) &&
              (p->df = upload(p, df, m, &st)) && (p->tan_s = upload(p, tan_s, m, &st)) &&
              (p->order = upload(p, order.data(), m, &st)) && (p->krp = upload(p, krp, grid, &st)) &&
              (p->kcp = upload(p, kcp, grid, &st)) && (p->hamr = upload(p, hamr, grid, &st)) &&
              (p->hamc = upload(p, hamc, grid, &st));
    if (ok && sinc_tab) ok = (p->sinc = upload(p, sinc_tab, 256 * 32, &st)) != nullptr;
    if (ok) {                                     // the FFT passes' twiddles, set up with the plan
        std::vector<int32_t> codes(grid);
        pfa_twiddle_codes(grid, 16, codes.data());
        std::vector<int16_t> tw(grid);
        for (int k = 0; k < grid; ++k) tw[k] = (int16_t)codes[k];
        ok = (p->tw16 = upload(p, tw.data(), grid, &st)) != nullptr;
    }
    if (!ok) {
        pfa_plan_destroy(p);
        return st;
    }
    return plan_register(p, out);
    PFA_CATCH(ctx)
}

extern "C" void pfa_plan_destroy(pfa_plan *p) {
    if (!p) return;
    if (p->ctx) {                                 // null: orphaned by pfa_destroy
        hipStreamSynchronize(p->ctx->stream);
        plan_release(p);
        p->ctx->plans.erase(p);
    }
    delete p;
}

// The plan is registered with its context once built (pfa_destroy orphans the
// plans it still holds).
static int plan_register(pfa_plan *p, pfa_plan **out) {
    try {
        p->ctx->plans.insert(p);
    } catch (...) {
        pfa_ctx *ctx = p->ctx;
        pfa_plan_destroy(p);
        return pfa_fail(ctx, PFA_ERR_NOMEM, "host memory exhausted");
    }
    *out = p;
    return PFA_OK;
}

static int ldp_of(int m) { return (m + 63) / 64 * 64; }

// resample only (stage drop-in); g2 device [grid][grid]
static int resample_fixed_dev(pfa_plan *p, const void *sig, void *g2) {
    pfa_ctx *ctx = p->ctx;
    int st = PFA_OK;
    const int G = p->grid, ldp = ldp_of(p->m);
    void *g1t = pfa_ws(ctx, "g1t", (size_t)G * G * 4, &st);
    if (!g1t) return st;
    CK(launch_fixed_range(p, sig, g1t, ldp));
    CK(launch_fixed_azimuth_fft(p, g1t, ldp, g2, nullptr, nullptr, false, false));
    return PFA_OK;
}

extern "C" int pfa_resample_fixed(pfa_plan *p, const int16_t *sig, int16_t *g2) {
    PFA_TRY
    if (!p || !p->ctx || p->kind != PLAN_FIXED) return PFA_ERR_STATE;
    pfa_ctx *ctx = p->ctx;
    if (!sig || !g2) return pfa_fail(ctx, PFA_ERR_ARG, "resample: null pointer");
    int st = PFA_OK;
    const size_t in_b = (size_t)p->m * p->n * 4, out_b = (size_t)p->grid * p->grid * 4;
    void *ds = pfa_ws(ctx, "io_in", in_b, &st);
    if (!ds) return st;
    void *dg = pfa_ws(ctx, "io_out", out_b, &st);
    if (!dg) return st;
    CK(pfa_io_h2d(ctx, ds, sig, in_b));
    CK(resample_fixed_dev(p, ds, dg));
    CK(pfa_io_d2h(ctx, g2, dg, out_b));
    return PFA_OK;
    PFA_CATCH((p ? p->ctx : nullptr))
}

extern "C" int pfa_form_image_fixed_dev(pfa_plan *p, const int16_t *sig, uint16_t *out,
                                        int32_t *exp_a, int32_t *exp_b) {
    PFA_TRY
    if (!p || !p->ctx || p->kind != PLAN_FIXED) return PFA_ERR_STATE;
    pfa_ctx *ctx = p->ctx;
    if (!sig || !out) return pfa_fail(ctx, PFA_ERR_ARG, "form_image: null pointer");
    int st = PFA_OK;
    const int G = p->grid, ldp = ldp_of(p->m);
    void *g1t = pfa_ws(ctx, "g1t", (size_t)G * G * 4, &st);
    if (!g1t) return st;
    void *ya = pfa_ws(ctx, "ya", (size_t)G * G * 4, &st);
    if (!ya) return st;
    int32_t *ea = exp_a ? exp_a : (int32_t *)pfa_ws(ctx, "ea", (size_t)G * 4, &st);
    if (!ea) return st;
    int32_t *eb = exp_b ? exp_b : (int32_t *)pfa_ws(ctx, "eb", (size_t)G * 4, &st);
    if (!eb) return st;
    int *emax = ctx->d_counters;          // [0] pass A, [1] pass B
    CK(launch_zero_i32(ctx, emax, 2));
    CK(launch_fixed_range(p, sig, g1t, ldp));
    CK(launch_fixed_azimuth_fft(p, g1t, ldp, ya, ea, emax, true, true));
    void *yb = g1t;                       // the pass-1 buffer is dead after pass 2
    CK(launch_fixed_fft_b(p, ya, ea, emax, yb, eb, emax + 1));
    CK(launch_fixed_detect_t(p, yb, eb, emax + 1, out));
    return PFA_OK;
    PFA_CATCH((p ? p->ctx : nullptr))
}

extern "C" int pfa_form_image_fixed(pfa_plan *p, const int16_t *sig, uint16_t *out,
                                    int32_t *exp_a, int32_t *exp_b) {
    PFA_TRY
    if (!p || !p->ctx || p->kind != PLAN_FIXED) return PFA_ERR_STATE;
    pfa_ctx *ctx = p->ctx;
    if (!sig || !out) return pfa_fail(ctx, PFA_ERR_ARG, "form_image: null pointer");
    int st = PFA_OK;
    const int G = p->grid;
    const size_t in_b = (size_t)p->m * p->n * 4, out_b = (size_t)G * G * 2;
    void *ds = pfa_ws(ctx, "io_in", in_b, &st);
    if (!ds) return st;
    void *dout = pfa_ws(ctx, "io_out", out_b, &st);
    if (!dout) return st;
    int32_t *de = (int32_t *)pfa_ws(ctx, "io_e", (size_t)G * 8, &st);
    if (!de) return st;
    CK(pfa_io_h2d(ctx, ds, sig, in_b));
    CK(pfa_form_image_fixed_dev(p, (const int16_t *)ds, (uint16_t *)dout, de, de + G));
    if (exp_a) PFA_HIP(ctx, hipMemcpyAsync(exp_a, de, (size_t)G * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (exp_b) PFA_HIP(ctx, hipMemcpyAsync(exp_b, de + G, (size_t)G * 4, hipMemcpyDeviceToHost, ctx->stream));
    CK(pfa_io_d2h(ctx, out, dout, out_b));
    return PFA_OK;
    PFA_CATCH((p ? p->ctx : nullptr))
}

// ------------------------------------------------------------------ float: plan + pipeline
static bool linspace_params(const double *g, int cnt, double out[3]);

static int pow2_ceil(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

extern "C" int pfa_plan_float(pfa_ctx *ctx, pfa_plan **out, int m, int n, int sgn,
                              const double *pr, const int32_t *order, const double *tan_s,
                              const double *kr_grid, const double *kc_grid, const double *f0,
                              const double *df, const double *ham_c, const double *ham_r,
                              const int16_t *sinc_tab, int flags) {
    PFA_TRY
    if (!ctx || !out) return PFA_ERR_ARG;
    *out = nullptr;
    if (!pr || !order || !tan_s || !kr_grid || !kc_grid)
        return pfa_fail(ctx, PFA_ERR_ARG, "plan_float: null table");
    if ((f0 == nullptr) != (df == nullptr)) return pfa_fail(ctx, PFA_ERR_ARG, "plan_float: give both f0 and df or neither");
    if ((ham_c == nullptr) != (ham_r == nullptr)) return pfa_fail(ctx, PFA_ERR_ARG, "plan_float: give both tapers or neither");
    if (flags & ~(PFA_FLOAT_SINC32 | PFA_FLOAT_FFT_BF16 | PFA_FLOAT_FFT_FP16 | PFA_FLOAT_FFT_SPLIT))
        return pfa_fail(ctx, PFA_ERR_ARG, "plan_float: unknown flag");
    if ((flags & PFA_FLOAT_FFT_BF16) && (flags & PFA_FLOAT_FFT_FP16))
        return pfa_fail(ctx, PFA_ERR_ARG, "plan_float: choose one of FFT_BF16 and FFT_FP16");
    if ((flags & PFA_FLOAT_FFT_SPLIT) && !(flags & PFA_FLOAT_FFT_BF16))
        return pfa_fail(ctx, PFA_ERR_ARG, "plan_float: FFT_SPLIT applies to FFT_BF16");
    if ((flags & PFA_FLOAT_SINC32) && !sinc_tab) return pfa_fail(ctx, PFA_ERR_ARG, "plan_float: SINC32 needs the tap table");
    if (sgn == 0) return pfa_fail(ctx, PFA_ERR_ARG, "plan_float: sgn must be -1 or +1");
    if (m < 2 || n < 2) return pfa_fail(ctx, PFA_ERR_DIMS, "plan_float: need m, n >= 2");
    const int m2 = std::max(64, pow2_ceil(m)), n2 = std::max(64, pow2_ceil(n));
    // the cross line (m2) is transformed in LDS with its resample; the range
    // transform (n2) is one line per workgroup and reaches 16384 (FP32 form)
    if (m2 > 16384) return pfa_fail(ctx, PFA_ERR_DIMS, "plan_float: padded pulse (cross) size above 16384");
    if (m2 > 8192 && (flags & (PFA_FLOAT_FFT_BF16 | PFA_FLOAT_FFT_FP16)))
        return pfa_fail(ctx, PFA_ERR_DIMS, "plan_float: matrix-core FFT sizes are 256, 512, 1024, 4096 and 8192");
    if (n2 > 16384) return pfa_fail(ctx, PFA_ERR_DIMS, "plan_float: padded sample (range) size above 16384");
    if (pow2_ceil(m) != m2 || pow2_ceil(n) != n2)
        return pfa_fail(ctx, PFA_ERR_DIMS, "plan_float: m and n must be at least 33 (padded size >= 64)");
    if (flags & (PFA_FLOAT_FFT_BF16 | PFA_FLOAT_FFT_FP16)) {
        auto ok_size = [](int v) { return v == 256 || v == 512 || v == 1024 || v == 4096 || v == 8192; };
        if (!ok_size(m2) || !ok_size(n2))
            return pfa_fail(ctx, PFA_ERR_DIMS, "plan_float: matrix-core FFT sizes are 256, 512, 1024, 4096 and 8192");
    }
    std::vector<int> seen(m, 0);
    for (int i = 0; i < m; ++i) {
        if (order[i] < 0 || order[i] >= m || seen[order[i]]++) return pfa_fail(ctx, PFA_ERR_GEOMETRY, "order is not a permutation");
        if (!std::isfinite(pr[i]) || pr[i] == 0.0 || !std::isfinite(tan_s[i]) || !std::isfinite(kc_grid[i]))
            return pfa_fail(ctx, PFA_ERR_GEOMETRY, "plan_float: non-finite or zero per-pulse geometry");
        if (i && tan_s[i] < tan_s[i - 1]) return pfa_fail(ctx, PFA_ERR_GEOMETRY, "tan_s must be ascending");
        if (f0 && (!std::isfinite(f0[i]) || !std::isfinite(df[i]))) return pfa_fail(ctx, PFA_ERR_GEOMETRY, "plan_float: non-finite frequencies");
    }
    for (int j = 0; j < n; ++j) {
        if (!std::isfinite(kr_grid[j]) || kr_grid[j] <= 0.0) return pfa_fail(ctx, PFA_ERR_GEOMETRY, "KR must be finite and positive");
        if (j && kr_grid[j] < kr_grid[j - 1]) return pfa_fail(ctx, PFA_ERR_GEOMETRY, "KR must be non-decreasing");
    }
    for (int i = 1; i < m; ++i)
        if (kc_grid[i] < kc_grid[i - 1]) return pfa_fail(ctx, PFA_ERR_GEOMETRY, "KC must be non-decreasing");
    pfa_plan *p = new (std::nothrow) pfa_plan();
    if (!p) return PFA_ERR_NOMEM;
    p->kind = PLAN_FLOAT;
    p->ctx = ctx;
    p->m = m; p->n = n; p->m2 = m2; p->n2 = n2; p->sgn = sgn < 0 ? -1 : 1; p->flags = flags;
    int st = PFA_OK;
    std::vector<double> a64(m);               // 2 pr / c per pulse (range positions)
    for (int i = 0; i < m; ++i) a64[i] = (2.0 * pr[i]) / 299792458.0;
    bool ok = (p->pr64 = upload(p, pr, m, &st)) && (p->a64 = upload(p, a64.data(), m, &st)) &&
              (p->order = upload(p, order, m, &st)) &&
              (p->tan_s64 = upload(p, tan_s, m, &st)) && (p->KR = upload(p, kr_grid, n, &st)) &&
              (p->KC = upload(p, kc_grid, m, &st));
    if (ok && f0) ok = (p->f0d = upload(p, f0, m, &st)) && (p->dfd = upload(p, df, m, &st));
    if (ok && ham_c) ok = (p->hm = upload(p, ham_c, m, &st)) && (p->hn = upload(p, ham_r, n, &st));
    if (ok && sinc_tab) ok = (p->sinc = upload(p, sinc_tab, 256 * 32, &st)) != nullptr;
    if (ok && sinc_tab && m2 > 8192) {        // 16384-point cross lines read packed taps from global memory:
        // the layout load_taps2 stages in LDS: word h TPH + 256 r + ph holds rows
        // (2r + h, 2r + h + 1) of the taps padded by 4 zero rows either side,
        // the h = 1 half 32 words past 20 rows
        constexpr int TPR = 20, TPH = PFA_TAP2_HALF;
        std::vector<uint32_t> tp(2 * TPH, 0u);
        for (int i = 0; i < 2 * TPR * 256; ++i) {
            const int h = i / (TPR * 256), r = (i / 256) % TPR, ph = i & 255;
            const int ra = 2 * r + h - 4, rb = ra + 1;
            const int a = (ra >= 0 && ra < 32) ? sinc_tab[ph * 32 + ra] : 0;
            const int b = (rb >= 0 && rb < 32) ? sinc_tab[ph * 32 + rb] : 0;
            tp[h * TPH + r * 256 + ph] = ((uint32_t)a & 0xffffu) | ((uint32_t)b << 16);
        }
        ok = (p->taps2 = upload(p, tp.data(), tp.size(), &st)) != nullptr;
    }
    if (ok) {                                 // tan bins: 2 per pulse over [tan_s[0], tan_s[m-1]]
        const double lo = tan_s[0], hi = tan_s[m - 1];
        int bins = 2 * m;
        double inv_w = (hi > lo) ? (double)bins / (hi - lo) : 0.0;
        if (!(inv_w > 0.0) || !std::isfinite(inv_w)) { bins = 1; inv_w = 0.0; }
        std::vector<int32_t> tb(bins + 1);
        for (int b = 0, c = 0; b <= bins; ++b) {
            const double edge = inv_w > 0.0 ? lo + (double)b / inv_w : (b ? hi : lo);
            while (c < m && tan_s[c] < edge) ++c;
            tb[b] = c;
        }
        p->tb_bins = bins;
        p->tb_lo = lo;
        p->tb_inv_w = inv_w;
        ok = (p->tan_bin = upload(p, tb.data(), tb.size(), &st)) != nullptr;
    }
    if (!ok) {
        pfa_plan_destroy(p);
        return st;
    }
    if (linspace_params(kr_grid, n, p->kr_lin_v)) p->kr_lin = p->kr_lin_v;
    if (linspace_params(kc_grid, m, p->kc_lin_v)) p->kc_lin = p->kc_lin_v;
    return plan_register(p, out);
    PFA_CATCH(ctx)
}

static int ldp32(int m) { return (m + 31) / 32 * 32; }

// Is g[0..cnt) exactly np.linspace(g[0], g[cnt-1], cnt)?  Then {start, step, stop}.
static bool linspace_params(const double *g, int cnt, double out[3]) {
    if (cnt < 2) return false;
    const double start = g[0], stop = g[cnt - 1];
    const volatile double delta = stop - start;
    const double step = delta / (double)(cnt - 1);
    if (step == 0.0) return false;
    for (int k = 0; k < cnt - 1; ++k) {
        const volatile double prod = (double)k * step;
        if (prod + start != g[k]) return false;
    }
    out[0] = start;
    out[1] = step;
    out[2] = stop;
    return true;
}

static int float_check(pfa_plan *p, const void *sig, const double *freq, const void *out, int out_kind) {
    if (!p || !p->ctx || p->kind != PLAN_FLOAT) return PFA_ERR_STATE;
    if (!sig || !out) return pfa_fail(p->ctx, PFA_ERR_ARG, "float: null pointer");
    if (out_kind != PFA_OUT_C64 && out_kind != PFA_OUT_ABS_F32) return pfa_fail(p->ctx, PFA_ERR_ARG, "float: bad out_kind");
    if (!freq && !p->f0d) return pfa_fail(p->ctx, PFA_ERR_ARG, "float: plan has no compact frequencies; pass freq");
    return PFA_OK;
}

extern "C" int pfa_focus_f32_dev(pfa_plan *p, const float *sig, const double *freq, void *out,
                                 int out_kind) {
    PFA_TRY
    CK(float_check(p, sig, freq, out, out_kind));
    pfa_ctx *ctx = p->ctx;
    int st = PFA_OK;
    const int ldp = ldp32(p->m);
    void *g1t = pfa_ws(ctx, "f_g1t", (size_t)p->n * ldp * 8, &st);
    if (!g1t) return st;
    const int fk = (p->flags & PFA_FLOAT_FFT_BF16) ? ((p->flags & PFA_FLOAT_FFT_SPLIT) ? 3 : 1)
                   : (p->flags & PFA_FLOAT_FFT_FP16) ? 2 : 0;
    if (fk) {                                  // precision form: A as 16-bit pairs + row exponents
        unsigned *A16 = (unsigned *)pfa_ws(ctx, "f_A16", (size_t)(p->n + 15) / 16 * 16 * p->m2 * 4, &st);
        if (!A16) return st;
        int *expA = (int *)pfa_ws(ctx, "f_expA", (size_t)p->n * 4, &st);
        if (!expA) return st;
        CK(launch_float_range(p, sig, freq, g1t, ldp));
        CK(launch_float_azimuth_mfma(p, g1t, ldp, A16, expA, fk));
        CK(launch_float_fft_b_mfma(p, A16, expA, out, out_kind, fk));
        return PFA_OK;
    }
    void *A = pfa_ws(ctx, "f_A", (size_t)(p->n + 15) / 16 * 16 * p->m2 * 8, &st);
    if (!A) return st;
    CK(launch_float_range(p, sig, freq, g1t, ldp));
    CK(launch_float_azimuth(p, 1, g1t, ldp, nullptr, A));
    CK(launch_float_fft_b(p, A, out, out_kind));
    return PFA_OK;
    PFA_CATCH((p ? p->ctx : nullptr))
}

// ------------------------------------------------------------------ autofocus
static int af_check(pfa_plan *p) {
    if (!p || !p->ctx || p->kind != PLAN_FLOAT) return PFA_ERR_STATE;
    if (p->flags & (PFA_FLOAT_FFT_BF16 | PFA_FLOAT_FFT_FP16))
        return pfa_fail(p->ctx, PFA_ERR_ARG, "autofocus: needs the FP32 FFT form");
    if (p->m2 > 4096) return pfa_fail(p->ctx, PFA_ERR_DIMS, "autofocus: padded cross size above 4096");
    if (p->n2 > 8192) return pfa_fail(p->ctx, PFA_ERR_DIMS, "autofocus: padded range size above 8192");
    return PFA_OK;
}

extern "C" int pfa_af_checkpoint(pfa_plan *p, const float *sig, const double *freq) {
    PFA_TRY
    CK(af_check(p));
    if (!sig) return pfa_fail(p->ctx, PFA_ERR_ARG, "af_checkpoint: null signal");
    if (!freq && !p->f0d) return pfa_fail(p->ctx, PFA_ERR_ARG, "af_checkpoint: plan has no compact frequencies; pass freq");
    const int ldp = ldp32(p->m);
    if (!p->af_g1t) {                             // the plan's own buffer: plans never share it
        void *d = nullptr;
        const hipError_t e = hipMalloc(&d, (size_t)p->n * ldp * 8);
        if (e != hipSuccess) return pfa_hip_fail(p->ctx, e, "autofocus checkpoint");
        try {
            p->owned.push_back(d);
        } catch (...) {
            hipFree(d);
            throw;
        }
        p->af_g1t = d;
    }
    p->af_ready = false;
    CK(launch_float_range(p, sig, freq, p->af_g1t, ldp));
    p->af_ready = true;
    return PFA_OK;
    PFA_CATCH((p ? p->ctx : nullptr))
}

static int af_buffers(pfa_plan *p, void **g1t, void **rot, const float *phi) {
    int st = PFA_OK;
    *g1t = p->af_g1t;
    *rot = pfa_ws(p->ctx, "af_rot", (size_t)p->m * 8, &st);
    if (!*rot) return st;
    return launch_af_rot(p, phi, *rot);
}

extern "C" int pfa_af_forward(pfa_plan *p, const float *phi, float *img, double *stats) {
    PFA_TRY
    CK(af_check(p));
    if (!p->af_ready) return pfa_fail(p->ctx, PFA_ERR_STATE, "af_forward: no checkpoint (call pfa_af_checkpoint)");
    if (!phi || !img || !stats) return pfa_fail(p->ctx, PFA_ERR_ARG, "af_forward: null pointer");
    int st = PFA_OK;
    void *g1t, *rot;
    CK(af_buffers(p, &g1t, &rot, phi));
    void *A = pfa_ws(p->ctx, "f_A", (size_t)(p->n + 15) / 16 * 16 * p->m2 * 8, &st);
    if (!A) return st;
    constexpr int parts = 1024;
    double *part = (double *)pfa_ws(p->ctx, "af_stat", parts * 2 * sizeof(double), &st);
    if (!part) return st;
    CK(launch_float_azimuth(p, 1, g1t, ldp32(p->m), nullptr, A, rot));
    CK(launch_float_fft_b(p, A, img, PFA_OUT_C64));
    return launch_af_stats(p, img, part, parts, stats);
    PFA_CATCH((p ? p->ctx : nullptr))
}

extern "C" int pfa_af_backward(pfa_plan *p, const float *phi, const float *img,
                               const double *stats, float *grad) {
    PFA_TRY
    CK(af_check(p));
    if (!p->af_ready) return pfa_fail(p->ctx, PFA_ERR_STATE, "af_backward: no checkpoint (call pfa_af_checkpoint)");
    if (!phi || !img || !stats || !grad) return pfa_fail(p->ctx, PFA_ERR_ARG, "af_backward: null pointer");
    int st = PFA_OK;
    void *g1t, *rot;
    CK(af_buffers(p, &g1t, &rot, phi));
    void *B = pfa_ws(p->ctx, "af_B", (size_t)p->m2 * ((p->n + 7) & ~7) * 8, &st);   // 8-bin blocks
    if (!B) return st;
    const size_t max_parts = (size_t)std::max(p->ctx->num_cu, 8) + 8;   // >= the adjoint's grid
    float *part = (float *)pfa_ws(p->ctx, "af_part", max_parts * p->m * 4, &st);
    if (!part) return st;
    CK(launch_af_fftb_adj(p, img, stats, B));
    int nparts = 0;
    CK(launch_af_az_adj(p, g1t, ldp32(p->m), rot, B, part, &nparts));
    if ((size_t)nparts > max_parts) return pfa_fail(p->ctx, PFA_ERR_STATE, "af_backward: partial buffer too small");
    return launch_af_grad_reduce(p, part, nparts, stats, grad);
    PFA_CATCH((p ? p->ctx : nullptr))
}

extern "C" int pfa_focus_f32(pfa_plan *p, const float *sig, const double *freq, void *out,
                             int out_kind) {
    PFA_TRY
    CK(float_check(p, sig, freq, out, out_kind));
    pfa_ctx *ctx = p->ctx;
    int st = PFA_OK;
    const size_t in_b = (size_t)p->m * p->n * 8;
    const size_t out_b = (size_t)p->m2 * p->n2 * (out_kind == PFA_OUT_C64 ? 8 : 4);
    void *ds = pfa_ws(ctx, "io_in", in_b, &st);
    if (!ds) return st;
    void *dout = pfa_ws(ctx, "io_out", out_b, &st);
    if (!dout) return st;
    double *dfreq = nullptr;
    if (freq) {
        dfreq = (double *)pfa_ws(ctx, "io_freq", (size_t)p->m * p->n * 8, &st);
        if (!dfreq) return st;
        CK(pfa_io_h2d(ctx, dfreq, freq, (size_t)p->m * p->n * 8));
    }
    CK(pfa_io_h2d(ctx, ds, sig, in_b));
    CK(pfa_focus_f32_dev(p, (const float *)ds, dfreq, dout, out_kind));
    CK(pfa_io_d2h(ctx, out, dout, out_b));
    return PFA_OK;
    PFA_CATCH((p ? p->ctx : nullptr))
}

extern "C" int pfa_resample_kspace_f32(pfa_plan *p, const float *sig, const double *freq,
                                       float *g2) {
    PFA_TRY
    CK(float_check(p, sig, freq, g2, PFA_OUT_C64));
    pfa_ctx *ctx = p->ctx;
    int st = PFA_OK;
    const size_t in_b = (size_t)p->m * p->n * 8;
    void *ds = pfa_ws(ctx, "io_in", in_b, &st);
    if (!ds) return st;
    void *dg = pfa_ws(ctx, "io_out", in_b, &st);
    if (!dg) return st;
    double *dfreq = nullptr;
    if (freq) {
        dfreq = (double *)pfa_ws(ctx, "io_freq", (size_t)p->m * p->n * 8, &st);
        if (!dfreq) return st;
        CK(pfa_io_h2d(ctx, dfreq, freq, (size_t)p->m * p->n * 8));
    }
    const int ldp = ldp32(p->m);
    void *g1t = pfa_ws(ctx, "f_g1t", (size_t)p->n * ldp * 8, &st);
    if (!g1t) return st;
    CK(pfa_io_h2d(ctx, ds, sig, in_b));
    CK(launch_float_range(p, ds, dfreq, g1t, ldp));
    CK(launch_float_azimuth(p, 0, g1t, ldp, nullptr, dg));
    CK(pfa_io_d2h(ctx, g2, dg, in_b));
    return PFA_OK;
    PFA_CATCH((p ? p->ctx : nullptr))
}

extern "C" int pfa_kr_extrema_dev(pfa_ctx *ctx, const double *freq, const double *pr, int m,
                                  int n, double *row_max, double *row_min) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    if (m < 0 || n < 0 || ((size_t)m * n && (!freq || !pr || !row_max || !row_min)))
        return pfa_fail(ctx, PFA_ERR_ARG, "kr_extrema: bad argument");
    return launch_kr_extrema(ctx, freq, pr, m, n, row_max, row_min);
    PFA_CATCH(ctx)
}

extern "C" int pfa_kr_extrema(pfa_ctx *ctx, const double *freq, const double *pr, int m, int n,
                              double *row_max, double *row_min) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    if (m < 0 || n < 0 || ((size_t)m * n && (!freq || !pr || !row_max || !row_min)))
        return pfa_fail(ctx, PFA_ERR_ARG, "kr_extrema: bad argument");
    if (!m || !n) return PFA_OK;
    int st = PFA_OK;
    double *df = (double *)pfa_ws(ctx, "io_in", (size_t)m * n * 8, &st);
    if (!df) return st;
    double *dp = (double *)pfa_ws(ctx, "io_pr", (size_t)m * 3 * 8, &st);
    if (!dp) return st;
    CK(pfa_io_h2d(ctx, df, freq, (size_t)m * n * 8));
    CK(pfa_io_h2d(ctx, dp, pr, (size_t)m * 8));
    CK(launch_kr_extrema(ctx, df, dp, m, n, dp + m, dp + 2 * m));
    CK(pfa_io_d2h(ctx, row_max, dp + m, (size_t)m * 8));
    CK(pfa_io_d2h(ctx, row_min, dp + 2 * m, (size_t)m * 8));
    return PFA_OK;
    PFA_CATCH(ctx)
}

extern "C" int pfa_transform_f32(pfa_plan *p, const float *g2, void *out, int out_kind) {
    PFA_TRY
    if (!p || !p->ctx || p->kind != PLAN_FLOAT) return PFA_ERR_STATE;
    pfa_ctx *ctx = p->ctx;
    if (!g2 || !out || (out_kind != PFA_OUT_C64 && out_kind != PFA_OUT_ABS_F32))
        return pfa_fail(ctx, PFA_ERR_ARG, "transform: bad argument");
    int st = PFA_OK;
    const size_t in_b = (size_t)p->m * p->n * 8;
    const size_t out_b = (size_t)p->m2 * p->n2 * (out_kind == PFA_OUT_C64 ? 8 : 4);
    void *dg = pfa_ws(ctx, "io_in", in_b, &st);
    if (!dg) return st;
    void *dout = pfa_ws(ctx, "io_out", out_b, &st);
    if (!dout) return st;
    void *A = pfa_ws(ctx, "f_A", (size_t)(p->n + 15) / 16 * 16 * p->m2 * 8, &st);
    if (!A) return st;
    CK(pfa_io_h2d(ctx, dg, g2, in_b));
    CK(launch_float_azimuth(p, 2, nullptr, 0, dg, A));
    CK(launch_float_fft_b(p, A, dout, out_kind));
    CK(pfa_io_d2h(ctx, out, dout, out_b));
    return PFA_OK;
    PFA_CATCH((p ? p->ctx : nullptr))
}
