// pfa_float.hip -- gfx950 kernels of the float (FP32) Polar Format focus.
//
// Contract: src/form_image_pfa.py (resample_kspace :165-222, focus :225-240,
// transform2d :144-147).  Layouts: signal [pulse][sample] complex64; g1t
// [range bin j][sorted pulse p]; A [range bin j][cross frequency r]; image
// [cross r][range c] (complex64 or |.| float32), centred as fftshift(fft2(
// ifftshift(pad))) -- folded here into (-1)^(q+j) on the input and (-1)^(r+c)
// on the output.
//
// Kernels
//   k_range_f32    pass 1: per pulse, FP64 k-space positions kr = (2 f / c) pr
//                  computed on the fly, np.interp bracket/formula (LINEAR) or
//                  the 32-tap polyphase gather (SINC32); pulses in tan(phi)
//                  order written corner-turned through an LDS tile.
//   k_azimuth_f32  pass 2 per range bin: sources KR_j tan_s (FP64), same
//                  interpolators, float32 Hamming product, sign fold, zero pad
//                  and the FFT along cross, LDS-resident.
//   k_fft_b_f32    FFT along range per cross bin (column load = corner-turn),
//                  output sign fold and detection fused into the store.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "pfa_float_dev.hpp"



#define PFA_F32_SWITCH(L, MACRO) \
    switch (L) {                 \
        MACRO(6) MACRO(7) MACRO(8) MACRO(9) MACRO(10) MACRO(11) MACRO(12) MACRO(13) \
        default: return pfa_fail(ctx, PFA_ERR_DIMS, "padded size must be a power of two in 64..8192"); \
    }
// the azimuth pass with its cross transform: 16384-point lines as well
#define PFA_F32_SWITCH_A(L, MACRO) \
    switch (L) {                   \
        MACRO(6) MACRO(7) MACRO(8) MACRO(9) MACRO(10) MACRO(11) MACRO(12) MACRO(13) MACRO(14) \
        default: return pfa_fail(ctx, PFA_ERR_DIMS, "padded cross size must be a power of two in 64..16384"); \
    }
// the range transform (FFT-B) also takes 16384-point lines
#define PFA_F32_SWITCH_B(L, MACRO) \
    switch (L) {                   \
        MACRO(6) MACRO(7) MACRO(8) MACRO(9) MACRO(10) MACRO(11) MACRO(12) MACRO(13) MACRO(14) \
        default: return pfa_fail(ctx, PFA_ERR_DIMS, "padded range size must be a power of two in 64..16384"); \
    }

int launch_float_range(pfa_plan *p, const void *sig, const double *freq, void *g1t, int ldp) {
    pfa_ctx *ctx = p->ctx;
    const int tiles_p = (p->m + RP - 1) / RP;
    dim3 grid(std::max(1, std::min(tiles_p, ctx->num_cu * 2)));
    const bool sinc = p->flags & PFA_FLOAT_SINC32;
    const bool full = freq != nullptr;
    pfa_timer_begin(ctx, PFA_TIMER_RANGE);
#define ARGS (const float2 *)sig, p->m, p->n, p->order, p->a64, p->pr64, p->f0d, p->dfd, freq, grid_of(p->KR, p->kr_lin, p->n), p->sinc, \
             (float2 *)g1t, ldp, (unsigned *)(ctx->d_counters + 32)
    if (sinc && full) hipLaunchKernelGGL((k_range_f32<true, true>), grid, dim3(RT), 0, ctx->stream, ARGS);
    else if (sinc) hipLaunchKernelGGL((k_range_f32<true, false>), grid, dim3(RT), 0, ctx->stream, ARGS);
    else if (full) hipLaunchKernelGGL((k_range_f32<false, true>), grid, dim3(RT), 0, ctx->stream, ARGS);
    else hipLaunchKernelGGL((k_range_f32<false, false>), grid, dim3(RT), 0, ctx->stream, ARGS);
#undef ARGS
    pfa_timer_end(ctx, PFA_TIMER_RANGE);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

int launch_float_azimuth(pfa_plan *p, int mode, const void *g1t, int ldp, const void *g2in,
                         void *out, const void *rot) {
    pfa_ctx *ctx = p->ctx;
    int st = PFA_OK;
    const int m2 = p->m2;
    const float2 *tw = twiddles_f32(ctx, m2, &st);
    if (!tw) return st;
    const bool sinc = p->flags & PFA_FLOAT_SINC32;
    const bool inv = p->sgn > 0;
    const int L = log2_exact(m2);
    if (L >= 14 && sinc && !p->taps2) return pfa_fail(ctx, PFA_ERR_STATE, "azimuth: no global tap table");
    // persistent: one resident workgroup per CU, a multiple of 8 (one per XCD slot)
    const int grid_az = std::max(8, std::min((p->n + 7) / 8 * 8, ctx->num_cu / 8 * 8));
    pfa_timer_begin(ctx, PFA_TIMER_AZIMUTH);
    const TanBins tb{p->tan_bin, p->tb_lo, p->tb_inv_w, p->tb_bins};
#define ARGS (const float2 *)g1t, ldp, p->m, p->n, grid_of(p->KR, p->kr_lin, p->n), p->tan_s64, tb, grid_of(p->KC, p->kc_lin, p->m), p->hm, p->hn, p->sinc, \
             (const float2 *)g2in, (float2 *)out, tw, (unsigned *)(ctx->d_counters + 32), nullptr, nullptr, (const float2 *)rot, nullptr, \
             (const unsigned *)p->taps2
#define K4(LL, MODE)                                                                                 \
    if (sinc && inv) hipLaunchKernelGGL((k_azimuth_f32<LL, MODE, true, true>), dim3(grid_az), dim3(az_nb<LL>()), 0, ctx->stream, ARGS); \
    else if (sinc) hipLaunchKernelGGL((k_azimuth_f32<LL, MODE, true, false>), dim3(grid_az), dim3(az_nb<LL>()), 0, ctx->stream, ARGS); \
    else if (inv) hipLaunchKernelGGL((k_azimuth_f32<LL, MODE, false, true>), dim3(grid_az), dim3(az_nb<LL>()), 0, ctx->stream, ARGS); \
    else hipLaunchKernelGGL((k_azimuth_f32<LL, MODE, false, false>), dim3(grid_az), dim3(az_nb<LL>()), 0, ctx->stream, ARGS);
#define K4R(LL)                                                                                      \
    if (sinc && inv) hipLaunchKernelGGL((k_azimuth_f32<LL, 1, true, true, FK_F32, true>), dim3(grid_az), dim3(az_nb<LL>()), 0, ctx->stream, ARGS); \
    else if (sinc) hipLaunchKernelGGL((k_azimuth_f32<LL, 1, true, false, FK_F32, true>), dim3(grid_az), dim3(az_nb<LL>()), 0, ctx->stream, ARGS); \
    else if (inv) hipLaunchKernelGGL((k_azimuth_f32<LL, 1, false, true, FK_F32, true>), dim3(grid_az), dim3(az_nb<LL>()), 0, ctx->stream, ARGS); \
    else hipLaunchKernelGGL((k_azimuth_f32<LL, 1, false, false, FK_F32, true>), dim3(grid_az), dim3(az_nb<LL>()), 0, ctx->stream, ARGS);
#define LAUNCH(LL)                          \
    case LL:                                \
        if (mode == 0) { K4(LL, 0) }        \
        else if (mode == 1 && rot) { K4R(LL) } \
        else if (mode == 1) { K4(LL, 1) }   \
        else { K4(LL, 2) }                  \
        break;
    PFA_F32_SWITCH_A(L, LAUNCH)
#undef LAUNCH
#undef K4R
#undef K4
#undef ARGS
    pfa_timer_end(ctx, PFA_TIMER_AZIMUTH);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

int launch_float_fft_b(pfa_plan *p, const void *A, void *out, int out_kind) {
    pfa_ctx *ctx = p->ctx;
    int st = PFA_OK;
    const int n2 = p->n2;
    const float2 *tw = twiddles_f32(ctx, n2, &st);
    if (!tw) return st;
    const bool inv = p->sgn > 0;
    const int L = log2_exact(n2);
    pfa_timer_begin(ctx, PFA_TIMER_FFT_B);
#define ARGS (const float2 *)A, p->n, p->m2, out, tw
#define LAUNCH(LL)                                                                                   \
    case LL:                                                                                         \
        if (inv && out_kind) hipLaunchKernelGGL((k_fft_b2_f32<LL, true, 1>), dim3(p->m2 / 2), dim3(fb2_nb<LL>()), 0, ctx->stream, ARGS); \
        else if (inv) hipLaunchKernelGGL((k_fft_b2_f32<LL, true, 0>), dim3(p->m2 / 2), dim3(fb2_nb<LL>()), 0, ctx->stream, ARGS); \
        else if (out_kind) hipLaunchKernelGGL((k_fft_b2_f32<LL, false, 1>), dim3(p->m2 / 2), dim3(fb2_nb<LL>()), 0, ctx->stream, ARGS); \
        else hipLaunchKernelGGL((k_fft_b2_f32<LL, false, 0>), dim3(p->m2 / 2), dim3(fb2_nb<LL>()), 0, ctx->stream, ARGS); \
        break;
    PFA_F32_SWITCH_B(L, LAUNCH)
#undef LAUNCH
#undef ARGS
    pfa_timer_end(ctx, PFA_TIMER_FFT_B);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

// ------------------------------------------------------------------ scene geometry
// Per-pulse extrema of kr = (2 f / c) pr (the reference's numpy order of
// operations, each rounded): one workgroup per pulse, loads of a row issued
// eight at a time from clamped columns.
namespace {
constexpr int KX_NB = 256;
__global__ __launch_bounds__(KX_NB) void k_kr_extrema(const double *__restrict__ freq,
                                                      const double *__restrict__ pr, int n,
                                                      double *__restrict__ row_max,
                                                      double *__restrict__ row_min) {
    __shared__ double smx[KX_NB / 64], smn[KX_NB / 64];
    const int i = blockIdx.x, t = threadIdx.x;
    const double *row = freq + (size_t)i * n;
    const double a = pr[i];
    double mx = -INFINITY, mn = INFINITY;
    for (int j0 = 0; j0 < n; j0 += 8 * KX_NB) {
        double f[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) f[k] = row[min(j0 + t + k * KX_NB, n - 1)];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const double kr = __dmul_rn(__ddiv_rn(2.0 * f[k], C_LIGHT), a);
            if (j0 + t + k * KX_NB < n) {
                mx = fmax(mx, kr);
                mn = fmin(mn, kr);
            }
        }
    }
    for (int off = 32; off >= 1; off >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, off, 64));
        mn = fmin(mn, __shfl_xor(mn, off, 64));
    }
    if ((t & 63) == 0) {
        smx[t >> 6] = mx;
        smn[t >> 6] = mn;
    }
    __syncthreads();
    if (t == 0) {
        for (int w = 1; w < KX_NB / 64; ++w) {
            mx = fmax(mx, smx[w]);
            mn = fmin(mn, smn[w]);
        }
        row_max[i] = mx;
        row_min[i] = mn;
    }
}
}  // namespace

int launch_kr_extrema(pfa_ctx *ctx, const double *freq, const double *pr, int m, int n,
                      double *row_max, double *row_min) {
    if (m <= 0 || n <= 0) return PFA_OK;
    pfa_timer_begin(ctx, PFA_TIMER_OTHER);
    hipLaunchKernelGGL(k_kr_extrema, dim3(m), dim3(KX_NB), 0, ctx->stream, freq, pr, n, row_max,
                       row_min);
    pfa_timer_end(ctx, PFA_TIMER_OTHER);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}
