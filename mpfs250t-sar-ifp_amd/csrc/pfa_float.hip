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

#include "pfa_fft_f32.hpp"
#include "pfa_internal.hpp"

namespace {

constexpr double C_LIGHT = 299792458.0;

// np.interp bracket count on an ascending view: #{u < S : xa(u) <= Q}
template <typename XA>
__device__ __forceinline__ int count_le(XA &&xa, int S, double Q, int c) {
    if (c < 0) {
        int lo = 0, hi = S;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (xa(mid) <= Q) lo = mid + 1; else hi = mid;
        }
        return lo;
    }
    int steps = 0;
    while (c < S && xa(c) <= Q && steps < 64) { ++c; ++steps; }
    while (c > 0 && !(xa(c - 1) <= Q) && steps < 128) { --c; ++steps; }
    if (steps >= 64) return count_le(xa, S, Q, -1);
    return c;
}

// One output of np.interp(Q, xa, fa, left=0, right=0) on the ascending view
// (or the SINC32 gather), given c = #{xa <= Q}.  Double-precision formula
// slope * (Q - x0) + f0 with each operation rounded, stored as float.
template <bool SINC, typename XA, typename FA>
__device__ __forceinline__ float2 interp_one(XA &&xa, FA &&fa, int S, double Q, int c,
                                             const int16_t *__restrict__ sinc) {
    if (c <= 0) return make_float2(0.f, 0.f);
    if (c >= S) {
        if (Q == xa(S - 1)) return fa(S - 1);
        return make_float2(0.f, 0.f);
    }
    const int J = c - 1;
    const double x0 = xa(J), x1 = xa(J + 1);
    const float2 f0 = fa(J);
    if (SINC && J >= 15 && J <= S - 17) {
        const double frac = __ddiv_rn(__dsub_rn(Q, x0), __dsub_rn(x1, x0));
        const int wq = (int)fmin(fmax(rint(frac * 32768.0), 0.0), 32767.0);
        const int16_t *c32 = sinc + (wq >> 7) * 32;
        float ar = 0.f, ai = 0.f;
#pragma unroll
        for (int t = 0; t < 32; ++t) {
            const float cf = (float)c32[t] * (1.0f / 32768.0f);
            const float2 s = fa(J - 15 + t);
            ar = fmaf(cf, s.x, ar);
            ai = fmaf(cf, s.y, ai);
        }
        return make_float2(ar, ai);
    }
    if (x0 == Q) return f0;
    const float2 f1 = fa(J + 1);
    const double dx = __dsub_rn(x1, x0), dq = __dsub_rn(Q, x0);
    const double sr = __ddiv_rn(__dsub_rn((double)f1.x, (double)f0.x), dx);
    const double si = __ddiv_rn(__dsub_rn((double)f1.y, (double)f0.y), dx);
    return make_float2((float)__dadd_rn(__dmul_rn(sr, dq), (double)f0.x),
                       (float)__dadd_rn(__dmul_rn(si, dq), (double)f0.y));
}

// ------------------------------------------------------------------ pass 1 (range)
constexpr int RP = 32;   // pulses per tile
constexpr int RQ = 64;   // range bins per tile

template <bool SINC, bool FULLFREQ>
__global__ __launch_bounds__(256) void k_range_f32(
    const float2 *__restrict__ sig, int m, int n, const int32_t *__restrict__ order,
    const double *__restrict__ pr, const double *__restrict__ f0v, const double *__restrict__ dfv,
    const double *__restrict__ freq, const double *__restrict__ KR,
    const int16_t *__restrict__ sinc, float2 *__restrict__ g1t, int ldp) {
    __shared__ float2 tile[RQ][RP + 1];
    const int p0 = blockIdx.x * RP, q0 = blockIdx.y * RQ;
    const int t = threadIdx.x;
    const int pl = t >> 3, ch = t & 7;
    const int p = p0 + pl;
    if (p < m) {
        const int i = order[p];
        const double prr = pr[i];
        const double f0 = FULLFREQ ? 0.0 : f0v[i], df = FULLFREQ ? 0.0 : dfv[i];
        const double *frow = FULLFREQ ? freq + (size_t)i * n : nullptr;
        const float2 *row = sig + (size_t)i * n;
        auto kr = [&](int j) {
            const double f = FULLFREQ ? frow[j] : __dadd_rn(f0, __dmul_rn((double)j, df));
            return __dmul_rn(__ddiv_rn(__dmul_rn(2.0, f), C_LIGHT), prr);
        };
        const bool asc = !(kr(0) > kr(n - 1));
        auto xa = [&](int u) { return kr(asc ? u : n - 1 - u); };
        auto fa = [&](int u) { return row[asc ? u : n - 1 - u]; };
        int c = -1;
#pragma unroll 1
        for (int u = 0; u < RQ / 8; ++u) {
            const int q = q0 + ch * (RQ / 8) + u;
            if (q >= n) break;
            const double Q = KR[q];
            if (c < 0 && !FULLFREQ) {        // closed-form guess from the affine frequency
                const double js = ((Q / prr) * (C_LIGHT / 2.0) - f0) / df;
                double g = asc ? js : (double)(n - 1) - js;
                g = fmin(fmax(floor(g) + 1.0, 0.0), (double)n);
                c = (g == g) ? (int)g : 0;
            }
            c = count_le(xa, n, Q, c);
            tile[q - q0][pl] = interp_one<SINC>(xa, fa, n, Q, c, sinc);
        }
    }
    __syncthreads();
    for (int r = t; r < RQ * RP; r += 256) {
        const int qq = r / RP, pp = r % RP;
        if (q0 + qq < n && p0 + pp < m) g1t[(size_t)(q0 + qq) * ldp + p0 + pp] = tile[qq][pp];
    }
}

// ------------------------------------------------------------------ pass 2 (+ window + FFT A)
// MODE 0: store the windowed k-space g2[q][j] (resample_kspace stage output)
// MODE 1: fused -- window, sign fold, pad, FFT along cross -> A[j][r]
// MODE 2: FFT along cross of a given g2 (focus() stage) -> A[j][r]
template <int L, int MODE, bool SINC, bool INV>
__global__ __launch_bounds__(pfa32::Cfg<L>::NT < 64 ? 64 : pfa32::Cfg<L>::NT) void k_azimuth_f32(
    const float2 *__restrict__ g1t, int ldp, int m, int n, const double *__restrict__ KR,
    const double *__restrict__ tan_s, const double *__restrict__ KC,
    const double *__restrict__ hm, const double *__restrict__ hn,
    const int16_t *__restrict__ sinc, const float2 *__restrict__ g2in,
    float2 *__restrict__ out, const float2 *__restrict__ tw) {
    using Cf = pfa32::Cfg<L>;
    constexpr int M2 = Cf::N;
    __shared__ float2 buf[Cf::PADN];
    const int j = blockIdx.x;
    const int t = threadIdx.x;
    const int nb = blockDim.x;
    if (MODE == 2) {
        for (int q = t; q < M2; q += nb) {
            float2 v = q < m ? g2in[(size_t)q * n + j] : make_float2(0.f, 0.f);
            if ((q + j) & 1) v = make_float2(-v.x, -v.y);
            buf[pfa32::padf(q)] = v;
        }
    } else {
        const double krj = KR[j];
        const float2 *row = g1t + (size_t)j * ldp;
        auto xa = [&](int u) { return __dmul_rn(krj, tan_s[u]); };
        auto fa = [&](int u) { return row[u]; };
        const int per = (M2 + nb - 1) / nb;
        const double hj = hn ? hn[j] : 1.0;
        int c = -1;
#pragma unroll 1
        for (int u = 0; u < per; ++u) {
            const int q = t * per + u;
            if (q >= M2) break;
            float2 v = make_float2(0.f, 0.f);
            if (q < m) {
                const double Q = KC[q];
                c = count_le(xa, m, Q, c);
                v = interp_one<SINC>(xa, fa, m, Q, c, sinc);
                if (hm) {
                    const float w = (float)__dmul_rn(hm[q], hj);
                    v = make_float2(v.x * w, v.y * w);
                }
            }
            if (MODE == 0) {
                if (q < m) out[(size_t)q * n + j] = v;
            } else {
                if ((q + j) & 1) v = make_float2(-v.x, -v.y);
                buf[pfa32::padf(q)] = v;
            }
        }
    }
    if constexpr (MODE != 0) {
        __syncthreads();
        float2 *dst = out + (size_t)j * M2;
        const float scale = INV ? 1.0f / (float)M2 : 1.0f;
        pfa32::fft<L, INV>(buf, tw, [&](int idx, float2 v) {
            if (INV) v = make_float2(v.x * scale, v.y * scale);
            dst[idx] = v;
        });
    }
}

// ------------------------------------------------------------------ FFT B (range) + detect
// OUT 0: complex64 image, 1: |image| float32.  Workgroup b -> cross bin r with
// consecutive bins on the same XCD so the strided column reads share L2 lines.
template <int L, bool INV, int OUT>
__global__ __launch_bounds__(pfa32::Cfg<L>::NT < 64 ? 64 : pfa32::Cfg<L>::NT) void k_fft_b_f32(
    const float2 *__restrict__ A, int n, int m2, void *__restrict__ outp,
    const float2 *__restrict__ tw) {
    using Cf = pfa32::Cfg<L>;
    constexpr int N2 = Cf::N;
    __shared__ float2 buf[Cf::PADN];
    const int b = blockIdx.x;
    const int r = (m2 >= 8) ? (b % 8) * (m2 / 8) + b / 8 : b;
    for (int jj = threadIdx.x; jj < N2; jj += blockDim.x)
        buf[pfa32::padf(jj)] = jj < n ? A[(size_t)jj * m2 + r] : make_float2(0.f, 0.f);
    __syncthreads();
    const float scale = INV ? 1.0f / (float)N2 : 1.0f;
    pfa32::fft<L, INV>(buf, tw, [&](int c, float2 v) {
        if (INV) v = make_float2(v.x * scale, v.y * scale);
        if ((r + c) & 1) v = make_float2(-v.x, -v.y);
        if (OUT == 0) reinterpret_cast<float2 *>(outp)[(size_t)r * N2 + c] = v;
        else reinterpret_cast<float *>(outp)[(size_t)r * N2 + c] = __fsqrt_rn(v.x * v.x + v.y * v.y);
    });
}

const float2 *twiddles_f32(pfa_ctx *ctx, int n, int *st) {
    char slot[32];
    snprintf(slot, sizeof slot, "twf_%d", n);
    auto it = ctx->ws.find(slot);
    if (it != ctx->ws.end() && it->second.bytes) return (const float2 *)it->second.ptr;
    std::vector<float2> h((size_t)n);
    for (int k = 0; k < n; ++k) {
        const double th = 2.0 * M_PI * (double)k / (double)n;
        h[k] = make_float2((float)std::cos(th), (float)-std::sin(th));
    }
    void *d = pfa_ws(ctx, slot, h.size() * sizeof(float2), st);
    if (!d) return nullptr;
    if (hipMemcpy(d, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) {
        *st = pfa_fail(ctx, PFA_ERR_HIP, "twiddle upload failed");
        return nullptr;
    }
    return (const float2 *)d;
}

int log2_exact(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return (1 << l) == n ? l : -1;
}

template <int L>
constexpr int nb_of() { return pfa32::Cfg<L>::NT < 64 ? 64 : pfa32::Cfg<L>::NT; }

}  // namespace

#define PFA_F32_SWITCH(L, MACRO) \
    switch (L) {                 \
        MACRO(6) MACRO(7) MACRO(8) MACRO(9) MACRO(10) MACRO(11) MACRO(12) MACRO(13) \
        default: return pfa_fail(ctx, PFA_ERR_DIMS, "padded size must be a power of two in 64..8192"); \
    }

int launch_float_range(pfa_plan *p, const void *sig, const double *freq, void *g1t, int ldp) {
    pfa_ctx *ctx = p->ctx;
    dim3 grid((p->m + RP - 1) / RP, (p->n + RQ - 1) / RQ);
    const bool sinc = p->flags & PFA_FLOAT_SINC32;
    const bool full = freq != nullptr;
    pfa_timer_begin(ctx, PFA_TIMER_RANGE);
#define ARGS (const float2 *)sig, p->m, p->n, p->order, p->pr64, p->f0d, p->dfd, freq, p->KR, p->sinc, \
             (float2 *)g1t, ldp
    if (sinc && full) hipLaunchKernelGGL((k_range_f32<true, true>), grid, dim3(256), 0, ctx->stream, ARGS);
    else if (sinc) hipLaunchKernelGGL((k_range_f32<true, false>), grid, dim3(256), 0, ctx->stream, ARGS);
    else if (full) hipLaunchKernelGGL((k_range_f32<false, true>), grid, dim3(256), 0, ctx->stream, ARGS);
    else hipLaunchKernelGGL((k_range_f32<false, false>), grid, dim3(256), 0, ctx->stream, ARGS);
#undef ARGS
    pfa_timer_end(ctx, PFA_TIMER_RANGE);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

int launch_float_azimuth(pfa_plan *p, int mode, const void *g1t, int ldp, const void *g2in,
                         void *out) {
    pfa_ctx *ctx = p->ctx;
    int st = PFA_OK;
    const int m2 = p->m2;
    const float2 *tw = twiddles_f32(ctx, m2, &st);
    if (!tw) return st;
    const bool sinc = p->flags & PFA_FLOAT_SINC32;
    const bool inv = p->sgn > 0;
    const int L = log2_exact(m2);
    pfa_timer_begin(ctx, PFA_TIMER_AZIMUTH);
#define ARGS (const float2 *)g1t, ldp, p->m, p->n, p->KR, p->tan_s64, p->KC, p->hm, p->hn, p->sinc, \
             (const float2 *)g2in, (float2 *)out, tw
#define K4(LL, MODE)                                                                                 \
    if (sinc && inv) hipLaunchKernelGGL((k_azimuth_f32<LL, MODE, true, true>), dim3(p->n), dim3(nb_of<LL>()), 0, ctx->stream, ARGS); \
    else if (sinc) hipLaunchKernelGGL((k_azimuth_f32<LL, MODE, true, false>), dim3(p->n), dim3(nb_of<LL>()), 0, ctx->stream, ARGS); \
    else if (inv) hipLaunchKernelGGL((k_azimuth_f32<LL, MODE, false, true>), dim3(p->n), dim3(nb_of<LL>()), 0, ctx->stream, ARGS); \
    else hipLaunchKernelGGL((k_azimuth_f32<LL, MODE, false, false>), dim3(p->n), dim3(nb_of<LL>()), 0, ctx->stream, ARGS);
#define LAUNCH(LL)                       \
    case LL:                             \
        if (mode == 0) { K4(LL, 0) }     \
        else if (mode == 1) { K4(LL, 1) } \
        else { K4(LL, 2) }               \
        break;
    PFA_F32_SWITCH(L, LAUNCH)
#undef LAUNCH
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
        if (inv && out_kind) hipLaunchKernelGGL((k_fft_b_f32<LL, true, 1>), dim3(p->m2), dim3(nb_of<LL>()), 0, ctx->stream, ARGS); \
        else if (inv) hipLaunchKernelGGL((k_fft_b_f32<LL, true, 0>), dim3(p->m2), dim3(nb_of<LL>()), 0, ctx->stream, ARGS); \
        else if (out_kind) hipLaunchKernelGGL((k_fft_b_f32<LL, false, 1>), dim3(p->m2), dim3(nb_of<LL>()), 0, ctx->stream, ARGS); \
        else hipLaunchKernelGGL((k_fft_b_f32<LL, false, 0>), dim3(p->m2), dim3(nb_of<LL>()), 0, ctx->stream, ARGS); \
        break;
    PFA_F32_SWITCH(L, LAUNCH)
#undef LAUNCH
#undef ARGS
    pfa_timer_end(ctx, PFA_TIMER_FFT_B);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}
