// pfa_display.hip -- the display / product passes that follow detection:
//
//  * the 65536-bin histogram and the look-up-table gather of the normal-score
//    contrast stretch (mpfs/host/sar_display.py:24-60: np.bincount of the
//    uint16 magnitudes, then lut[img - lo]);
//  * order statistics of a float32 image (radix select over the float bits)
//    and the 99.7-percentile uint8 stretch of the GeoTIFF writer
//    (src/form_image_pfa.py:286-291: clip(amp / (hi + 1e-12) * 255, 0, 255)
//    converted to uint8).
//
// All of it is byte / integer work bound by HBM: each pass reads its input
// once with 16-byte loads and keeps its tables in LDS.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <algorithm>
#include <cstring>

#include "pfa_internal.hpp"

namespace {

// ---------------------------------------------------------------- histogram
// Persistent: one 1024-thread workgroup per CU keeps 32-bit counters for the
// values below 32768 in LDS (128 KB) over its whole share of the image and
// adds them to the global histogram once at the end.  Values of 32768 and
// above (the cores of bright targets in a detected image) are counted with
// global atomics directly.
constexpr int HT = 1024;                  // threads per workgroup
constexpr int HLDS = 32768;               // values counted in LDS

template <bool VEC>
__global__ __launch_bounds__(HT) void k_hist_u16(const uint16_t *__restrict__ img, int64_t count,
                                                 uint32_t *__restrict__ hist) {
    __shared__ __align__(16) uint32_t h[HLDS];
    const int t = threadIdx.x;
    for (int w = t * 4; w < HLDS; w += HT * 4) *reinterpret_cast<uint4 *>(&h[w]) = make_uint4(0, 0, 0, 0);
    __syncthreads();
    auto add = [&](unsigned v) {
        if (v < (unsigned)HLDS) atomicAdd(&h[v], 1u);
        else atomicAdd(&hist[v], 1u);
    };
    const int64_t stride = (int64_t)gridDim.x * HT;
    int64_t i = (int64_t)blockIdx.x * HT + t;
    if (VEC) {
        const int64_t n8 = count / 8;
        const uint4 *img8 = reinterpret_cast<const uint4 *>(img);
        for (; i + 3 * stride < n8; i += 4 * stride) {   // four 16-byte loads in flight
            uint4 d[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) d[k] = img8[i + k * stride];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t w4[4] = {d[k].x, d[k].y, d[k].z, d[k].w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    add(w4[q] & 0xFFFFu);
                    add(w4[q] >> 16);
                }
            }
        }
        for (; i < n8; i += stride) {
            const uint4 d = img8[i];
            const uint32_t w4[4] = {d.x, d.y, d.z, d.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                add(w4[q] & 0xFFFFu);
                add(w4[q] >> 16);
            }
        }
        i = n8 * 8 + (int64_t)blockIdx.x * HT + t;
    }
    for (; i < count; i += stride) add(img[i]);
    __syncthreads();
    for (int w = t; w < HLDS; w += HT) {
        const uint32_t c = h[w];
        if (c) atomicAdd(&hist[w], c);
    }
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, int count, uint32_t v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

// ---------------------------------------------------------------- LUT gather
// out[i] = lut[img[i] - lo]; values outside [lo, lo + nb) take the nearest
// table end.  The table sits in LDS; each workgroup covers LCHUNK values.
constexpr int LT = 1024;
constexpr int LCHUNK = 1 << 18;
static_assert(LCHUNK % (LT * 8) == 0, "whole 16-byte loads per thread");

template <bool VEC>
__global__ __launch_bounds__(LT) void k_lut_u16_u8(const uint16_t *__restrict__ img, int64_t count,
                                                   const uint8_t *__restrict__ lut, int lo, int nb,
                                                   uint8_t *__restrict__ out) {
    __shared__ uint8_t tab[65536];
    const int t = threadIdx.x;
    for (int i = t; i < nb; i += LT) tab[i] = lut[i];
    __syncthreads();
    auto map = [&](unsigned v) -> unsigned {
        const int i = (int)v - lo;
        return tab[i < 0 ? 0 : (i >= nb ? nb - 1 : i)];
    };
    const int64_t base = (int64_t)blockIdx.x * LCHUNK;
    const int64_t end = base + LCHUNK < count ? base + LCHUNK : count;
    if (VEC) {
        constexpr int STEP = LT * 8, UNR = 4;
        int64_t i0 = base + (int64_t)t * 8;
        for (; i0 + (UNR - 1) * (int64_t)STEP + 8 <= end; i0 += UNR * (int64_t)STEP) {
            uint4 d[UNR];
#pragma unroll
            for (int k = 0; k < UNR; ++k) d[k] = *reinterpret_cast<const uint4 *>(img + i0 + k * STEP);
#pragma unroll
            for (int k = 0; k < UNR; ++k) {
                const uint32_t w4[4] = {d[k].x, d[k].y, d[k].z, d[k].w};
                uint32_t o[2];
#pragma unroll
                for (int q = 0; q < 2; ++q)
                    o[q] = map(w4[2 * q] & 0xFFFFu) | (map(w4[2 * q] >> 16) << 8) |
                           (map(w4[2 * q + 1] & 0xFFFFu) << 16) | (map(w4[2 * q + 1] >> 16) << 24);
                *reinterpret_cast<uint2 *>(out + i0 + k * STEP) = make_uint2(o[0], o[1]);
            }
        }
        for (; i0 < end; i0 += STEP) {               // the remaining groups of 8 (and the tail)
            if (i0 + 8 <= end) {
                const uint4 d = *reinterpret_cast<const uint4 *>(img + i0);
                const uint32_t w4[4] = {d.x, d.y, d.z, d.w};
                uint32_t o[2];
#pragma unroll
                for (int q = 0; q < 2; ++q)
                    o[q] = map(w4[2 * q] & 0xFFFFu) | (map(w4[2 * q] >> 16) << 8) |
                           (map(w4[2 * q + 1] & 0xFFFFu) << 16) | (map(w4[2 * q + 1] >> 16) << 24);
                *reinterpret_cast<uint2 *>(out + i0) = make_uint2(o[0], o[1]);
            } else {
                for (int64_t i = i0; i < end; ++i) out[i] = (uint8_t)map(img[i]);
            }
        }
    } else {
        for (int64_t i = base + t; i < end; i += LT) out[i] = (uint8_t)map(img[i]);
    }
}

// ---------------------------------------------------------------- uint8 stretch
// out = uint8(trunc(clip(x / d * 255, 0, 255))), each operation rounded in
// float32 as numpy evaluates it; NaN gives 0.
__device__ __forceinline__ uint32_t stretch1(float x, float d) {
    const float v = __fmul_rn(__fdiv_rn(x, d), 255.0f);
    if (v != v) return 0u;
    return (uint32_t)(int)fminf(fmaxf(v, 0.0f), 255.0f);
}

constexpr int ST = 256;

template <bool VEC>
__global__ __launch_bounds__(ST) void k_stretch_u8(const float *__restrict__ x, int64_t count, float d,
                                                   uint8_t *__restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * ST;
    if (VEC) {
        const int64_t n4 = count / 4;
        for (int64_t i = (int64_t)blockIdx.x * ST + threadIdx.x; i < n4; i += stride) {
            const float4 v = reinterpret_cast<const float4 *>(x)[i];
            reinterpret_cast<uint32_t *>(out)[i] = stretch1(v.x, d) | (stretch1(v.y, d) << 8) |
                                                   (stretch1(v.z, d) << 16) | (stretch1(v.w, d) << 24);
        }
        for (int64_t i = n4 * 4 + (int64_t)blockIdx.x * ST + threadIdx.x; i < count; i += stride)
            out[i] = (uint8_t)stretch1(x[i], d);
    } else {
        for (int64_t i = (int64_t)blockIdx.x * ST + threadIdx.x; i < count; i += stride)
            out[i] = (uint8_t)stretch1(x[i], d);
    }
}

// ---------------------------------------------------------------- order statistics
// Radix select on the order-preserving key of the float bits (NaN sorted
// last, as np.partition / np.sort place it): three histogram passes over the
// key's bits 31..21, 20..10 and 9..0, each restricted to the values whose
// higher bits equal the prefix found so far, for two ranks at once.  One
// histogram workgroup per CU keeps the flushes of its counters to the global
// histogram few.  A one-workgroup scan between passes picks each rank's
// bucket on the device.
struct SelState {
    uint32_t prefix[2];                   // key bits found so far
    uint32_t pad[2];
    unsigned long long rank[2];           // rank left within the prefix
    unsigned long long nan_count;
    uint32_t key[2];                      // final keys
};
constexpr int SEL_BINS = 2048;
constexpr int SEL_T = 256;                // scan workgroup
constexpr int SEL_HT = 1024;              // histogram workgroups: one per CU

__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t b = __float_as_uint(f);
    if ((b & 0x7FFFFFFFu) > 0x7F800000u) return 0xFFFFFFFFu;   // every NaN: one largest key
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ void k_sel_init(SelState *st, unsigned long long k0, unsigned long long k1,
                           uint32_t *hist) {
    for (int i = threadIdx.x; i < 2 * SEL_BINS; i += blockDim.x) hist[i] = 0;
    if (threadIdx.x == 0) {
        st->prefix[0] = st->prefix[1] = 0;
        st->rank[0] = k0;
        st->rank[1] = k1;
        st->nan_count = 0;
    }
}

// level 0: bins = key >> 21 of every value (one histogram for both ranks);
// level 1: (key >> 10) & 2047 where key >> 21 == prefix[r];
// level 2: key & 1023 where key >> 10 == prefix[r].
template <int LEVEL, bool VEC>
__global__ __launch_bounds__(SEL_HT) void k_sel_hist(const float *__restrict__ x, int64_t count,
                                                     const SelState *__restrict__ st,
                                                     uint32_t *__restrict__ hist) {
    __shared__ uint32_t h[2][SEL_BINS];
    const int t = threadIdx.x;
    for (int i = t; i < 2 * SEL_BINS; i += SEL_HT) h[i / SEL_BINS][i % SEL_BINS] = 0;
    const uint32_t p0 = LEVEL ? st->prefix[0] : 0u, p1 = LEVEL ? st->prefix[1] : 0u;
    const bool two = LEVEL && p0 != p1;
    __syncthreads();
    constexpr int SHIFT = LEVEL == 0 ? 21 : (LEVEL == 1 ? 10 : 0);
    constexpr uint32_t MASK = LEVEL == 2 ? 1023u : 2047u;
    constexpr int PSHIFT = LEVEL == 1 ? 21 : 10;
    auto add = [&](float f) {
        const uint32_t k = fkey(f);
        const uint32_t b = (k >> SHIFT) & MASK;
        if (LEVEL == 0) {
            atomicAdd(&h[0][b], 1u);
        } else {
            const uint32_t p = k >> PSHIFT;
            if (p == p0) atomicAdd(&h[0][b], 1u);
            else if (two && p == p1) atomicAdd(&h[1][b], 1u);
        }
    };
    const int64_t stride = (int64_t)gridDim.x * SEL_HT;
    if (VEC) {
        const int64_t n4 = count / 4;
        const float4 *x4 = reinterpret_cast<const float4 *>(x);
        int64_t i = (int64_t)blockIdx.x * SEL_HT + t;
        for (; i + 3 * stride < n4; i += 4 * stride) {    // four 16-byte loads in flight
            float4 v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) v[k] = x4[i + k * stride];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                add(v[k].x);
                add(v[k].y);
                add(v[k].z);
                add(v[k].w);
            }
        }
        for (; i < n4; i += stride) {
            const float4 v = x4[i];
            add(v.x);
            add(v.y);
            add(v.z);
            add(v.w);
        }
        for (int64_t j = n4 * 4 + (int64_t)blockIdx.x * SEL_HT + t; j < count; j += stride) add(x[j]);
    } else {
        for (int64_t i = (int64_t)blockIdx.x * SEL_HT + t; i < count; i += stride) add(x[i]);
    }
    __syncthreads();
    for (int i = t; i < 2 * SEL_BINS; i += SEL_HT) {
        const uint32_t c = h[i / SEL_BINS][i % SEL_BINS];
        if (c) atomicAdd(&hist[i], c);
    }
}

// One workgroup: for each rank, the bucket whose cumulative count passes it;
// the prefix grows by the bucket and the rank drops by the counts below it.
// The histogram is cleared for the next pass.
template <int LEVEL>
__global__ __launch_bounds__(SEL_T) void k_sel_scan(SelState *__restrict__ st, uint32_t *__restrict__ hist) {
    constexpr int PER = SEL_BINS / SEL_T;    // bins per thread
    constexpr int BITS = LEVEL == 2 ? 10 : 11;
    constexpr int NB = 1 << BITS;
    __shared__ unsigned long long wsum[SEL_T / 64];
    __shared__ uint32_t pick[2];
    __shared__ unsigned long long below[2];
    const int t = threadIdx.x;
    const bool two = LEVEL && st->prefix[0] != st->prefix[1];
    if (LEVEL == 0 && t == 0) st->nan_count = hist[SEL_BINS - 1];
    for (int r = 0; r < 2; ++r) {
        const uint32_t *hh = hist + (two && r == 1 ? SEL_BINS : 0);
        unsigned long long c[PER], s = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int b = t * PER + k;
            c[k] = b < NB ? hh[b] : 0;
            s += c[k];
        }
        // exclusive prefix over threads: wave scan, then across waves
        unsigned long long incl = s;
        for (int off = 1; off < 64; off <<= 1) {
            const unsigned long long o = __shfl_up(incl, off, 64);
            if ((t & 63) >= off) incl += o;
        }
        if ((t & 63) == 63) wsum[t >> 6] = incl;
        __syncthreads();
        unsigned long long before = incl - s;
        for (int w = 0; w < (t >> 6); ++w) before += wsum[w];
        const unsigned long long rank = st->rank[r];
        if (rank >= before && rank < before + s) {
            unsigned long long acc = before;
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                if (rank >= acc && rank < acc + c[k]) {
                    pick[r] = (uint32_t)(t * PER + k);
                    below[r] = acc;
                }
                acc += c[k];
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        for (int r = 0; r < 2; ++r) {
            st->prefix[r] = (st->prefix[r] << BITS) | pick[r];
            st->rank[r] -= below[r];
            if (LEVEL == 2) st->key[r] = st->prefix[r];
        }
    }
    __syncthreads();
    for (int i = t; i < 2 * SEL_BINS; i += SEL_T) hist[i] = 0;
}

inline bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

float key_to_float(uint32_t k) {
    const uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

}  // namespace

// ---------------------------------------------------------------- launchers
static int launch_hist_u16(pfa_ctx *ctx, const uint16_t *img, int64_t count, uint32_t *hist) {
    hipLaunchKernelGGL(k_fill_u32, dim3(65536 / 256), dim3(256), 0, ctx->stream, hist, 65536, 0u);
    if (count > 0) {
        const int64_t need = (count + (int64_t)HT * 64 - 1) / ((int64_t)HT * 64);
        const unsigned wg = (unsigned)std::max<int64_t>(1, std::min<int64_t>(need, ctx->num_cu));
        if (aligned16(img))
            hipLaunchKernelGGL(k_hist_u16<true>, dim3(wg), dim3(HT), 0, ctx->stream, img, count, hist);
        else
            hipLaunchKernelGGL(k_hist_u16<false>, dim3(wg), dim3(HT), 0, ctx->stream, img, count, hist);
    }
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

static int launch_lut_u16_u8(pfa_ctx *ctx, const uint16_t *img, int64_t count, const uint8_t *lut,
                             int lo, int nb, uint8_t *out) {
    if (count <= 0) return PFA_OK;
    const int64_t wg = (count + LCHUNK - 1) / LCHUNK;
    if (aligned16(img) && ((uintptr_t)out & 7u) == 0)
        hipLaunchKernelGGL(k_lut_u16_u8<true>, dim3((unsigned)wg), dim3(LT), 0, ctx->stream, img, count, lut, lo, nb, out);
    else
        hipLaunchKernelGGL(k_lut_u16_u8<false>, dim3((unsigned)wg), dim3(LT), 0, ctx->stream, img, count, lut, lo, nb, out);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

static unsigned stream_blocks(pfa_ctx *ctx, int64_t count, int threads, int per_thread) {
    const int64_t need = (count + (int64_t)threads * per_thread - 1) / ((int64_t)threads * per_thread);
    const int64_t cap = (int64_t)ctx->num_cu * 8;
    return (unsigned)(need < 1 ? 1 : (need < cap ? need : cap));
}

static int launch_stretch_u8(pfa_ctx *ctx, const float *x, int64_t count, float d, uint8_t *out) {
    if (count <= 0) return PFA_OK;
    const unsigned g = stream_blocks(ctx, count, ST, 16);
    if (aligned16(x) && ((uintptr_t)out & 3u) == 0)
        hipLaunchKernelGGL(k_stretch_u8<true>, dim3(g), dim3(ST), 0, ctx->stream, x, count, d, out);
    else
        hipLaunchKernelGGL(k_stretch_u8<false>, dim3(g), dim3(ST), 0, ctx->stream, x, count, d, out);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

static int launch_select_f32(pfa_ctx *ctx, const float *x, int64_t count, int64_t k0, int64_t k1,
                             SelState *st, uint32_t *hist) {
    const int64_t need = (count + (int64_t)SEL_HT * 64 - 1) / ((int64_t)SEL_HT * 64);
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(need, ctx->num_cu));
    const bool vec = aligned16(x);
    hipLaunchKernelGGL(k_sel_init, dim3(1), dim3(SEL_T), 0, ctx->stream, st, (unsigned long long)k0,
                       (unsigned long long)k1, hist);
#define SEL_LEVEL(L)                                                                                  \
    if (vec) hipLaunchKernelGGL((k_sel_hist<L, true>), dim3(g), dim3(SEL_HT), 0, ctx->stream, x, count, st, hist); \
    else hipLaunchKernelGGL((k_sel_hist<L, false>), dim3(g), dim3(SEL_HT), 0, ctx->stream, x, count, st, hist); \
    hipLaunchKernelGGL(k_sel_scan<L>, dim3(1), dim3(SEL_T), 0, ctx->stream, st, hist);
    SEL_LEVEL(0)
    SEL_LEVEL(1)
    SEL_LEVEL(2)
#undef SEL_LEVEL
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

// ---------------------------------------------------------------- entry points
extern "C" int pfa_hist_u16_dev(pfa_ctx *ctx, const uint16_t *img, int64_t count, uint32_t *hist) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    if (count < 0 || !hist || (count && !img)) return pfa_fail(ctx, PFA_ERR_ARG, "hist_u16: bad argument");
    pfa_timer_begin(ctx, PFA_TIMER_OTHER);
    const int st = launch_hist_u16(ctx, img, count, hist);
    pfa_timer_end(ctx, PFA_TIMER_OTHER);
    return st;
    PFA_CATCH(ctx)
}

extern "C" int pfa_hist_u16(pfa_ctx *ctx, const uint16_t *img, int64_t count, uint32_t *hist) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    if (count < 0 || !hist || (count && !img)) return pfa_fail(ctx, PFA_ERR_ARG, "hist_u16: bad argument");
    int st = PFA_OK;
    uint16_t *di = (uint16_t *)pfa_ws(ctx, "io_in", (size_t)(count ? count : 1) * 2, &st);
    if (!di) return st;
    uint32_t *dh = (uint32_t *)pfa_ws(ctx, "disp_hist", 65536 * 4, &st);
    if (!dh) return st;
    if (count) CK(pfa_io_h2d(ctx, di, img, (size_t)count * 2));
    CK(pfa_hist_u16_dev(ctx, di, count, dh));
    CK(pfa_io_d2h(ctx, hist, dh, 65536 * 4));
    return PFA_OK;
    PFA_CATCH(ctx)
}

extern "C" int pfa_lut_u16_u8_dev(pfa_ctx *ctx, const uint16_t *img, int64_t count, const uint8_t *lut,
                                  int lo, int nb, uint8_t *out) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    if (count < 0 || nb < 1 || nb > 65536 || lo < 0 || lo > 65535 || !lut || (count && (!img || !out)))
        return pfa_fail(ctx, PFA_ERR_ARG, "lut_u16_u8: bad argument");
    pfa_timer_begin(ctx, PFA_TIMER_OTHER);
    const int st = launch_lut_u16_u8(ctx, img, count, lut, lo, nb, out);
    pfa_timer_end(ctx, PFA_TIMER_OTHER);
    return st;
    PFA_CATCH(ctx)
}

extern "C" int pfa_lut_u16_u8(pfa_ctx *ctx, const uint16_t *img, int64_t count, const uint8_t *lut,
                              int lo, int nb, uint8_t *out) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    if (count < 0 || nb < 1 || nb > 65536 || lo < 0 || lo > 65535 || !lut || (count && (!img || !out)))
        return pfa_fail(ctx, PFA_ERR_ARG, "lut_u16_u8: bad argument");
    if (!count) return PFA_OK;
    int st = PFA_OK;
    uint16_t *di = (uint16_t *)pfa_ws(ctx, "io_in", (size_t)count * 2, &st);
    if (!di) return st;
    uint8_t *dl = (uint8_t *)pfa_ws(ctx, "disp_lut", 65536, &st);
    if (!dl) return st;
    uint8_t *dout = (uint8_t *)pfa_ws(ctx, "io_out", (size_t)count, &st);
    if (!dout) return st;
    CK(pfa_io_h2d(ctx, di, img, (size_t)count * 2));
    CK(pfa_io_h2d(ctx, dl, lut, (size_t)nb));
    CK(pfa_lut_u16_u8_dev(ctx, di, count, dl, lo, nb, dout));
    CK(pfa_io_d2h(ctx, out, dout, (size_t)count));
    return PFA_OK;
    PFA_CATCH(ctx)
}

extern "C" int pfa_stretch_u8_dev(pfa_ctx *ctx, const float *x, int64_t count, float d, uint8_t *out) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    if (count < 0 || (count && (!x || !out))) return pfa_fail(ctx, PFA_ERR_ARG, "stretch_u8: bad argument");
    pfa_timer_begin(ctx, PFA_TIMER_OTHER);
    const int st = launch_stretch_u8(ctx, x, count, d, out);
    pfa_timer_end(ctx, PFA_TIMER_OTHER);
    return st;
    PFA_CATCH(ctx)
}

extern "C" int pfa_stretch_u8(pfa_ctx *ctx, const float *x, int64_t count, float d, uint8_t *out) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    if (count < 0 || (count && (!x || !out))) return pfa_fail(ctx, PFA_ERR_ARG, "stretch_u8: bad argument");
    if (!count) return PFA_OK;
    int st = PFA_OK;
    float *dx = (float *)pfa_ws(ctx, "io_in", (size_t)count * 4, &st);
    if (!dx) return st;
    uint8_t *dout = (uint8_t *)pfa_ws(ctx, "io_out", (size_t)count, &st);
    if (!dout) return st;
    CK(pfa_io_h2d(ctx, dx, x, (size_t)count * 4));
    CK(pfa_stretch_u8_dev(ctx, dx, count, d, dout));
    CK(pfa_io_d2h(ctx, out, dout, (size_t)count));
    return PFA_OK;
    PFA_CATCH(ctx)
}

extern "C" int pfa_select_f32_dev(pfa_ctx *ctx, const float *x, int64_t count, int64_t k0, int64_t k1,
                                  float *v0, float *v1, int64_t *nan_count) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    if (count < 1 || !x || !v0 || !v1 || k0 < 0 || k1 < 0 || k0 >= count || k1 >= count)
        return pfa_fail(ctx, PFA_ERR_ARG, "select_f32: ranks must lie in [0, count)");
    int st = PFA_OK;
    auto *sst = (SelState *)pfa_ws(ctx, "disp_sel", sizeof(SelState), &st);
    if (!sst) return st;
    auto *hist = (uint32_t *)pfa_ws(ctx, "disp_selhist", 2 * SEL_BINS * 4, &st);
    if (!hist) return st;
    pfa_timer_begin(ctx, PFA_TIMER_OTHER);
    CK(launch_select_f32(ctx, x, count, k0, k1, sst, hist));
    pfa_timer_end(ctx, PFA_TIMER_OTHER);
    SelState h;
    PFA_HIP(ctx, hipMemcpyAsync(&h, sst, sizeof(SelState), hipMemcpyDeviceToHost, ctx->stream));
    PFA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *v0 = key_to_float(h.key[0]);
    *v1 = key_to_float(h.key[1]);
    if (nan_count) *nan_count = (int64_t)h.nan_count;
    return PFA_OK;
    PFA_CATCH(ctx)
}

extern "C" int pfa_select_f32(pfa_ctx *ctx, const float *x, int64_t count, int64_t k0, int64_t k1,
                              float *v0, float *v1, int64_t *nan_count) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    if (count < 1 || !x || !v0 || !v1 || k0 < 0 || k1 < 0 || k0 >= count || k1 >= count)
        return pfa_fail(ctx, PFA_ERR_ARG, "select_f32: ranks must lie in [0, count)");
    int st = PFA_OK;
    float *dx = (float *)pfa_ws(ctx, "io_in", (size_t)count * 4, &st);
    if (!dx) return st;
    CK(pfa_io_h2d(ctx, dx, x, (size_t)count * 4));
    return pfa_select_f32_dev(ctx, dx, count, k0, k1, v0, v1, nan_count);
    PFA_CATCH(ctx)
}
