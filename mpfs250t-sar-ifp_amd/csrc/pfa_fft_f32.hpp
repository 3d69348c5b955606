// pfa_fft_f32.hpp -- LDS-resident FP32 FFT for one line of N = 2^L points.
//
// Stockham autosort passes of radix R <= 32 held in registers: each thread
// reads its R-point butterflies from the LDS line, applies the inter-pass
// twiddles W_{Ls R}^{j r} (two-level tables in LDS, see Tw), runs the R-point
// DFT in registers and writes the results back to the same LDS line after a
// barrier.  The last pass may hand
// its outputs to a caller-provided sink instead of LDS.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace pfa32 {

constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v >> 1); }

// exp(-2 pi i k / 32) for k < 16, as (cos, sin) of +2 pi k / 32
__device__ constexpr float kC32[16][2] = {
    {1.0f, 0.0f},
    {0.9807852804032304f, 0.19509032201612825f},
    {0.9238795325112867f, 0.3826834323650898f},
    {0.8314696123025452f, 0.5555702330196022f},
    {0.7071067811865476f, 0.7071067811865475f},
    {0.5555702330196023f, 0.8314696123025452f},
    {0.38268343236508984f, 0.9238795325112867f},
    {0.19509032201612833f, 0.9807852804032304f},
    {0.0f, 1.0f},
    {-0.1950903220161282f, 0.9807852804032304f},
    {-0.3826834323650897f, 0.9238795325112867f},
    {-0.555570233019602f, 0.8314696123025455f},
    {-0.7071067811865475f, 0.7071067811865476f},
    {-0.8314696123025453f, 0.5555702330196022f},
    {-0.9238795325112867f, 0.3826834323650899f},
    {-0.9807852804032304f, 0.1950903220161286f},
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// In-register R-point DFT (R = 2^p <= 32), forward sign exp(-i...) or inverse.
template <int R, bool INV>
__device__ __forceinline__ void dft(float2 (&x)[R]) {
    constexpr int P = ilog2c(R);
    // bit-reverse the order (compile-time swaps)
#pragma unroll
    for (int i = 0; i < R; ++i) {
        int r = 0;
#pragma unroll
        for (int b = 0; b < P; ++b) r |= ((i >> b) & 1) << (P - 1 - b);
        if (r > i) {
            const float2 t = x[i];
            x[i] = x[r];
            x[r] = t;
        }
    }
#pragma unroll
    for (int s = 1; s <= P; ++s) {
        const int m = 1 << s, h = m >> 1;
#pragma unroll
        for (int base = 0; base < R; base += m) {
#pragma unroll
            for (int j = 0; j < h; ++j) {
                const float2 a = x[base + j], b = x[base + j + h];
                float2 t;
                if (j == 0) {
                    t = b;
                } else if (4 * j == m) {          // W = -i (forward) or +i (inverse)
                    t = INV ? make_float2(-b.y, b.x) : make_float2(b.y, -b.x);
                } else {
                    const int k = j * (32 / m);
                    const float c = kC32[k][0], sn = INV ? kC32[k][1] : -kC32[k][1];
                    t = make_float2(fmaf(b.x, c, -b.y * sn), fmaf(b.x, sn, b.y * c));
                }
                x[base + j] = cadd(a, t);
                x[base + j + h] = csub(a, t);
            }
        }
    }
}

// Line of N = 2^L points held by NT threads with E = N / NT elements each
// (E <= EMAX, NT >= min(64, N/2)).
template <int L, int EMAX = 32>
struct Cfg {
    static constexpr int N = 1 << L;
    static constexpr int NT_RAW = (N / EMAX > 64) ? N / EMAX : (N / 2 < 64 ? N / 2 : 64);
    static constexpr int NT = NT_RAW;
    static constexpr int NB = NT < 64 ? 64 : NT;      // block size
    static constexpr int E = N / NT;                  // elements per thread
    static constexpr int PADN = N + N / 32;           // padded float2 slots
    // radix of pass k: min(E, remaining)
    static constexpr int radix(int k) {
        int rem = N, r = 1;
        for (int i = 0; i <= k; ++i) {
            r = rem < E ? rem : E;
            rem /= r;
        }
        return r;
    }
    static constexpr int passes() {
        int rem = N, c = 0;
        while (rem > 1) {
            const int r = rem < E ? rem : E;
            rem /= r;
            ++c;
        }
        return c;
    }
};

__device__ __forceinline__ int padf(int i) { return i + (i >> 5); }

// Twiddle exp(-2 pi i t / N) from two small tables staged in LDS:
// hi[t >> 6] = exp(-2 pi i 64 (t >> 6) / N), lo[t & 63] = exp(-2 pi i (t & 63) / N),
// both rounded from double on the host; one complex product per twiddle.
struct Tw {
    const float2 *hi, *lo;
    __device__ __forceinline__ float2 operator()(int t) const { return cmul(hi[t >> 6], lo[t & 63]); }
};

// Stage the two twiddle tables of an N-point line (N / 64 + 64 float2).
template <int N>
__device__ __forceinline__ Tw load_tw(float2 *lds, const float2 *__restrict__ g) {
    for (int i = threadIdx.x; i < N / 64 + 64; i += blockDim.x) lds[i] = g[i];
    return Tw{lds, lds + N / 64};
}

// A first-pass source is called as src(e) or, when it takes two arguments, as
// src(e, s) with the compile-time slot s: e = threadIdx.x + s NT, so a source
// holding a thread's elements in a register array can index it statically.
template <typename Src>
__device__ __forceinline__ float2 call_src(Src &src, int e, int s) {
    if constexpr (std::is_invocable_v<Src &, int, int>) return src(e, s);
    else return src(e);
}

// One Stockham pass over the LDS line `buf`; LS = product of earlier radices.
// On the last pass the outputs go to sink(index, value) instead of LDS.  That
// pass writes nothing to LDS itself, so its reads need a barrier only when the
// sink writes the line (SINK_LDS), and its closing barrier only when the
// caller writes the line next without one of its own (END_SYNC).
template <int L, int EMAX, int R, int LS, bool INV, bool LAST, typename Src, typename Sink,
          bool SINK_LDS = true, bool END_SYNC = true>
__device__ __forceinline__ void pass(float2 *buf, const Tw &tw, Src *src, Sink &sink) {
    using C = Cfg<L, EMAX>;
    constexpr int N = C::N, NT = C::NT, B = C::E / R;   // butterflies per thread
    // opaque per pass: in a loop over lines, the twiddles and indices would
    // otherwise be hoisted and held in registers across the whole loop
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    t &= 1023;                                         // known non-negative: k / LS, k % LS are shifts
    // For power-of-two N, R and LS the padded index splits exactly:
    // padf(k + r N/R) = padf(k) + r N/R + (r N/R) / 32 for k < N/R, and
    // padf(o + r LS) = padf(o) + r LS + (r LS) / 32 for the outputs o of a
    // butterfly, so every element of a butterfly sits at a compile-time offset
    // from one base address (immediate offsets of the LDS instructions).
    constexpr int NR = N / R;
    float2 x[B][R];
    if (t < NT) {
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int k = t + b * NT;
            const float2 *rb = buf + padf(k);
#pragma unroll
            for (int r = 0; r < R; ++r)
                x[b][r] = src ? call_src(*src, k + r * NR, b + r * B) : rb[r * NR + (r * NR) / 32];
        }
    }
    if (!src && (!LAST || SINK_LDS)) __syncthreads();
    if (t < NT) {
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int k = t + b * NT;
            const int j = k % LS;
            if (LS > 1) {
                constexpr int step = N / (LS * R);
                if constexpr (LS == 2 && R == 16) {
                    // W^(j r N/32) = W_32^(j r), j = k % 2: exact constants for j = 1
                    if (j) {
#pragma unroll
                        for (int r = 1; r < R; ++r) {
                            const float c = kC32[r][0], sn = INV ? kC32[r][1] : -kC32[r][1];
                            x[b][r] = make_float2(fmaf(x[b][r].x, c, -x[b][r].y * sn),
                                                  fmaf(x[b][r].x, sn, x[b][r].y * c));
                        }
                    }
                } else if constexpr (R >= 8) {
                    // W^(j r step) for r = 4 a + c from two lookups, w1 = W^(j step)
                    // and w4 = W^(4 j step) (< N / 2), and at most four products deep:
                    // w2 = w1 w1, w3 = w2 w1, W^(4a) = w4^a, W^r = W^(4a) w_c.
                    float2 w1 = tw(j * step), w4 = tw(4 * j * step);
                    if (INV) {
                        w1.y = -w1.y;
                        w4.y = -w4.y;
                    }
                    const float2 w2 = cmul(w1, w1), w3 = cmul(w2, w1);
                    const float2 wc[4] = {make_float2(1.f, 0.f), w1, w2, w3};
                    float2 wa = w4;
#pragma unroll
                    for (int a = 0; a < R / 4; ++a) {
                        if (a >= 2) wa = cmul(wa, w4);
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            const int r = 4 * a + c;
                            if (r == 0) continue;
                            const float2 w = a == 0 ? wc[c] : (c == 0 ? wa : cmul(wa, wc[c]));
                            x[b][r] = cmul(x[b][r], w);
                        }
                    }
                } else {
#pragma unroll
                    for (int r = 1; r < R; ++r) {
                        float2 w = tw(j * r * step);
                        if (INV) w.y = -w.y;
                        x[b][r] = cmul(x[b][r], w);
                    }
                }
            }
            dft<R, INV>(x[b]);
            const int o = (k / LS) * LS * R + j;
            float2 *wb = buf + padf(o);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (LAST) sink(o + r * LS, x[b][r]);
                else wb[r * LS + (r * LS) / 32] = x[b][r];
            }
        }
    }
    if (!LAST || END_SYNC) __syncthreads();
}

template <int L, int EMAX, bool INV, int K, int LS, bool SINK_LDS, bool END_SYNC, typename Src,
          typename Sink>
__device__ __forceinline__ void fft_from(float2 *buf, const Tw &tw, Src *src, Sink &sink) {
    using C = Cfg<L, EMAX>;
    if constexpr (K < C::passes()) {
        constexpr int R = C::radix(K);
        pass<L, EMAX, R, LS, INV, K == C::passes() - 1, Src, Sink, SINK_LDS, END_SYNC>(
            buf, tw, K == 0 ? src : (Src *)nullptr, sink);
        fft_from<L, EMAX, INV, K + 1, LS * R, SINK_LDS, END_SYNC>(buf, tw, src, sink);
    }
}

struct NoSrc {
    __device__ float2 operator()(int) const { return make_float2(0.f, 0.f); }
};

// Full transform of the padded natural-order line in `buf`; the last pass
// delivers (index, value) to `sink` (which may write back to buf).
template <int L, bool INV, int EMAX = 32, typename Sink>
__device__ __forceinline__ void fft(float2 *buf, const Tw &tw, Sink &&sink) {
    fft_from<L, EMAX, INV, 0, 1, true, true>(buf, tw, (NoSrc *)nullptr, sink);
}

// The passes after a radix-G first pass that the caller ran in registers and
// stored to the line (outputs G k + r of the butterfly over k + r N/G): radix
// min(E, N / LS) from LS = G on.  The last pass runs without barriers
// (the sink writes global memory, and the caller synchronises before it
// writes the line again).
template <int L, int EMAX, bool INV, int LS, typename Sink>
__device__ __forceinline__ void fft_rest(float2 *buf, const Tw &tw, Sink &sink) {
    constexpr int N = 1 << L, E = Cfg<L, EMAX>::E;
    if constexpr (LS < N) {
        constexpr int R = N / LS < E ? N / LS : E;
        pass<L, EMAX, R, LS, INV, LS * R == N, NoSrc, Sink, false, false>(buf, tw, (NoSrc *)nullptr, sink);
        fft_rest<L, EMAX, INV, LS * R>(buf, tw, sink);
    }
}

// As fft, but the first pass loads element e with src(e) (e.g. from global
// memory) instead of from the LDS line.  SINK_LDS / END_SYNC as for pass: a
// sink that writes only global memory and a caller that synchronises before
// it next writes the line may drop the last pass's barriers.
template <int L, bool INV, int EMAX = 32, bool SINK_LDS = true, bool END_SYNC = true, typename Src,
          typename Sink>
__device__ __forceinline__ void fft_src(float2 *buf, const Tw &tw, Src &&src, Sink &&sink) {
    fft_from<L, EMAX, INV, 0, 1, SINK_LDS, END_SYNC>(buf, tw, &src, sink);
}

}  // namespace pfa32
