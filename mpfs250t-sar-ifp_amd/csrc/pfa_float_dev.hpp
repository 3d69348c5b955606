// pfa_float_dev.hpp -- device code of the float (FP32) Polar Format focus,
// shared by the FP32 launchers (pfa_float.hip) and the matrix-core FFT
// launchers (pfa_mfma.hip).  See pfa_float.hip for the kernel list.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "pfa_fft_f32.hpp"
#include "pfa_fft_mfma.hpp"
#include "pfa_div.hpp"
#include "pfa_internal.hpp"

namespace {

constexpr double C_LIGHT = 299792458.0;

// ------------------------------------------------------------------ shared machinery
// A uniform grid as np.linspace forms it -- g[k] = k * step + start (each op
// rounded), g[cnt-1] = stop -- or an explicit table when the grid is not one.
struct Grid1 {
    const double *tab;
    double start, step, stop;
    int cnt;
    __device__ __forceinline__ double operator()(int k) const {
        if (tab) return tab[k];
        return k == cnt - 1 ? stop : __dadd_rn(__dmul_rn((double)k, step), start);
    }
};

typedef float f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ld_stream(const float2 *p) {      // read-once data
    const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v *>(p));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void st_stream(float2 *p, float2 v) {     // write-once data
    f2v w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<f2v *>(p));
}
// Taps of the Q15 sinc table as float (exact: c / 2^15; the autofocus adjoint's
// table), staged in LDS tap-major: tap t of phase ph at tapl[(t + 4) * 256 + scr(ph)], with four zero
// rows on either side (t in [-4, 36)) so a window shifted by up to KD samples
// reads zeros instead of needing masks.  The LDS has 64 banks of 4 bytes, so a
// row's word for phase ph sits in bank ph % 64: lanes whose phases step
// through a line (the azimuth pass) or differ by pulse (the range pass) spread
// over the banks.  Measured at 8192^2 against the earlier ph ^ (ph >> 3)
// column order: bank-conflict cycles of the range pass 34.8 M -> 27.4 M per
// launch (with the half offset below), images identical.
constexpr int TROWS = 40;
__device__ __forceinline__ int scr(int ph) { return ph; }
__device__ __forceinline__ void load_taps(float *tapl, const int16_t *__restrict__ sinc) {
    for (int i = threadIdx.x; i < TROWS * 256; i += blockDim.x) {
        const int row = i >> 8, ph = i & 255, t = row - 4;
        tapl[row * 256 + scr(ph)] = (t >= 0 && t < 32) ? (float)sinc[ph * 32 + t] * (1.0f / 32768.0f) : 0.f;
    }
}

// The same taps packed as Q15 integer pairs for the sinc gathers: word
// tp[h * TPH + r * 256 + scr(ph)] holds rows (2r + h, 2r + h + 1) of the
// zero-padded tap rows (low half first), h = 0 or 1, r < 20, so a window
// shifted by any d reads aligned pairs.  Half the LDS bytes of float taps; one
// conversion per tap; sums of integer taps scaled by 2^-15 at the end equal
// the float-tap sums exactly (a power-of-two scale commutes with rounding).
// The rows stay 256 words apart (the gathers read two rows per instruction
// at offsets in 64-word units), so every row of a half is on the same banks;
// the h = 1 half starts 32 words further on, so outputs of one phase whose
// windows are shifted by an odd and an even d read different banks.
constexpr int TPR = TROWS / 2, TPH = TPR * 256 + 32, TPWORDS = 2 * TPH;
static_assert(TPH == PFA_TAP2_HALF, "the host builds the same layout");
__device__ __forceinline__ void load_taps2(unsigned *tp, const int16_t *__restrict__ sinc) {
    for (int i = threadIdx.x; i < 2 * TPR * 256; i += blockDim.x) {
        const int h = i / (TPR * 256), r = (i / 256) % TPR, ph = i & 255;
        const int ra = 2 * r + h - 4, rb = ra + 1;           // tap indices of the two rows
        const int a = (ra >= 0 && ra < 32) ? sinc[ph * 32 + ra] : 0;
        const int b = (rb >= 0 && rb < 32) ? sinc[ph * 32 + rb] : 0;
        tp[h * TPH + r * 256 + scr(ph)] = ((unsigned)a & 0xffffu) | ((unsigned)b << 16);
    }
}
__device__ __forceinline__ float tap_lo(unsigned w) { return (float)(short)(w & 0xffffu); }
__device__ __forceinline__ float tap_hi(unsigned w) { return (float)((int)w >> 16); }

constexpr int KD = 4;            // bracket advance covered by one 4-output window
constexpr int KW = 32 + KD;      // samples in that window

// Four consecutive outputs whose brackets J_u = J_0 + d_u (0 <= d_u <= KD) share
// one window w[k] = row[J_0 - 15 + k], k < KW; output u uses taps shifted by
// d_u: sum_k tap[ph_u][k - d_u] w[k].  Each window sample is read once per four
// outputs instead of once per output.
// The row is an unpadded LDS array: every window read is an LDS load with an
// immediate offset, so the loop is pure loads and FMAs.
// The gather over the packed integer taps (load_taps2): two window samples
// per tap word.
// PAD16: the row is stored with one spare slot per 16 samples (pad16), so the
// windows of lanes four samples apart start in different LDS banks; a batch of
// BATCH <= 16 samples crosses at most one spare slot.
__device__ __forceinline__ int pad16(int s) { return s + (s >> 4); }

template <int BATCH = 6, bool PAD16 = false>
__device__ __forceinline__ void sinc4p(const float2 *row, int start, const unsigned *tp,
                                       const int (&ph)[4], const int (&d)[4], float2 (&out)[4]) {
    static_assert(KW % BATCH == 0 && BATCH % 2 == 0 && BATCH <= 16, "window splits into batches of pairs");
    const unsigned *col[4];
    float ar[4], ai[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int base = 4 - d[u];                     // first tap row of output u
        col[u] = tp + (base & 1) * TPH + (base >> 1) * 256 + scr(ph[u]);
        ar[u] = 0.f;
        ai[u] = 0.f;
    }
    const float2 *ws = row + start;
#pragma unroll
    for (int kb = 0; kb < KW; kb += BATCH) {
        const float2 *wa = ws, *wb = ws;               // before / after the spare slot
        int kbrk = BATCH;
        if constexpr (PAD16) {
            const int sb = start + kb;
            wa = row + pad16(sb);
            wb = wa + 1;
            kbrk = 16 - (sb & 15);
        }
#pragma unroll
        for (int k = 0; k < BATCH; k += 2) {
            const float2 w0 = (k < kbrk ? wa : wb)[k], w1 = (k + 1 < kbrk ? wa : wb)[k + 1];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const unsigned v = col[u][(k / 2) * 256];
                const float c0 = tap_lo(v), c1 = tap_hi(v);
                ar[u] = fmaf(c0, w0.x, ar[u]);
                ai[u] = fmaf(c0, w0.y, ai[u]);
                ar[u] = fmaf(c1, w1.x, ar[u]);
                ai[u] = fmaf(c1, w1.y, ai[u]);
            }
        }
        ws += BATCH;
#pragma unroll
        for (int u = 0; u < 4; ++u) col[u] += (BATCH / 2) * 256;
    }
    constexpr float Q = 1.0f / 32768.0f;
#pragma unroll
    for (int u = 0; u < 4; ++u) out[u] = make_float2(ar[u] * Q, ai[u] * Q);
}

// SINC32 value of one interior output with the packed taps.
template <typename FA>
__device__ __forceinline__ float2 sinc1p(FA &&fa, int J, int ph, const unsigned *tp) {
    const unsigned *col = tp + 2 * 256 + scr(ph);      // rows 4 + 2i, 5 + 2i = taps 2i, 2i + 1
    float ar = 0.f, ai = 0.f;
#pragma unroll 4
    for (int i = 0; i < 16; ++i) {
        const unsigned v = col[i * 256];
        const float2 s0 = fa(J - 15 + 2 * i), s1 = fa(J - 14 + 2 * i);
        ar = fmaf(tap_lo(v), s0.x, ar);
        ai = fmaf(tap_lo(v), s0.y, ai);
        ar = fmaf(tap_hi(v), s1.x, ar);
        ai = fmaf(tap_hi(v), s1.y, ai);
    }
    return make_float2(ar * (1.0f / 32768.0f), ai * (1.0f / 32768.0f));
}

// Sum a per-thread count over the wave and add it once (fallback statistics).
__device__ __forceinline__ void count_wave(unsigned *dst, unsigned v) {
    if (!dst) return;
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
}

// ------------------------------------------------------------------ pass 1 (range)
constexpr int RP = 16;                   // pulses per tile
constexpr int RQ = 256;                  // range bins per tile
constexpr int RT = 512;                  // threads: 16 pulses x 32 groups of 4 bins, twice
constexpr int RS = 304;                  // staged samples per pulse (RQ + window + slack)
constexpr int RSP = RS + 1;              // odd pitch: a wave's 16 pulses hit different banks
constexpr int RLD = (RS + 31) / 32;      // strip samples loaded per thread (32 threads a pulse)
static_assert(RT == 32 * RP, "32 strip-loading threads per pulse");

// Exact count c = #{u : xa(u) <= Q} over an ascending view of S positions,
// walking from a guess (a few steps), bisecting when the guess is far off;
// leaves xl = xa(c - 1), xh = xa(c) (+-inf past the ends).
template <typename XA>
__device__ __forceinline__ int bracket_from(XA &xa, int S, double Q, int g, double &xl, double &xh) {
    int c = min(max(g, 0), S);
    xl = c > 0 ? xa(c - 1) : -INFINITY;
    xh = c < S ? xa(c) : INFINITY;
    if ((xl <= Q) && !(xh <= Q)) return c;
    int steps = 0;
    while (c > 0 && !(xl <= Q) && steps < 8) { --c; ++steps; xh = xl; xl = c > 0 ? xa(c - 1) : -INFINITY; }
    while (c < S && xh <= Q && steps < 8) { ++c; ++steps; xl = xh; xh = c < S ? xa(c) : INFINITY; }
    if ((c == 0 || xl <= Q) && (c == S || !(xh <= Q))) return c;
    int lo = 0, hi = S;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (xa(mid) <= Q) lo = mid + 1; else hi = mid;
    }
    c = lo;
    xl = c > 0 ? xa(c - 1) : -INFINITY;
    xh = c < S ? xa(c) : INFINITY;
    return c;
}

// Persistent: each workgroup stages the tap table once, then walks tiles of
// RP sorted pulses x RQ range bins (its pulse groups in turn, each swept
// along range so strip halos are L2 hits).  Per tile: every pulse's source
// strip goes to LDS while each thread finds the FP64 brackets of its 8 bins
// (closed-form guess from the uniform frequency grid, exact comparisons);
// then four-bin shared-window sinc sums (or np.interp) stored straight to
// the corner-turned g1t rows.
// Lane layout: thread t serves pulse t % 16 and bin groups t / 16 and
// t / 16 + 32, so the lanes of a wave read 16 different strips and each
// store writes 128 contiguous bytes of four g1t rows.
template <bool SINC, bool FULLFREQ>
__global__ __launch_bounds__(RT, 2) void k_range_f32(
    const float2 *__restrict__ sig, int m, int n, const int32_t *__restrict__ order,
    const double *__restrict__ a64, const double *__restrict__ pr64, const double *__restrict__ f0v,
    const double *__restrict__ dfv, const double *__restrict__ freq, const Grid1 KR,
    const int16_t *__restrict__ sinc, float2 *__restrict__ g1t, int ldp,
    unsigned *__restrict__ nslow) {
    static_assert(RQ == 4 * 2 * (RT / RP), "two groups of four bins per thread");
    __shared__ float2 strip[RP * RSP];                 // source strips
    __shared__ unsigned tp[SINC ? TPWORDS : 1];
    __shared__ int lo_s[RP], hi_s[RP], st_s[RP], dir_s[RP];
    __shared__ int ps_i[RP];                          // per pulse of the current group: sorted index
    __shared__ double ps_a[RP], ps_f0[RP], ps_df[RP], ps_pr[RP];  // 2 pr / c, f0, df, pr
    __shared__ double ps_x0[RP], ps_iad[RP];          // a f0, 1 / (a df) (bracket guesses)
    __shared__ int ps_asc[RP];                        // ascending positions
    if (SINC) load_taps2(tp, sinc);
    const int tiles_p = (m + RP - 1) / RP, tiles_q = (n + RQ - 1) / RQ;
    const int my_groups = blockIdx.x < tiles_p ? (tiles_p - 1 - blockIdx.x) / gridDim.x + 1 : 0;
    unsigned slow = 0;
    for (int it = 0; it < my_groups * tiles_q; ++it) {
        const int p0 = (blockIdx.x + (it / tiles_q) * gridDim.x) * RP, q0 = (it % tiles_q) * RQ;
        const int qe = min(q0 + RQ, n);
        int t = threadIdx.x;                          // opaque: per-thread values stay in the loop
        asm volatile("" : "+v"(t));
        t &= RT - 1;                                  // ... known in [0, RT): shifts, not signed division
        const int pl = t % RP, gi = t / RP;           // pulse in tile, bin group
        const int p = p0 + pl;
        if (q0 == 0) {                                // first tile of a pulse group: its pulse values to LDS
            __syncthreads();                          // the previous group's tiles are done with them
            if (t < RP) {
                int ii = 0;
                double aa = 1.0, ff = 0.0, dd = 1.0, pv = 1.0;
                if (p0 + t < m) {
                    ii = order[p0 + t];
                    aa = a64[ii];
                    pv = pr64[ii];
                    if (FULLFREQ) {
                        ff = freq[(size_t)ii * n];
                        dd = __ddiv_rn(__dsub_rn(freq[(size_t)ii * n + n - 1], ff), (double)(n - 1));
                    } else {
                        ff = f0v[ii];
                        dd = dfv[ii];
                    }
                }
                ps_i[t] = ii;
                ps_a[t] = aa;
                ps_f0[t] = ff;
                ps_df[t] = dd;
                ps_pr[t] = pv;
                auto fqj = [&](int j) { return FULLFREQ ? freq[(size_t)ii * n + j] : __dadd_rn(ff, __dmul_rn((double)j, dd)); };
                auto knp = [&](int j) { return __dmul_rn(__ddiv_rn(2.0 * fqj(j), C_LIGHT), pv); };
                ps_asc[t] = p0 + t < m ? !(knp(0) > knp(n - 1)) : 1;
                ps_x0[t] = aa * ff;
                ps_iad[t] = 1.0 / (aa * dd);
            }
            __syncthreads();
        }
        // grid values of this thread's bins first: loads of a KR table complete
        // before the strip loads start, so the bracket search overlaps them
        double Qv[8];
#pragma unroll
        for (int g = 0; g < 2; ++g)
#pragma unroll
            for (int u = 0; u < 4; ++u) Qv[g * 4 + u] = KR(min(q0 + 4 * (gi + g * (RT / RP)) + u, n - 1));
        // per-pulse position model kr(j) = (2 pr / c) f(j), f(j) = f0 + j df
        const int i = ps_i[pl];
        const double ai = ps_a[pl], f0 = ps_f0[pl], df = ps_df[pl];
        const double *frow = freq + (size_t)i * n;
        auto fq = [&](int j) { return FULLFREQ ? frow[j] : __dadd_rn(f0, __dmul_rn((double)j, df)); };
        // kr as (2 pr / c) f: one product per position, within a few ulp of numpy's
        // (2 f / c) pr; kr_np is numpy's rounding exactly, used to decide near-ties
        // (the KR grid's end points are such ties by construction)
        auto kr = [&](int j) { return __dmul_rn(ai, fq(j)); };
        const double prv = ps_pr[pl];
        auto kr_np = [&](int j) { return __dmul_rn(__ddiv_rn(2.0 * fq(j), C_LIGHT), prv); };
        const bool asc = ps_asc[pl] != 0;
        auto xa = [&](int u) { return kr(asc ? u : n - 1 - u); };
        auto xe = [&](int u) { return kr_np(asc ? u : n - 1 - u); };
        const double x0 = ps_x0[pl], inv_adf = ps_iad[pl];
        auto guess = [&](double Q) {                  // ascending-view count estimate
            const double js = (Q - x0) * inv_adf;
            const double g = asc ? js : (double)(n - 1) - js;
            return (g == g) ? (int)fmin(fmax(floor(g) + 1.0, 0.0), (double)n) : 0;
        };
        auto bracket = [&](double Q, double &xl, double &xh) {
            int c = bracket_from(xa, n, Q, guess(Q), xl, xh);
            const double tol = fabs(Q) * 4e-15;
            if (fabs(xl - Q) <= tol || fabs(xh - Q) <= tol) c = bracket_from(xe, n, Q, c, xl, xh);
            return c;
        };
        const float2 *row = sig + (size_t)i * n;
        auto fa_g = [&](int u) { return row[asc ? u : n - 1 - u]; };
        __syncthreads();                              // previous tile fully consumed
        if (t < RP) {                                 // pulse t's strip: first window on, RS samples
            int lo = 0, hi = 0;
            if (p < m) {
                double xl, xh;
                const double Q = KR(q0);
                const int c = bracket(Q, xl, xh);
                lo = max(0, min(c - 1 - 15 - 2, n - RS));
                hi = min(lo + RS, n);
            }
            lo_s[t] = lo;
            hi_s[t] = hi;
            st_s[t] = asc ? lo : n - 1 - lo;          // sample index of strip slot 0
            dir_s[t] = asc ? 1 : -1;
        }
        __syncthreads();
        // strip loads in flight while the brackets are found: 32 threads per
        // pulse, thread samples u = t % 32 + 32 k (consecutive lanes, consecutive
        // samples), issued unconditionally from clamped slots; slots past the
        // strip are zeroed when stored (overwriting a register whose load is
        // in flight would wait for it)
        float2 pre[RLD];
        {
            const int sp_ = t / 32, su = t % 32;
            const int slo = lo_s[sp_], shi = hi_s[sp_], sst = st_s[sp_], sdir = dir_s[sp_];
            const float2 *srow = sig + (size_t)ps_i[sp_] * n + sst;
            const int ulast = max(shi - slo - 1, 0);
#pragma unroll
            for (int k = 0; k < RLD; ++k) pre[k] = ld_stream(srow + sdir * min(su + 32 * k, ulast));
        }
        // brackets of this thread's bins: (c << 8) | phase and the fraction.
        // The first of four consecutive bins starts from the closed-form guess;
        // the next three walk up from their predecessor's bracket (KR is
        // non-decreasing), then take the same near-tie re-decision.
        int code[8];
        float frq[8];
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            int c = 0;
            double xl = -INFINITY, xh = INFINITY;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = g * 4 + u;
                const int q = q0 + 4 * (gi + g * (RT / RP)) + u;
                code[k] = 0;
                frq[k] = 0.f;
                if (p >= m || q >= qe) continue;
                const double Q = Qv[k];
                if (u == 0) {
                    c = bracket(Q, xl, xh);
                } else {
                    int steps = 0;
                    while (c < n && xh <= Q && steps < 8) { ++c; ++steps; xl = xh; xh = c < n ? xa(c) : INFINITY; }
                    if (c < n && xh <= Q) {
                        c = bracket(Q, xl, xh);
                    } else {
                        const double tol = fabs(Q) * 4e-15;
                        if (fabs(xl - Q) <= tol || fabs(xh - Q) <= tol) c = bracket_from(xe, n, Q, c, xl, xh);
                    }
                }
                int ph = 0;
                if (c == n) {
                    ph = (Q == xl) ? 1 : 0;           // last sample, or outside
                } else if (c > 0) {
                    const double fr = div_bracket(__dsub_rn(Q, xl), __dsub_rn(xh, xl));
                    ph = (int)fmin(rint(fr * 32768.0), 32767.0) >> 7;
                    frq[k] = (float)fr;
                }
                code[k] = (c << 8) | ph;
            }
        }
        {
            const int sp_ = t / 32, slo = lo_s[sp_], shi = hi_s[sp_];
#pragma unroll
            for (int k = 0; k < RLD; ++k) {
                const int u = t % 32 + 32 * k;
                if (u < RS) strip[sp_ * RSP + u] = slo + u < shi ? pre[k] : make_float2(0.f, 0.f);
            }
        }
        __syncthreads();
        // gather
        float2 v[8];
        const int lo = lo_s[pl], hi = hi_s[pl];
        const float2 *sp = strip + pl * RSP;
        auto fa = [&](int u) { return (u >= lo && u < hi) ? sp[u - lo] : fa_g(u); };
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            int d[4], ph[4];
            const int J0 = (code[g * 4] >> 8) - 1;
            // every J in [max(15, J0), min(n - 17, J0 + KD)]: tested on the group's
            // smallest and largest J; the last bin (the largest q) inside the tile
            int jmin = J0, jmax = J0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int J = (code[g * 4 + u] >> 8) - 1;
                d[u] = J - J0;
                ph[u] = code[g * 4 + u] & 255;
                jmin = min(jmin, J);
                jmax = max(jmax, J);
            }
            const bool fast = SINC && p < m && q0 + 4 * (gi + g * (RT / RP)) + 3 < qe && jmin >= max(15, J0) &&
                              jmax <= min(n - 17, J0 + KD) && J0 - 15 >= lo && J0 - 15 + KW <= hi;
            if (fast) {
                float2 w4[4];
                sinc4p(sp, J0 - 15 - lo, tp, ph, d, w4);
#pragma unroll
                for (int u = 0; u < 4; ++u) v[g * 4 + u] = w4[u];
            } else {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = g * 4 + u;
                    const int q = q0 + 4 * (gi + g * (RT / RP)) + u;
                    const int c = code[k] >> 8;
                    float2 r = make_float2(0.f, 0.f);
                    if (p < m && q < qe && c > 0) {
                        if (c == n) {
                            if (code[k] & 255) r = fa(n - 1);
                        } else if (SINC && c - 1 >= 15 && c - 1 <= n - 17) {
                            ++slow;
                            r = sinc1p(fa, c - 1, code[k] & 255, tp);
                        } else {
                            const float2 a0 = fa(c - 1), a1 = fa(c);
                            const float fr = frq[k];
                            r = make_float2(fmaf(a1.x - a0.x, fr, a0.x), fmaf(a1.y - a0.y, fr, a0.y));
                        }
                    }
                    v[k] = r;
                }
            }
        }
        // lanes 16 i .. 16 i + 15 hold 16 consecutive pulses of one range bin:
        // each store writes 128 contiguous bytes of a g1t row
#pragma unroll
        for (int g = 0; g < 2; ++g)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int q = q0 + 4 * (gi + g * (RT / RP)) + u;
                if (q < qe && p < m) st_stream(g1t + (size_t)q * ldp + p, v[g * 4 + u]);
            }
    }
    count_wave(nslow, slow);
}

// ------------------------------------------------------------------ pass 2 (+ window + FFT A)
// A, the azimuth pass's output [range row j][cross bin r], is stored in blocks
// of 8 cross bins: element (j, r) at ((r / 8) nj + j) 8 + r % 8, with nj = the
// row count rounded up to 16.  A line's FFT writes whole 64-byte chunks
// (8 consecutive r), and the column pairs FFT-B reads lie 64 bytes apart for
// consecutive rows, so a wave's 16-byte loads touch 64-byte chunks of 32
// cache lines instead of one line per row.  Measured at 8192^2 (identical
// images): 16-bin blocks 0.302 ms FFT-B, 8-bin 0.239 ms with the azimuth pass
// unchanged, 4-bin 0.209 ms FFT-B but 1.16 ms azimuth (its 32-byte scattered
// writes), row-major 0.316 ms.
constexpr int A_BLK = 8;
__device__ __forceinline__ int a_rows(int n) { return (n + 15) & ~15; }
__device__ __forceinline__ size_t ablk(int j, int r, int nj) {
    return ((size_t)(r / A_BLK) * nj + j) * A_BLK + (r % A_BLK);
}
// The matrix-core forms' A16 (one 32-bit word per complex sample) in blocks of
// 16 cross bins: the same 64-byte chunks.
__device__ __forceinline__ size_t a16blk(int j, int r, int nj) {
    return ((size_t)(r >> 4) * nj + j) * 16 + (r & 15);
}

// Workgroup b -> consecutive items on the same XCD (b % 8 selects the XCD).
__device__ __forceinline__ int xcd_map(int b, int count) {
    const int per = (count + 7) / 8;
    return (b % 8) * per + b / 8;
}

// Elements per thread of the azimuth kernel's FFT and resample: 8, so an
// 8192-point line runs on 1024 threads (16 waves keep the CU busy while one
// workgroup holds the LDS).
// 16384 points: 16 a thread on 1024 threads; 4096 points: 4 a thread, so they
// too run on 1024 threads (the LDS holds one workgroup a CU either way)
template <int L>
constexpr int az_emax() { return L >= 14 ? 16 : (L == 12 ? 4 : 8); }
// The line FFT runs radix-16 passes (half the threads active) -- fewer LDS round
// trips than radix 8 on all 1024.
constexpr int AZ_FFT_EMAX = 16;
template <int L>
constexpr int az_nb() { return pfa32::Cfg<L, az_emax<L>()>::NB; }

// Bracket-guess table over tan(phi): idx[b] = #{p : tan_s[p] < lo + b / inv_w}.
struct TanBins {
    const int32_t *idx;
    double lo, inv_w;
    int bins;
    __device__ __forceinline__ int bin(double t) const {
        const double b = floor(__dmul_rn(__dsub_rn(t, lo), inv_w));
        return (b == b) ? (int)fmin(fmax(b, 0.0), (double)bins) : 0;
    }
    __device__ __forceinline__ int guess(double t) const { return idx[bin(t)]; }
};

// Persistent over range bins: iteration it of workgroup b (XCD x = b % 8, rank
// b / 8) takes j = x * ceil(n/8) + it * (grid/8) + rank.  One LDS region holds,
// in turn, the line's FP64 source positions KR_j tan_s[p] (bracket search), the
// sorted-pulse row (the gather's source, prefetched into registers during the
// previous line) and the windowed, sign-folded line the FFT runs on.
// Thread t owns outputs 4 (g NB + t) + u, u < 4: the four share one sinc window.
// MODE 0: store the windowed k-space g2[q][j] (resample_kspace stage output)
// MODE 1: fused -- window, sign fold, pad, FFT along cross -> A (8-bin blocks, ablk)
// MODE 2: FFT along cross of a given g2 (focus() stage; windowed when the plan has tapers:
//         quicklook()) -> A (8-bin blocks, ablk)
// FFT form of the line transform: 0 FP32 vector FFT (A complex64); 1 BF16,
// 2 FP16, 3 BF16 with hi/lo DFT matrices -- matrix-core FFT (pfa_fft_mfma.hpp),
// A stored as packed 16-bit pairs with a per-row exponent.
constexpr int FK_F32 = 0, FK_BF16 = 1, FK_FP16 = 2, FK_BF16S = 3;
template <int L, int MODE, bool SINC, bool INV, int FK = FK_F32, bool ROT = false>
__global__ __launch_bounds__(az_nb<L>()) void k_azimuth_f32(
    const float2 *__restrict__ g1t, int ldp, int m, int n, const Grid1 KR,
    const double *__restrict__ tan_s, const TanBins tb, const Grid1 KC,
    const double *__restrict__ hm, const double *__restrict__ hn,
    const int16_t *__restrict__ sinc, const float2 *__restrict__ g2in,
    float2 *__restrict__ out, const float2 *__restrict__ twg, unsigned *__restrict__ nslow,
    unsigned *__restrict__ A16, int *__restrict__ expA, const float2 *__restrict__ rot,
    const pfam::u16x8 *__restrict__ ops, const unsigned *__restrict__ tapg) {
    static_assert(FK == FK_F32 || MODE == 1, "matrix-core FFT only in the fused mode");
    // 16384-point lines (pulse counts above 8192): the positions, row and line
    // take 140 KB of LDS, so the packed taps are read from global memory
    // (tapg, the same layout) and the window is formed as it is applied
    constexpr bool BIG = L >= 14;
    static_assert(!BIG || FK == FK_F32, "16384-point lines in the FP32 form only");
    constexpr bool BF = FK == FK_BF16 || FK == FK_BF16S;
    constexpr bool SPLIT = FK == FK_BF16S;
    constexpr int EM = az_emax<L>();
    using Cf = pfa32::Cfg<L, EM>;
    constexpr int M2 = Cf::N;
    constexpr int NB = Cf::NB;
    constexpr int G = (M2 + 4 * NB - 1) / (4 * NB);   // 4-output groups per thread
    constexpr int PR = (M2 + NB - 1) / NB;             // row samples per thread
    constexpr bool RES = MODE != 2;                    // resampling modes
    // the fused FP32 form runs the line FFT's first pass (radix G) in registers
    constexpr bool R1 = MODE == 1 && FK == FK_F32 && G >= 2 && 4 * G * NB == M2;
    __shared__ float2 reg[(M2 + M2 / 16 + 64 > Cf::PADN + 64) ? M2 + M2 / 16 + 64 : Cf::PADN + 64];  // positions | row | FFT line
    __shared__ unsigned tp[(SINC && RES && !BIG) ? TPWORDS : 1];
    __shared__ float2 twl[M2 / 64 + 64];
    __shared__ float wmax[NB / 64];
    __shared__ float wl[(RES && !BIG) ? G * 4 * NB : 1];   // this line's window weights, thread-private slots
    double *xs = reinterpret_cast<double *>(reg);
    const int t = threadIdx.x;
    if (SINC && RES && !BIG) load_taps2(tp, sinc);
    const pfa32::Tw tw = pfa32::load_tw<M2>(twl, twg);
    const int per_x = (n + 7) / 8;
    const int ranks = (gridDim.x + 7) / 8;
    const int x = blockIdx.x % 8, rank = blockIdx.x / 8;
    auto line_of = [&](int base) { return x * per_x + base + rank; };
    auto active_at = [&](int base) { return base + rank < per_x && line_of(base) < n; };
    unsigned slow = 0;
    // next row, in flight: loaded unconditionally from clamped positions and
    // masked when stored (overwriting a register whose load is in flight would
    // wait for it); rot (ROT): per sorted pulse e^{j phi} of a phase correction
    // (autofocus), applied after the row is stored.  Loads on paths taken at
    // run time only (a null-pointer test) would leave registers the compiler
    // must wait for after every later join, the row prefetch included.
    float2 pre[PR];
    auto prefetch = [&](int j) {
        const float2 *row = g1t + (size_t)j * ldp;
        int tp_ = t;                                   // opaque: addresses are not held across lines
        asm volatile("" : "+v"(tp_));
#pragma unroll
        for (int k = 0; k < PR; ++k) pre[k] = ld_stream(row + min(tp_ + k * NB, m - 1));
    };
    if (RES && active_at(0)) prefetch(line_of(0));
    // the tan values of this thread's source positions are the same for every
    // line: held in registers where they fit (lines up to 8192 points, every
    // FFT form), loaded per line otherwise
    constexpr bool HOLD = RES && !BIG;
    double tsv[HOLD ? PR : 1];
    if constexpr (HOLD) {
#pragma unroll
        for (int k = 0; k < PR; ++k) tsv[k] = tan_s[min(t + k * NB, m - 1)];
    }
    // tan-bin guesses of each group's first output, looked up during the
    // previous line so the bracket phase starts without a global round trip
    // (16384-point lines, short of registers, look them up at the line's start)
    int gix[G];
    auto guesses = [&](int jj, int (&gx)[G]) {
        const double ik = 1.0 / KR(jj);
        int to = t;                                    // opaque: the KC values are not held across lines
        asm volatile("" : "+v"(to));
#pragma unroll
        for (int g = 0; g < G; ++g) gx[g] = tb.bin(__dmul_rn(KC(min((g * NB + to) * 4, m - 1)), ik));
#pragma unroll
        for (int g = 0; g < G; ++g) gx[g] = tb.idx[gx[g]];
    };
    if (RES && !BIG && active_at(0)) guesses(line_of(0), gix);
    for (int base = 0; base < per_x; base += ranks) {
        if (!active_at(base)) continue;               // uniform per workgroup
        const int j = line_of(base);
        // per-output values (KC(q), tapers) are recomputed per line: hoisted
        // out of the loop they would pin ~40 registers for its whole length
        int tl = t;
        asm volatile("" : "+v"(tl));
        float2 v[G * 4];
        if (MODE == 2) {
            // a given k-space (focus()), or with the plan's tapers the raw
            // history of quicklook(): f32(hm[q] hn[j]) as hamming2d forms it
            const double hj = hn ? hn[j] : 1.0;
            __syncthreads();                          // previous line fully consumed
            for (int q = t; q < M2; q += NB) {
                float2 w = q < m ? g2in[(size_t)q * n + j] : make_float2(0.f, 0.f);
                if (hm && q < m) {
                    const float wgt = (float)__dmul_rn(hm[q], hj);
                    w = make_float2(w.x * wgt, w.y * wgt);
                }
                if ((q + j) & 1) w = make_float2(-w.x, -w.y);
                reg[pfa32::padf(q)] = w;
            }
        } else {
            const double krj = KR(j);
            const double hj = hn ? hn[j] : 1.0;
            double hmv[BIG ? 1 : G * 4];              // the window's cross tapers, in flight
            if constexpr (!BIG) {                     // while the positions are formed
#pragma unroll
                for (int k = 0; k < G * 4; ++k) hmv[k] = hm ? hm[min(((k / 4) * NB + tl) * 4 + k % 4, m - 1)] : 1.0;
            }
            // (1) FP64 source positions of this line
            double tv[HOLD ? 1 : PR];
            if constexpr (!HOLD) {
#pragma unroll
                for (int k = 0; k < PR; ++k) tv[k] = tan_s[min(tl + k * NB, m - 1)];
            }
            __syncthreads();                          // previous line fully consumed
#pragma unroll
            for (int k = 0; k < PR; ++k)
                if (tl + k * NB < m) xs[tl + k * NB] = __dmul_rn(krj, HOLD ? tsv[k] : tv[k]);
            __syncthreads();
            int gixl[BIG ? G : 1];
            if constexpr (BIG) guesses(j, gixl);
            // (2) brackets: c = #{p : xs[p] <= Q} (np.interp / searchsorted-right),
            // kept as (c << 8) | sinc phase, and the interval fraction.  The
            // first output of each group of four starts from the tan-bin guess;
            // the next three walk up from their predecessor's bracket (KC is
            // non-decreasing).
            int code[G * 4];
            float frq[G * 4];
#pragma unroll
            for (int g = 0; g < G; ++g) {
                int c = 0;
                double xl = -INFINITY, xh = INFINITY;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int q = (g * NB + tl) * 4 + u;
                    const int k = g * 4 + u;
                    code[k] = 0;
                    frq[k] = 0.f;
                    if constexpr (!BIG) wl[k * NB + t] = (float)__dmul_rn(hmv[k], hj);
                    if (q >= m) continue;
                    const double Q = KC(q);
                    if (u == 0) {
                        if constexpr (BIG) c = gixl[g];
                        else c = gix[g];
                        xl = c > 0 ? xs[c - 1] : -INFINITY;
                        xh = c < m ? xs[c] : INFINITY;
                        if (xh <= Q) {                // the guess is low by one for ~1/4 of outputs
                            ++c;
                            xl = xh;
                            xh = c < m ? xs[c] : INFINITY;
                        }
                        if (!(xl <= Q) || xh <= Q) {
                            ++slow;
                            while (c > 0 && !(xl <= Q)) { --c; xh = xl; xl = c > 0 ? xs[c - 1] : -INFINITY; }
                            while (c < m && xh <= Q) { ++c; xl = xh; xh = c < m ? xs[c] : INFINITY; }
                        }
                    } else {
                        while (c < m && xh <= Q) { ++c; xl = xh; xh = c < m ? xs[c] : INFINITY; }
                    }
                    int ph = 0;
                    if (c == m) {
                        ph = (Q == xl) ? 1 : 0;                              // last sample, or outside
                    } else if (c > 0) {
                        const double fr = div_bracket(__dsub_rn(Q, xl), __dsub_rn(xh, xl));
                        ph = (int)fmin(rint(fr * 32768.0), 32767.0) >> 7;
                        frq[k] = (float)fr;
                    }
                    code[k] = (c << 8) | ph;
                }
            }
            __syncthreads();                          // positions dead: the row takes over
            // (3) the row (pad16 layout), zero past m so shifted windows read zeros
#pragma unroll
            for (int k = 0; k < PR; ++k) {
                const int p = t + k * NB;
                if (p < M2) reg[pad16(p)] = p < m ? pre[k] : make_float2(0.f, 0.f);
            }
            if constexpr (ROT) {                      // this thread's own slots: no barrier needed
                float2 rv[PR];
#pragma unroll
                for (int k = 0; k < PR; ++k) rv[k] = rot[min(tl + k * NB, m - 1)];
#pragma unroll
                for (int k = 0; k < PR; ++k) {
                    const int p = tl + k * NB;
                    if (p < m) reg[pad16(p)] = pfa32::cmul(reg[pad16(p)], rv[k]);
                }
            }
            if (t < 48) reg[pad16(m + t)] = make_float2(0.f, 0.f);
            __syncthreads();
            // (4) gather
            auto fa = [&](int u) { return reg[pad16(u)]; };
#pragma unroll
            for (int g = 0; g < G; ++g) {
                int d[4], ph[4];
                bool fast = SINC;
                const int J0 = (code[g * 4] >> 8) - 1;
                // every J in [max(15, J0), min(m - 17, J0 + KD)]: the FP32 form tests
                // the group's smallest and largest J; the matrix-core forms, at the
                // register limit, test each J (no two extra live registers)
                int jmin = J0, jmax = J0;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int J = (code[g * 4 + u] >> 8) - 1;
                    d[u] = J - J0;
                    ph[u] = code[g * 4 + u] & 255;
                    if constexpr (FK == FK_F32) {
                        jmin = min(jmin, J);
                        jmax = max(jmax, J);
                    } else {
                        fast = fast && J >= 15 && J <= m - 17 && d[u] >= 0 && d[u] <= KD;
                    }
                }
                if constexpr (FK == FK_F32) fast = fast && jmin >= max(15, J0) && jmax <= min(m - 17, J0 + KD);
                if (fast) {
                    float2 w4[4];
                    if constexpr (BIG) sinc4p<6, true>(reg, J0 - 15, tapg, ph, d, w4);
                    else sinc4p<6, true>(reg, J0 - 15, tp, ph, d, w4);
#pragma unroll
                    for (int u = 0; u < 4; ++u) v[g * 4 + u] = w4[u];
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int k = g * 4 + u;
                        const int q = (g * NB + tl) * 4 + u;
                        const int c = code[k] >> 8;
                        float2 r = make_float2(0.f, 0.f);
                        if (q < m && c > 0) {
                            if (c == m) {
                                if (ph[u]) r = fa(m - 1);
                            } else if (SINC && c - 1 >= 15 && c - 1 <= m - 17) {
                                if constexpr (BIG) r = sinc1p(fa, c - 1, ph[u], tapg);
                                else r = sinc1p(fa, c - 1, ph[u], tp);
                            } else {
                                const float2 f0 = fa(c - 1), f1 = fa(c);
                                const float fr = frq[k];
                                r = make_float2(fmaf(f1.x - f0.x, fr, f0.x), fmaf(f1.y - f0.y, fr, f0.y));
                            }
                        }
                        v[k] = r;
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
            {                                         // the next row, in flight during the FFT
                const int nb2 = base + ranks;
                if (nb2 < per_x && active_at(nb2)) {
                    prefetch(line_of(nb2));
                    if constexpr (!BIG) guesses(line_of(nb2), gix);
                }
            }
            // window
            if (hm) {
#pragma unroll
                for (int k = 0; k < G * 4; ++k) {
                    float w;
                    if constexpr (BIG) {
                        const int q = ((k / 4) * NB + tl) * 4 + k % 4;
                        w = q < m ? (float)__dmul_rn(hm[q], hj) : 0.f;
                    } else {
                        w = wl[k * NB + t];
                    }
                    v[k] = make_float2(v[k].x * w, v[k].y * w);
                }
            }
            if (MODE == 0) {
#pragma unroll
                for (int g = 0; g < G; ++g) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int q = (g * NB + tl) * 4 + u;
                        if (q < m) out[(size_t)q * n + j] = v[g * 4 + u];
                    }
                }
                continue;
            }
            if constexpr (FK != FK_F32) {
                // block scaling: the line's largest component sets a power-of-two exponent
                float mx = 0.f;
#pragma unroll
                for (int k = 0; k < G * 4; ++k) mx = fmaxf(mx, fmaxf(fabsf(v[k].x), fabsf(v[k].y)));
                for (int off = 32; off >= 1; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
                if ((t & 63) == 0) wmax[t >> 6] = mx;
                __syncthreads();                      // the row is dead: the FFT planes take over
                float lm = 0.f;
#pragma unroll
                for (int i = 0; i < NB / 64; ++i) lm = fmaxf(lm, wmax[i]);
                const int se = pfam::norm_exp(lm);
                const float sc = ldexpf(1.0f, -se);
                unsigned *planes = reinterpret_cast<unsigned *>(reg);
                static_assert(4 * G * NB == M2, "the threads' groups cover the line");
                if constexpr (pfam::MPlan<L>::LS0 == 2) {
                    // the radix-2 first pass in registers: groups g and g + G/2
                    // hold points k and k + N/2 (k = 4 (g NB + tl) + u)
#pragma unroll
                    for (int g = 0; g < G / 2; ++g) {
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const float s = ((u + j) & 1) ? -sc : sc;        // (-1)^(q + j) 2^-se
                            const float2 a = make_float2(v[g * 4 + u].x * s, v[g * 4 + u].y * s);
                            const float2 b = make_float2(v[(g + G / 2) * 4 + u].x * s, v[(g + G / 2) * 4 + u].y * s);
                            pfam::put_r2<L, BF, INV>(planes, 4 * (g * NB + tl) + u, a, b);
                        }
                    }
                } else {
                    unsigned *pb = pfam::put_base<L>(planes, 0, 4 * tl);
#pragma unroll
                    for (int g = 0; g < G; ++g) {
#pragma unroll
                        for (int u = 0; u < 4; ++u) {          // q = (g NB + tl) 4 + u
                            float2 w = make_float2(v[g * 4 + u].x * sc, v[g * 4 + u].y * sc);
                            if ((u + j) & 1) w = make_float2(-w.x, -w.y);   // (-1)^(q + j)
                            pfam::put_at<L, BF>(pb, 4 * g * NB + u, w);
                        }
                    }
                }
                __syncthreads();
                constexpr int NW = pfam::nw_for<L, 1, NB / 64>();
                const int nj = a_rows(n);
                pfam::fft<L, BF, INV, SPLIT, 1, NW>(planes, tw, pfam::OpsGlobal<L, BF, INV>{ops},
                    [&](int, int o, float2 y) {
                        __builtin_nontemporal_store(pfam::pack<BF>(y), A16 + a16blk(j, o, nj));
                    });
                if (t == 0) expA[j] = se + (INV ? 0 : L);
                continue;
            }
            __syncthreads();                          // the row is dead: the FFT line takes over
            if constexpr (R1) {
                // the FFT's first pass in registers: the thread's groups hold
                // points k + g N/G (k = 4 tl + u), the inputs of one radix-G
                // butterfly (LS = 1, no twiddles); outputs G k + r go to the line
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = 4 * tl + u;
                    float2 x[G];
#pragma unroll
                    for (int g = 0; g < G; ++g) {     // (-1)^(q + j): N/G is even
                        x[g] = v[g * 4 + u];
                        if ((k + j) & 1) x[g] = make_float2(-x[g].x, -x[g].y);
                    }
                    pfa32::dft<G, INV>(x);
#pragma unroll
                    for (int r = 0; r < G; ++r) reg[pfa32::padf(G * k + r)] = x[r];
                }
            } else {
#pragma unroll
                for (int g = 0; g < G; ++g) {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int q = (g * NB + tl) * 4 + u;
                        if (q < M2) {
                            float2 w = v[g * 4 + u];
                            if ((q + j) & 1) w = make_float2(-w.x, -w.y);
                            reg[pfa32::padf(q)] = w;
                        }
                    }
                }
            }
        }
        if constexpr (MODE != 0) {
            __syncthreads();
            const int nj = a_rows(n);
            const float scale = INV ? 1.0f / (float)M2 : 1.0f;
            auto sink = [&](int idx, float2 w) {
                if (INV) w = make_float2(w.x * scale, w.y * scale);
                st_stream(out + ablk(j, idx, nj), w);
            };
            // the last pass runs without barriers: the next line synchronises
            // before it writes the LDS region
            if constexpr (R1) pfa32::fft_rest<L, AZ_FFT_EMAX, INV, G>(reg, tw, sink);
            else pfa32::fft<L, INV, AZ_FFT_EMAX>(reg, tw, sink);
        }
    }
    count_wave(nslow ? nslow + 1 : nullptr, slow);
}

// ------------------------------------------------------------------ FFT B (range) + detect
// OUT 0: complex64 image, 1: |image| float32, image [cross r][range c]; the
// column loads of A are the corner-turn.  Consecutive workgroups on one XCD.
// Two adjacent cross bins r0, r0 + 1 per workgroup: every row of A is read as
// one 16-byte load (half the memory requests of one column per workgroup).
// Thread t holds elements t + s NT (s < E) of both columns -- exactly the
// inputs of its first Stockham pass -- transforms the first column through
// the LDS line, then the second from its registers through the same line.
// Radix-16 passes (16 elements a thread) keep a workgroup within 128 VGPRs,
// so two share a CU and one's loads overlap the other's passes.
constexpr int FB_EMAX = 16;
template <int L>
constexpr int fb2_nb() { return pfa32::Cfg<L, FB_EMAX>::NT < 64 ? 64 : pfa32::Cfg<L, FB_EMAX>::NT; }
template <int L, bool INV, int OUT>
__global__ __launch_bounds__(fb2_nb<L>(), 4) void k_fft_b2_f32(
    const float2 *__restrict__ A, int n, int m2, void *__restrict__ outp,
    const float2 *__restrict__ twg) {
    using Cf = pfa32::Cfg<L, FB_EMAX>;
    constexpr int N2 = Cf::N, NT = Cf::NT, E = Cf::E;
    __shared__ float2 buf[Cf::PADN];
    __shared__ float2 twl[N2 / 64 + 64];
    const pfa32::Tw tw = pfa32::load_tw<N2>(twl, twg);
    const int r0 = 2 * xcd_map(blockIdx.x, m2 / 2);
    const int t = threadIdx.x;
    const int nj = a_rows(n);
    const float scale = INV ? 1.0f / (float)N2 : 1.0f;
    float2 c0[E], c1[E];
#pragma unroll
    for (int s2 = 0; s2 < E; ++s2) {
        const int jj = t + s2 * NT;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (t < NT && jj < n) v = *reinterpret_cast<const float4 *>(A + ablk(jj, r0, nj));
        c0[s2] = make_float2(v.x, v.y);
        c1[s2] = make_float2(v.z, v.w);
    }
    // the first pass asks for element t + s NT as slot s (pass 0 has LS = 1)
    auto sink_for = [&](int r) {
        return [&, r](int c, float2 v) {
            if (INV) v = make_float2(v.x * scale, v.y * scale);
            if (OUT == 0) {
                if ((r + c) & 1) v = make_float2(-v.x, -v.y);
                reinterpret_cast<float2 *>(outp)[(size_t)r * N2 + c] = v;
            } else {                                   // |(-1)^(r+c) v| = |v|, bit for bit
                reinterpret_cast<float *>(outp)[(size_t)r * N2 + c] = __fsqrt_rn(v.x * v.x + v.y * v.y);
            }
        };
    };
    // the sinks write global memory: the last passes skip their read barrier,
    // and the second transform its closing one (the kernel ends)
    pfa32::fft_src<L, INV, FB_EMAX, false, true>(buf, tw, [&](int, int s2) { return c0[s2]; }, sink_for(r0));
    pfa32::fft_src<L, INV, FB_EMAX, false, false>(buf, tw, [&](int, int s2) { return c1[s2]; }, sink_for(r0 + 1));
}

// Two-level twiddle tables of an n-point line: hi[k] = W^(64 k), k < n/64,
// then lo[k] = W^k, k < 64, with W = exp(-2 pi i / n), rounded from double.
const float2 *twiddles_f32(pfa_ctx *ctx, int n, int *st) {
    char slot[32];
    snprintf(slot, sizeof slot, "twf2_%d", n);
    auto it = ctx->ws.find(slot);
    if (it != ctx->ws.end() && it->second.bytes) return (const float2 *)it->second.ptr;
    std::vector<float2> h((size_t)n / 64 + 64);
    for (int k = 0; k < n / 64; ++k) {
        const double th = 2.0 * M_PI * (double)(64 * k) / (double)n;
        h[k] = make_float2((float)std::cos(th), (float)-std::sin(th));
    }
    for (int k = 0; k < 64; ++k) {
        const double th = 2.0 * M_PI * (double)k / (double)n;
        h[n / 64 + k] = make_float2((float)std::cos(th), (float)-std::sin(th));
    }
    void *d = pfa_ws(ctx, slot, h.size() * sizeof(float2), st);
    if (!d) return nullptr;
    if (hipMemcpy(d, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) {
        *st = pfa_fail(ctx, PFA_ERR_HIP, "twiddle upload failed");
        return nullptr;
    }
    return (const float2 *)d;
}

Grid1 grid_of(const double *tab, const double *lin, int cnt) {
    if (lin) return Grid1{nullptr, lin[0], lin[1], lin[2], cnt};
    return Grid1{tab, 0.0, 0.0, 0.0, cnt};
}

int log2_exact(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return (1 << l) == n ? l : -1;
}

template <int L>
constexpr int nb_of() { return pfa32::Cfg<L>::NT < 64 ? 64 : pfa32::Cfg<L>::NT; }

}  // namespace
