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
// Taps of the Q15 sinc table as float (exact: c / 2^15), staged in LDS
// tap-major: tap t of phase ph at tapl[(t + 4) * 256 + scr(ph)], with four zero
// rows on either side (t in [-4, 36)) so a window shifted by up to KD samples
// reads zeros instead of needing masks, and scr(ph) = ph ^ (ph >> 3) spreads
// the phases of a wave's lanes over the LDS banks.
constexpr int TROWS = 40;
__device__ __forceinline__ int scr(int ph) { return ph ^ (ph >> 3); }
__device__ __forceinline__ void load_taps(float *tapl, const int16_t *__restrict__ sinc) {
    for (int i = threadIdx.x; i < TROWS * 256; i += blockDim.x) {
        const int row = i >> 8, ph = i & 255, t = row - 4;
        tapl[row * 256 + scr(ph)] = (t >= 0 && t < 32) ? (float)sinc[ph * 32 + t] * (1.0f / 32768.0f) : 0.f;
    }
}

// LDS index padding for sample strips: one spare float2 per 32, so per-thread
// windows that start 8, 16 or 32 samples apart do not share banks.
__device__ __forceinline__ int spad(int i) { return i + (i >> 5); }

constexpr int KD = 4;            // bracket advance covered by one 4-output window
constexpr int KW = 32 + KD;      // samples in that window

// Four consecutive outputs whose brackets J_u = J_0 + d_u (0 <= d_u <= KD) share
// one register window w[k] = fa(J_0 - 15 + k); output u uses taps shifted by d_u:
// sum_k tap[ph_u][k - d_u] * w[k].  Each window sample is read once per four
// outputs instead of once per output.
// PADDED: samples in an spad-padded LDS strip; otherwise a plain array (global).
template <bool PADDED = true>
__device__ __forceinline__ void sinc4(const float2 *strip, int start, const float *tapl,
                                      const int (&ph)[4], const int (&d)[4], float2 (&out)[4]) {
    const float *col[4];
    float ar[4], ai[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        col[u] = tapl + (4 - d[u]) * 256 + scr(ph[u]);   // col[u][k * 256] = tap k - d[u]
        ar[u] = 0.f;
        ai[u] = 0.f;
    }
    const float2 *ws = strip + (PADDED ? spad(start) : start);
    const int brk = 32 - (start & 31);             // window indices where the strip padding steps
#pragma unroll
    for (int k = 0; k < KW; ++k) {                 // one window sample feeds four outputs
        const float2 w = PADDED ? ws[k + (k >= brk ? 1 : 0) + (k >= brk + 32 ? 1 : 0)] : ws[k];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float c = col[u][k * 256];
            ar[u] = fmaf(c, w.x, ar[u]);
            ai[u] = fmaf(c, w.y, ai[u]);
        }
        if (k % 12 == 11) __builtin_amdgcn_sched_barrier(0);     // bound live loads
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) out[u] = make_float2(ar[u], ai[u]);
}

// np.interp / SINC32 value of one output from its bracket state.
// cls: 0 zero, 1 value fa(J) (exact hit or last sample), 2 linear, 3 sinc interior
struct Bracket {
    int cls, J, ph;
    double x0, x1;
};

__device__ __forceinline__ Bracket classify(bool sinc, int c, int S, double Q, double xprev,
                                            double xcur, double xlast) {
    Bracket b{0, 0, 0, 0.0, 0.0};
    if (c <= 0) return b;
    if (c >= S) {
        if (Q == xlast) { b.cls = 1; b.J = S - 1; }
        return b;
    }
    b.J = c - 1;
    b.x0 = xprev;
    b.x1 = xcur;
    if (sinc && b.J >= 15 && b.J <= S - 17) {
        const double frac = __ddiv_rn(__dsub_rn(Q, xprev), __dsub_rn(xcur, xprev));
        const int wq = (int)fmin(fmax(rint(frac * 32768.0), 0.0), 32767.0);
        b.cls = 3;
        b.ph = wq >> 7;
        return b;
    }
    b.cls = (xprev == Q) ? 1 : 2;
    return b;
}

template <typename FA>
__device__ __forceinline__ float2 value_slow(const Bracket &b, double Q, FA &&fa, const float *tapl) {
    if (b.cls == 0) return make_float2(0.f, 0.f);
    if (b.cls == 1) return fa(b.J);
    if (b.cls == 3) {
        const float *col = tapl + 4 * 256 + scr(b.ph);
        float ar = 0.f, ai = 0.f;
#pragma unroll 8
        for (int t = 0; t < 32; ++t) {
            const float2 s = fa(b.J - 15 + t);
            ar = fmaf(col[t * 256], s.x, ar);
            ai = fmaf(col[t * 256], s.y, ai);
        }
        return make_float2(ar, ai);
    }
    const float2 f0 = fa(b.J), f1 = fa(b.J + 1);
    const double dx = __dsub_rn(b.x1, b.x0), dq = __dsub_rn(Q, b.x0);
    const double sr = __ddiv_rn(__dsub_rn((double)f1.x, (double)f0.x), dx);
    const double si = __ddiv_rn(__dsub_rn((double)f1.y, (double)f0.y), dx);
    return make_float2((float)__dadd_rn(__dmul_rn(sr, dq), (double)f0.x),
                       (float)__dadd_rn(__dmul_rn(si, dq), (double)f0.y));
}

// Value of one output from its scattered bracket e = (J << 8) | ph.
template <bool SINC, typename FA>
__device__ __forceinline__ float2 value_scattered(int e, int S, double Q, double x0, double x1,
                                                  FA &&fa, const float *tapl) {
    Bracket b{0, e >> 8, e & 255, x0, x1};
    b.cls = (SINC && b.J >= 15 && b.J <= S - 17) ? 3 : (x0 == Q ? 1 : 2);
    return value_slow(b, Q, fa, tapl);
}

// Bracket walk over an ascending source view xa(0..S-1) for non-decreasing
// queries: c = #{xa <= Q}, with xa(c-1) and xa(c) cached so each output costs
// about one position evaluation.
template <typename XA>
struct Walk {
    XA &xa;
    int S;
    int c = -1;
    double xprev = -INFINITY, xcur = INFINITY;

    __device__ Walk(XA &f, int s) : xa(f), S(s) {}

    __device__ __forceinline__ void set(int c0) {
        c = c0;
        xprev = c > 0 ? xa(c - 1) : -INFINITY;
        xcur = c < S ? xa(c) : INFINITY;
    }
    __device__ __forceinline__ void seek(double Q, int guess) {
        if (c < 0) {
            int lo = 0, hi = S;
            if (guess >= 0) {                 // verify a closed-form guess, else bisect
                const int g = min(max(guess, 0), S);
                const bool left_ok = (g == 0) || xa(g - 1) <= Q;
                const bool right_ok = (g == S) || !(xa(g) <= Q);
                if (left_ok && right_ok) { set(g); return; }
                if (left_ok) lo = g; else hi = g;
            }
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (xa(mid) <= Q) lo = mid + 1; else hi = mid;
            }
            set(lo);
            return;
        }
        while (c < S && xcur <= Q) {
            ++c;
            xprev = xcur;
            xcur = c < S ? xa(c) : INFINITY;
        }
    }
    // Start from the bracket c0 of a neighbouring line (positions moved a
    // little): walk both ways from it, bisect if it is far off.
    __device__ __forceinline__ void start_near(double Q, int c0) {
        c = -1;
        if (c0 < 0) { seek(Q, -1); return; }
        int c1 = min(max(c0, 0), S), steps = 0;
        while (c1 > 0 && !(xa(c1 - 1) <= Q) && steps < 48) { --c1; ++steps; }
        while (c1 < S && xa(c1) <= Q && steps < 48) { ++c1; ++steps; }
        if (steps >= 48) { seek(Q, -1); return; }
        set(c1);
    }
};

// First query q in [qa, qb] with Qg(q) >= x (qb if none) on a non-decreasing
// query grid: an inverse-linspace estimate fixed up by exact comparisons, or a
// bisection when the grid is a table.
__device__ __forceinline__ int first_q_ge(const Grid1 &Qg, double inv_step, double x, int qa, int qb) {
    int q;
    if (!Qg.tab) {
        const double e = ceil(__dmul_rn(__dsub_rn(x, Qg.start), inv_step));
        q = (e == e) ? (int)fmin(fmax(e, (double)qa), (double)qb) : qa;
        int steps = 0;
        while (q > qa && Qg(q - 1) >= x && steps < 16) { --q; ++steps; }
        while (q < qb && Qg(q) < x && steps < 32) { ++q; ++steps; }
        if (steps < 16) return q;
    }
    int lo = qa, hi = qb;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (Qg(mid) < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Scattered bracket of one output: J (ascending-view source index) and the
// sinc phase packed as (J << 8) | ph; BR_NONE = not reached by the scatter,
// BR_ZERO = outside the sources (np.interp's left/right = 0), BR_LAST = the
// query equals the last source position (value of the last sample).
constexpr int BR_NONE = -1, BR_ZERO = -2, BR_LAST = -3;

// Queries of [qa, qb) below the first source x_first or above the last x_last.
__device__ __forceinline__ void scatter_outside(const Grid1 &Qg, double inv_step, bool first,
                                                bool last, double x_first, double x_last, int qa,
                                                int qb, int *brk, int qbase) {
    if (first) {
        const int qs = first_q_ge(Qg, inv_step, x_first, qa, qb);
        for (int q = qa; q < qs; ++q) brk[q - qbase] = BR_ZERO;
    }
    if (last) {
        for (int q = first_q_ge(Qg, inv_step, x_last, qa, qb); q < qb; ++q)
            brk[q - qbase] = (Qg(q) == x_last) ? BR_LAST : BR_ZERO;
    }
}

// Sum a per-thread count over the wave and add it once (fallback statistics).
__device__ __forceinline__ void count_wave(unsigned *dst, unsigned v) {
    if (!dst) return;
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
}

// Bracket scatter for the source interval [x0, x1) of ascending source u:
// every query of [qa, qb) inside it gets (u, phase) and, when frq is given,
// its fraction (Q - x0) / (x1 - x0).
__device__ __forceinline__ void scatter_interval(const Grid1 &Qg, double inv_step, int u, double x0,
                                                 double x1, int qa, int qb, int *brk, float *frq,
                                                 int qbase) {
    if (!(x0 < x1)) return;                                   // empty interval
    const int qs = first_q_ge(Qg, inv_step, x0, qa, qb);
    const double dx = __dsub_rn(x1, x0);
    for (int q = qs; q < qb; ++q) {
        const double Q = Qg(q);
        if (!(Q < x1)) break;
        const double frac = __ddiv_rn(__dsub_rn(Q, x0), dx);
        const int wq = (int)fmin(fmax(rint(frac * 32768.0), 0.0), 32767.0);
        brk[q - qbase] = (u << 8) | (wq >> 7);
        if (frq) frq[q - qbase] = (float)frac;
    }
}

// Exact per-output value of the azimuth pass (row edges, scatter misses): a
// bisection over the sources KR_j * tan_s and np.interp / SINC32 there.  Kept
// out of line so the unrolled gather stays small.
template <bool SINC>
__device__ __noinline__ float2 azimuth_slow(double krj, const double *__restrict__ tan_s, int m,
                                            double Q, const float2 *row, const float *tapl) {
    auto xa = [&](int u) { return __dmul_rn(krj, tan_s[u]); };
    auto fa = [&](int u) { return row[u]; };
    Walk<decltype(xa)> w(xa, m);
    w.seek(Q, -1);
    const Bracket bb = classify(SINC, w.c, m, Q, w.xprev, w.xcur, xa(m - 1));
    return value_slow(bb, Q, fa, tapl);
}

// ------------------------------------------------------------------ pass 1 (range)
constexpr int RP = 16;                   // pulses per tile
constexpr int RQ = 128;                  // range bins per tile (8 per thread)
constexpr int RS = 176;                  // staged samples per pulse (RQ + 37 + slack)
constexpr int RSP = 196;                 // padded strip row: spad(RS-1) < RSP and RSP = 4 mod 32,
                                         // so a wave's 4 pulses start 8 banks apart

// Persistent: each workgroup stages the tap table once, then walks tiles of
// RP sorted pulses x RQ range bins.  Per tile: (A) one FP64 position per source
// sample, each source interval scattering its bracket to the range bins inside
// it; (B) a branch-light gather over those brackets.  Outputs the scatter did
// not reach (edges, or a strip estimate that fell short) take the per-output
// walk.
template <bool SINC, bool FULLFREQ>
__global__ __launch_bounds__(256, 2) void k_range_f32(
    const float2 *__restrict__ sig, int m, int n, const int32_t *__restrict__ order,
    const double *__restrict__ a64, const double *__restrict__ f0v, const double *__restrict__ dfv,
    const double *__restrict__ freq, const Grid1 KR,
    const int16_t *__restrict__ sinc, float2 *__restrict__ g1t, int ldp,
    unsigned *__restrict__ nslow) {
    static_assert(sizeof(float2) * RQ * (RP + 1) <= sizeof(float2) * RP * RSP, "tile fits the strip");
    __shared__ float2 strip[RP][RSP];                  // source strips, then the output tile
    __shared__ int brk[RP * RQ];
    __shared__ float frq[SINC ? 1 : RP * RQ];
    __shared__ float tapl[SINC ? TROWS * 256 : 1];
    __shared__ int lo_s[RP], hi_s[RP];
    auto tile = reinterpret_cast<float2 (*)[RP + 1]>(&strip[0][0]);
    unsigned slow = 0;
    const int t = threadIdx.x;
    const int pl = t >> 4, ch = t & 15;
    if (SINC) load_taps(tapl, sinc);
    const double inv_step = KR.tab ? 0.0 : 1.0 / KR.step;
    const int tiles_p = (m + RP - 1) / RP, tiles_q = (n + RQ - 1) / RQ;
    // each workgroup owns whole pulse groups and sweeps their range tiles in
    // order: strip halos of consecutive tiles are L2 hits, and concurrently
    // running workgroups write different pulse columns of the same g1t rows
    const int my_groups = blockIdx.x < tiles_p ? (tiles_p - 1 - blockIdx.x) / gridDim.x + 1 : 0;
    for (int it = 0; it < my_groups * tiles_q; ++it) {
        const int p0 = (blockIdx.x + (it / tiles_q) * gridDim.x) * RP, q0 = (it % tiles_q) * RQ;
        const int qe = min(q0 + RQ, n);
        const int p = p0 + pl;
        const int qa = q0 + ch * 8;                   // this thread's first output
        __syncthreads();                              // previous tile fully consumed
        for (int r = t; r < RP * RQ; r += 256) brk[r] = BR_NONE;
        int i = 0;
        double ai = 1.0, f0 = 0.0, df = 1.0;
        bool asc = true;
        if (p < m) {
            i = order[p];
            ai = a64[i];
            if (!FULLFREQ) { f0 = f0v[i]; df = dfv[i]; }
        }
        const double *frow = FULLFREQ ? freq + (size_t)i * n : nullptr;
        // kr = (2 f / c) pr as (2 pr / c) * f: within 2 ulp of numpy's rounding
        auto kr = [&](int j) {
            const double f = FULLFREQ ? frow[j] : __dadd_rn(f0, __dmul_rn((double)j, df));
            return __dmul_rn(ai, f);
        };
        if (p < m) asc = !(kr(0) > kr(n - 1));
        auto xa = [&](int u) { return kr(asc ? u : n - 1 - u); };
        const float2 *row = sig + (size_t)i * n;
        auto fa_g = [&](int u) { return row[asc ? u : n - 1 - u]; };
        auto guess = [&](double Q) {                  // ascending-view count estimate
            if (FULLFREQ) return -1;
            const double js = (Q / ai - f0) / df;
            const double g = asc ? js : (double)(n - 1) - js;
            if (!(g == g)) return -1;
            return (int)fmin(fmax(floor(g) + 1.0, 0.0), (double)n);
        };
        if (ch == 0 && p < m) {                       // strip: first bracket minus taps ..
            const double Q = KR(min(q0, n - 1));
            Walk<decltype(xa)> w(xa, n);
            w.seek(Q, guess(Q));
            const int lo = max(0, min(w.c - 1 - 15 - 2, n - RS));
            int len = RS;                             // .. to the estimated last bracket
            const int gl = guess(KR(qe - 1));
            if (gl >= 0) len = min(RS, max(0, gl + 16 + 3 - lo));
            lo_s[pl] = lo;
            hi_s[pl] = min(lo + len, n);
        }
        __syncthreads();
        const int lo = lo_s[pl], hi = hi_s[pl];       // strip holds [lo, hi)
        if (p < m) {
            for (int u = ch; u < hi - lo; u += 16)
                strip[pl][spad(u)] = ld_stream(row + (asc ? lo + u : n - 1 - lo - u));
            // (A) scatter: source intervals [x_u, x_u+1) of the strip
            for (int u = lo + ch; u < hi - 1; u += 16)
                scatter_interval(KR, inv_step, u, xa(u), xa(u + 1), q0, qe, brk + pl * RQ,
                                 SINC ? nullptr : frq + pl * RQ, q0);
            if (ch == 0 && (lo == 0 || hi == n))
                scatter_outside(KR, inv_step, lo == 0, hi == n, xa(0), xa(n - 1), q0, qe,
                                brk + pl * RQ, q0);
        }
        __syncthreads();
        float2 vals[8];
        if (p < m) {
            auto fa = [&](int u) { return (u >= lo && u < hi) ? strip[pl][spad(u - lo)] : fa_g(u); };
            // (B) gather, four outputs per shared window
#pragma unroll
            for (int g = 0; g < 2; ++g) {
                int e[4], d[4], ph[4];
                bool fast = SINC;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int q = qa + g * 4 + u;
                    e[u] = q < qe ? brk[pl * RQ + q - q0] : BR_NONE;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int J = e[u] >> 8;
                    d[u] = J - (e[0] >> 8);
                    ph[u] = e[u] & 255;
                    fast = fast && e[u] != BR_NONE && J >= 15 && J <= n - 17 && d[u] >= 0 && d[u] <= KD;
                }
                const int s0 = (e[0] >> 8) - 15;
                fast = fast && s0 >= lo && s0 + KW <= hi;
                float2 v[4];
                if (fast) {
                    sinc4(strip[pl], s0 - lo, tapl, ph, d, v);
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int q = qa + g * 4 + u;
                        const int J = e[u] >> 8;
                        if (q >= qe || e[u] == BR_ZERO) {
                            v[u] = make_float2(0.f, 0.f);
                        } else if (e[u] == BR_LAST) {
                            v[u] = fa(n - 1);
                        } else if (!SINC && e[u] != BR_NONE) {        // linear from the scatter
                            const float fr = frq[pl * RQ + q - q0];
                            const float2 a0 = fa(J), a1 = fa(J + 1);
                            v[u] = make_float2(fmaf(a1.x - a0.x, fr, a0.x), fmaf(a1.y - a0.y, fr, a0.y));
                        } else if (e[u] >= 0) {                       // sinc off the shared window
                            v[u] = value_scattered<SINC>(e[u], n, KR(q), xa(J), xa(J + 1), fa, tapl);
                        } else {                                      // exact per-output path
                            ++slow;
                            const double Q = KR(q);
                            Walk<decltype(xa)> w(xa, n);
                            w.seek(Q, guess(Q));
                            const Bracket bb = classify(SINC, w.c, n, Q, w.xprev, w.xcur, xa(n - 1));
                            v[u] = value_slow(bb, Q, fa, tapl);
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) vals[g * 4 + u] = v[u];
            }
        }
        __syncthreads();                              // strips dead: the tile takes over
        if (p < m) {
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (qa + u < qe) tile[qa + u - q0][pl] = vals[u];
        }
        __syncthreads();
        for (int r = t; r < RQ * RP; r += 256) {
            const int qq = r / RP, pp = r % RP;
            if (q0 + qq < n && p0 + pp < m) st_stream(g1t + (size_t)(q0 + qq) * ldp + p0 + pp, tile[qq][pp]);
        }
    }
    count_wave(nslow, slow);
}

// ------------------------------------------------------------------ pass 2 (+ window + FFT A)
// A is row-major [range bin j][cross r]: every workgroup of the azimuth
// kernel writes whole 128-byte lines.
__device__ __forceinline__ size_t ablk(int j, int r, int m2) { return (size_t)j * m2 + r; }

// Workgroup b -> consecutive items on the same XCD (b % 8 selects the XCD).
__device__ __forceinline__ int xcd_map(int b, int count) {
    const int per = (count + 7) / 8;
    return (b % 8) * per + b / 8;
}

// Elements per thread of the azimuth kernel's FFT: 16 (512 threads) for the
// 8192-point line so all waves stay busy, 32 below.
template <int L>
constexpr int az_emax() { return 32; }
template <int L>
constexpr int az_nb() { return pfa32::Cfg<L, az_emax<L>()>::NB; }

// Persistent over range bins: iteration it of workgroup b (XCD x = b % 8, rank
// b / 8) takes j = x * ceil(n/8) + it * (grid/8) + rank.
// Per range bin: stage the sorted-pulse row, (A) scatter each source interval
// [KR_j tan_s[p], KR_j tan_s[p+1]) to the KC queries inside it, (B) gather
// four outputs per shared window, window and sign-fold into the LDS line, FFT.
// MODE 0: store the windowed k-space g2[q][j] (resample_kspace stage output)
// MODE 1: fused -- window, sign fold, pad, FFT along cross -> A (row-major)
// MODE 2: FFT along cross of a given g2 (focus() stage) -> A (row-major)
template <int L, int MODE, bool SINC, bool INV>
__global__ __launch_bounds__(az_nb<L>()) void k_azimuth_f32(
    const float2 *__restrict__ g1t, int ldp, int m, int n, const Grid1 KR,
    const double *__restrict__ tan_s, const Grid1 KC,
    const double *__restrict__ hm, const double *__restrict__ hn,
    const int16_t *__restrict__ sinc, const float2 *__restrict__ g2in,
    float2 *__restrict__ out, const float2 *__restrict__ twg, unsigned *__restrict__ nslow) {
    constexpr int EM = az_emax<L>();
    using Cf = pfa32::Cfg<L, EM>;
    constexpr int M2 = Cf::N;
    constexpr int NB = Cf::NB;
    constexpr bool RES = MODE != 2;                    // resampling modes
    __shared__ float2 buf[Cf::PADN];                   // the FFT line
    __shared__ float tapl[(SINC && RES) ? TROWS * 256 : 1];
    __shared__ float2 twl[M2 / 64 + 64];
    __shared__ int brk[RES ? M2 : 1];
    __shared__ float frq[(RES && !SINC) ? M2 : 1];
    const int t = threadIdx.x;
    if (SINC && RES) load_taps(tapl, sinc);
    const pfa32::Tw tw = pfa32::load_tw<M2>(twl, twg);
    const double inv_step = KC.tab ? 0.0 : 1.0 / KC.step;
    const int per_x = (n + 7) / 8;
    const int ranks = (gridDim.x + 7) / 8;
    const int x = blockIdx.x % 8, rank = blockIdx.x / 8;
    unsigned slow = 0;
    for (int base = 0; base < per_x; base += ranks) {
        const int jl = base + rank;
        const int j = x * per_x + jl;
        const bool active = jl < per_x && j < n;     // uniform per workgroup
        if (!active) continue;
        __syncthreads();                              // previous line fully consumed
        if (MODE == 2) {
            for (int q = t; q < M2; q += NB) {
                float2 v = q < m ? g2in[(size_t)q * n + j] : make_float2(0.f, 0.f);
                if ((q + j) & 1) v = make_float2(-v.x, -v.y);
                buf[pfa32::padf(q)] = v;
            }
        } else {
            // the sorted-pulse row is read through L1: each window sample feeds
            // four outputs, and outputs go straight into the FFT line
            const float2 *row = g1t + (size_t)j * ldp;
            for (int u = t; u < m; u += NB) brk[u] = BR_NONE;
            __syncthreads();
            const double krj = KR(j);
            auto xa = [&](int u) { return __dmul_rn(krj, tan_s[u]); };
            auto fa = [&](int u) { return row[u]; };
            // (A) scatter the source intervals to the KC queries
            for (int pp = t; pp < m - 1; pp += NB)
                scatter_interval(KC, inv_step, pp, xa(pp), xa(pp + 1), 0, m, brk,
                                 SINC ? nullptr : frq, 0);
            if (t == 0) scatter_outside(KC, inv_step, true, false, xa(0), 0.0, 0, m, brk, 0);
            if (t == NB - 1) scatter_outside(KC, inv_step, false, true, 0.0, xa(m - 1), 0, m, brk, 0);
            __syncthreads();
            const double hj = hn ? hn[j] : 1.0;
            // (B) gather: sub-group g of thread t covers outputs 4 (g NB + t) .. +3,
            // so a wave's lanes read neighbouring windows
            for (int g = 0; g * NB * 4 < M2; ++g) {
                const int qa = (g * NB + t) * 4;
                int e[4], d[4], ph[4];
                bool fast = SINC;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int q = qa + u;
                    e[u] = q < m ? brk[q] : BR_NONE;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int J = e[u] >> 8;
                    d[u] = J - (e[0] >> 8);
                    ph[u] = e[u] & 255;
                    fast = fast && e[u] != BR_NONE && J >= 15 && J <= m - 17 && d[u] >= 0 && d[u] <= KD;
                }
                const int s0 = (e[0] >> 8) - 15;
                fast = fast && s0 + KW <= m;            // the whole window lies in the row
                float2 v[4];
                if (fast) {
                    sinc4<false>(row, s0, tapl, ph, d, v);
                } else {
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int q = qa + u;
                        const int J = e[u] >> 8;
                        if (q >= m || e[u] == BR_ZERO) {
                            v[u] = make_float2(0.f, 0.f);
                        } else if (e[u] == BR_LAST) {
                            v[u] = fa(m - 1);
                        } else if (!SINC && e[u] != BR_NONE) {        // linear from the scatter
                            const float fr = frq[q];
                            const float2 a0 = fa(J), a1 = fa(J + 1);
                            v[u] = make_float2(fmaf(a1.x - a0.x, fr, a0.x), fmaf(a1.y - a0.y, fr, a0.y));
                        } else if (e[u] >= 0) {                       // sinc off the shared window
                            v[u] = value_scattered<SINC>(e[u], m, KC(q), xa(J), xa(J + 1), fa, tapl);
                        } else {                                      // exact per-output path
                            ++slow;
                            v[u] = azimuth_slow<SINC>(krj, tan_s, m, KC(q), row, tapl);
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int q = qa + u;
                    if (q >= M2) break;
                    float2 xv = v[u];
                    if (q < m && hm) {
                        const float wgt = (float)__dmul_rn(hm[q], hj);
                        xv = make_float2(xv.x * wgt, xv.y * wgt);
                    }
                    if (MODE == 0) {
                        if (q < m) out[(size_t)q * n + j] = xv;
                    } else {
                        if ((q + j) & 1) xv = make_float2(-xv.x, -xv.y);
                        buf[pfa32::padf(q)] = xv;
                    }
                }
            }
        }
        if constexpr (MODE != 0) {
            __syncthreads();
            float2 *dst = out + ablk(j, 0, M2);
            const float scale = INV ? 1.0f / (float)M2 : 1.0f;
            pfa32::fft<L, INV, EM>(buf, tw, [&](int idx, float2 v) {
                if (INV) v = make_float2(v.x * scale, v.y * scale);
                st_stream(dst + idx, v);
            });
        }
    }
    count_wave(nslow ? nslow + 1 : nullptr, slow);
}

// ------------------------------------------------------------------ FFT B (range) + detect
// OUT 0: complex64 image, 1: |image| float32.  Workgroup b -> cross bin r with
// consecutive bins on the same XCD.
template <int L, bool INV, int OUT>
__global__ __launch_bounds__(pfa32::Cfg<L>::NT < 64 ? 64 : pfa32::Cfg<L>::NT) void k_fft_b_f32(
    const float2 *__restrict__ A, int n, int m2, void *__restrict__ outp,
    const float2 *__restrict__ twg) {
    using Cf = pfa32::Cfg<L>;
    constexpr int N2 = Cf::N;
    __shared__ float2 buf[Cf::PADN];
    __shared__ float2 twl[N2 / 64 + 64];
    const pfa32::Tw tw = pfa32::load_tw<N2>(twl, twg);
    const int r = xcd_map(blockIdx.x, m2);
    const float scale = INV ? 1.0f / (float)N2 : 1.0f;
    // the first pass reads its butterflies straight from A (the corner-turn)
    auto src = [&](int jj) { return jj < n ? A[ablk(jj, r, m2)] : make_float2(0.f, 0.f); };
    pfa32::fft_src<L, INV>(buf, tw, src, [&](int c, float2 v) {
        if (INV) v = make_float2(v.x * scale, v.y * scale);
        if ((r + c) & 1) v = make_float2(-v.x, -v.y);
        if (OUT == 0) reinterpret_cast<float2 *>(outp)[(size_t)r * N2 + c] = v;
        else reinterpret_cast<float *>(outp)[(size_t)r * N2 + c] = __fsqrt_rn(v.x * v.x + v.y * v.y);
    });
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

#define PFA_F32_SWITCH(L, MACRO) \
    switch (L) {                 \
        MACRO(6) MACRO(7) MACRO(8) MACRO(9) MACRO(10) MACRO(11) MACRO(12) MACRO(13) \
        default: return pfa_fail(ctx, PFA_ERR_DIMS, "padded size must be a power of two in 64..8192"); \
    }

int launch_float_range(pfa_plan *p, const void *sig, const double *freq, void *g1t, int ldp) {
    pfa_ctx *ctx = p->ctx;
    const int tiles = ((p->m + RP - 1) / RP) * ((p->n + RQ - 1) / RQ);
    const int tiles_p = (p->m + RP - 1) / RP;
    dim3 grid(std::max(1, std::min(tiles_p, ctx->num_cu * 2)));
    (void)tiles;
    const bool sinc = p->flags & PFA_FLOAT_SINC32;
    const bool full = freq != nullptr;
    pfa_timer_begin(ctx, PFA_TIMER_RANGE);
#define ARGS (const float2 *)sig, p->m, p->n, p->order, p->a64, p->f0d, p->dfd, freq, grid_of(p->KR, p->kr_lin, p->n), p->sinc, \
             (float2 *)g1t, ldp, (unsigned *)(ctx->d_counters + 32)
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
    // persistent: one resident workgroup per CU, a multiple of 8 (one per XCD slot)
    const int grid_az = std::max(8, std::min((p->n + 7) / 8 * 8, ctx->num_cu / 8 * 8));
    pfa_timer_begin(ctx, PFA_TIMER_AZIMUTH);
#define ARGS (const float2 *)g1t, ldp, p->m, p->n, grid_of(p->KR, p->kr_lin, p->n), p->tan_s64, grid_of(p->KC, p->kc_lin, p->m), p->hm, p->hn, p->sinc, \
             (const float2 *)g2in, (float2 *)out, tw, (unsigned *)(ctx->d_counters + 32)
#define K4(LL, MODE)                                                                                 \
    if (sinc && inv) hipLaunchKernelGGL((k_azimuth_f32<LL, MODE, true, true>), dim3(grid_az), dim3(az_nb<LL>()), 0, ctx->stream, ARGS); \
    else if (sinc) hipLaunchKernelGGL((k_azimuth_f32<LL, MODE, true, false>), dim3(grid_az), dim3(az_nb<LL>()), 0, ctx->stream, ARGS); \
    else if (inv) hipLaunchKernelGGL((k_azimuth_f32<LL, MODE, false, true>), dim3(grid_az), dim3(az_nb<LL>()), 0, ctx->stream, ARGS); \
    else hipLaunchKernelGGL((k_azimuth_f32<LL, MODE, false, false>), dim3(grid_az), dim3(az_nb<LL>()), 0, ctx->stream, ARGS);
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
