// pfa_quicklook.hip -- form_image_pfa.quicklook for any shape
// (src/form_image_pfa.py:158-162): the raw history times hamming2d
// (:154-155), then transform2d (:144-147): fftshift(fft2(ifftshift(x))) for
// sgn < 0, the same with ifft2 otherwise, with no padding, so the transform
// lengths are the scene's own (5634 x 4319 for the board scene).
//
// Two passes of one line per workgroup, each a 1-D transform of length N:
//   rows    (N = n samples): input row (r + m/2) mod m, columns (i + n/2) mod n
//           (the two ifftshifts), window f32(hm[row] hn[col]); output T[r][k]
//           stored at column (k + n/2) mod n (fftshift along range);
//   columns (N = m pulses): column b of T, output row (k + m/2) mod m.
// A power-of-two N is one LDS FFT of N points.  Any other N is Bluestein's
// chirp-z transform on an M-point FFT, M = 2^ceil(log2(2N - 1)):
//   X_k = c_k sum_i (x_i c_i) conj(c_{k-i}),  c_i = exp(-+ pi i (i^2 mod 2N) / N),
// i.e. a cyclic convolution of a_i = x_i c_i with B (B_j = conj(c_j) for
// j < N, B_{M-j} = conj(c_j) for 0 < j < N), done as FFT_M, a product with
// Bhat = FFT_M(B) / M (and / N for the inverse transform), and an inverse
// FFT_M.  The chirp and Bhat are formed on the host in double precision.  The
// inverse FFT is the forward one on conjugated data: IFFT(y) = conj(FFT(conj(y))).
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "pfa_float_dev.hpp"
#include "pfa_internal.hpp"

namespace {

using pfa32::cmul;
constexpr int QL_EMAX = 16;

__device__ __forceinline__ float2 conjf2(float2 v) { return make_float2(v.x, -v.y); }

struct LineArgs {
    const float2 *src;      // rows: the history [m][n]; columns: T [m][n]
    float2 *dst;            // rows: T; columns: the image (complex) or |image| (float)
    int m, n;
    int cols;               // 0: rows pass, 1: columns pass
    int N;                  // transform length of this pass
    int inv;                // 1: ifft (sgn > 0)
    int blue;               // 1: Bluestein (N not a power of two, or below 64)
    int out_abs;            // columns pass: |image| float32
    const double *hm, *hn;  // tapers (rows pass; null: no window)
    const float2 *chirp;    // c_i, i < N (Bluestein)
    const float2 *bhat;     // Bhat, M entries (Bluestein)
    const float2 *twg;      // M-point twiddle tables
};

template <int LM>
constexpr int ql_nb() { return pfa32::Cfg<LM, QL_EMAX>::NB; }

template <int LM>
__global__ __launch_bounds__(ql_nb<LM>()) void k_ql_line(const LineArgs a) {
    using Cf = pfa32::Cfg<LM, QL_EMAX>;
    constexpr int M = Cf::N, NB = Cf::NB;
    __shared__ float2 buf[Cf::PADN];
    __shared__ float2 twl[M / 64 + 64];
    const pfa32::Tw tw = pfa32::load_tw<M>(twl, a.twg);
    const int line = blockIdx.x, t = threadIdx.x;
    const int N = a.N;
    // element i of this line's (shifted) input
    auto load = [&](int i) -> float2 {
        if (!a.cols) {
            const int r = (line + a.m / 2) % a.m, c = (i + a.n / 2) % a.n;
            float2 v = a.src[(size_t)r * a.n + c];
            if (a.hm) {                                   // hamming2d: f32(hm[r] hn[c]) as numpy forms it
                const float w = (float)__dmul_rn(a.hm[r], a.hn[c]);
                v = make_float2(v.x * w, v.y * w);
            }
            return v;
        }
        return a.src[(size_t)i * a.n + line];
    };
    auto store = [&](int k, float2 v) {
        if (!a.cols) {
            a.dst[(size_t)line * a.n + (k + a.n / 2) % a.n] = v;
        } else {
            const size_t o = (size_t)((k + a.m / 2) % a.m) * a.n + line;
            if (a.out_abs) reinterpret_cast<float *>(a.dst)[o] = __fsqrt_rn(v.x * v.x + v.y * v.y);
            else a.dst[o] = v;
        }
    };
    for (int i = t; i < M; i += NB) {
        float2 v = make_float2(0.f, 0.f);
        if (i < N) {
            v = load(i);
            if (a.blue) v = cmul(v, a.chirp[i]);
            else if (a.inv) v = conjf2(v);               // ifft by the forward FFT of conj(x)
        }
        buf[pfa32::padf(i)] = v;
    }
    __syncthreads();
    if (a.blue) {
        // A = FFT(a); A Bhat, conjugated for the inverse FFT that follows
        pfa32::fft<LM, false, QL_EMAX>(buf, tw, [&](int o, float2 v) {
            buf[pfa32::padf(o)] = conjf2(cmul(v, a.bhat[o]));
        });
        pfa32::fft<LM, false, QL_EMAX>(buf, tw, [&](int k, float2 v) {
            if (k < N) store(k, cmul(conjf2(v), a.chirp[k]));
        });
    } else {
        const float scale = 1.0f / (float)N;
        pfa32::fft<LM, false, QL_EMAX>(buf, tw, [&](int k, float2 v) {
            if (a.inv) v = make_float2(v.x * scale, -v.y * scale);
            store(k, v);
        });
    }
}

// Host tables of one Bluestein length: chirp c[N] and Bhat[M], in a
// workspace slot per (N, direction).
struct BlueTables {
    const float2 *chirp = nullptr, *bhat = nullptr;
};

void fft_double(std::vector<std::complex<double>> &x) {
    const size_t n = x.size();
    for (size_t i = 1, j = 0; i < n; ++i) {              // bit reversal
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(x[i], x[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        for (size_t k = 0; k < len / 2; ++k) {
            const double th = -2.0 * M_PI * (double)k / (double)len;
            const std::complex<double> w(std::cos(th), std::sin(th));
            for (size_t s = 0; s < n; s += len) {
                const std::complex<double> u = x[s + k], v = x[s + k + len / 2] * w;
                x[s + k] = u + v;
                x[s + k + len / 2] = u - v;
            }
        }
    }
}

int blue_tables(pfa_ctx *ctx, int N, int M, bool inv, BlueTables *out) {
    char slot[48];
    snprintf(slot, sizeof slot, "ql_blue_%d_%d_%d", N, M, (int)inv);
    auto it = ctx->ws.find(slot);
    if (it == ctx->ws.end() || !it->second.bytes) {
        std::vector<float2> h((size_t)N + M);
        std::vector<std::complex<double>> B((size_t)M, 0.0);
        const double sgn = inv ? 1.0 : -1.0;
        for (int i = 0; i < N; ++i) {
            const long long r = ((long long)i * i) % (2LL * N);
            const double th = sgn * M_PI * (double)r / (double)N;
            h[i] = make_float2((float)std::cos(th), (float)std::sin(th));
            const std::complex<double> b(std::cos(th), -std::sin(th));     // conj(c_i)
            B[i] = b;
            if (i > 0) B[M - i] = b;
        }
        fft_double(B);
        const double scale = 1.0 / (double)M / (inv ? (double)N : 1.0);
        for (int j = 0; j < M; ++j)
            h[(size_t)N + j] = make_float2((float)(B[j].real() * scale), (float)(B[j].imag() * scale));
        int st = PFA_OK;
        void *d = pfa_ws(ctx, slot, h.size() * sizeof(float2), &st);
        if (!d) return st;
        PFA_HIP(ctx, hipMemcpy(d, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice));
        it = ctx->ws.find(slot);
    }
    out->chirp = (const float2 *)it->second.ptr;
    out->bhat = out->chirp + N;
    return PFA_OK;
}

int pow2_ceil(int v) {
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

int launch_line_pass(pfa_ctx *ctx, LineArgs a, int lines) {
    const bool p2 = (a.N & (a.N - 1)) == 0 && a.N >= 64;
    const int M = p2 ? a.N : std::max(64, pow2_ceil(2 * a.N - 1));
    a.blue = !p2;
    int st = PFA_OK;
    a.twg = twiddles_f32(ctx, M, &st);
    if (!a.twg) return st;
    if (a.blue) {
        BlueTables bt;
        CK(blue_tables(ctx, a.N, M, a.inv != 0, &bt));
        a.chirp = bt.chirp;
        a.bhat = bt.bhat;
    }
    const int LM = log2_exact(M);
#define QL_CASE(L)                                                                          \
    case L:                                                                                 \
        hipLaunchKernelGGL(k_ql_line<L>, dim3(lines), dim3(ql_nb<L>()), 0,                 \
                           ctx->stream, a);                                                 \
        break;
    switch (LM) {
        QL_CASE(6) QL_CASE(7) QL_CASE(8) QL_CASE(9) QL_CASE(10) QL_CASE(11) QL_CASE(12)
        QL_CASE(13) QL_CASE(14)
        default: return pfa_fail(ctx, PFA_ERR_DIMS, "quicklook: transform length above 16384");
    }
#undef QL_CASE
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

// Largest supported length: powers of two up to 16384, others up to 8192
// (their M = 2N - 1 rounded up must stay within 16384 points of LDS).
bool ql_len_ok(int N) {
    if (N < 1) return false;
    if ((N & (N - 1)) == 0) return N <= 16384;
    return N <= 8192;
}

int quicklook_dev(pfa_ctx *ctx, const float *sig, int m, int n, int sgn, const double *hm,
                  const double *hn, void *out, int out_kind) {
    int st = PFA_OK;
    float2 *T = (float2 *)pfa_ws(ctx, "ql_T", (size_t)m * n * 8, &st);
    if (!T) return st;
    LineArgs a{};
    a.m = m;
    a.n = n;
    a.inv = sgn > 0;
    a.hm = hm;
    a.hn = hn;
    pfa_timer_begin(ctx, PFA_TIMER_OTHER);
    a.src = (const float2 *)sig;
    a.dst = T;
    a.cols = 0;
    a.N = n;
    CK(launch_line_pass(ctx, a, m));
    a.src = T;
    a.dst = (float2 *)out;
    a.cols = 1;
    a.N = m;
    a.hm = a.hn = nullptr;
    a.out_abs = out_kind == PFA_OUT_ABS_F32;
    CK(launch_line_pass(ctx, a, n));
    pfa_timer_end(ctx, PFA_TIMER_OTHER);
    return PFA_OK;
}

int ql_check(pfa_ctx *ctx, const float *sig, int m, int n, int sgn, void *out, int out_kind) {
    if (!ctx) return PFA_ERR_ARG;
    if (!sig || !out || sgn == 0 || (out_kind != PFA_OUT_C64 && out_kind != PFA_OUT_ABS_F32))
        return pfa_fail(ctx, PFA_ERR_ARG, "quicklook: bad argument");
    if (!ql_len_ok(m) || !ql_len_ok(n))
        return pfa_fail(ctx, PFA_ERR_DIMS, "quicklook: sizes are 1..8192, or powers of two up to 16384");
    return PFA_OK;
}

std::vector<double> hamming_np(int M) {
    std::vector<double> h(M, 1.0);
    if (M == 1) return h;
    for (int i = 0; i < M; ++i) {                      // np.hamming: 0.54 + 0.46 cos(pi k / (M - 1))
        const double k = (1.0 - (double)M) + 2.0 * (double)i;
        const volatile double x = M_PI * k;
        const volatile double y = x / ((double)M - 1.0);
        const volatile double c = 0.46 * std::cos(y);
        h[i] = 0.54 + c;
    }
    return h;
}

int ql_tapers(pfa_ctx *ctx, int m, int n, const double **hm, const double **hn) {
    int st = PFA_OK;
    double *d = (double *)pfa_ws(ctx, "ql_taper", (size_t)(m + n) * 8, &st);
    if (!d) return st;
    std::vector<double> h = hamming_np(m), g = hamming_np(n);
    h.insert(h.end(), g.begin(), g.end());
    PFA_HIP(ctx, hipMemcpyAsync(d, h.data(), h.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    PFA_HIP(ctx, hipStreamSynchronize(ctx->stream));   // h is a host temporary
    *hm = d;
    *hn = d + m;
    return PFA_OK;
}

}  // namespace

extern "C" int pfa_quicklook_f32_dev(pfa_ctx *ctx, const float *sig_c64, int m, int n, int sgn,
                                     int window, void *out, int out_kind) {
    PFA_TRY
    CK(ql_check(ctx, sig_c64, m, n, sgn, out, out_kind));
    const double *hm = nullptr, *hn = nullptr;
    if (window) CK(ql_tapers(ctx, m, n, &hm, &hn));
    return quicklook_dev(ctx, sig_c64, m, n, sgn, hm, hn, out, out_kind);
    PFA_CATCH(ctx)
}

extern "C" int pfa_quicklook_f32(pfa_ctx *ctx, const float *sig_c64, int m, int n, int sgn,
                                 int window, void *out, int out_kind) {
    PFA_TRY
    CK(ql_check(ctx, sig_c64, m, n, sgn, out, out_kind));
    int st = PFA_OK;
    const size_t in_b = (size_t)m * n * 8, out_b = (size_t)m * n * (out_kind == PFA_OUT_C64 ? 8 : 4);
    void *ds = pfa_ws(ctx, "io_in", in_b, &st);
    if (!ds) return st;
    void *dout = pfa_ws(ctx, "io_out", out_b, &st);
    if (!dout) return st;
    const double *hm = nullptr, *hn = nullptr;
    if (window) CK(ql_tapers(ctx, m, n, &hm, &hn));
    CK(pfa_io_h2d(ctx, ds, sig_c64, in_b));
    CK(quicklook_dev(ctx, (const float *)ds, m, n, sgn, hm, hn, dout, out_kind));
    CK(pfa_io_d2h(ctx, out, dout, out_b));
    return PFA_OK;
    PFA_CATCH(ctx)
}
