// pfa_geometry.cpp -- the float focus's O(m + n) scene geometry from the
// reference's own arguments (freq, ax, ay), on the host, each operation
// rounded as the reference's numpy / Python expression rounds it
// (src/form_image_pfa.py:178-193, 205-209, 219-221, 235-238), and the
// plan / one-shot entry points built on it.
//
// The numpy semantics reproduced here:
//   ndarray.mean   add.reduce in blocks of 8192 elements (the ufunc buffer),
//                  each block summed pairwise (8-way unrolled below 128
//                  elements, halves rounded down to a multiple of 8 above),
//                  the blocks added in order, then one true division by m;
//   math.hypot     correctly rounded (CPython 3.10's vector_norm is, to
//                  within a few 1e-5 of cases we have not met in 20000
//                  random pairs; a double-double residual decides here);
//   a * b + c * d  every product and the sum rounded separately (no FMA);
//   np.linspace    k * step + start, step = (stop - start) / (cnt - 1), last
//                  element = stop;
//   np.argsort     the pulses sorted by tan_phi; ties (pulses with the same
//                  look angle) keep their index order, where numpy's
//                  default sort may order them either way;
//   np.hamming     0.54 + 0.46 cos(pi k / (M - 1)), k = 1 - M, 3 - M, ..., with
//                  the C library's cos (the routine numpy's float64 cos calls
//                  on this platform).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "pfa_internal.hpp"

namespace {

constexpr double C_LIGHT = 299792458.0;

// numpy's pairwise summation of one block (numpy/_core/src/umath/loops_utils.h)
double pairwise_sum(const double *a, int64_t n) {
    if (n < 8) {
        double r = 0.0;
        for (int64_t i = 0; i < n; ++i) r += a[i];
        return r;
    }
    if (n <= 128) {
        double r[8];
        for (int j = 0; j < 8; ++j) r[j] = a[j];
        int64_t i = 8;
        for (; i < n - (n % 8); i += 8)
            for (int j = 0; j < 8; ++j) r[j] += a[i + j];
        double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) res += a[i];
        return res;
    }
    int64_t n2 = n / 2;
    n2 -= n2 % 8;
    return pairwise_sum(a, n2) + pairwise_sum(a + n2, n - n2);
}

double np_mean(const double *a, int64_t n) {
    constexpr int64_t BUF = 8192;
    double s = 0.0;
    for (int64_t i = 0; i < n; i += BUF) s += pairwise_sum(a + i, std::min(BUF, n - i));
    return s / (double)n;
}

// sqrt(x^2 + y^2) correctly rounded: the sum of squares as a double-double
// (exact products by fma), a first square root, then neighbours checked
// against the exact residual s - r^2.
double hypot_cr(double x, double y) {
    x = std::fabs(x);
    y = std::fabs(y);
    if (std::isinf(x) || std::isinf(y)) return INFINITY;
    if (std::isnan(x) || std::isnan(y)) return NAN;
    if (x < y) std::swap(x, y);
    if (x == 0.0) return 0.0;
    if (y == 0.0) return x;
    const int e = std::ilogb(x);
    const double xs = std::ldexp(x, -e), ys = std::ldexp(y, -e);
    const double p1 = xs * xs, e1 = std::fma(xs, xs, -p1);
    const double p2 = ys * ys, e2 = std::fma(ys, ys, -p2);
    const double sh0 = p1 + p2;
    const double bb = sh0 - p1, sl0 = (p1 - (sh0 - bb)) + (p2 - bb);   // two-sum
    const double lo = sl0 + e1 + e2;
    const double sh = sh0 + lo, sl = lo - (sh - sh0);
    double r = std::sqrt(sh);
    for (int it = 0; it < 4; ++it) {
        const double ch = r * r, cl = std::fma(r, r, -ch);
        const double res = ((sh - ch) - cl) + sl;                        // s - r^2
        const double up = std::nextafter(r, INFINITY) - r;
        const double dn = r - std::nextafter(r, 0.0);
        if (res > r * up + 0.25 * up * up) r += up;
        else if (res < -(r * dn) + 0.25 * dn * dn) r -= dn;
        else break;
    }
    return std::ldexp(r, e);
}

// np.linspace(start, stop, cnt)
void linspace(double start, double stop, int cnt, double *out) {
    if (cnt == 1) {
        out[0] = start;
        return;
    }
    const volatile double delta = stop - start;
    const double step = delta / (double)(cnt - 1);
    if (step == 0.0) {                            // numpy's denormal branch: (k / div) * delta
        for (int k = 0; k < cnt; ++k) {
            const volatile double q = (double)k / (double)(cnt - 1);
            const volatile double v = q * delta;
            out[k] = v + start;
        }
    } else {
        for (int k = 0; k < cnt; ++k) {
            const volatile double v = (double)k * step;
            out[k] = v + start;
        }
    }
    out[cnt - 1] = stop;
}

void hamming(int M, double *out) {
    if (M == 1) {
        out[0] = 1.0;
        return;
    }
    const double Mf = (double)M;
    for (int i = 0; i < M; ++i) {
        const double k = (1.0 - Mf) + 2.0 * (double)i;
        const volatile double a = M_PI * k;
        const volatile double b = a / (Mf - 1.0);
        const volatile double c = 0.46 * std::cos(b);
        out[i] = 0.54 + c;
    }
}

}  // namespace

// The tables of pfa_plan_float and the unpadded geo of resample_kspace.
extern "C" int pfa_float_geometry(int m, int n, const double *freq, const double *f0,
                                  const double *df, const double *ax, const double *ay,
                                  double *pr, int32_t *order, double *tan_s, double *kr_grid,
                                  double *kc_grid, double *ham_c, double *ham_r, double geo[6]) {
    PFA_TRY
    if (m < 2 || n < 2) return PFA_ERR_DIMS;
    if (!ax || !ay || !pr || !order || !tan_s || !kr_grid || !kc_grid || !geo) return PFA_ERR_ARG;
    if ((ham_c == nullptr) != (ham_r == nullptr)) return PFA_ERR_ARG;
    if (!freq && (!f0 || !df)) return PFA_ERR_ARG;
    for (int i = 0; i < m; ++i)
        if (!std::isfinite(ax[i]) || !std::isfinite(ay[i])) return PFA_ERR_GEOMETRY;
    // rotated frame (form_image_pfa.py:180-186)
    double dx = np_mean(ax, m), dy = np_mean(ay, m);
    const double dn = hypot_cr(dx, dy);
    if (!(dn > 0.0) || !std::isfinite(dn)) return PFA_ERR_GEOMETRY;
    dx = dx / dn;
    dy = dy / dn;
    const double cx = -dy, cy = dx;
    std::vector<double> tan_phi(m);
    for (int i = 0; i < m; ++i) {
        const volatile double a1 = ax[i] * dx, a2 = ay[i] * dy;
        const volatile double c1 = ax[i] * cx, c2 = ay[i] * cy;
        pr[i] = a1 + a2;
        const double pc = c1 + c2;
        tan_phi[i] = pc / pr[i];
        if (pr[i] == 0.0 || !std::isfinite(tan_phi[i])) return PFA_ERR_GEOMETRY;
    }
    // KR over the common overlap of the pulses' kr rows (:187-193):
    // kr[i][k] = (2 freq[i][k] / C) pr[i]
    auto krv = [&](double f, double p) {
        const volatile double k = (2.0 * f) / C_LIGHT;
        return k * p;
    };
    double lo = INFINITY, hi = -INFINITY, gmin = INFINITY, gmax = -INFINITY;
    for (int i = 0; i < m; ++i) {
        double rmax = -INFINITY, rmin = INFINITY;
        if (freq) {
            const double *row = freq + (size_t)i * n;
            for (int k = 0; k < n; ++k) {
                const double v = krv(row[k], pr[i]);
                if (!std::isfinite(v)) return PFA_ERR_GEOMETRY;
                rmax = std::max(rmax, v);
                rmin = std::min(rmin, v);
            }
        } else {                                  // freq[i][k] = f0 + k df, each op rounded:
            for (int k : {0, n - 1}) {            // monotone in k, extrema at the ends
                const volatile double kd = (double)k * df[i];
                const double v = krv(f0[i] + kd, pr[i]);
                if (!std::isfinite(v)) return PFA_ERR_GEOMETRY;
                rmax = std::max(rmax, v);
                rmin = std::min(rmin, v);
            }
        }
        lo = std::min(lo, rmax);                  // kr.max(axis=1).min()
        hi = std::max(hi, rmin);                  // kr.min(axis=1).max()
        gmin = std::min(gmin, rmin);
        gmax = std::max(gmax, rmax);
    }
    if (lo > hi) {
        lo = gmin;
        hi = gmax;
    }
    linspace(lo, hi, n, kr_grid);
    // KC (:205-206): max |KR_j tan_phi_i| = max |KR| max |tan_phi| (one rounding,
    // monotone), KR's largest magnitude at an end of the linspace
    double tmax = 0.0;
    for (int i = 0; i < m; ++i) tmax = std::max(tmax, std::fabs(tan_phi[i]));
    const double kc_max = std::max(std::fabs(kr_grid[0]), std::fabs(kr_grid[n - 1])) * tmax;
    linspace(-kc_max, kc_max, m, kc_grid);
    // pulse order (:208-209)
    std::vector<int32_t> ord(m);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t a, int32_t b) { return tan_phi[a] < tan_phi[b]; });
    for (int i = 0; i < m; ++i) {
        order[i] = ord[i];
        tan_s[i] = tan_phi[ord[i]];
    }
    if (ham_c) {
        hamming(m, ham_c);
        hamming(n, ham_r);
    }
    // pixel spacing (:219-221) and the rotated frame
    geo[0] = kc_grid[m - 1] != kc_grid[0] ? 1.0 / (kc_grid[m - 1] - kc_grid[0]) : NAN;   // dc
    geo[1] = kr_grid[n - 1] != kr_grid[0] ? 1.0 / (kr_grid[n - 1] - kr_grid[0]) : NAN;   // dr
    geo[2] = dx;
    geo[3] = dy;
    geo[4] = cx;
    geo[5] = cy;
    return PFA_OK;
    PFA_CATCH(nullptr)
}

// pfa_plan_float from the reference's arguments; geo is pfa's (focus scales
// the spacing to the padded grid, form_image_pfa.py:237-238).
extern "C" int pfa_plan_float_pfa(pfa_ctx *ctx, pfa_plan **out, int m, int n, int sgn,
                                  const double *freq, const double *f0, const double *df,
                                  const double *ax, const double *ay, int window,
                                  const int16_t *sinc_tab, int flags, double geo[6]) {
    PFA_TRY
    if (!ctx || !out) return PFA_ERR_ARG;
    *out = nullptr;
    if (!geo) return pfa_fail(ctx, PFA_ERR_ARG, "plan_float_pfa: null geo");
    if (m < 2 || n < 2) return pfa_fail(ctx, PFA_ERR_DIMS, "plan_float_pfa: need m, n >= 2");
    if (!freq && (!f0 || !df)) return pfa_fail(ctx, PFA_ERR_ARG, "plan_float_pfa: pass freq or both f0 and df");
    std::vector<double> pr(m), tan_s(m), kr(n), kc(m), hc(window ? m : 0), hr(window ? n : 0);
    std::vector<int32_t> order(m);
    const int st = pfa_float_geometry(m, n, freq, f0, df, ax, ay, pr.data(), order.data(),
                                      tan_s.data(), kr.data(), kc.data(),
                                      window ? hc.data() : nullptr, window ? hr.data() : nullptr, geo);
    if (st == PFA_ERR_GEOMETRY) return pfa_fail(ctx, st, "plan_float_pfa: degenerate or non-finite geometry");
    if (st != PFA_OK) return pfa_fail(ctx, st, "plan_float_pfa: bad argument");
    CK(pfa_plan_float(ctx, out, m, n, sgn, pr.data(), order.data(), tan_s.data(), kr.data(),
                      kc.data(), freq ? nullptr : f0, freq ? nullptr : df,
                      window ? hc.data() : nullptr, window ? hr.data() : nullptr, sinc_tab, flags));
    const int m2 = (*out)->m2, n2 = (*out)->n2;
    geo[1] *= (double)n / (double)n2;
    geo[0] *= (double)m / (double)m2;
    return PFA_OK;
    PFA_CATCH(ctx)
}

// form_image_pfa.pfa(sig, freq, ax, ay, sgn) in one call: plan, focus, release.
extern "C" int pfa_pfa_f32(pfa_ctx *ctx, const float *sig_c64, int m, int n, const double *freq,
                           const double *ax, const double *ay, int sgn, int flags, void *out,
                           int out_kind, double geo[6]) {
    PFA_TRY
    if (!ctx) return PFA_ERR_ARG;
    if (!freq) return pfa_fail(ctx, PFA_ERR_ARG, "pfa_f32: null freq");
    pfa_plan *p = nullptr;
    CK(pfa_plan_float_pfa(ctx, &p, m, n, sgn, freq, nullptr, nullptr, ax, ay, 1, nullptr, flags, geo));
    const int s = pfa_focus_f32(p, sig_c64, freq, out, out_kind);
    pfa_plan_destroy(p);
    return s;
    PFA_CATCH(ctx)
}
