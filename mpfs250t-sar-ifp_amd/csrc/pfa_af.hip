// pfa_af.hip -- gradient autofocus on the float focus path (no counterpart in
// the reference).  Per-pulse phase corrections phi give sig'[i][k] =
// sig[i][k] e^{j phi_i}; the image is I = PFA(sig'), the sharpness
//     S = sum |I|^4 / (sum |I|^2)^2,
// and dS/dphi comes from hand-written adjoints of each linear stage.
//
// The range pass acts per pulse, so the phase commutes with it: the
// corner-turned range-pass output g1t is computed once (the checkpoint) and
// every forward recomputes g1' = g1t e^{j phi} inside the azimuth kernel's row
// load.  With G = dS/dI* = (2 I / E2^2)(|I|^2 - E4 / E2) (Wirtinger), the
// backward is
//   k_af_fftb_adj   per image row r: conjugate-direction FFT along range of
//                   the sign-folded G -> B[r][j] (8-bin blocks along j)
//   k_af_az_adj     per range bin j: conjugate-direction FFT along cross of
//                   column j of B (the corner-turn in its load), sign fold,
//                   window, then the transpose of the azimuth gather as a
//                   segmented sum: every source sample sums the outputs whose
//                   32-tap (or 2-tap) windows cover it -- a contiguous run of
//                   outputs, found by bisection over their brackets -- so no
//                   atomics; then Im(H conj(g1')) accumulated per pulse
//   k_af_grad       dS/dphi_i = 2 sum_j Im(H[j][p] conj(g1'[j][p])), p = sorted i
// G is carried scaled by E2 (so it stays in FP32 range); the reduction divides.
#include "pfa_float_dev.hpp"

namespace {

__global__ __launch_bounds__(256) void k_af_rot(const float *__restrict__ phi,
                                                const int32_t *__restrict__ order, int m,
                                                float2 *__restrict__ rot) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < m) {
        float s, c;
        sincosf(phi[order[p]], &s, &c);
        rot[p] = make_float2(c, s);
    }
}

// E2 = sum |z|^2, E4 = sum |z|^4 in double: per-block partials, then one block.
__global__ __launch_bounds__(256) void k_af_stats1(const float2 *__restrict__ img, size_t count,
                                                   double *__restrict__ partial) {
    __shared__ double s2[256], s4[256];
    double e2 = 0.0, e4 = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (size_t)gridDim.x * 256) {
        const float2 z = img[i];
        const double a = (double)z.x * z.x + (double)z.y * z.y;
        e2 += a;
        e4 += a * a;
    }
    s2[threadIdx.x] = e2;
    s4[threadIdx.x] = e4;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s2[threadIdx.x] += s2[threadIdx.x + h];
            s4[threadIdx.x] += s4[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s2[0];
        partial[2 * blockIdx.x + 1] = s4[0];
    }
}

__global__ __launch_bounds__(256) void k_af_stats2(const double *__restrict__ partial, int nparts,
                                                   double *__restrict__ stats) {
    __shared__ double s2[256], s4[256];
    double e2 = 0.0, e4 = 0.0;
    for (int b = threadIdx.x; b < nparts; b += 256) {
        e2 += partial[2 * b];
        e4 += partial[2 * b + 1];
    }
    s2[threadIdx.x] = e2;
    s4[threadIdx.x] = e4;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            s2[threadIdx.x] += s2[threadIdx.x + h];
            s4[threadIdx.x] += s4[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        stats[0] = s2[0];
        stats[1] = s4[0];
        stats[2] = s2[0] > 0.0 ? s4[0] / (s2[0] * s2[0]) : 0.0;
    }
}

// B, the adjoint of FFT-B [cross row r][range bin j], is stored in blocks of 8
// range bins: element (r, j) at ((j / 8) m2 + r) 8 + j % 8.  A row's transform
// writes 64-byte chunks, and the azimuth adjoint's reads of one range bin down
// consecutive rows are 64 bytes apart rather than a full row (n * 8 bytes).
__device__ __forceinline__ size_t bblk(int r, int j, int m2) {
    return ((size_t)(j >> 3) * m2 + r) * 8 + (j & 7);
}

// Adjoint of FFT-B + output sign fold for image row r (one workgroup per row,
// gridDim.x = m2).  INVF: the forward transform was the inverse one (sgn > 0,
// with its 1/N).  Output B (bblk layout), j < n.
// 16 elements per thread (radix-16 passes): twice the threads of the default
// configuration, so a CU holds enough waves to cover the row loads.
constexpr int AF_FB_EMAX = 16;
template <int L>
constexpr int af_fb_nb() { return pfa32::Cfg<L, AF_FB_EMAX>::NT < 64 ? 64 : pfa32::Cfg<L, AF_FB_EMAX>::NT; }

template <int L, bool INVF>
__global__ __launch_bounds__(af_fb_nb<L>()) void k_af_fftb_adj(const float2 *__restrict__ img, int n,
                                                               const double *__restrict__ stats,
                                                               float2 *__restrict__ B,
                                                               const float2 *__restrict__ twg) {
    using Cf = pfa32::Cfg<L, AF_FB_EMAX>;
    constexpr int N = Cf::N;
    __shared__ float2 buf[Cf::PADN];
    __shared__ float2 twl[N / 64 + 64];
    const pfa32::Tw tw = pfa32::load_tw<N>(twl, twg);
    const int r = blockIdx.x;
    const double E2 = stats[0], E4 = stats[1];
    // G scaled by E2: (2 / E2) (|z|^2 - E4 / E2) z
    const float c1 = E2 > 0.0 ? (float)(2.0 / E2) : 0.f;
    const float c2 = E2 > 0.0 ? (float)(E4 / E2) : 0.f;
    const float scale = INVF ? 1.0f / (float)N : 1.0f;
    auto src = [&](int c) {
        const float2 z = img[(size_t)r * N + c];
        const float g = c1 * (z.x * z.x + z.y * z.y - c2);
        float2 v = make_float2(g * z.x, g * z.y);
        if ((r + c) & 1) v = make_float2(-v.x, -v.y);
        return v;
    };
    // the sink writes global memory and the kernel ends: no barriers in the last pass
    pfa32::fft_src<L, !INVF, AF_FB_EMAX, false, false>(buf, tw, src, [&](int j, float2 v) {
        if (j < n) B[bblk(r, j, gridDim.x)] = make_float2(v.x * scale, v.y * scale);
    });
}

// Adjoint of FFT-A + sign fold + window + azimuth gather, and the phase
// gradient partials, persistent over range bins as the forward azimuth kernel.
// Thread t owns sources p = 4 (g NB + t) + u and accumulates their gradient
// over its lines in registers; partial[blockIdx][p] at the end.
// M2 / 4 threads (1024 at 4096 points): one group of four sources and four
// outputs a thread, 16 waves a CU to cover the LDS latencies of the
// segmented sums; the line FFT runs on its own thread count within them.
template <int L>
constexpr int af_adj_nb() { return (1 << L) / 4 < 64 ? 64 : ((1 << L) / 4 > 1024 ? 1024 : (1 << L) / 4); }
template <int L, bool SINC, bool INVF>
__global__ __launch_bounds__(af_adj_nb<L>()) void k_af_az_adj(
    const float2 *__restrict__ g1t, int ldp, int m, int n, const Grid1 KR,
    const double *__restrict__ tan_s, const TanBins tb, const Grid1 KC,
    const double *__restrict__ hm, const double *__restrict__ hn,
    const int16_t *__restrict__ sinc, const float2 *__restrict__ rot,
    const float2 *__restrict__ B, const float2 *__restrict__ twg, float *__restrict__ partial) {
    constexpr int EM = L >= 14 ? 16 : 8;              // the line FFT's elements a thread
    using Cf = pfa32::Cfg<L, EM>;
    constexpr int M2 = Cf::N;
    constexpr int NB = af_adj_nb<L>();
    constexpr int G = (M2 + 4 * NB - 1) / (4 * NB);
    constexpr int PQ = (M2 + NB - 1) / NB;
    // first[p + FO] = the first output q whose bracket count c(q) >= p, for
    // p in [-FO, M2 + FO): the outputs whose windows cover sources p0 .. p0 + 3
    // are q in [first[p0 - 15], first[p0 + 20]) (c non-decreasing in q)
    constexpr int FO = 32;
    __shared__ float2 reg[Cf::PADN + 64];              // positions | adjoint line
    __shared__ float tapl[SINC ? TROWS * 256 : 1];
    __shared__ float2 twl[M2 / 64 + 64];
    __shared__ int code[M2];                           // (c << 8) | phase of every output
    __shared__ float frq[M2];
    __shared__ int first[M2 + 2 * FO];
    double *xs = reinterpret_cast<double *>(reg);
    const int t = threadIdx.x;
    if (SINC) load_taps(tapl, sinc);
    const pfa32::Tw tw = pfa32::load_tw<M2>(twl, twg);
    const int per_x = (n + 7) / 8;
    const int ranks = (gridDim.x + 7) / 8;
    const int x = blockIdx.x % 8, rank = blockIdx.x / 8;
    const float scale = INVF ? 1.0f / (float)M2 : 1.0f;
    constexpr int NTF = Cf::NT, EF = Cf::E;            // the line FFT's threads and elements a thread
    float acc[G * 4];
#pragma unroll
    for (int k = 0; k < G * 4; ++k) acc[k] = 0.f;
    // line-independent values held in registers: the tan values of this
    // thread's source positions and the phase rotation of its sources.
    // Loads are unconditional, from clamped indexes (a load on a path taken
    // at run time is waited for where the paths join).
    double tsv[PQ];
#pragma unroll
    for (int k = 0; k < PQ; ++k) tsv[k] = tan_s[min(t + k * NB, m - 1)];
    float2 rv[G * 4];
#pragma unroll
    for (int k = 0; k < G * 4; ++k) rv[k] = rot[min(4 * ((k / 4) * NB + t) + k % 4, m - 1)];
    for (int base = 0; base < per_x; base += ranks) {
        const int jl = base + rank;
        const int j = x * per_x + jl;
        if (!(jl < per_x && j < n)) continue;          // uniform per workgroup
        int tl = t;
        asm volatile("" : "+v"(tl));
        const double krj = KR(j);
        const double inv_krj = 1.0 / krj;
        // this line's global reads, issued before its first barrier: the
        // bracket guesses, the FFT's column of B and the g1t row
        int gix[PQ];
#pragma unroll
        for (int k = 0; k < PQ; ++k) gix[k] = tb.guess(__dmul_rn(KC(min(tl + k * NB, m - 1)), inv_krj));
        float2 bv[EF];
#pragma unroll
        for (int s2 = 0; s2 < EF; ++s2) bv[s2] = B[bblk(min(tl, NTF - 1) + s2 * NTF, j, M2)];
        float2 gv[G * 4];
#pragma unroll
        for (int k = 0; k < G * 4; ++k) gv[k] = g1t[(size_t)j * ldp + min(4 * ((k / 4) * NB + tl) + k % 4, m - 1)];
        __syncthreads();                               // previous line fully consumed
#pragma unroll
        for (int k = 0; k < PQ; ++k)
            if (tl + k * NB < m) xs[tl + k * NB] = __dmul_rn(krj, tsv[k]);
        __syncthreads();
        // brackets of every output, as the forward computes them
#pragma unroll
        for (int k = 0; k < PQ; ++k) {
            const int q = tl + k * NB;
            if (q >= M2) continue;
            if (q >= m) {
                code[q] = (m + 64) << 8;               // beyond every source window
                frq[q] = 0.f;
                continue;
            }
            const double Q = KC(q);
            int c = gix[k];
            double xl = c > 0 ? xs[c - 1] : -INFINITY;
            double xh = c < m ? xs[c] : INFINITY;
            while (c > 0 && !(xl <= Q)) { --c; xh = xl; xl = c > 0 ? xs[c - 1] : -INFINITY; }
            while (c < m && xh <= Q) { ++c; xl = xh; xh = c < m ? xs[c] : INFINITY; }
            int ph = 0;
            float fr = 0.f;
            if (c == m) {
                ph = (Q == xl) ? 1 : 0;
            } else if (c > 0) {
                const double f = div_bracket(__dsub_rn(Q, xl), __dsub_rn(xh, xl));
                ph = (int)fmin(rint(f * 32768.0), 32767.0) >> 7;
                fr = (float)f;
            }
            code[q] = (c << 8) | ph;
            frq[q] = fr;
        }
        __syncthreads();                               // positions dead: the FFT line takes over
        // first[]: output q >= 1 fills the counts in (c(q - 1), c(q)] (a few
        // each); the counts up to c(0) (output 0) and past c(M2 - 1) (none:
        // M2), which can be long runs, are filled by all threads together
        {
            const int c0 = min(code[0] >> 8, M2 + FO - 1), cl = min(code[M2 - 1] >> 8, M2 + FO - 1);
            for (int i = tl; i < M2 + 2 * FO; i += NB) {
                const int pp = i - FO;
                if (pp <= c0) first[i] = 0;
                else if (pp > cl) first[i] = M2;
            }
        }
#pragma unroll
        for (int k = 0; k < PQ; ++k) {
            const int q = tl + k * NB;
            if (q < 1 || q >= M2) continue;
            const int c = min(code[q] >> 8, M2 + FO - 1);
            const int cp = min(code[q - 1] >> 8, M2 + FO - 1);
            for (int pp = cp + 1; pp <= c; ++pp) first[pp + FO] = q;
        }
        const double hj = hn ? hn[j] : 1.0;
        pfa32::fft_src<L, !INVF, EM>(reg, tw, [&](int, int s2) { return bv[s2]; },
            [&](int q, float2 v) {
                if (q < m) {
                    float w = scale;
                    if (hm) w *= (float)__dmul_rn(hm[q], hj);
                    v = make_float2(v.x * w, v.y * w);
                    if ((q + j) & 1) v = make_float2(-v.x, -v.y);
                } else {
                    v = make_float2(0.f, 0.f);
                }
                reg[pfa32::padf(q)] = v;
            });
        // transpose gather: each source sums the outputs whose windows cover it
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int p0 = 4 * (g * NB + tl);
            if (p0 >= m) continue;
            // outputs with bracket count c in [p0 - 15, p0 + 19] (J in [p0 - 16, p0 + 18])
            const int qa = first[p0 - 15 + FO], qb = first[p0 + 20 + FO];
            float2 H[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) H[u] = make_float2(0.f, 0.f);
            // Outputs in batches of 4: every batch's code, value and fraction loads
            // are issued before any is used (a slot past qb reads as c = 0 and is
            // skipped), and the sums still run over q in increasing order.
            for (int q0 = qa; q0 < qb; q0 += 4) {
                int cwb[4];
                float2 gqb[4];
                float frb[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int q = min(q0 + i, qb - 1);
                    cwb[i] = q0 + i < qb ? code[q] : 0;
                    gqb[i] = reg[pfa32::padf(q)];
                    frb[i] = frq[q];
                }
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int c = cwb[i] >> 8, ph = cwb[i] & 255;
                    if (c == 0) continue;
                    const float2 gq = gqb[i];
                    if (c == m) {                      // np.interp's x == x_last: value of the last sample
                        if (ph) {
#pragma unroll
                            for (int u = 0; u < 4; ++u)
                                if (p0 + u == m - 1) H[u] = make_float2(H[u].x + gq.x, H[u].y + gq.y);
                        }
                        continue;
                    }
                    const int J = c - 1;
                    if (SINC && J >= 15 && J <= m - 17) {
                        const float *col = tapl + scr(ph);
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const float h = col[(p0 + u - J + 15 + 4) * 256];   // zero rows outside the taps
                            H[u] = make_float2(fmaf(h, gq.x, H[u].x), fmaf(h, gq.y, H[u].y));
                        }
                    } else {
                        const float fr = frb[i];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int pu = p0 + u;
                            const float w = pu == J ? 1.0f - fr : (pu == J + 1 ? fr : 0.f);
                            H[u] = make_float2(fmaf(w, gq.x, H[u].x), fmaf(w, gq.y, H[u].y));
                        }
                    }
                }
            }
            // dS/dphi contribution: Im(H conj(g1')), g1' = g1t e^{j phi}
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int p = p0 + u;
                if (p < m) {
                    const float2 g1 = pfa32::cmul(gv[g * 4 + u], rv[g * 4 + u]);
                    acc[g * 4 + u] += H[u].y * g1.x - H[u].x * g1.y;
                }
            }
        }
    }
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int p = 4 * (g * NB + t) + u;
            if (p < m) partial[(size_t)blockIdx.x * m + p] = acc[g * 4 + u];
        }
}

// grad[order[p]] = (2 / E2) sum_b partial[b][p]   (the 1/E2 undoes G's scaling)
// 32 sources per workgroup, their sums split over 8 slices of the partials
// (coalesced 128-byte rows, 8 independent chains), combined in LDS in a fixed
// order.
constexpr int GR_P = 32, GR_S = 8;
__global__ __launch_bounds__(GR_P * GR_S) void k_af_grad(const float *__restrict__ partial, int nparts,
                                                         int m, const int32_t *__restrict__ order,
                                                         const double *__restrict__ stats,
                                                         float *__restrict__ grad) {
    __shared__ double red[GR_S][GR_P];
    const int pl = threadIdx.x % GR_P, sl = threadIdx.x / GR_P;
    const int p = blockIdx.x * GR_P + pl;
    double s = 0.0;
    if (p < m)
        for (int b = sl; b < nparts; b += GR_S) s += partial[(size_t)b * m + p];
    red[sl][pl] = s;
    __syncthreads();
    if (sl == 0 && p < m) {
        double tot = 0.0;
#pragma unroll
        for (int k = 0; k < GR_S; ++k) tot += red[k][pl];
        grad[order[p]] = stats[0] > 0.0 ? (float)(2.0 * tot / stats[0]) : 0.f;
    }
}

}  // namespace

int launch_af_rot(pfa_plan *p, const float *phi, void *rot) {
    pfa_ctx *ctx = p->ctx;
    hipLaunchKernelGGL(k_af_rot, dim3((p->m + 255) / 256), dim3(256), 0, ctx->stream, phi, p->order,
                       p->m, (float2 *)rot);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

int launch_af_stats(pfa_plan *p, const void *img, double *partial, int nparts, double *stats) {
    pfa_ctx *ctx = p->ctx;
    hipLaunchKernelGGL(k_af_stats1, dim3(nparts), dim3(256), 0, ctx->stream, (const float2 *)img,
                       (size_t)p->m2 * p->n2, partial);
    hipLaunchKernelGGL(k_af_stats2, dim3(1), dim3(256), 0, ctx->stream, (const double *)partial,
                       nparts, stats);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

#define PFA_AF_SWITCH(L, MACRO)                                                                     \
    switch (L) {                                                                                    \
        MACRO(6) MACRO(7) MACRO(8) MACRO(9) MACRO(10) MACRO(11) MACRO(12)                          \
        default: return pfa_fail(ctx, PFA_ERR_DIMS, "autofocus: padded cross size must be 64..4096"); \
    }
#define PFA_AF_SWITCH_N(L, MACRO)                                                                   \
    switch (L) {                                                                                    \
        MACRO(6) MACRO(7) MACRO(8) MACRO(9) MACRO(10) MACRO(11) MACRO(12) MACRO(13)                \
        default: return pfa_fail(ctx, PFA_ERR_DIMS, "autofocus: padded range size must be 64..8192"); \
    }

int launch_af_fftb_adj(pfa_plan *p, const void *img, const double *stats, void *B) {
    pfa_ctx *ctx = p->ctx;
    int st = PFA_OK;
    const float2 *tw = twiddles_f32(ctx, p->n2, &st);
    if (!tw) return st;
    const bool inv = p->sgn > 0;
    const int L = log2_exact(p->n2);
#define LAUNCH(LL)                                                                                  \
    case LL:                                                                                        \
        if (inv) hipLaunchKernelGGL((k_af_fftb_adj<LL, true>), dim3(p->m2), dim3(af_fb_nb<LL>()), 0, \
                                    ctx->stream, (const float2 *)img, p->n, stats, (float2 *)B, tw); \
        else hipLaunchKernelGGL((k_af_fftb_adj<LL, false>), dim3(p->m2), dim3(af_fb_nb<LL>()), 0,    \
                                ctx->stream, (const float2 *)img, p->n, stats, (float2 *)B, tw);     \
        break;
    PFA_AF_SWITCH_N(L, LAUNCH)
#undef LAUNCH
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

int launch_af_az_adj(pfa_plan *p, const void *g1t, int ldp, const void *rot, const void *B,
                     float *partial, int *nparts) {
    pfa_ctx *ctx = p->ctx;
    int st = PFA_OK;
    const float2 *tw = twiddles_f32(ctx, p->m2, &st);
    if (!tw) return st;
    const bool sinc = p->flags & PFA_FLOAT_SINC32;
    const bool inv = p->sgn > 0;
    const int L = log2_exact(p->m2);
    const int grid = std::max(8, std::min((p->n + 7) / 8 * 8, ctx->num_cu / 8 * 8));
    *nparts = grid;
    const TanBins tb{p->tan_bin, p->tb_lo, p->tb_inv_w, p->tb_bins};
#define ARGS (const float2 *)g1t, ldp, p->m, p->n, grid_of(p->KR, p->kr_lin, p->n), p->tan_s64, tb, \
             grid_of(p->KC, p->kc_lin, p->m), p->hm, p->hn, p->sinc, (const float2 *)rot,           \
             (const float2 *)B, tw, partial
#define K(LL, S, I) hipLaunchKernelGGL((k_af_az_adj<LL, S, I>), dim3(grid), dim3(af_adj_nb<LL>()), 0, ctx->stream, ARGS)
#define LAUNCH(LL)                          \
    case LL:                                \
        if (sinc && inv) K(LL, true, true);  \
        else if (sinc) K(LL, true, false);   \
        else if (inv) K(LL, false, true);    \
        else K(LL, false, false);            \
        break;
    PFA_AF_SWITCH(L, LAUNCH)
#undef LAUNCH
#undef K
#undef ARGS
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

int launch_af_grad_reduce(pfa_plan *p, const float *partial, int nparts, const double *stats,
                          float *grad) {
    pfa_ctx *ctx = p->ctx;
    hipLaunchKernelGGL(k_af_grad, dim3((p->m + GR_P - 1) / GR_P), dim3(GR_P * GR_S), 0, ctx->stream,
                       partial, nparts, p->m, p->order, stats, grad);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}
