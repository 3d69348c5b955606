// pfa_mfma.hip -- the precision form of the float focus: both 1-D FFT passes as
// dense DFT-matrix stages on the gfx950 matrix cores (BF16 or FP16 operands,
// FP32 accumulation, per-line power-of-two block scaling, optional hi/lo DFT
// matrices; pfa_fft_mfma.hpp), with the intermediate A stored as packed
// 16-bit complex pairs plus one exponent per range row.  The resample passes
// are the FP32 kernels of pfa_float_dev.hpp.
//
// Kernels
//   k_azimuth_f32<.., FK>  azimuth pass + window + sign fold (FP32), then the
//                          FFT along cross on the matrix cores -> A16 (16-bin blocks), expA[j]
//   k_fft_b_mfma           FFT along range for FBM_NCOL = 2 adjacent cross bins per
//                          workgroup (8-byte loads of A16 rows = the corner-turn),
//                          output sign fold (complex output) or detection in the store
#include "pfa_float_dev.hpp"

#include <climits>

namespace {

constexpr int FBM_NCOL = 2;                 // cross bins (lines) per FFT-B workgroup
template <int L>
constexpr int fftb_mfma_nb() { return 64 * pfam::nw_for<L, FBM_NCOL, 8>(); }

// DFT-matrix operand table of an L-point matrix-core FFT (pfam::OpsGlobal layout).
template <int L, bool BF, bool INV>
__global__ __launch_bounds__(256) void k_build_ops(pfam::u16x8 *tab) {
    pfam::build_ops_stage<L, BF, INV, 0>(tab);
}

// OUT 0: complex64 image, 1: |image| float32; image [cross r][range c].
template <int L, bool INV, int OUT, int FK>
__global__ __launch_bounds__(fftb_mfma_nb<L>()) void k_fft_b_mfma(
    const unsigned *__restrict__ A16, const int *__restrict__ expA, int n, int m2,
    void *__restrict__ outp, const float2 *__restrict__ twg, const pfam::u16x8 *__restrict__ ops) {
    constexpr bool BF = FK == FK_BF16 || FK == FK_BF16S;
    constexpr bool SPLIT = FK == FK_BF16S;
    constexpr int N2 = 1 << L, NCOL = FBM_NCOL, NB = fftb_mfma_nb<L>(), BW = NB / 64;
    constexpr int NW = pfam::nw_for<L, NCOL, BW>();
    constexpr int PL = pfam::line_words<L>();
    __shared__ unsigned planes[NCOL * PL];
    __shared__ float2 twl[N2 / 64 + 64];
    __shared__ int emx[BW];
    const pfa32::Tw tw = pfa32::load_tw<N2>(twl, twg);
    const int t = threadIdx.x;
    const int r0 = xcd_map(blockIdx.x, m2 / NCOL) * NCOL;
    // the corner-turn: NCOL adjacent cross bins of every range row.  A16 is in
    // 16-bin blocks, so consecutive rows of a block are 64 bytes apart and
    // consecutive lanes take consecutive rows: a wave's 8-byte loads (NCOL = 2)
    // span 64 rows x 64 bytes, 32 cache lines with two rows' runs in each.  All
    // loads are issued first (clamped rows); the row exponents loaded with them
    // also give the block's common exponent.
    constexpr int NI = N2 / NB;
    static_assert(N2 % NB == 0, "row assignment covers the line");
    const int nj = a_rows(n);
    auto row_of = [&](int i) { return t + i * NB; };
    unsigned a[NI][NCOL];
    int ex[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const int jc = min(row_of(i), n - 1);
        const unsigned *src = A16 + a16blk(jc, r0, nj);
        if constexpr (NCOL == 4) {
            const uint4 v = *reinterpret_cast<const uint4 *>(src);
            a[i][0] = v.x; a[i][1] = v.y; a[i][2] = v.z; a[i][3] = v.w;
        } else {
            const uint2 v = *reinterpret_cast<const uint2 *>(src);
            a[i][0] = v.x; a[i][1] = v.y;
        }
        ex[i] = expA[jc];
    }
    // common exponent: |A16| < 2^2, so 2^-(max expA + 2) bounds every input by 1
    int e = INT_MIN;
#pragma unroll
    for (int i = 0; i < NI; ++i) e = max(e, ex[i]);  // clamped rows repeat row n - 1
    for (int off = 32; off >= 1; off >>= 1) e = max(e, __shfl_xor(e, off, 64));
    if ((t & 63) == 0) emx[t >> 6] = e;
    __syncthreads();
    int S = INT_MIN;
#pragma unroll
    for (int i = 0; i < BW; ++i) S = max(S, emx[i]);
    S = (S == INT_MIN) ? 0 : S + 2;
    auto value = [&](int i, int c) {
        const float sc = ldexpf(1.0f, ex[i] - S);
        const float2 v = pfam::unpack<BF>(a[i][c]);
        return row_of(i) < n ? make_float2(v.x * sc, v.y * sc) : make_float2(0.f, 0.f);
    };
    if constexpr (pfam::MPlan<L>::LS0 == 2) {
        // the radix-2 first pass in registers: rows i and i + NI/2 are points
        // k and k + N/2 (k = t + i NB)
#pragma unroll
        for (int i = 0; i < NI / 2; ++i)
#pragma unroll
            for (int c = 0; c < NCOL; ++c)
                pfam::put_r2<L, BF, INV>(planes + c * PL, row_of(i), value(i, c), value(i + NI / 2, c));
    } else {
        unsigned *pb = pfam::put_base<L>(planes, 0, t);
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int c = 0; c < NCOL; ++c) pfam::put_at<L, BF>(pb + c * PL, i * NB, value(i, c));
    }
    __syncthreads();
    const float osc = ldexpf(1.0f, S + (INV ? 0 : L));
    pfam::fft<L, BF, INV, SPLIT, NCOL, NW>(planes, tw, pfam::OpsGlobal<L, BF, INV>{ops},
        [&](int col, int c, float2 y) {
            const int r = r0 + col;
            float2 v = make_float2(y.x * osc, y.y * osc);
            if (OUT == 0) {
                if ((r + c) & 1) v = make_float2(-v.x, -v.y);
                reinterpret_cast<float2 *>(outp)[(size_t)r * N2 + c] = v;
            } else {                                   // |(-1)^(r+c) v| = |v|, bit for bit
                reinterpret_cast<float *>(outp)[(size_t)r * N2 + c] = __fsqrt_rn(v.x * v.x + v.y * v.y);
            }
        });
}

}  // namespace

#define PFA_MFMA_SWITCH(L, MACRO)                                                                   \
    switch (L) {                                                                                  \
        MACRO(8) MACRO(9) MACRO(10) MACRO(12) MACRO(13)                                           \
        default: return pfa_fail(ctx, PFA_ERR_DIMS, "matrix-core FFT sizes are 256, 512, 1024, 4096 and 8192"); \
    }

// The operand table of (L, precision, direction), built on first use and kept
// in the context's workspace.
static const pfam::u16x8 *dft_ops(pfa_ctx *ctx, int L, bool bf, bool inv, int *st) {
    char slot[48];
    snprintf(slot, sizeof slot, "dftops_%d_%d_%d", L, (int)bf, (int)inv);
    auto it = ctx->ws.find(slot);
    if (it != ctx->ws.end() && it->second.bytes) return (const pfam::u16x8 *)it->second.ptr;
    int nt = 0;
    switch (L) {
        case 8: nt = pfam::ntiles<8>(); break;
        case 9: nt = pfam::ntiles<9>(); break;
        case 10: nt = pfam::ntiles<10>(); break;
        case 12: nt = pfam::ntiles<12>(); break;
        case 13: nt = pfam::ntiles<13>(); break;
        default: *st = pfa_fail(ctx, PFA_ERR_DIMS, "matrix-core FFT sizes are 256, 512, 1024, 4096 and 8192"); return nullptr;
    }
    auto *tab = (pfam::u16x8 *)pfa_ws(ctx, slot, (size_t)2 * nt * 64 * sizeof(pfam::u16x8), st);
    if (!tab) return nullptr;
#define B(LL)                                                                                              case LL:                                                                                                   if (bf && inv) hipLaunchKernelGGL((k_build_ops<LL, true, true>), dim3(1), dim3(256), 0, ctx->stream, tab);          else if (bf) hipLaunchKernelGGL((k_build_ops<LL, true, false>), dim3(1), dim3(256), 0, ctx->stream, tab);           else if (inv) hipLaunchKernelGGL((k_build_ops<LL, false, true>), dim3(1), dim3(256), 0, ctx->stream, tab);          else hipLaunchKernelGGL((k_build_ops<LL, false, false>), dim3(1), dim3(256), 0, ctx->stream, tab);                  break;
    switch (L) { B(8) B(9) B(10) B(12) B(13) default: break; }
#undef B
    if (hipGetLastError() != hipSuccess) {
        *st = pfa_fail(ctx, PFA_ERR_HIP, "DFT operand table build failed");
        return nullptr;
    }
    return tab;
}

int launch_float_azimuth_mfma(pfa_plan *p, const void *g1t, int ldp, unsigned *A16, int *expA,
                              int fk) {
    pfa_ctx *ctx = p->ctx;
    int st = PFA_OK;
    const int m2 = p->m2;
    const float2 *tw = twiddles_f32(ctx, m2, &st);
    if (!tw) return st;
    const bool sinc = p->flags & PFA_FLOAT_SINC32;
    const bool inv = p->sgn > 0;
    const int L = log2_exact(m2);
    const pfam::u16x8 *ops = dft_ops(ctx, L, fk != FK_FP16, inv, &st);
    if (!ops) return st;
    const int grid_az = std::max(8, std::min((p->n + 7) / 8 * 8, ctx->num_cu / 8 * 8));
    const TanBins tb{p->tan_bin, p->tb_lo, p->tb_inv_w, p->tb_bins};
    pfa_timer_begin(ctx, PFA_TIMER_AZIMUTH);
#define ARGS (const float2 *)g1t, ldp, p->m, p->n, grid_of(p->KR, p->kr_lin, p->n), p->tan_s64, tb, \
             grid_of(p->KC, p->kc_lin, p->m), p->hm, p->hn, p->sinc, nullptr, nullptr, tw,          \
             (unsigned *)(ctx->d_counters + 32), A16, expA, nullptr, ops, nullptr
#define K(LL, S, I, F) hipLaunchKernelGGL((k_azimuth_f32<LL, 1, S, I, F>), dim3(grid_az), dim3(az_nb<LL>()), 0, ctx->stream, ARGS)
#define KF(LL, F)                                 \
    if (sinc && inv) K(LL, true, true, F);        \
    else if (sinc) K(LL, true, false, F);         \
    else if (inv) K(LL, false, true, F);          \
    else K(LL, false, false, F);
#define LAUNCH(LL)                                 \
    case LL:                                       \
        if (fk == FK_BF16) { KF(LL, FK_BF16) }     \
        else if (fk == FK_FP16) { KF(LL, FK_FP16) } \
        else { KF(LL, FK_BF16S) }                  \
        break;
    PFA_MFMA_SWITCH(L, LAUNCH)
#undef LAUNCH
#undef KF
#undef K
#undef ARGS
    pfa_timer_end(ctx, PFA_TIMER_AZIMUTH);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

int launch_float_fft_b_mfma(pfa_plan *p, const unsigned *A16, const int *expA, void *out,
                            int out_kind, int fk) {
    pfa_ctx *ctx = p->ctx;
    int st = PFA_OK;
    const int n2 = p->n2;
    const float2 *tw = twiddles_f32(ctx, n2, &st);
    if (!tw) return st;
    const bool inv = p->sgn > 0;
    const int L = log2_exact(n2);
    const pfam::u16x8 *ops = dft_ops(ctx, L, fk != FK_FP16, inv, &st);
    if (!ops) return st;
    const dim3 grid(p->m2 / FBM_NCOL);
    pfa_timer_begin(ctx, PFA_TIMER_FFT_B);
#define ARGS A16, expA, p->n, p->m2, out, tw, ops
#define K(LL, I, O, F) hipLaunchKernelGGL((k_fft_b_mfma<LL, I, O, F>), grid, dim3(fftb_mfma_nb<LL>()), 0, ctx->stream, ARGS)
#define KF(LL, F)                              \
    if (inv && out_kind) K(LL, true, 1, F);    \
    else if (inv) K(LL, true, 0, F);           \
    else if (out_kind) K(LL, false, 1, F);     \
    else K(LL, false, 0, F);
#define LAUNCH(LL)                                 \
    case LL:                                       \
        if (fk == FK_BF16) { KF(LL, FK_BF16) }     \
        else if (fk == FK_FP16) { KF(LL, FK_FP16) } \
        else { KF(LL, FK_BF16S) }                  \
        break;
    PFA_MFMA_SWITCH(L, LAUNCH)
#undef LAUNCH
#undef KF
#undef K
#undef ARGS
    pfa_timer_end(ctx, PFA_TIMER_FFT_B);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}
