// pfa_fft_mfma.hpp -- N-point FFT (N = 2^L, L in {8, 9, 10, 12, 13}) as dense
// DFT-matrix stages on the gfx950 matrix cores, 16-bit operands (BF16 or FP16)
// with FP32 accumulation.
//
// A radix-R stage (R = 16 or 32) of a Stockham FFT is, for each group of 16
// vectors, the real GEMM  [Yr; Yi] = M X  with the 2R x 2R real embedding of
// the R-point DFT; the K index of X interleaves the components,
// X = [xr_0, xi_0, xr_1, xi_1, ...], so a lane's 8 K-values are 4 complex
// samples stored side by side:
//     Yr[m] = sum_r  c xr_r + s' xi_r,    Yi[m] = sum_r  -s' xr_r + c xi_r,
//     c = cos(2 pi m r / R),  s' = sin(2 pi m r / R) (forward) or -sin (inverse),
// run as v_mfma_f32_16x16x32_{bf16,f16}: A = 16 rows of M, B = 32 K-values of
// 16 vectors (one 16-byte LDS read per lane), C = 16 x 16 FP32 outputs.  Lane
// layouts (checked on the device):
//     A[l % 16][8 (l / 16) + i],  B[8 (l / 16) + i][l % 16],  C[4 (l / 16) + i][l % 16].
//
// Data between stages lives in LDS as 32-bit words of (re, im) 16-bit pairs,
// stored vector-major for the stage that reads it: element e = k + r N/R of a
// radix-R stage sits at word k R + r, padded (pp) so the 16-byte operand reads
// and the scattered result writes of a wave use distinct banks.
// The twiddle of the next stage is applied in FP32 to the accumulators before
// they are rounded and stored (one 32-bit LDS write per complex output).
//
// Block scaling: every stage multiplies by 1/R (exact), so the transform is
// DFT/N; the caller normalises each line to max |component| <= 1 with a
// power-of-two exponent s and restores 2^(s + L) (forward) or 2^s (inverse,
// numpy's ifft already divides by N).  Magnitudes stay below 2^1.5 inside.
// SPLIT: the DFT matrix as hi + lo 16-bit parts (two products per tile), so
// its rounding error drops from 2^-9 (BF16) to below 2^-16.
#pragma once

#include <hip/hip_runtime.h>

#include "pfa_fft_f32.hpp"

namespace pfam {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

template <bool BF>
struct H16 {
    using T = __bf16;
    using V = bf16x8;
};
template <>
struct H16<false> {
    using T = _Float16;
    using V = f16x8;
};

__device__ __forceinline__ f32x4 mma(const bf16x8 &a, const bf16x8 &b, const f32x4 &c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mma(const f16x8 &a, const f16x8 &b, const f32x4 &c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// 16-bit storage of a float (round to nearest even) and back; a complex value
// as one 32-bit word (re in the low half).
template <bool BF>
__device__ __forceinline__ unsigned short to16(float x) {
    const typename H16<BF>::T h = (typename H16<BF>::T)x;
    return __builtin_bit_cast(unsigned short, h);
}
template <bool BF>
__device__ __forceinline__ float from16(unsigned short b) {
    return (float)__builtin_bit_cast(typename H16<BF>::T, b);
}
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
// both halves in one conversion (v_cvt_pk_bf16_f32 for BF16), round to nearest even
template <bool BF>
__device__ __forceinline__ unsigned pack(float2 v) {
    const f32x2 x = {v.x, v.y};
    if constexpr (BF) return __builtin_bit_cast(unsigned, __builtin_convertvector(x, bf16x2));
    else return __builtin_bit_cast(unsigned, __builtin_convertvector(x, f16x2));
}
template <bool BF>
__device__ __forceinline__ float2 unpack(unsigned w) {
    return make_float2(from16<BF>((unsigned short)(w & 0xffffu)), from16<BF>((unsigned short)(w >> 16)));
}

// Radix plan: stage radices (first stage first).
// LS0: the product of the radices the caller has already applied in
// registers before it puts the line (1: none; 2: a radix-2 first pass,
// see put_r2).
template <int L>
struct MPlan;
template <> struct MPlan<8>  { static constexpr int NS = 2, LS0 = 1; static constexpr int R[3] = {16, 16, 1}; };
template <> struct MPlan<9>  { static constexpr int NS = 2, LS0 = 1; static constexpr int R[3] = {32, 16, 1}; };
template <> struct MPlan<10> { static constexpr int NS = 2, LS0 = 1; static constexpr int R[3] = {32, 32, 1}; };
template <> struct MPlan<12> { static constexpr int NS = 3, LS0 = 1; static constexpr int R[3] = {16, 16, 16}; };
template <> struct MPlan<13> { static constexpr int NS = 3, LS0 = 2; static constexpr int R[3] = {16, 16, 16}; };

template <int L>
constexpr bool supported() { return L == 8 || L == 9 || L == 10 || L == 12 || L == 13; }

// Waves that share NCOL lines: at most the block's, at most the 16-vector
// groups of the stage with the fewest.
template <int L, int NCOL, int BLOCK_WAVES>
constexpr int nw_for() {
    int g = 1 << 30;
    for (int s = 0; s < MPlan<L>::NS; ++s) {
        const int gs = NCOL * ((1 << L) / MPlan<L>::R[s]) / 16;
        g = gs < g ? gs : g;
    }
    return BLOCK_WAVES < g ? BLOCK_WAVES : g;
}

template <int L, int S>
constexpr int ls_of() {                      // product of the radices before stage S
    int p = MPlan<L>::LS0;
    for (int s = 0; s < S; ++s) p *= MPlan<L>::R[s];
    return p;
}

// Padded word of position p in a line; line length in words incl. padding
// (a multiple of 4 words, so every line starts 16-byte aligned).  4 words every
// 32 spread the 16-byte operand reads of a wave over the banks; 4 more every
// 512 spread the scattered 32-bit result writes, whose lanes land 512 words
// apart (same bank) under the first rule alone.
__device__ __forceinline__ int pp(int p) { return p + ((p >> 5) << 2) + ((p >> 9) << 2); }
template <int L>
constexpr int line_words() { return ((1 << L) + ((1 << L) >> 3) + (((1 << L) >> 9) << 2) + 3) & ~3; }

// Position of element e in the plane layout read by stage S.
template <int L, int S>
__device__ __forceinline__ constexpr unsigned pos_of(unsigned e) {
    constexpr unsigned R = MPlan<L>::R[S], NV = (1u << L) / R;
    return (e % NV) * R + e / NV;
}

// cos / sin of 2 pi t / 32 for t in [0, 32), from the 16-entry table of the FP32 FFT.
__device__ __forceinline__ float cos32(int t) {
    t &= 31;
    return t < 16 ? pfa32::kC32[t][0] : -pfa32::kC32[t - 16][0];
}
__device__ __forceinline__ float sin32(int t) {
    t &= 31;
    return t < 16 ? pfa32::kC32[t][1] : -pfa32::kC32[t - 16][1];
}

// A operand of tile (mt, ks) of the radix-R real DFT matrix for this lane:
// hi part and (SPLIT) the lo remainder.
template <int R, bool BF, bool INV>
__device__ __forceinline__ void dft_operand(int mt, int ks, int l, typename H16<BF>::V &hi,
                                            typename H16<BF>::V &lo) {
    const int row = 16 * mt + (l & 15);
    const int br = row / R, mm = row % R;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int col = 32 * ks + 8 * (l >> 4) + i;
        const int bc = col & 1, rr = col >> 1;          // interleaved (re, im) K order
        const int t = (mm * rr) % R * (32 / R);
        const float c = cos32(t), s = sin32(t);
        float v;
        if (br == bc) v = c;
        else if (br == 0) v = INV ? -s : s;            // (0, 1)
        else v = INV ? s : -s;                         // (1, 0)
        // the stage's 1/R scale, folded in: a power of two, so the 16-bit
        // entries and every FP32 product and sum are exactly 1/R of the
        // unscaled ones (all entries stay far above the 16-bit normal range)
        v *= 1.0f / (float)R;
        const typename H16<BF>::T h = (typename H16<BF>::T)v;
        hi[i] = h;
        lo[i] = (typename H16<BF>::T)(v - (float)h);
    }
}

// Tiles (mt, ks) of the DFT matrices of all stages, stage S's first tile.
template <int L>
constexpr int ntiles() {
    int t = 0;
    for (int s = 0; s < MPlan<L>::NS; ++s) t += (2 * MPlan<L>::R[s] / 16) * (2 * MPlan<L>::R[s] / 32);
    return t;
}
template <int L, int S>
constexpr int tile0() {
    int t = 0;
    for (int s = 0; s < S; ++s) t += (2 * MPlan<L>::R[s] / 16) * (2 * MPlan<L>::R[s] / 32);
    return t;
}

// DFT-matrix operands from a table in global memory, built once per size,
// precision and direction (build_ops_stage), read through the caches: 16 bytes
// per lane and tile, tab[t * 64 + lane] = hi of tile t, tab[(ntiles + t) * 64 +
// lane] = lo.
template <int L, bool BF, bool INV>
struct OpsGlobal {
    const u16x8 *__restrict__ tab;
    template <int S>
    __device__ __forceinline__ void get(int mt, int ks, int lane, typename H16<BF>::V &hi,
                                        typename H16<BF>::V &lo) const {
        constexpr int KS = 2 * MPlan<L>::R[S] / 32;
        const int t = tile0<L, S>() + mt * KS + ks;
        hi = __builtin_bit_cast(typename H16<BF>::V, tab[t * 64 + lane]);
        lo = __builtin_bit_cast(typename H16<BF>::V, tab[(ntiles<L>() + t) * 64 + lane]);
    }
};

template <int L, bool BF, bool INV, int S>
__device__ __forceinline__ void build_ops_stage(u16x8 *tab) {
    if constexpr (S < MPlan<L>::NS) {
        constexpr int R = MPlan<L>::R[S], MT = 2 * R / 16, KS = 2 * R / 32;
        for (int i = threadIdx.x; i < MT * KS * 64; i += blockDim.x) {
            const int tt = i / 64, lane = i % 64;
            typename H16<BF>::V hi, lo;
            dft_operand<R, BF, INV>(tt / KS, tt % KS, lane, hi, lo);
            tab[(tile0<L, S>() + tt) * 64 + lane] = __builtin_bit_cast(u16x8, hi);
            tab[(ntiles<L>() + tile0<L, S>() + tt) * 64 + lane] = __builtin_bit_cast(u16x8, lo);
        }
        build_ops_stage<L, BF, INV, S + 1>(tab);
    }
}

// One stage over NCOL lines held in `planes` (line_words<L>() words per line).  Stage S reads its layout, applies the
// DFT (its 1/R scale folded into the operands), then either (not last) multiplies by the next stage's
// twiddle and writes the next layout, or (last) hands (line, index, value) to
// sink.  NW waves share the work; results wait in registers for the barrier.
template <int L, bool BF, bool INV, bool SPLIT, int NCOL, int NW, int S, typename Ops, typename Sink>
__device__ __forceinline__ void stage(unsigned *planes, const pfa32::Tw &tw, const Ops &ops,
                                      Sink &sink) {
    constexpr int N = 1 << L;
    constexpr int R = MPlan<L>::R[S];
    constexpr int NV = N / R;
    constexpr int LS = ls_of<L, S>();
    constexpr bool LAST = S == MPlan<L>::NS - 1;
    constexpr int MT = 2 * R / 16, KS = 2 * R / 32;
    constexpr int GROUPS = NCOL * NV / 16;             // 16-vector groups over all lines
    static_assert(NW >= 1 && GROUPS % NW == 0, "groups split evenly over the waves");
    constexpr int GPW = GROUPS / NW;
    constexpr int PL = line_words<L>();
    using V = typename H16<BF>::V;
    int lane = threadIdx.x & 63;
    asm volatile("" : "+v"(lane));
    const int w = threadIdx.x >> 6;
    const int n = lane & 15, q = lane >> 4;
    V Ah[MT][KS], Al[MT][KS];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) ops.template get<S>(mt, ks, lane, Ah[mt][ks], Al[mt][ks]);
    f32x4 acc[GPW][MT];
    const bool active = w < NW;                        // waves beyond NW only keep the barriers
#pragma unroll
    for (int g = 0; g < GPW; ++g) {
        if (!active) break;
        const int G = w + g * NW;                      // global group
        const int col = G / (NV / 16), k = (G % (NV / 16)) * 16 + n;
        const unsigned *pl = planes + col * PL;
        V b[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {              // K rows 32 ks + 8 q .. +7 = 4 complex samples
            const u32x4 raw = *reinterpret_cast<const u32x4 *>(pl + pp(k * R + 16 * ks + 4 * q));
            b[ks] = __builtin_bit_cast(V, raw);
        }
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            f32x4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                c = mma(Ah[mt][ks], b[ks], c);
                if (SPLIT) c = mma(Al[mt][ks], b[ks], c);
            }
            acc[g][mt] = c;
        }
    }
    if (!LAST) __syncthreads();                        // every operand read before any write
#pragma unroll
    for (int g = 0; g < GPW; ++g) {
        if (!active) break;
        const unsigned G = w + g * NW;
        const unsigned col = G / (NV / 16), k = (G % (NV / 16)) * 16 + n;
        const unsigned j = k % LS;
        if constexpr (LAST) {
            const unsigned obase = (k / LS) * LS * R + j;
#pragma unroll
            for (int mt = 0; mt < MT / 2; ++mt)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const unsigned m = 16 * mt + 4 * q + i;    // output radix index
                    sink((int)col, (int)(obase + m * LS),
                         make_float2(acc[g][mt][i], acc[g][mt + MT / 2][i]));   // 1/R is in the operands
                }
        } else {
            // Output m = 16 mt + 4 q + i sits at o = (k / LS) LS R + j + m LS.  The
            // next stage (radix R2, NV2 = N / R2 vectors, LS2 = LS R) reads it as
            // vector k2 = o % NV2, radix index r2 = o / NV2, twiddle exponent
            // (k2 % LS2) r2 N / (LS2 R2).  NV2 is a multiple of LS2 and m LS < LS2,
            // so r2 is the lane's own and, over (mt, i), the word position
            // k2 R2 + r2 and the exponent step by compile-time multiples:
            //     position  = p0 + (16 mt + i) LS R2,  pp(p0 + d) = pp(p0) + pp(d)
            //     exponent  = T0 + (16 mt + i) D,      D = LS r2 N / (LS2 R2)
            // (no carry crosses the pad boundaries of pp; both identities checked
            // exhaustively for every size and stage).  The twiddle of exponent 0
            // is exactly (1, 0), so it is applied unconditionally.
            constexpr int R2 = MPlan<L>::R[S + 1];
            constexpr unsigned NV2 = N / R2, LS2 = LS * R, KB = NV2 / LS2, C = N / (LS2 * R2);
            static_assert(NV2 % LS2 == 0, "the next stage's vectors hold whole twiddle classes");
            const unsigned h = k / LS, r2 = h / KB;
            const unsigned kq = j + 4 * q * LS;
            unsigned *dst = planes + col * PL + pp(((h % KB) * LS2 + kq) * R2 + r2);
            const unsigned T0 = kq * r2 * C, D = LS * r2 * C;
            // W^(T0 + (16 mt + i) D) as products of W^T0, W^D and (R = 32) W^(16 D),
            // at most four deep: a few FP32 ulp, far below the 16-bit rounding
            // that follows.  16 D < N, since every exponent above is.
            float2 w0 = tw((int)T0), wd = tw((int)D), w16 = make_float2(1.f, 0.f);
            if (MT / 2 > 1) w16 = tw((int)(16 * D));
            if (INV) {
                w0.y = -w0.y;
                wd.y = -wd.y;
                w16.y = -w16.y;
            }
#pragma unroll
            for (int mt = 0; mt < MT / 2; ++mt) {
                float2 wv = mt == 0 ? w0 : pfa32::cmul(w0, w16);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (i) wv = pfa32::cmul(wv, wd);
                    const float2 y = pfa32::cmul(make_float2(acc[g][mt][i], acc[g][mt + MT / 2][i]), wv);
                    dst[pp((16 * mt + i) * LS * R2)] = pack<BF>(y);
                }
            }
        }
    }
    if (!LAST) __syncthreads();
}

template <int L, bool BF, bool INV, bool SPLIT, int NCOL, int NW, int S, typename Ops, typename Sink>
__device__ __forceinline__ void stages_from(unsigned *planes, const pfa32::Tw &tw,
                                            const Ops &ops, Sink &sink) {
    if constexpr (S < MPlan<L>::NS) {
        stage<L, BF, INV, SPLIT, NCOL, NW, S>(planes, tw, ops, sink);
        stages_from<L, BF, INV, SPLIT, NCOL, NW, S + 1>(planes, tw, ops, sink);
    }
}

// Stores of a lane's elements e0 + d of line col into the first-stage layout
// (values already scaled), as a lane base and compile-time offsets:
//     pp(pos(e0 + d)) = pp(pos(e0)) + pp(pos(d))
// holds for the two patterns the kernels use -- e0 = 4 t, d = 4 g NB + u
// (u < 4) in the azimuth pass and e0 = t, d = i NB in FFT-B, NB the block's
// threads -- checked exhaustively for every size.
template <int L>
__device__ __forceinline__ unsigned *put_base(unsigned *planes, int col, unsigned e0) {
    return planes + col * line_words<L>() + pp((int)pos_of<L, 0>(e0));
}
template <int L, bool BF>
__device__ __forceinline__ void put_at(unsigned *base, unsigned d, float2 v) {
    base[pp((int)pos_of<L, 0>(d))] = pack<BF>(v);
}
// Plans with LS0 = 2: the radix-2 first pass of points a = x[k], b = x[k + N/2]
// (k < N/2), run by the caller in FP32 and put as the first stage's input:
// outputs (a + b) / 2 and (a - b) / 2 (the 1/2 keeps the transform at DFT/N)
// at o = 2k and 2k + 1, the odd one times the first stage's twiddle
// W_N^(r N / 32) = W_32^r, r = o / (N / 16) (j = o % 2 = 1 of LS = 2, step N / 32).
template <int L, bool BF, bool INV>
__device__ __forceinline__ void put_r2(unsigned *line, unsigned k, float2 a, float2 b) {
    static_assert(MPlan<L>::LS0 == 2 && MPlan<L>::R[0] == 16, "a radix-2 first pass ahead of radix 16");
    constexpr unsigned NV = (1u << L) / 16;
    const float2 y0 = make_float2((a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f);
    float2 y1 = make_float2((a.x - b.x) * 0.5f, (a.y - b.y) * 0.5f);
    const unsigned o1 = 2 * k + 1;
    const int r = (int)(o1 / NV);
    const float c = pfa32::kC32[r][0], sn = INV ? pfa32::kC32[r][1] : -pfa32::kC32[r][1];
    y1 = make_float2(fmaf(y1.x, c, -y1.y * sn), fmaf(y1.x, sn, y1.y * c));
    line[pp((int)pos_of<L, 0>(2 * k))] = pack<BF>(y0);
    line[pp((int)pos_of<L, 0>(o1))] = pack<BF>(y1);
}

// Transform NCOL lines whose first-stage layout is in `planes` (after a
// barrier); sink(col, index, value) receives DFT/N (or IDFT/N) in FP32.
// The last stage has no barriers: the sink must not write `planes`, and the
// caller synchronises before it writes `planes` again.
template <int L, bool BF, bool INV, bool SPLIT, int NCOL, int NW, typename Ops, typename Sink>
__device__ __forceinline__ void fft(unsigned *planes, const pfa32::Tw &tw, const Ops &ops,
                                    Sink &&sink) {
    static_assert(supported<L>(), "MFMA FFT sizes: 256, 512, 1024, 4096, 8192");
    stages_from<L, BF, INV, SPLIT, NCOL, NW, 0>(planes, tw, ops, sink);
}

// Power-of-two exponent s with max_component * 2^-s <= 1 (0 for an all-zero line).
__device__ __forceinline__ int norm_exp(float mx) {
    if (!(mx > 0.f)) return 0;
    int e;
    frexpf(mx, &e);                                    // mx = f 2^e, f in [0.5, 1)
    return e;
}

}  // namespace pfam
