// pfa_fixed.hip -- gfx950 kernels of the bit-exact fixed-point focus path.
//
// Arithmetic contract: mpfs/host/silicon_emulator.py (form_image :216-253 and
// its stage functions), src/fixedpoint.py:266-289, serialize_inputs.py:67-94,
// gen_resample_vectors.py:257-296.  Complex int16 data is short2 (x = I, y = Q).
//
// Kernels
//   k_range_resample    pass 1 (F3+F4+F5): FP64 positions, bracket, 2-tap lerp or
//                       32-tap sinc, pulse-sorted rows written corner-turned
//                       from registers: g1t[range q][sorted pulse p].
//   k_azimuth_fft       pass 2 (F6+F7) + window (F8) + BFP FFT along cross (F9a),
//                       one range row per workgroup, everything LDS-resident.
//   k_fft_b             BFP FFT along range (F9b) for one cross column per
//                       workgroup, pass-A renorm folded into its column load.
//   k_detect_t          pass-B renorm + detect (F11) + LDS-tiled transpose back
//                       to [range][cross].
//   k_fft_rows          stand-alone per-row BFP FFT (fft_pass_bfp / hw_perrow).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "pfa_div.hpp"
#include "pfa_internal.hpp"

namespace {

constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v >> 1); }

__device__ __forceinline__ int sat16(int v) { return min(max(v, -32768), 32767); }
__device__ __forceinline__ int wrap16(int v) { return (int)(short)(v & 0xffff); }
__device__ __forceinline__ int pad32(int i) { return i + (i >> 5); }

// ------------------------------------------------------------------ BFP FFT core
// fixedpoint.fft1d_bfp_hw_perrow for one line of N = 2^L, 16-bit data and
// 16-bit truncated twiddles.  NT threads hold E = N/NT elements each; stages
// whose butterfly partner bit is among a thread's E local index bits run in
// registers, and the local bit window moves through LDS between them.  Every
// stage ends with the line-wide max that sets the shift (0/1/2) the reference
// applies before the next stage.
// Wave-wide max on DPP row operations: quad swaps and row rotations leave each
// 16-lane row's max in all its lanes, two row broadcasts carry it to lane 63.
__device__ __forceinline__ int wave_max_dpp(int v) {
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x124, 0xF, 0xF, false));  // row_ror:4
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x128, 0xF, 0xF, false));  // row_ror:8
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x142, 0xA, 0xF, false));  // row_bcast:15
    v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return v;                                          // the wave's max in lane 63
}

template <int L>
struct Fft {
    static constexpr int N = 1 << L;
    // threads per line: 16 elements each from 1024 points up (512 threads at
    // 8192, so a CU holds 16 waves of two lines), two each below
    static constexpr int NT = N >= 1024 ? (N / 16 < 512 ? N / 16 : 512) : (N / 2 < 256 ? N / 2 : 256);
    static constexpr int E = N / NT;
    static constexpr int P = ilog2c(E);
    static constexpr int NB = NT < 64 ? 64 : NT;
    static constexpr int NW = NB / 64;
    static constexpr int DATA = N + N / 32;          // padded short2 slots

    static __device__ __forceinline__ int idx(int t, int l, int b) {
        const int lo = t & ((1 << b) - 1);
        const int hi = t >> b;
        return lo | (l << b) | (hi << (b + P));
    }
    static constexpr int base_of(int s) {                     // layout of stage s
        const int b = ((s - 1) / P) * P;
        return b < L - P ? b : L - P;
    }

    // data: padded natural-order input (read), reused as exchange buffer.
    // On return xr/xi hold the output at idx(t, l, L - P); e = summed shifts.
    static __device__ __forceinline__ void run(short2 *data, const short2 *tw, int *red,
                                               int (&xr)[E], int (&xi)[E], int &e) {
        const int t = threadIdx.x;
        const bool act = t < NT;
        if (act) {
#pragma unroll
            for (int l = 0; l < E; ++l) {
                const int i = idx(t, l, 0);
                const int src = (int)(__brev((unsigned)i) >> (32 - L));
                const short2 v = data[pad32(src)];
                xr[l] = v.x;
                xi[l] = v.y;
            }
        }
        e = 0;
        stages<1>(data, tw, red, xr, xi, e, t, act);
    }

    // Stage S as straight-line code: its partner bit, half size and bit window
    // are compile-time constants, so each butterfly names its registers
    // directly (a run-time stage index would make the register array dynamic).
    template <int S>
    static __device__ __forceinline__ void stages(short2 *data, const short2 *tw, int *red,
                                                  int (&xr)[E], int (&xi)[E], int &e, int t,
                                                  bool act) {
        if constexpr (S <= L) {
            constexpr int b = S == 1 ? 0 : base_of(S - 1);
            constexpr int nb = base_of(S);
            if constexpr (nb != b) {                // move the local bit window
                if (act) {
#pragma unroll
                    for (int l = 0; l < E; ++l)
                        data[pad32(idx(t, l, b))] = make_short2((short)xr[l], (short)xi[l]);
                }
                __syncthreads();
                if (act) {
#pragma unroll
                    for (int l = 0; l < E; ++l) {
                        const short2 v = data[pad32(idx(t, l, nb))];
                        xr[l] = v.x;
                        xi[l] = v.y;
                    }
                }
            }
            constexpr int q = (S - 1) - nb;
            // twiddle index of local butterfly l: j = idx(t, l, nb) & (h - 1) =
            // (t's low nb bits) + ((l's bits below q) << nb), so the thread's
            // twiddles lie at one base plus compile-time offsets
            const short2 *twb = tw + ((t & ((1 << nb) - 1)) << (L - S));
            int vmax = 0, vmin = 0;
            if (act) {
#pragma unroll
                for (int l = 0; l < E; ++l) {
                    if (l & (1 << q)) continue;
                    const int l1 = l | (1 << q);
                    const short2 w = twb[((l & ((1 << q) - 1)) << nb) << (L - S)];
                    const int br = xr[l1], bi = xi[l1];
                    // |w| <= 2^15 and |b| <= 2^15 at a stage's start: 24-bit products are exact
                    const int tr = sat16((__mul24(w.x, br) - __mul24(w.y, bi)) >> 15);
                    const int ti = sat16((__mul24(w.x, bi) + __mul24(w.y, br)) >> 15);
                    const int ar = xr[l], ai = xi[l];
                    const int y0r = ar + tr, y0i = ai + ti, y1r = ar - tr, y1i = ai - ti;
                    xr[l] = y0r;
                    xi[l] = y0i;
                    xr[l1] = y1r;
                    xi[l1] = y1i;
                    vmax = max(max(vmax, y0r), max(y0i, max(y1r, y1i)));
                    vmin = min(min(vmin, y0r), min(y0i, min(y1r, y1i)));
                }
            }
            int lmax = wave_max_dpp(max(vmax, -vmin));   // max |y| of the wave
            if ((t & 63) == 63) red[(S & 1) * NW + (t >> 6)] = lmax;
            __syncthreads();
            int mx = red[(S & 1) * NW];
#pragma unroll
            for (int w = 1; w < NW; ++w) mx = max(mx, red[(S & 1) * NW + w]);
            const int sh = mx <= 32766 ? 0 : (mx <= 65533 ? 1 : 2);
            // The reference clips floor(y / 2^sh) to int16.  That clip never
            // acts: a stage's inputs are int16 and t is saturated, so |y| <=
            // 65536; sh = 0 leaves |y| <= 32766, sh = 1 gives floor(y / 2) in
            // [-32767, 32766] (|y| <= 65533), sh = 2 floor(y / 4) in
            // [-16384, 16383].
            if (sh && act) {
#pragma unroll
                for (int l = 0; l < E; ++l) {
                    xr[l] >>= sh;
                    xi[l] >>= sh;
                }
            }
            e += sh;
            stages<S + 1>(data, tw, red, xr, xi, e, t, act);
        }
    }
};

// The pass-A output ya [range row j][cross bin k] (G x G) in blocks of 16
// cross bins: element (j, k) at ((k / 16) G + j) 16 + k % 16.  FFT-B reads a
// column down consecutive rows 64 bytes apart (16 columns share each chunk, and
// their workgroups run on one XCD), not a full row (4 G bytes) apart; pass A
// still writes whole 64-byte chunks.
__device__ __forceinline__ size_t yablk(int j, int k, int G) {
    return ((size_t)(k >> 4) * G + j) * 16 + (k & 15);
}

template <int L>
__device__ __forceinline__ void load_twiddles(short2 *tw, const short2 *twg) {
    constexpr int H = (1 << L) / 2;
    for (int i = threadIdx.x; i < H; i += blockDim.x) tw[i] = twg[i];
}

// ------------------------------------------------------------------ stand-alone row FFT
template <int L>
__global__ __launch_bounds__(Fft<L>::NB) void k_fft_rows(const short2 *__restrict__ in, int ld,
                                                        short2 *__restrict__ out,
                                                        int32_t *__restrict__ e_row,
                                                        int *__restrict__ emax,
                                                        const short2 *__restrict__ twg) {
    using F = Fft<L>;
    __shared__ short2 data[F::DATA];
    __shared__ short2 tw[F::N / 2];
    __shared__ int red[2 * F::NW];
    const int row = blockIdx.x;
    for (int i = threadIdx.x; i < F::N; i += blockDim.x) data[pad32(i)] = in[(size_t)row * ld + i];
    load_twiddles<L>(tw, twg);
    __syncthreads();
    int xr[F::E], xi[F::E], e;
    F::run(data, tw, red, xr, xi, e);
    if (threadIdx.x < F::NT) {
#pragma unroll
        for (int l = 0; l < F::E; ++l)
            out[(size_t)row * F::N + F::idx(threadIdx.x, l, L - F::P)] =
                make_short2((short)xr[l], (short)xi[l]);
    }
    if (threadIdx.x == 0) {
        e_row[row] = e;
        if (emax) atomicMax(emax, e);
    }
}

__global__ void k_renorm_rows(short2 *y, const int32_t *e_row, const int *emax, int n,
                              int64_t total) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int sh = *emax - e_row[i / n];
    short2 v = y[i];
    v.x = (short)sat16(v.x >> sh);
    v.y = (short)sat16(v.y >> sh);
    y[i] = v;
}

// ------------------------------------------------------------------ elementwise stages
__global__ void k_window(const short2 *g2, int rows, int cols, const int32_t *hamr,
                         const int32_t *hamc, short2 *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)rows * cols) return;
    const int j = (int)(i / cols), k = (int)(i % cols);
    const int cw = wrap16((hamr[j] * hamc[k]) >> 15);
    const short2 v = g2[i];
    out[i] = make_short2((short)wrap16((v.x * cw) >> 15), (short)wrap16((v.y * cw) >> 15));
}

// floor(sqrt(I^2 + Q^2)) of int16 codes, exact (the sum s is at most 2^31).
// The estimate is the hardware square root of float(s): float(s) is within 64
// of s (exact below 2^24) and the root within 2^-22 relative, so the estimate
// is within 0.02 of sqrt(s) and its truncation at most one off the integer
// root; one step down and one step up fix it.  d <= 46341, so d^2 and
// (d + 1)^2 are 24-bit products below 2^32.
__device__ __forceinline__ uint16_t detect_one(short2 v) {
    const unsigned s = (unsigned)((int)v.x * v.x) + (unsigned)((int)v.y * v.y);
    unsigned d = (unsigned)__builtin_amdgcn_sqrtf((float)s);
    d -= __umul24(d, d) > s ? 1u : 0u;
    d += __umul24(d + 1, d + 1) <= s ? 1u : 0u;
    return (uint16_t)d;
}

__global__ void k_detect(const short2 *z, int64_t count, uint16_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = detect_one(z[i]);
}

__global__ void k_zero_i32(int *p, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0;
}

// fixedpoint.fit_scale's max|re|, max|im| over a complex64 array (float bits of
// non-negative values order like unsigned integers).
// Scalar forms, for arrays not aligned for the vector forms below.
__global__ void k_absmax1(const float *x, int64_t count, unsigned *maxbits) {
    unsigned m = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        m = max(m, __float_as_uint(fabsf(x[i])));
    for (int off = 32; off >= 1; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(maxbits, m);
}

__global__ void k_quantize1(const float *x, int64_t count, float lsb, int16_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = (int16_t)fminf(fmaxf(floorf(x[i] / lsb), -32768.f), 32767.f);
}

// 16-byte loads over the whole float4 part of the array; the count % 4 tail
// values are taken by the first threads.
__global__ void k_absmax(const float *x, int64_t count, unsigned *maxbits) {
    unsigned m = 0;
    const int64_t n4 = count / 4;
    const float4 *x4 = reinterpret_cast<const float4 *>(x);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) {
        const float4 v = x4[i];
        m = max(max(max(m, __float_as_uint(fabsf(v.x))), __float_as_uint(fabsf(v.y))),
                max(__float_as_uint(fabsf(v.z)), __float_as_uint(fabsf(v.w))));
    }
    const int64_t tail = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tail < count - 4 * n4) m = max(m, __float_as_uint(fabsf(x[4 * n4 + tail])));
    for (int off = 32; off >= 1; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(maxbits, m);
}

// x / 2^e is exact, so the quotient, floor and clamp give fit_scale's codes.
__device__ __forceinline__ int16_t quant1(float v, float lsb) {
    return (int16_t)fminf(fmaxf(floorf(v / lsb), -32768.f), 32767.f);
}
__global__ void k_quantize(const float *x, int64_t count, float lsb, int16_t *out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;   // one group of four
    if (4 * i + 3 < count) {
        const float4 v = reinterpret_cast<const float4 *>(x)[i];
        short4 q;
        q.x = quant1(v.x, lsb);
        q.y = quant1(v.y, lsb);
        q.z = quant1(v.z, lsb);
        q.w = quant1(v.w, lsb);
        reinterpret_cast<short4 *>(out)[i] = q;
    } else {
        for (int64_t k = 4 * i; k < count; ++k) out[k] = quant1(x[k], lsb);
    }
}

// ------------------------------------------------------------------ resample helpers
// 2-tap lerp of serialize_inputs' (idx, wq) contract: a + floor((b-a)*wq / 2^15).
// f32: the product is rounded to float32 first (form_image's complex64 line).
__device__ __forceinline__ short2 lerp_q15(const short2 *row, int idx, int wq, bool f32) {
    const short2 a = row[idx], bb = row[idx + 1];
    int pr = (bb.x - a.x) * wq, pi = (bb.y - a.y) * wq;
    if (f32) {
        pr = (int)__int2float_rn(pr);
        pi = (int)__int2float_rn(pi);
    }
    return make_short2((short)sat16(a.x + (pr >> 15)), (short)sat16(a.y + (pi >> 15)));
}

// 32-tap polyphase sinc (gen_resample_vectors.gather_sinc): taps idx-15..idx+16,
// exact integer dot product, round half up, sat16.  Each sample is split into
// a signed high byte and an unsigned low byte so both partial sums stay in
// int32 (|c*xh| <= 2^22, |c*xl| <= 2^23 over 32 taps).  Two taps at a time:
// the taps of a phase are int16 pairs as stored (a table row is 64 bytes,
// 64-byte aligned, read as four 16-byte words); one byte permute builds each
// of the two samples' high-byte or low-byte pairs as int16 lanes, and a packed
// 16-bit dot product with 32-bit accumulation adds both products.
__device__ __forceinline__ unsigned byte_pair(unsigned s1, unsigned s0, unsigned sel) {
    return __builtin_amdgcn_perm(s1, s0, sel);
}
// selectors over {s1 : s0} (bytes 0..3 = s0 = (re0, im0), 4..7 = s1 = (re1, im1));
// 8, 9, 10, 11 give the sign of bytes 1, 3, 5, 7 and 12 gives zero
constexpr unsigned SEL_RE_HI = 0x0A050801u;   // (re0 >> 8, re1 >> 8)
constexpr unsigned SEL_RE_LO = 0x0C040C00u;   // (re0 & 255, re1 & 255)
constexpr unsigned SEL_IM_HI = 0x0B070903u;   // (im0 >> 8, im1 >> 8)
constexpr unsigned SEL_IM_LO = 0x0C060C02u;   // (im0 & 255, im1 & 255)

__device__ __forceinline__ int dot2(unsigned a, unsigned b, int c) {
    typedef short v2s __attribute__((ext_vector_type(2)));
    return __builtin_amdgcn_sdot2(__builtin_bit_cast(v2s, a), __builtin_bit_cast(v2s, b), c, false);
}

__device__ __forceinline__ short2 sinc_q15w(const short2 *row, int idx, const int16_t *taps) {
    const uint4 *tw = reinterpret_cast<const uint4 *>(taps);
    uint4 w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = tw[i];
    int hr = 0, lr = 0, hi = 0, li = 0;
    const unsigned *s = reinterpret_cast<const unsigned *>(row + idx - 15);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const unsigned cw = (&w[k >> 2].x)[k & 3];    // taps 2k (low half), 2k + 1
        const unsigned s0 = s[2 * k], s1 = s[2 * k + 1];
        hr = dot2(cw, byte_pair(s1, s0, SEL_RE_HI), hr);
        lr = dot2(cw, byte_pair(s1, s0, SEL_RE_LO), lr);
        hi = dot2(cw, byte_pair(s1, s0, SEL_IM_HI), hi);
        li = dot2(cw, byte_pair(s1, s0, SEL_IM_LO), li);
    }
    const int64_t ar = (int64_t)hr * 256 + lr + 16384;
    const int64_t ai = (int64_t)hi * 256 + li + 16384;
    return make_short2((short)max(min(ar >> 15, (int64_t)32767), (int64_t)-32768),
                       (short)max(min(ai >> 15, (int64_t)32767), (int64_t)-32768));
}

// The same gather with one dot product per tap pair and component: the
// plan has checked that every phase's taps 0..15 and taps 16..31 each have
// sum |c| <= 65535, so each group's partial sums stay within
// 65535 * 2^15 < 2^31 for any int16 samples and the two int32 accumulators
// of a component are exact.  Their sum plus 2^14, shifted by 15, is formed
// from 15-bit halves so nothing wraps.  One byte permute per component and
// pair builds the (x[2k], x[2k+1]) int16 lanes.
constexpr unsigned SEL_RE = 0x05040100u;      // (re0, re1)
constexpr unsigned SEL_IM = 0x07060302u;      // (im0, im1)
__device__ __forceinline__ int round_pair15(int a, int b) {
    const int lo = ((a & 32767) + (b & 32767) + 16384) >> 15;
    return sat16((a >> 15) + (b >> 15) + lo);
}
__device__ __forceinline__ short2 sinc_q15s(const short2 *row, int idx, const int16_t *taps) {
    const uint4 *tw = reinterpret_cast<const uint4 *>(taps);
    uint4 w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) w[i] = tw[i];
    int r0 = 0, r1 = 0, i0 = 0, i1 = 0;
    const unsigned *s = reinterpret_cast<const unsigned *>(row + idx - 15);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        const unsigned cw = (&w[k >> 2].x)[k & 3];    // taps 2k (low half), 2k + 1
        const unsigned s0 = s[2 * k], s1 = s[2 * k + 1];
        const unsigned pr = byte_pair(s1, s0, SEL_RE), pi = byte_pair(s1, s0, SEL_IM);
        if (k < 8) {
            r0 = dot2(cw, pr, r0);
            i0 = dot2(cw, pi, i0);
        } else {
            r1 = dot2(cw, pr, r1);
            i1 = dot2(cw, pi, i1);
        }
    }
    return make_short2((short)round_pair15(r0, r1), (short)round_pair15(i0, i1));
}

__device__ __forceinline__ int q15_weight(double w) {
    const double r = rint(w * 32768.0);
    return (int)fmin(fmax(r, 0.0), 32767.0);
}

// pass-1 source position x_j = f64(a) * (f64(f0) + j * f64(df)), each op rounded
// (silicon_emulator.py:144 under NumPy 2 promotion).
__device__ __forceinline__ double xpos(double a, double f0, double df, int j) {
    return __dmul_rn(a, __dadd_rn(f0, __dmul_rn((double)j, df)));
}

// ------------------------------------------------------------------ pass 1
constexpr int RT = 64;   // pulses per tile
constexpr int QT = 64;   // range outputs per tile
constexpr int SW = 112;  // staged source samples per pulse (64 outputs + window + slack)
constexpr int SWP = SW + 1;  // odd pitch: the pulses of a wave start in different banks

// Tile of RT sorted pulses x QT range outputs; thread t serves pulse
// 16 (t / 64) + t % 16 and the 16 consecutive outputs 16 ((t / 16) % 4) + u, so
// the 32 lanes a paired-word LDS read (ds_read2) serves per cycle are 16
// pulses x two chunks: with the odd strip pitch their windows start in
// distinct banks of the 64, where four chunks of one pulse among those 32
// lanes put chunks c and c + 2, 32 samples apart, on the same banks (measured
// at 8167 x 8192: bank-conflict cycles 130 M -> 49 M per launch, 0.60 ->
// 0.555 ms).  Three phases:
//  (1) brackets and Q15 weights of the thread's outputs, walking the FP64
//      positions with the two that bracket the previous output kept;
//  (2) each pulse's source strip -- from its first window on, SW samples --
//      staged in LDS by coalesced loads (the four threads of a pulse, lanes
//      16 apart, agree on its start with two lane swaps);
//  (3) the gathers from the strip (from the row in global memory for the
//      rare window that leaves it), stored straight to g1t: 16 pulses of one
//      output are one 64-byte run of a row.
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_range_resample(
    const short2 *__restrict__ sig, int m, int n, int grid, const float *__restrict__ a32,
    const float *__restrict__ f0v, const float *__restrict__ dfv,
    const int32_t *__restrict__ order, const float *__restrict__ krp,
    const int16_t *__restrict__ sinc, int flags, short2 *__restrict__ g1t, int ldp) {
    __shared__ short2 strip[RT * SWP];             // source strips
    __shared__ uint4 stl[256 * 4];                 // the 256 x 32 Q15 tap table, 16 KB
    __shared__ int lo_s[RT], row_s[RT];
    const int p0 = blockIdx.x * RT, q0 = blockIdx.y * QT;
    const int t = threadIdx.x;
    const int pl = ((t >> 6) << 4) | (t & 15), ch = (t >> 4) & 3;
    const int p = p0 + pl;
    const bool use_sinc = flags & PFA_FIXED_RANGE_SINC;
    if (use_sinc)
        for (int i = t; i < 256 * 4; i += 256) stl[i] = reinterpret_cast<const uint4 *>(sinc)[i];
    const int16_t *sincl = reinterpret_cast<const int16_t *>(stl);
    const bool f32 = !(flags & PFA_FIXED_LERP_EXACT);
    // (1) brackets: code = idx << 15 | wq (idx < 2^15, wq < 2^15), -1 = zero output
    constexpr int PER = QT / 4;
    int code[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) code[u] = -1;
    int i = 0;
    int lo = 1 << 30;
    if (p < m) {
        i = order[p];
        const double a = a32[i], f0 = f0v[i], df = dfv[i];
        const bool asc = xpos(a, f0, df, n - 1) >= xpos(a, f0, df, 0);
        // ascending view xa(u) = x(asc ? u : n-1-u); c = #{u : xa(u) < Q}
        auto xa = [&](int u) { return xpos(a, f0, df, asc ? u : n - 1 - u); };
        int c = 0;
        double xl = -INFINITY, xh = INFINITY;      // xa(c - 1), xa(c)
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = q0 + ch * PER + u;
            const double Q = krp[q];
            if (u == 0) {                          // closed-form guess, then the exact count
                const double js = (Q / a - f0) / df;
                double g = asc ? js : (double)(n - 1) - js;
                g = fmin(fmax(floor(g) + 1.0, 0.0), (double)n);
                c = (g == g) ? (int)g : 0;
                int steps = 0;
                xl = c > 0 ? xa(c - 1) : -INFINITY;
                xh = c < n ? xa(c) : INFINITY;
                while (c < n && xh < Q && steps < 64) { ++c; ++steps; xl = xh; xh = c < n ? xa(c) : INFINITY; }
                while (c > 0 && !(xl < Q) && steps < 128) { --c; ++steps; xh = xl; xl = c > 0 ? xa(c - 1) : -INFINITY; }
                if (steps >= 64) {                 // far from the guess: binary search
                    int bl = 0, bh = n;
                    while (bl < bh) {
                        const int mid = (bl + bh) >> 1;
                        if (xa(mid) < Q) bl = mid + 1; else bh = mid;
                    }
                    c = bl;
                    xl = c > 0 ? xa(c - 1) : -INFINITY;
                    xh = c < n ? xa(c) : INFINITY;
                }
            } else {                               // KRp is non-decreasing: walk up
                int steps = 0;
                while (c < n && xh < Q && steps < 64) { ++c; ++steps; xl = xh; xh = c < n ? xa(c) : INFINITY; }
                if (c < n && xh < Q) {
                    int bl = c, bh = n;
                    while (bl < bh) {
                        const int mid = (bl + bh) >> 1;
                        if (xa(mid) < Q) bl = mid + 1; else bh = mid;
                    }
                    c = bl;
                    xl = c > 0 ? xa(c - 1) : -INFINITY;
                    xh = c < n ? xa(c) : INFINITY;
                }
            }
            const int k = c - 1;
            int cd = -1;
            if (k >= 0 && k <= n - 2) {
                const double frac = div_bracket(Q - xl, xh - xl);
                const int idx = asc ? k : n - 2 - k;
                const int wq = q15_weight(asc ? frac : 1.0 - frac);
                cd = idx << 15 | wq;
                lo = min(lo, idx);
            }
            code[u] = cd;
        }
    }
    // (2) the pulse's strip starts at its lowest window: min over its four threads
    lo = min(lo, __shfl_xor(lo, 16, 64));
    lo = min(lo, __shfl_xor(lo, 32, 64));
    if (ch == 0) {
        lo_s[pl] = max(0, min(lo - 15, n - SW));
        row_s[pl] = p < m ? i : -1;
    }
    __syncthreads();
    {                                              // loads in flight in two batches, then the stores
        constexpr int NL = RT * SW / 256, NH = NL / 2;
        static_assert(NL * 256 == RT * SW && NH * 2 == NL, "strips split evenly over the threads");
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            short2 v[NH];
#pragma unroll
            for (int k = 0; k < NH; ++k) {
                const int f = t + 256 * (h * NH + k), pp = f / SW, u = f % SW;
                const int r = row_s[pp], src = min(lo_s[pp] + u, n - 1);
                v[k] = sig[(size_t)max(r, 0) * n + src];  // clamped: always in bounds
            }
#pragma unroll
            for (int k = 0; k < NH; ++k) {
                const int f = t + 256 * (h * NH + k), pp = f / SW, u = f % SW;
                const bool ok = row_s[pp] >= 0 && lo_s[pp] + u < n;
                strip[pp * SWP + u] = ok ? v[k] : make_short2(0, 0);
            }
        }
    }
    __syncthreads();
    // (3) gathers
    short2 out[PER];
    {
        const int base = lo_s[pl];
        const short2 *sp = strip + pl * SWP;
        const short2 *row = sig + (size_t)i * n;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int cd = code[u];
            short2 v = make_short2(0, 0);
            if (cd >= 0) {
                const int idx = cd >> 15, wq = cd & 0x7fff;
                const int16_t *taps = sincl + (wq >> 7) * 32;
                // strip reads stay LDS loads: the strip and the global row are
                // separate calls, never one pointer that could be either
                if (use_sinc && idx >= 15 && idx <= n - 17) {
                    if (idx - 15 >= base && idx + 16 < base + SW)
                        v = SPLIT ? sinc_q15s(sp, idx - base, taps) : sinc_q15w(sp, idx - base, taps);
                    else
                        v = SPLIT ? sinc_q15s(row, idx, taps) : sinc_q15w(row, idx, taps);
                } else if (idx >= base && idx + 1 < base + SW) {
                    v = lerp_q15(sp, idx - base, wq, f32 && !use_sinc);
                } else {
                    v = lerp_q15(row, idx, wq, f32 && !use_sinc);
                }
            }
            out[u] = v;
        }
    }
    // lanes 16 i .. 16 i + 15 hold 16 consecutive pulses of one range output:
    // each store writes four 64-byte runs of g1t rows (the LDS transpose of the
    // output tile this replaced cost two barriers and 2.5 % of the pass)
#pragma unroll
    for (int u = 0; u < PER; ++u)
        if (p0 + pl < m) g1t[(size_t)(q0 + ch * PER + u) * ldp + p0 + pl] = out[u];
}

// ------------------------------------------------------------------ pass 2 + window + FFT A
// Two phases per line, each with its own lane assignment:
//  (A) brackets: thread t owns PER consecutive outputs (the walk from one output
//      to the next is a step or two); each output's gather recipe -- kind,
//      sample index, Q15 weight -- is parked in its own slot of `data`;
//  (B) gathers: the sorted-pulse row is staged in LDS over the dead positions,
//      and thread t takes outputs t, t + NB, ... so the lanes of a wave read
//      neighbouring windows (distinct banks) and coalesced tapers; each
//      output's sinc taps are read as four 16-byte loads.
// The arithmetic of every output is unchanged: the same bracket, weight and
// integer gather in the same order.
__device__ __forceinline__ unsigned gather_code(int kind, int idx, int wq) {
    return (unsigned)kind << 30 | (unsigned)idx << 15 | (unsigned)wq;   // idx < 2^15, wq < 2^15
}

template <int L, bool DO_FFT, bool SPLIT>
__global__ __launch_bounds__(Fft<L>::NB) void k_azimuth_fft(
    const short2 *__restrict__ g1t, int ldp, int m, const float *__restrict__ krp,
    const float *__restrict__ tan_s, const float *__restrict__ kcp,
    const int16_t *__restrict__ sinc, int flags, const int32_t *__restrict__ hamr,
    const int32_t *__restrict__ hamc, int do_window, short2 *__restrict__ out,
    int32_t *__restrict__ e_row, int *__restrict__ emax, const short2 *__restrict__ twg) {
    using F = Fft<L>;
    constexpr int G = F::N;
    constexpr int NB = F::NB;
    constexpr int PER = G / NB > 0 ? G / NB : 1;
    static_assert(PER * NB == G, "the threads' outputs cover the line");
    __shared__ short2 data[F::DATA];
    // sources for the bracket search, then the row, then twiddles.  The float32
    // positions sit at si(p) = p + p / 16: each lane walks 16 consecutive outputs,
    // so the lanes of a wave read positions about 16 floats apart, which unpadded
    // would put a whole wave's reads on two of the 32 banks.
    __shared__ float srcv[G + G / 16];
    __shared__ int red[2 * F::NW];
    unsigned *codes = reinterpret_cast<unsigned *>(data);
    short2 *rowl = reinterpret_cast<short2 *>(srcv);
    const int j = blockIdx.x;
    const int t = threadIdx.x;
    const float kr = krp[j];
    const short2 *row = g1t + (size_t)j * ldp;
    // everything the line reads from global memory is requested up front
    // (clamped slots, masked when used), so the loads overlap one another and
    // the bracket phase instead of waiting one by one
    float ts[PER], kq[PER];
    short2 rw[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        ts[i] = tan_s[min(t + i * NB, m - 1)];
        kq[i] = kcp[t * PER + i];
        rw[i] = row[min(t + i * NB, m - 1)];
    }
    auto si = [](int p) { return p + (p >> 4); };
#pragma unroll
    for (int i = 0; i < PER; ++i)
        if (t + i * NB < m) srcv[si(t + i * NB)] = __fmul_rn(kr, ts[i]);
    __syncthreads();
    const bool asc = srcv[si(m - 1)] >= srcv[0];
    auto sa = [&](int u) { return srcv[si(asc ? u : m - 1 - u)]; };
    const bool use_sinc = flags & PFA_FIXED_AZ_SINC;
    // (A) brackets and weights: c = #{u : sa(u) < Q}, with xl = sa(c - 1) and
    // xh = sa(c) kept from one output to the next (KCp is non-decreasing).  The
    // first output starts from the linear guess between the end positions.
    int c;
    float xl, xh;
    {
        const float s0 = sa(0), s1 = sa(m - 1), Q = kq[0];
        const float g = (s1 > s0) ? (Q - s0) / (s1 - s0) * (float)(m - 1) : 0.f;
        c = (g == g) ? (int)fminf(fmaxf(g + 1.0f, 0.f), (float)m) : 0;
        xl = c > 0 ? sa(c - 1) : -INFINITY;
        xh = c < m ? sa(c) : INFINITY;
        int steps = 0;
        while (c < m && xh < Q && steps < 32) { ++c; ++steps; xl = xh; xh = c < m ? sa(c) : INFINITY; }
        while (c > 0 && !(xl < Q) && steps < 64) { --c; ++steps; xh = xl; xl = c > 0 ? sa(c - 1) : -INFINITY; }
        if (!((c == 0 || xl < Q) && (c == m || !(xh < Q)))) {   // far from the guess: bisect
            int lo = 0, hi = m;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (sa(mid) < Q) lo = mid + 1; else hi = mid;
            }
            c = lo;
            xl = c > 0 ? sa(c - 1) : -INFINITY;
            xh = c < m ? sa(c) : INFINITY;
        }
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int q = t * PER + u;
        const float Q = kq[u];
        if (u > 0) {
            int steps = 0;
            while (c < m && xh < Q && steps < 64) { ++c; ++steps; xl = xh; xh = c < m ? sa(c) : INFINITY; }
            if (c < m && xh < Q) {               // far: bisect above c
                int lo = c, hi = m;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (sa(mid) < Q) lo = mid + 1; else hi = mid;
                }
                c = lo;
                xl = c > 0 ? sa(c - 1) : -INFINITY;
                xh = c < m ? sa(c) : INFINITY;
            }
        }
        const int k = c - 1;
        unsigned code = 0;                                   // kind 0: zero output
        if (k >= 0 && k <= m - 2) {
            const double x0 = xl, x1 = xh;
            const double frac = div_bracket((double)Q - x0, x1 - x0);
            const int idx = asc ? k : m - 2 - k;
            const int wq = q15_weight(asc ? frac : 1.0 - frac);
            code = gather_code(use_sinc && idx >= 15 && idx <= m - 17 ? 1 : 2, idx, wq);
        }
        codes[pad32(q)] = code;
    }
    int hc[PER];                         // the window's cross tapers, in flight during the row store
    const int hr = hamr[j];
#pragma unroll
    for (int u = 0; u < PER; ++u) hc[u] = do_window ? hamc[t + u * NB] : 0;
    __syncthreads();                     // positions dead: the row takes over
#pragma unroll
    for (int i = 0; i < PER; ++i)
        if (t + i * NB < m) rowl[t + i * NB] = rw[i];
    __syncthreads();
    // (B) gathers and window
    constexpr int TWP = DO_FFT ? (G / 2 + NB - 1) / NB : 1;
    short2 twr[TWP];                     // the FFT's twiddles, in flight during the gathers
    if constexpr (DO_FFT) {
#pragma unroll
        for (int i = 0; i < TWP; ++i) twr[i] = twg[min(t + i * NB, G / 2 - 1)];
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
        const int q = t + u * NB;
        const unsigned code = codes[pad32(q)];
        const int kind = (int)(code >> 30), idx = (int)((code >> 15) & 0x7fffu), wq = (int)(code & 0x7fffu);
        short2 v = make_short2(0, 0);
        if (kind == 1) v = SPLIT ? sinc_q15s(rowl, idx, sinc + (wq >> 7) * 32)
                                 : sinc_q15w(rowl, idx, sinc + (wq >> 7) * 32);
        else if (kind == 2) v = lerp_q15(rowl, idx, wq, false);
        if (do_window) {
            const int cw = wrap16((hr * hc[u]) >> 15);
            v = make_short2((short)wrap16((v.x * cw) >> 15), (short)wrap16((v.y * cw) >> 15));
        }
        if (DO_FFT) data[pad32(q)] = v;
        else out[(size_t)j * G + q] = v;
    }
    if constexpr (DO_FFT) {
        __syncthreads();                 // the row is dead: the twiddles take over
        short2 *tw = reinterpret_cast<short2 *>(srcv);
#pragma unroll
        for (int i = 0; i < TWP; ++i)
            if (t + i * NB < G / 2) tw[t + i * NB] = twr[i];
        __syncthreads();
        int xr[F::E], xi[F::E], e;
        F::run(data, tw, red, xr, xi, e);
        if (t < F::NT) {
#pragma unroll
            for (int l = 0; l < F::E; ++l)
                out[yablk(j, F::idx(t, l, L - F::P), G)] = make_short2((short)xr[l], (short)xi[l]);
        }
        if (t == 0) {
            e_row[j] = e;
            atomicMax(emax, e);
        }
    }
}

// ------------------------------------------------------------------ FFT B (range axis)
// Workgroup b handles cross column k; consecutive columns go to the same XCD
// (workgroups are dispatched round-robin over the 8 XCDs) so their strided
// column reads share that XCD's L2 lines.
template <int L>
__global__ __launch_bounds__(Fft<L>::NB) void k_fft_b(const short2 *__restrict__ ya,
                                                     const int32_t *__restrict__ ea,
                                                     const int *__restrict__ emax_a,
                                                     short2 *__restrict__ yb,
                                                     int32_t *__restrict__ eb,
                                                     int *__restrict__ emax_b,
                                                     const short2 *__restrict__ twg) {
    using F = Fft<L>;
    constexpr int G = F::N;
    __shared__ short2 data[F::DATA];
    __shared__ short2 tw[G / 2];
    __shared__ int red[2 * F::NW];
    const int b = blockIdx.x;
    const int per_xcd = G / 8;
    const int k = (G >= 8) ? (b % 8) * per_xcd + b / 8 : b;
    const int ema = *emax_a;
    // every row load of the column in flight before the first is used
    constexpr int PERB = G / F::NB;
    static_assert(PERB * F::NB == G, "the threads' rows cover the column");
    const int t = threadIdx.x;
    short2 v[PERB];
    int er[PERB];
#pragma unroll
    for (int i = 0; i < PERB; ++i) {
        v[i] = ya[yablk(t + i * F::NB, k, G)];
        er[i] = ea[t + i * F::NB];
    }
#pragma unroll
    for (int i = 0; i < PERB; ++i) {
        const int sh = ema - er[i];
        data[pad32(t + i * F::NB)] = make_short2((short)sat16(v[i].x >> sh), (short)sat16(v[i].y >> sh));
    }
    load_twiddles<L>(tw, twg);
    __syncthreads();
    int xr[F::E], xi[F::E], e;
    F::run(data, tw, red, xr, xi, e);
    if (threadIdx.x < F::NT) {
#pragma unroll
        for (int l = 0; l < F::E; ++l)
            yb[(size_t)k * G + F::idx(threadIdx.x, l, L - F::P)] =
                make_short2((short)xr[l], (short)xi[l]);
    }
    if (threadIdx.x == 0) {
        eb[k] = e;
        atomicMax(emax_b, e);
    }
}

// pass-B renorm + detect, transposing yb[cross k][range j] to out[range j][cross k]
// through a 64 x 64 LDS tile.  Each thread reads two neighbouring range bins of
// a cross row (one 8-byte load) and writes two neighbouring cross bins of an
// output row (one 4-byte store); the tile's odd word pitch (33) puts the rows a
// wave touches on different banks.
__global__ __launch_bounds__(256) void k_detect_t(const short2 *__restrict__ yb,
                                                  const int32_t *__restrict__ eb,
                                                  const int *__restrict__ emax_b, int G,
                                                  uint16_t *__restrict__ out) {
    __shared__ uint16_t tile[64][66];
    const int k0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
    const int emb = *emax_b;
#pragma unroll
    for (int r = threadIdx.x; r < 64 * 32; r += 256) {
        const int kk = r / 32, jj = (r % 32) * 2;
        const uint2 w = *reinterpret_cast<const uint2 *>(yb + (size_t)(k0 + kk) * G + j0 + jj);
        const int sh = emb - eb[k0 + kk];
        const short2 a = __builtin_bit_cast(short2, w.x), b = __builtin_bit_cast(short2, w.y);
        tile[jj][kk] = detect_one(make_short2((short)sat16(a.x >> sh), (short)sat16(a.y >> sh)));
        tile[jj + 1][kk] = detect_one(make_short2((short)sat16(b.x >> sh), (short)sat16(b.y >> sh)));
    }
    __syncthreads();
#pragma unroll
    for (int r = threadIdx.x; r < 64 * 32; r += 256) {
        const int jj = r / 32, kk = (r % 32) * 2;
        const unsigned pr = (unsigned)tile[jj][kk] | ((unsigned)tile[jj][kk + 1] << 16);
        *reinterpret_cast<unsigned *>(out + (size_t)(j0 + jj) * G + k0 + kk) = pr;
    }
}

inline int grid1(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

const short2 *twiddles16(pfa_ctx *ctx, int n, int *st) {
    char slot[32];
    snprintf(slot, sizeof slot, "tw16_%d", n);
    auto it = ctx->ws.find(slot);
    if (it != ctx->ws.end() && it->second.bytes) return (const short2 *)it->second.ptr;
    std::vector<int32_t> codes((size_t)n);
    pfa_twiddle_codes(n, 16, codes.data());
    std::vector<short2> h((size_t)n / 2);
    for (int k = 0; k < n / 2; ++k) h[k] = make_short2((short)codes[2 * k], (short)codes[2 * k + 1]);
    void *d = pfa_ws(ctx, slot, h.size() * sizeof(short2), st);
    if (!d) return nullptr;
    if (hipMemcpy(d, h.data(), h.size() * sizeof(short2), hipMemcpyHostToDevice) != hipSuccess) {
        *st = pfa_fail(ctx, PFA_ERR_HIP, "twiddle upload failed");
        return nullptr;
    }
    return (const short2 *)d;
}

}  // namespace

// ------------------------------------------------------------------ launchers
#define PFA_SWITCH_L(L, MACRO) \
    switch (L) {               \
        MACRO(6) MACRO(7) MACRO(8) MACRO(9) MACRO(10) MACRO(11) MACRO(12) MACRO(13) \
        default: return pfa_fail(ctx, PFA_ERR_DIMS, "FFT length must be a power of two in 64..8192"); \
    }

static int log2_exact(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return (1 << l) == n ? l : -1;
}

int launch_fft_bfp_rows(pfa_ctx *ctx, const FixedFFTArgs &a) {
    int st = PFA_OK;
    const short2 *tw = twiddles16(ctx, a.n, &st);
    if (!tw) return st;
    const int L = log2_exact(a.n);
    if (a.rows <= 0) return PFA_OK;
#define LAUNCH(LL)                                                                          \
    case LL:                                                                                \
        hipLaunchKernelGGL(k_fft_rows<LL>, dim3(a.rows), dim3(Fft<LL>::NB), 0, ctx->stream, \
                           (const short2 *)a.in, a.in_ld, (short2 *)a.out, a.e_row, a.emax, tw); \
        break;
    PFA_SWITCH_L(L, LAUNCH)
#undef LAUNCH
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

int launch_renorm_rows(pfa_ctx *ctx, void *y, const int32_t *e_row, const int *emax, int rows,
                       int n) {
    const int64_t total = (int64_t)rows * n;
    if (total == 0) return PFA_OK;
    hipLaunchKernelGGL(k_renorm_rows, dim3(grid1(total, 256)), dim3(256), 0, ctx->stream,
                       (short2 *)y, e_row, emax, n, total);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

int launch_window_fixed(pfa_ctx *ctx, const void *g2, int rows, int cols, const int32_t *hamr,
                        const int32_t *hamc, void *out) {
    const int64_t total = (int64_t)rows * cols;
    if (total == 0) return PFA_OK;
    hipLaunchKernelGGL(k_window, dim3(grid1(total, 256)), dim3(256), 0, ctx->stream,
                       (const short2 *)g2, rows, cols, hamr, hamc, (short2 *)out);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

int launch_detect_fixed(pfa_ctx *ctx, const void *z, int64_t count, uint16_t *out) {
    if (count == 0) return PFA_OK;
    hipLaunchKernelGGL(k_detect, dim3(grid1(count, 256)), dim3(256), 0, ctx->stream,
                       (const short2 *)z, count, out);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

int launch_zero_i32(pfa_ctx *ctx, int *p, int count) {
    hipLaunchKernelGGL(k_zero_i32, dim3(grid1(count, 64)), dim3(64), 0, ctx->stream, p, count);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

int launch_quantize(pfa_ctx *ctx, const float *sig, int64_t count, int16_t *out, int *exp_out) {
    // count = number of float values (2 per complex sample)
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, ctx->stream, (int *)ctx->d_maxbits, 1);
    const int blocks = (int)std::min<int64_t>(4096, grid1(count, 256));
    const bool vec = ((uintptr_t)sig & 15) == 0 && ((uintptr_t)out & 7) == 0;
    if (count > 0)
        hipLaunchKernelGGL(vec ? k_absmax : k_absmax1, dim3(blocks), dim3(256), 0, ctx->stream, sig,
                           count, ctx->d_maxbits);
    PFA_HIP(ctx, hipGetLastError());
    unsigned bits = 0;
    PFA_HIP(ctx, hipMemcpyAsync(&bits, ctx->d_maxbits, 4, hipMemcpyDeviceToHost, ctx->stream));
    PFA_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float mxf;
    memcpy(&mxf, &bits, 4);
    const double mx = mxf;
    int e = 0;
    if (mx > 0.0) {   // smallest e with mx <= 32767 * 2^e (fixedpoint.fit_scale)
        e = (int)ceil(log2(mx / 32767.0));
        while (mx > 32767.0 * ldexp(1.0, e)) ++e;
        while (mx <= 32767.0 * ldexp(1.0, e - 1)) --e;
    }
    if (e < -126 || e > 127)
        return pfa_fail(ctx, PFA_ERR_GEOMETRY, "signal scale outside the float32 exponent range");
    *exp_out = e;
    if (count > 0 && vec)
        hipLaunchKernelGGL(k_quantize, dim3(grid1((count + 3) / 4, 256)), dim3(256), 0, ctx->stream,
                           sig, count, ldexpf(1.0f, e), out);
    else if (count > 0)
        hipLaunchKernelGGL(k_quantize1, dim3(grid1(count, 256)), dim3(256), 0, ctx->stream, sig,
                           count, ldexpf(1.0f, e), out);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

int launch_fixed_range(pfa_plan *p, const void *sig, void *g1t, int ldp) {
    pfa_ctx *ctx = p->ctx;
    dim3 grid((p->m + RT - 1) / RT, p->grid / QT);
    pfa_timer_begin(ctx, PFA_TIMER_RANGE);
    hipLaunchKernelGGL((p->flags & FIXED_SINC_SPLIT) ? k_range_resample<true> : k_range_resample<false>,
                       grid, dim3(256), 0, ctx->stream, (const short2 *)sig,
                       p->m, p->n, p->grid, p->a32, p->f0, p->df, p->order, p->krp, p->sinc,
                       p->flags, (short2 *)g1t, ldp);
    pfa_timer_end(ctx, PFA_TIMER_RANGE);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

int launch_fixed_azimuth_fft(pfa_plan *p, const void *g1t, int ldp, void *ya, int32_t *ea,
                             int *emax_a, bool do_fft, bool do_window) {
    pfa_ctx *ctx = p->ctx;
    int st = PFA_OK;
    const short2 *tw = p->tw16 ? (const short2 *)p->tw16 : twiddles16(ctx, p->grid, &st);
    if (!tw) return st;
    const int L = log2_exact(p->grid);
    pfa_timer_begin(ctx, PFA_TIMER_AZIMUTH);
#define LAUNCH(LL)                                                                             \
    case LL: {                                                                                 \
        const bool sp = p->flags & FIXED_SINC_SPLIT;                                           \
        auto kern = do_fft ? (sp ? k_azimuth_fft<LL, true, true> : k_azimuth_fft<LL, true, false>) \
                           : (sp ? k_azimuth_fft<LL, false, true> : k_azimuth_fft<LL, false, false>); \
        hipLaunchKernelGGL(kern, dim3(p->grid), dim3(Fft<LL>::NB), 0,                          \
                           ctx->stream, (const short2 *)g1t, ldp, p->m, p->krp, p->tan_s,      \
                           p->kcp, p->sinc, p->flags, p->hamr, p->hamc, (int)do_window,        \
                           (short2 *)ya, ea, emax_a, tw);                                      \
        break;                                                                                 \
    }
    PFA_SWITCH_L(L, LAUNCH)
#undef LAUNCH
    pfa_timer_end(ctx, PFA_TIMER_AZIMUTH);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

int launch_fixed_fft_b(pfa_plan *p, const void *ya, const int32_t *ea, const int *emax_a,
                       void *yb, int32_t *eb, int *emax_b) {
    pfa_ctx *ctx = p->ctx;
    int st = PFA_OK;
    const short2 *tw = p->tw16 ? (const short2 *)p->tw16 : twiddles16(ctx, p->grid, &st);
    if (!tw) return st;
    const int L = log2_exact(p->grid);
    pfa_timer_begin(ctx, PFA_TIMER_FFT_B);
#define LAUNCH(LL)                                                                           \
    case LL:                                                                                 \
        hipLaunchKernelGGL(k_fft_b<LL>, dim3(p->grid), dim3(Fft<LL>::NB), 0, ctx->stream,    \
                           (const short2 *)ya, ea, emax_a, (short2 *)yb, eb, emax_b, tw);    \
        break;
    PFA_SWITCH_L(L, LAUNCH)
#undef LAUNCH
    pfa_timer_end(ctx, PFA_TIMER_FFT_B);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}

int launch_fixed_detect_t(pfa_plan *p, const void *yb, const int32_t *eb, const int *emax_b,
                          uint16_t *out) {
    pfa_ctx *ctx = p->ctx;
    const int G = p->grid;
    pfa_timer_begin(ctx, PFA_TIMER_DETECT);
    hipLaunchKernelGGL(k_detect_t, dim3(G / 64, G / 64), dim3(256), 0, ctx->stream,
                       (const short2 *)yb, eb, emax_b, G, out);
    pfa_timer_end(ctx, PFA_TIMER_DETECT);
    PFA_HIP(ctx, hipGetLastError());
    return PFA_OK;
}
