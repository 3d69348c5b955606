/*
 * mi355x_pfa.h -- C ABI of the MI355X (gfx950) Polar Format focus library.
 *
 * The library replaces the focus stage of the reference SAR processor: the
 * keystone resample (range pass, corner-turn, azimuth pass), the separable
 * Hamming window, the two 1-D FFT passes with their corner-turn, and detection.
 * Two arithmetic modes:
 *   - fixed : bit-exact with mpfs/host/silicon_emulator.py (form_image and its
 *             stage functions), integer arithmetic on int16 I/Q codes;
 *   - float : src/form_image_pfa.py (pfa / resample_kspace / focus) in FP32,
 *             with a float64 GPU geometry stage.
 *
 * Conventions
 *   - Complex int16 data is interleaved I,Q (int16 I then int16 Q), the
 *     reference's on-board sig.bin layout (serialize_inputs.py:181,
 *     ddr_layout.py:346-364).  Complex float data is interleaved float32 re,im
 *     (numpy complex64).  All arrays are C-contiguous, row-major.
 *   - Functions without a _dev suffix take HOST pointers and are synchronous.
 *     Functions with a _dev suffix take DEVICE pointers for the bulk arrays and
 *     enqueue on the context's stream (see pfa_set_stream); small tables are
 *     always host pointers and are copied by the library.
 *   - Every function returns PFA_OK (0) or a negative pfa_status; the message
 *     is available from pfa_last_error(ctx).  Nothing aborts or throws across
 *     the ABI: a host allocation or thread failure inside a call returns
 *     PFA_ERR_NOMEM.  One context per host thread.
 */
#ifndef MI355X_PFA_H
#define MI355X_PFA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PFA_ABI_VERSION 1

typedef enum pfa_status {
    PFA_OK = 0,
    PFA_ERR_ARG = -1,        /* null pointer, bad flag, inconsistent sizes          */
    PFA_ERR_DIMS = -2,       /* sizes outside what the kernels support             */
    PFA_ERR_GEOMETRY = -3,   /* non-finite or degenerate geometry / tables         */
    PFA_ERR_HIP = -4,        /* HIP runtime error (message carries hipGetErrorString) */
    PFA_ERR_NOMEM = -5,      /* device allocation failed                           */
    PFA_ERR_STATE = -6       /* wrong object kind for the call                     */
} pfa_status;

typedef struct pfa_ctx pfa_ctx;
typedef struct pfa_plan pfa_plan;

/* ------------------------------------------------------------------ context */
int pfa_abi_version(void);
/* Create a context on HIP device `device`.  It owns a stream and a grow-only
 * device workspace that is reused across calls. */
int pfa_create(pfa_ctx **ctx, int device);
void pfa_destroy(pfa_ctx *ctx);
const char *pfa_last_error(const pfa_ctx *ctx);
/* Use an external hipStream_t (NULL restores the context's own stream). */
int pfa_set_stream(pfa_ctx *ctx, void *hip_stream);
int pfa_synchronize(pfa_ctx *ctx);
/* Device time of the last kernel of `kind` enqueued by this context (ms,
 * hipEvent pair on the context stream); kind: PFA_TIMER_*. */
int pfa_timer_enable(pfa_ctx *ctx, int on);
int pfa_timer_read(pfa_ctx *ctx, int kind, float *ms_total, int *launches);
int pfa_timer_reset(pfa_ctx *ctx);
enum { PFA_TIMER_RANGE = 0, PFA_TIMER_AZIMUTH = 1, PFA_TIMER_FFT_B = 2,
       PFA_TIMER_DETECT = 3, PFA_TIMER_OTHER = 4, PFA_TIMER_KINDS = 5 };

/* Diagnostic counters since the last reset: [0] range-pass sinc outputs whose
 * 32-tap window left the staged strip (read from global memory instead), [1]
 * azimuth-pass brackets started from the table guess (the first of each four
 * consecutive outputs) that needed the exact fix-up walk beyond one step. */
int pfa_fallback_counts(pfa_ctx *ctx, uint32_t out[2], int reset);

/* ------------------------------------------------------------------ tables */
/* Truncated Q(bits-1) twiddles W_n^k, k < n/2, as fixedpoint.py:58-60 /
 * :281-282 builds them: floor(cos(2*pi*k/n) * 2^(bits-1)) clipped to bits,
 * floor(-sin(..) * 2^(bits-1)); out is n/2 interleaved (re, im) int32. */
int pfa_twiddle_codes(int n, int bits, int32_t *out);

/* ------------------------------------------------------------------ fixed mode
 * Stage entry points, each a drop-in for the silicon_emulator.py function
 * named.  Arrays of int16 codes are interleaved I,Q. */

/* fixedpoint.fit_scale + floor + sat16 of a complex64 array
 * (silicon_emulator.py:226-229): out_iq[count*2], *exp_out = block exponent. */
int pfa_quantize_iq16(pfa_ctx *ctx, const float *sig_c64, int64_t count,
                      int16_t *out_iq, int *exp_out);
int pfa_quantize_iq16_dev(pfa_ctx *ctx, const float *sig_c64, int64_t count,
                          int16_t *out_iq, int *exp_out);

/* fft_pass_bfp (silicon_emulator.py:198-205 -> fixedpoint.py:266-289): per-row
 * 16-bit BFP radix-2 FFT of `rows` rows of length n (power of two, 64..8192),
 * then global renorm y >> (emax - e_row).  e_row[rows] receives the exponents. */
int pfa_fft_pass_bfp(pfa_ctx *ctx, const int16_t *x_iq, int rows, int n,
                     int16_t *y_iq, int32_t *e_row);
int pfa_fft_pass_bfp_dev(pfa_ctx *ctx, const int16_t *x_iq, int rows, int n,
                         int16_t *y_iq, int32_t *e_row);
/* fft1d_bfp_hw_perrow alone (no global renorm). */
int pfa_fft_bfp_perrow(pfa_ctx *ctx, const int16_t *x_iq, int rows, int n,
                       int16_t *y_iq, int32_t *e_row);

/* window_fixed (silicon_emulator.py:165-195) on g2 [rows][cols] (range rows,
 * cross columns); hamr[rows], hamc[cols] are the floor(hamming*32768) tapers
 * (int32: the centre tap may be 32768). */
int pfa_window_fixed(pfa_ctx *ctx, const int16_t *g2_iq, int rows, int cols,
                     const int32_t *hamr, const int32_t *hamc, int16_t *out_iq);

/* detect_fixed (silicon_emulator.py:208-213): floor(sqrt(I^2+Q^2)) -> uint16. */
int pfa_detect_fixed(pfa_ctx *ctx, const int16_t *z_iq, int64_t count, uint16_t *out);
int pfa_detect_fixed_dev(pfa_ctx *ctx, const int16_t *z_iq, int64_t count, uint16_t *out);

/* Scene tables of silicon_emulator.py:128-140 (the board's DDR roles,
 * serialize_inputs.py:203-230): f0, df, pr, tan_s float32[m]; inv_order
 * int32[m]; krp, kcp float32[grid]; hamr, hamc int32[grid]; sinc_tab int16
 * [256][32] (gen_resample_vectors.sinc_table_q15) or NULL when neither pass
 * uses the sinc gather. */
#define PFA_FIXED_RANGE_SINC  0x1   /* pass 1: 32-tap sinc (else 2-tap lerp)       */
#define PFA_FIXED_AZ_SINC     0x2   /* pass 2: 32-tap sinc (else 2-tap lerp)       */
#define PFA_FIXED_LERP_EXACT  0x4   /* pass-1 lerp product exact (the stage fed a
                                       complex128 line); default is form_image's
                                       float32 product of a complex64 line        */
int pfa_plan_fixed(pfa_ctx *ctx, pfa_plan **plan, int m, int n, int grid,
                   const float *f0, const float *df, const float *pr,
                   const float *tan_s, const int32_t *inv_order,
                   const float *krp, const float *kcp,
                   const int32_t *hamr, const int32_t *hamc,
                   const int16_t *sinc_tab, int flags);

/* resample_fixed (silicon_emulator.py:125-156): sig_iq[m][n] -> g2_iq
 * [grid range][grid cross]. */
int pfa_resample_fixed(pfa_plan *plan, const int16_t *sig_iq, int16_t *g2_iq);

/* form_image after quantisation (silicon_emulator.py:233-249): resample ->
 * window -> FFT along cross (per range row) -> corner-turn -> FFT along range
 * -> detect.  out[grid range][grid cross] uint16; exp_a[grid], exp_b[grid]
 * receive the per-row exponents of the two FFT passes (may be NULL). */
int pfa_form_image_fixed(pfa_plan *plan, const int16_t *sig_iq, uint16_t *out,
                         int32_t *exp_a, int32_t *exp_b);
int pfa_form_image_fixed_dev(pfa_plan *plan, const int16_t *sig_iq, uint16_t *out,
                             int32_t *exp_a, int32_t *exp_b);
/* pfa_form_image_fixed_dev without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * intermediates live in a static device array of the library (about 32 MB
 * per GPU, held until the process ends), so grid <= PFA_NOALLOC_MAX_GRID
 * (else PFA_ERR_DIMS); the result equals pfa_form_image_fixed_dev's bit for
 * bit.  Calls on one device share that array: calls that overlap on one GPU
 * (two streams, two host threads) race.  No timer events are recorded. */
#define PFA_NOALLOC_MAX_GRID 2048
int pfa_form_image_fixed_dev_noalloc(pfa_plan *plan, const int16_t *sig_iq, uint16_t *out,
                                     int32_t *exp_a, int32_t *exp_b);

/* ------------------------------------------------------------------ float mode
 * form_image_pfa.py in FP32.  The O(m + n) scene geometry (form_image_pfa.py
 * :178-193, 205-209: rotated frame, pr, tan(phi), pulse order, KR/KC grids) is
 * prepared by the caller exactly as the reference computes it with numpy; the
 * O(m n) k-space positions kr = (2 f / c) pr and every resample position are
 * computed on the GPU in FP64.
 *   pr[m]          per-pulse range projection (float64)
 *   order[m]       argsort(tan_phi) (pulse index of each sorted position)
 *   tan_s[m]       tan_phi[order] (float64, ascending)
 *   kr_grid[n]     KR (np.linspace over the union extent, finding :190-193)
 *   kc_grid[m]     KC
 *   f0, df [m]     compact frequencies: freq[i][k] = f0[i] + k * df[i] (each
 *                  operation rounded, as sc0 + k*scss builds them); both NULL
 *                  when the full freq[m][n] array is passed at execution
 *   ham_c[m], ham_r[n]  np.hamming(m), np.hamming(n) (NULL: no window)
 *   sinc_tab       256x32 Q15 table, needed with PFA_FLOAT_SINC32
 *   sgn            <0: fft2 (SGN=-1), >0: ifft2 (transform2d :144-147)
 * The image is the zero-padded (to_pow2) centred transform, [cross][range],
 * m2 = max(64, pow2(m)) <= 8192 (the cross line is transformed in LDS with its
 * resample) and n2 = max(64, pow2(n)) <= 16384 (the range transform takes one
 * line per workgroup); larger scenes return PFA_ERR_DIMS. */
#define PFA_FLOAT_SINC32 0x1        /* 32-tap polyphase gather both passes     */
/* Precision form of pfa_focus_f32[_dev] (no reference counterpart; the FP32
 * form above is the reference's arithmetic): both FFT passes as dense
 * DFT-matrix stages on the matrix cores with 16-bit operands and FP32
 * accumulation, per-line power-of-two block scaling, the intermediate stored
 * as 16-bit complex pairs.  Padded sizes m2, n2 in {256, 512, 1024, 4096,
 * 8192}.  pfa_transform_f32 and pfa_resample_kspace_f32 stay FP32. */
#define PFA_FLOAT_FFT_BF16 0x2      /* BF16 operands                           */
#define PFA_FLOAT_FFT_FP16 0x4      /* FP16 operands                           */
#define PFA_FLOAT_FFT_SPLIT 0x8     /* with BF16: DFT matrices as hi + lo parts */
enum { PFA_OUT_C64 = 0, PFA_OUT_ABS_F32 = 1 };
int pfa_plan_float(pfa_ctx *ctx, pfa_plan **plan, int m, int n, int sgn,
                   const double *pr, const int32_t *order, const double *tan_s,
                   const double *kr_grid, const double *kc_grid,
                   const double *f0, const double *df,
                   const double *ham_c, const double *ham_r,
                   const int16_t *sinc_tab, int flags);
/* The O(m + n) scene geometry above from the reference's own arguments, on
 * the host (replaces form_image_pfa.resample_kspace's geometry,
 * form_image_pfa.py:178-193, 205-209, 219-221): the rotated frame from the
 * mean look direction (ndarray.mean's blocked pairwise sum, math.hypot
 * correctly rounded), pr, tan(phi), the pulse order (ties in tan(phi) keep
 * pulse index order), KR over the pulses' common kr overlap (np.linspace),
 * KC, np.hamming tapers, and geo = {dc, dr, dhat_x, dhat_y, chat_x, chat_y}
 * of the unpadded grid.  freq is the full float64 [m][n] array (its O(m n)
 * kr extrema are scanned here), or NULL with compact f0, df [m] (freq[i][k] =
 * f0[i] + k df[i], each operation rounded).  ham_c[m] and ham_r[n] may both be
 * NULL.  Needs no device. */
int pfa_float_geometry(int m, int n, const double *freq, const double *f0, const double *df,
                       const double *ax, const double *ay, double *pr, int32_t *order,
                       double *tan_s, double *kr_grid, double *kc_grid, double *ham_c,
                       double *ham_r, double geo[6]);
/* pfa_plan_float from those arguments (form_image_pfa.pfa's own: freq or
 * f0/df, ax, ay, sgn); window != 0 applies the np.hamming tapers.  geo
 * receives pfa's geo: the spacing scaled to the padded grid (focus,
 * form_image_pfa.py:237-238). */
int pfa_plan_float_pfa(pfa_ctx *ctx, pfa_plan **plan, int m, int n, int sgn,
                       const double *freq, const double *f0, const double *df,
                       const double *ax, const double *ay, int window,
                       const int16_t *sinc_tab, int flags, double geo[6]);
/* form_image_pfa.pfa(sig, freq, ax, ay, sgn) (form_image_pfa.py:243-246) in
 * one synchronous call on host buffers: the windowed np.interp focus (flags
 * may select the FFT form only), out [m2][n2] as out_kind, geo as above. */
int pfa_pfa_f32(pfa_ctx *ctx, const float *sig_c64, int m, int n, const double *freq,
                const double *ax, const double *ay, int sgn, int flags, void *out,
                int out_kind, double geo[6]);
/* pfa (form_image_pfa.py:243-246) / NumpyBackend.focus (accel.py:39-44):
 * sig complex64 [m][n] -> out [m2][n2] complex64 (PFA_OUT_C64) or |img|
 * float32 (PFA_OUT_ABS_F32).  freq: full float64 [m][n] or NULL (compact). */
int pfa_focus_f32(pfa_plan *plan, const float *sig_c64, const double *freq, void *out,
                  int out_kind);
int pfa_focus_f32_dev(pfa_plan *plan, const float *sig_c64, const double *freq, void *out,
                      int out_kind);
/* resample_kspace (form_image_pfa.py:165-222): windowed k-space g2
 * complex64 [m][n] (rows cross, columns range). */
int pfa_resample_kspace_f32(pfa_plan *plan, const float *sig_c64, const double *freq,
                            float *g2_c64);
/* focus (form_image_pfa.py:225-240) of a given g2 complex64 [m][n]. */
int pfa_transform_f32(pfa_plan *plan, const float *g2_c64, void *out, int out_kind);
/* quicklook (form_image_pfa.py:158-162): sig complex64 [m][n] times
 * hamming2d (window != 0; np.hamming tapers, f32(hm[i] hn[k])), then
 * transform2d (:144-147) -- fftshift(fft2(ifftshift(x))) for sgn < 0, ifft2
 * for sgn > 0 -- with no padding: out [m][n] as out_kind.  Any m, n in
 * 1..8192, or powers of two up to 16384 (other lengths run as Bluestein
 * transforms on power-of-two LDS FFTs).  Host pointers (synchronous) or
 * device pointers (_dev, enqueued on the context stream). */
int pfa_quicklook_f32(pfa_ctx *ctx, const float *sig_c64, int m, int n, int sgn, int window,
                      void *out, int out_kind);
int pfa_quicklook_f32_dev(pfa_ctx *ctx, const float *sig_c64, int m, int n, int sgn,
                          int window, void *out, int out_kind);
/* The O(m n) part of the scene geometry (form_image_pfa.py:187-192,
 * sar_pipeline.prepare_tables :92-98): kr = (2 freq / c) pr[:, None] in FP64,
 * each operation rounded as numpy evaluates it, reduced per pulse to
 * row_max[i] = kr[i].max() and row_min[i] = kr[i].min().  The caller forms
 * KR's extent from them (kr.max(axis=1).min(), kr.min(axis=1).max(), with
 * the global extrema as numpy's fallback) -- the same values numpy gives.
 * Host pointers (synchronous; freq staged like the focus inputs) or device
 * pointers (_dev, enqueued on the context stream). */
int pfa_kr_extrema(pfa_ctx *ctx, const double *freq, const double *pr, int m, int n,
                   double *row_max, double *row_min);
int pfa_kr_extrema_dev(pfa_ctx *ctx, const double *freq, const double *pr, int m, int n,
                       double *row_max, double *row_min);

/* ------------------------------------------------------------------ display
 * The passes after detection that turn an image into a product.  Host
 * pointers (synchronous, staged copies) or device pointers (_dev, enqueued on
 * the context stream unless stated).
 *
 * Normal-score stretch of a uint16 magnitude image (sar_display.normal_score,
 * mpfs/host/sar_display.py:24-60):
 *   pfa_hist_u16    hist[65536] = np.bincount(img, minlength=65536) as uint32
 *                   (the counts behind lo, hi and the midpoint CDF, :30-33)
 *   pfa_lut_u16_u8  out[i] = lut[img[i] - lo], lut[nb] uint8, 1 <= nb <= 65536
 *                   (:59-60); values outside [lo, lo + nb) take the nearest
 *                   table end, where numpy would wrap or raise
 * The LUT itself (CDF, inverse normal CDF, clip) is O(nb) host work. */
int pfa_hist_u16(pfa_ctx *ctx, const uint16_t *img, int64_t count, uint32_t *hist);
int pfa_hist_u16_dev(pfa_ctx *ctx, const uint16_t *img, int64_t count, uint32_t *hist);
int pfa_lut_u16_u8(pfa_ctx *ctx, const uint16_t *img, int64_t count, const uint8_t *lut,
                   int lo, int nb, uint8_t *out);
int pfa_lut_u16_u8_dev(pfa_ctx *ctx, const uint16_t *img, int64_t count, const uint8_t *lut,
                       int lo, int nb, uint8_t *out);
/* Order statistics for np.percentile (the 99.7 % clip level of the uint8
 * GeoTIFF product, src/form_image_pfa.py:288): v0 and v1 are the k0-th and
 * k1-th smallest of x[count] (0-based ranks, NaN ordered last as np.partition
 * does; a -0.0 / +0.0 tie may return either zero), nan_count the number of
 * NaNs (may be NULL).  v0, v1 and nan_count are host pointers in both forms;
 * the _dev form reads x from the device and synchronises the stream. */
int pfa_select_f32(pfa_ctx *ctx, const float *x, int64_t count, int64_t k0, int64_t k1,
                   float *v0, float *v1, int64_t *nan_count);
int pfa_select_f32_dev(pfa_ctx *ctx, const float *x, int64_t count, int64_t k0, int64_t k1,
                       float *v0, float *v1, int64_t *nan_count);
/* uint8 product (src/form_image_pfa.py:289): out[i] = uint8(clip(x[i] / d *
 * 255, 0, 255)) with d = hi + 1e-12 formed by the caller in float32; each
 * operation rounded in float32 as numpy rounds it, truncation to uint8, NaN
 * -> 0. */
int pfa_stretch_u8(pfa_ctx *ctx, const float *x, int64_t count, float d, uint8_t *out);
int pfa_stretch_u8_dev(pfa_ctx *ctx, const float *x, int64_t count, float d, uint8_t *out);

/* ------------------------------------------------------------------ autofocus
 * Gradient autofocus on the float path.  No counterpart in the reference: for
 * callers that focus with per-pulse phase corrections phi[m] (original pulse
 * order, radians): sig'[i][k] = sig[i][k] e^{j phi[i]}, I = pfa(sig') (complex,
 * [m2][n2]), sharpness S = sum |I|^4 / (sum |I|^2)^2.  FP32 FFT form only;
 * padded cross size m2 <= 4096.  All pointers are device pointers; work is
 * enqueued on the context stream.
 *   pfa_af_checkpoint  range pass of sig once (the phase commutes with it); its
 *                      corner-turned output stays in a buffer of this plan's
 *                      own (plans on one context never share it)
 *   pfa_af_forward     img[m2][n2] complex64 of the corrected history and
 *                      stats[3] = {sum |I|^2, sum |I|^4, S} (double)
 *   pfa_af_backward    grad[m] = dS/dphi from img and stats of that forward,
 *                      by the adjoints of the azimuth pass, window and both
 *                      FFT passes (the transpose gather as a segmented sum) */
int pfa_af_checkpoint(pfa_plan *plan, const float *sig_c64, const double *freq);
int pfa_af_forward(pfa_plan *plan, const float *phi, float *img_c64, double *stats);
int pfa_af_backward(pfa_plan *plan, const float *phi, const float *img_c64,
                    const double *stats, float *grad);

/* Releases a plan.  Plans are normally destroyed before their context; a
 * plan still alive when pfa_destroy runs loses its device tables there, every
 * later call on it returns PFA_ERR_STATE, and pfa_plan_destroy then frees
 * only the plan object. */
void pfa_plan_destroy(pfa_plan *plan);

#ifdef __cplusplus
}
#endif
#endif /* MI355X_PFA_H */
