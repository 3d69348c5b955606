"""Bit-exact numpy restatement of the reference's fixed-point PFA datapath.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``): the GPU library is
checked against these functions; it never calls them.

Complex fixed-point arrays are carried as a pair ``(re, im)`` of int64 arrays
holding the integer codes.  Every function cites the reference lines whose
arithmetic it restates; the reference does the same arithmetic on
integer-valued float64/complex128 arrays, which is exact for these magnitudes
(products stay below 2**53), so integer shifts and floors reproduce it.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np

C_LIGHT = 299792458.0
INT16_MIN, INT16_MAX = -32768, 32767
SINC_TAPS = 32
SINC_PHASES = 256
SINC_LO = -(SINC_TAPS // 2 - 1)          # first tap offset relative to idx (-15)


# --------------------------------------------------------------------------- helpers
def sat16(x):
    """clip to int16 -- silicon_emulator.py:51-52 (_sat16) on integer arrays."""
    return np.clip(x, INT16_MIN, INT16_MAX)


def wrap16(x):
    """C (int16_t) cast: keep the low 16 bits -- silicon_emulator.py:159-162."""
    x = np.asarray(x, dtype=np.int64) & 0xFFFF
    return np.where(x >= 0x8000, x - 0x10000, x)


def to_iq(re, im, dtype=np.int16):
    """Pack (re, im) codes into an interleaved [..., 2] array (I first)."""
    out = np.empty(np.shape(re) + (2,), dtype=dtype)
    out[..., 0] = re
    out[..., 1] = im
    return out


def from_iq(iq):
    iq = np.asarray(iq)
    return iq[..., 0].astype(np.int64), iq[..., 1].astype(np.int64)


def bitrev_perm(n):
    """fixedpoint.py:44-46: index i -> i with its log2(n) bits reversed."""
    bits = n.bit_length() - 1
    idx = np.arange(n, dtype=np.int64)
    rev = np.zeros(n, dtype=np.int64)
    for b in range(bits):
        rev |= ((idx >> b) & 1) << (bits - 1 - b)
    return rev


def twiddle_codes(m, nbits_tw):
    """Truncated twiddles W_m^k, k < m/2, as integer codes at 2**-(nbits_tw-1).

    fixedpoint.py:58-60 / :281-282: quant(exp(-2j*pi*k/m), 2**-(b-1), b, 'trunc'),
    i.e. floor(cos * 2**(b-1)) and floor(-sin * 2**(b-1)), clipped to b bits.
    The angle is evaluated exactly as numpy evaluates it there.
    """
    w = np.exp(-2j * np.pi * np.arange(m // 2) / m)
    full = (1 << (nbits_tw - 1)) - 1
    scale = float(1 << (nbits_tw - 1))
    wr = np.clip(np.floor(w.real * scale), -full - 1, full).astype(np.int64)
    wi = np.clip(np.floor(w.imag * scale), -full - 1, full).astype(np.int64)
    return wr, wi


def fit_exp_exact(mag_max):
    """fixedpoint.py:19-26 exponent: the smallest E with mag_max <= 32767 * 2**E.

    ceil(log2(M/32767)) in float64 equals this exact rule for every float32 M
    (the quotient is never within float64 rounding of a power of two unless it
    is one), so the rule is stated in exact arithmetic here."""
    if mag_max == 0.0:
        return 0
    e = math.ceil(math.log2(mag_max / 32767.0))
    while mag_max > 32767.0 * math.ldexp(1.0, e):
        e += 1
    while mag_max <= 32767.0 * math.ldexp(1.0, e - 1):
        e -= 1
    return e


# --------------------------------------------------------------------------- F1
def quantize_iq16(signal):
    """silicon_emulator.py:227-229 with fixedpoint.fit_scale (:19-26).

    signal: complex64 (m, n).  Returns (re, im, exp): int16-valued codes
    floor(x / 2**exp) saturated to int16, and the block exponent."""
    sig = np.asarray(signal, dtype=np.complex64)
    mx = max(float(np.abs(sig.real).max(initial=0.0)), float(np.abs(sig.imag).max(initial=0.0)))
    e = fit_exp_exact(mx)
    lsb = np.float32(math.ldexp(1.0, e))
    re = sat16(np.floor(sig.real / lsb).astype(np.int64))
    im = sat16(np.floor(sig.imag / lsb).astype(np.int64))
    return re, im, e


# --------------------------------------------------------------------------- F2
def fixed_tables(ax, ay, freq, KR, KC, tan_phi, m, n, grid):
    """The per-scene tables of silicon_emulator.py:128-140 (the board's DDR roles).

    Returns dict f0, df, pr (float32 [m]), tan_s (float32 [m]), order,
    inv_order (int64 [m]), krp, kcp (float32 [grid])."""
    ax = np.asarray(ax, np.float64); ay = np.asarray(ay, np.float64)
    freq = np.asarray(freq, np.float64)
    KR = np.asarray(KR, np.float64); KC = np.asarray(KC, np.float64)
    tan_phi = np.asarray(tan_phi, np.float64)
    f0 = freq[:, 0].astype(np.float32)
    df = (freq[:, 1] - freq[:, 0]).astype(np.float32)
    dx, dy = ax.mean(), ay.mean()
    dn = math.hypot(dx, dy)
    pr = (ax * (dx / dn) + ay * (dy / dn)).astype(np.float32)
    order = np.argsort(tan_phi)
    tan_s = tan_phi[order].astype(np.float32)
    inv_order = np.argsort(order)
    oob_r = float(KR.max() + (KR.max() - KR.min()) + 1.0)
    oob_c = float(KC.max() + (KC.max() - KC.min()) + 1.0)
    krp = np.full(grid, oob_r, np.float32); krp[:n] = KR
    kcp = np.full(grid, oob_c, np.float32); kcp[:m] = KC
    return dict(f0=f0, df=df, pr=pr, tan_s=tan_s, order=order, inv_order=inv_order,
                krp=krp, kcp=kcp)


def range_positions(pr, f0, df, n):
    """Pass-1 source positions, silicon_emulator.py:144 under NumPy>=2 promotion:
    a = f32(f32(2*pr)/f32(C)); x_j = f64(a) * (f64(f0) + j * f64(df))."""
    pr = np.asarray(pr, np.float32)
    a = (np.float32(2.0) * pr) / np.float32(C_LIGHT)
    j = np.arange(n, dtype=np.float64)
    return (a.astype(np.float64)[:, None]
            * (np.asarray(f0, np.float32).astype(np.float64)[:, None]
               + j[None, :] * np.asarray(df, np.float32).astype(np.float64)[:, None]))


# --------------------------------------------------------------------------- F3/F6
def interp_coeffs(query, xp):
    """serialize_inputs.py:67-94: bracket + Q15 weight for a 2-tap resample.

    idx is in xp's natural order; out-of-range queries give idx=-1, wq=0."""
    xp = np.asarray(xp, np.float64)
    q = np.asarray(query, np.float64)
    n = xp.size
    ascending = xp[-1] >= xp[0]
    xa = xp if ascending else xp[::-1]
    k = np.searchsorted(xa, q, side="left") - 1
    valid = (k >= 0) & (k < n - 1)
    kk = np.clip(k, 0, n - 2)
    frac = (q - xa[kk]) / (xa[kk + 1] - xa[kk])
    if ascending:
        idx, w = kk, frac
    else:
        idx, w = (n - 2 - kk), 1.0 - frac
    idx = np.where(valid, idx, -1).astype(np.int64)
    wq = np.where(valid, np.clip(np.round(w * 32768.0), 0, 32767), 0).astype(np.int64)
    return idx, wq


# --------------------------------------------------------------------------- F4
def sinc_table_q15(phases=SINC_PHASES, taps=SINC_TAPS):
    """DC-normalised polyphase sinc in Q15, [phase][tap].

    Restates gen_resample_vectors.py:206-254: exact rational normalisation of
    sin(pi d)/(pi d) (libm sin on the float of d), round-half-up to Q15, clamp
    to int16, then spread the residual to reach exactly 2**15 starting from the
    largest-magnitude tap (stable order) within each tap's int16 headroom."""
    lo = -(taps // 2 - 1)
    half = Fraction(1, 2)
    out = np.zeros((phases, taps), dtype=np.int64)
    for ph in range(phases):
        mu = Fraction(ph, phases)
        raw = []
        for t in range(taps):
            d = mu - (lo + t)
            if d == 0:
                raw.append(Fraction(1))
            else:
                raw.append(Fraction(math.sin(math.pi * float(d))) / (Fraction(math.pi) * d))
        total = sum(raw)
        q = [math.floor(c / total * (1 << 15) + half) for c in raw]
        q = [min(max(v, INT16_MIN), INT16_MAX) for v in q]
        resid = (1 << 15) - sum(q)
        for t in sorted(range(taps), key=lambda u: -abs(q[u])):
            if resid == 0:
                break
            room = (INT16_MAX - q[t]) if resid > 0 else (INT16_MIN - q[t])
            step = resid if abs(resid) <= abs(room) else room
            q[t] += step
            resid -= step
        assert resid == 0
        out[ph] = q
    return out


def _prod(d, w, f32):
    """(b-a)*wq as the emulator evaluates it: exact for complex128 input, rounded
    to float32 (24-bit significand, nearest-even) when the line is complex64."""
    p = d * w
    return p if not f32 else p.astype(np.float32).astype(np.int64)


def lerp_fixed(xr, xi, idx, wq, f32=False):
    """silicon_emulator.py:60-71 (apply_fixed): a + floor((b-a)*wq / 2**15), sat16;
    idx<0 -> 0.  ``f32`` selects the float32 product form_image's complex64
    signal gets in pass 1 (silicon_emulator.py:226-229, :146-147)."""
    sn = xr.size
    v = idx >= 0
    j = np.clip(idx, 0, sn - 2)
    outr = sat16(xr[j] + (_prod(xr[j + 1] - xr[j], wq, f32) >> 15))
    outi = sat16(xi[j] + (_prod(xi[j + 1] - xi[j], wq, f32) >> 15))
    return np.where(v, outr, 0), np.where(v, outi, 0)


def sinc_fixed(xr, xi, idx, wq, sn, table):
    """silicon_emulator.py:88-122 (apply_sinc) = gen_resample_vectors.gather_sinc.

    Interior brackets (15 <= idx <= sn-17) use the 32 taps idx-15..idx+16 with
    phase wq>>7 and round-half-up >>15; other valid brackets use the 2-tap lerp."""
    outr = np.zeros(idx.shape, np.int64)
    outi = np.zeros(idx.shape, np.int64)
    interior = (idx >= -SINC_LO) & (idx <= sn - (SINC_TAPS + SINC_LO))
    edge = (idx >= 0) & ~interior
    if interior.any():
        k = idx[interior]
        cols = k[:, None] + np.arange(SINC_LO, SINC_LO + SINC_TAPS)[None, :]
        c = table[wq[interior] >> 7]
        accr = (c * xr[cols]).sum(axis=1)
        acci = (c * xi[cols]).sum(axis=1)
        outr[interior] = sat16((accr + (1 << 14)) >> 15)
        outi[interior] = sat16((acci + (1 << 14)) >> 15)
    if edge.any():
        k = np.clip(idx[edge], 0, xr.size - 2)
        w = wq[edge]
        outr[edge] = sat16(xr[k] + (((xr[k + 1] - xr[k]) * w) >> 15))
        outi[edge] = sat16(xi[k] + (((xi[k + 1] - xi[k]) * w) >> 15))
    return outr, outi


def resample_fixed(sig_re, sig_im, tabs, grid, range_sinc=False, az_sinc=False,
                   table=None, pass1_f32=True):
    """silicon_emulator.py:125-156: pass-1 range resample into pulse-sorted rows,
    corner-turn, pass-2 azimuth resample.  Returns g2 (re, im) [grid range][grid cross].

    pass1_f32=True is form_image's contract (complex64 quantised signal, so the
    pass-1 lerp product is a float32 product); False is the stage function fed
    a complex128 signal.  The sinc path and pass 2 are exact either way."""
    m, n = sig_re.shape
    if (range_sinc or az_sinc) and table is None:
        table = sinc_table_q15()
    x = range_positions(tabs["pr"], tabs["f0"], tabs["df"], n)
    krp = tabs["krp"].astype(np.float64)
    s_re = np.zeros((grid, grid), np.int64)       # [sorted pulse][range]
    s_im = np.zeros((grid, grid), np.int64)
    inv = tabs["inv_order"]
    for i in range(m):
        idx, wq = interp_coeffs(krp, x[i])
        if range_sinc:
            r, q = sinc_fixed(sig_re[i], sig_im[i], idx, wq, n, table)
        else:
            r, q = lerp_fixed(sig_re[i], sig_im[i], idx, wq, f32=pass1_f32)
        s_re[inv[i]] = r
        s_im[inv[i]] = q
    t_re = np.ascontiguousarray(s_re.T)            # [range][sorted pulse]
    t_im = np.ascontiguousarray(s_im.T)
    del s_re, s_im
    g_re = np.zeros((grid, grid), np.int64)
    g_im = np.zeros((grid, grid), np.int64)
    kcp = tabs["kcp"].astype(np.float64)
    for j in range(grid):
        src = (tabs["krp"][j] * tabs["tan_s"]).astype(np.float32)     # float32 product
        idx, wq = interp_coeffs(kcp, src)
        if az_sinc:
            r, q = sinc_fixed(t_re[j, :m], t_im[j, :m], idx, wq, m, table)
        else:
            r, q = lerp_fixed(t_re[j, :m], t_im[j, :m], idx, wq)
        g_re[j] = r
        g_im[j] = q
    return g_re, g_im


# --------------------------------------------------------------------------- F8
def hamming_q15(length, extent):
    """floor(np.hamming(length) * 32768) in an extent-long int array, zero beyond
    `length` (silicon_emulator.py:185-188).  May hold 32768 at an odd centre."""
    h = np.zeros(extent)
    h[:length] = np.hamming(length)
    return np.floor(h * 32768).astype(np.int64)


def window_fixed(g_re, g_im, m, n, hr=None, hc=None):
    """silicon_emulator.py:165-195: cw = wrap16((hr[j]*hc[k])>>15);
    out = wrap16(floor(x*cw / 2**15)) per component, rows = range j, cols = cross k."""
    Np, Mp = g_re.shape
    hr = hamming_q15(n, Np) if hr is None else np.asarray(hr, np.int64)
    hc = hamming_q15(m, Mp) if hc is None else np.asarray(hc, np.int64)
    cw = wrap16((hr[:, None] * hc[None, :]) >> 15)
    return wrap16((g_re * cw) >> 15), wrap16((g_im * cw) >> 15)


# --------------------------------------------------------------------------- F9
def bfp_shift_rule(mx):
    """fixedpoint.py:284: max(0, ceil(log2((mx+1)/32767))) for integer mx <= 65536."""
    return np.where(mx <= 32766, 0, np.where(mx <= 65533, 1, 2)).astype(np.int64)


def fft_bfp_perrow(x_re, x_im, nbits_tw=16):
    """fixedpoint.py:266-289 (fft1d_bfp_hw_perrow, nbits=16) on the last axis.

    Radix-2 DIT over bit-reversed input; each stage: t = sat16((w*b)>>15),
    a+-t, per-row shift by the max rule, floor, sat16.  Returns (re, im, e_row)."""
    N = x_re.shape[-1]
    stages = N.bit_length() - 1
    perm = bitrev_perm(N)
    xr = sat16(x_re[..., perm].astype(np.int64))
    xi = sat16(x_im[..., perm].astype(np.int64))
    lead = xr.shape[:-1]
    e = np.zeros(lead, np.int64)
    sh_tw = nbits_tw - 1
    for s in range(1, stages + 1):
        m = 1 << s
        h = m >> 1
        wr, wi = twiddle_codes(m, nbits_tw)
        ar = xr.reshape(lead + (N // m, m))[..., :h]
        ai = xi.reshape(lead + (N // m, m))[..., :h]
        br = xr.reshape(lead + (N // m, m))[..., h:]
        bi = xi.reshape(lead + (N // m, m))[..., h:]
        tr = sat16((wr * br - wi * bi) >> sh_tw)
        ti = sat16((wr * bi + wi * br) >> sh_tw)
        yr = np.concatenate([ar + tr, ar - tr], axis=-1).reshape(lead + (N,))
        yi = np.concatenate([ai + ti, ai - ti], axis=-1).reshape(lead + (N,))
        mx = np.maximum(np.abs(yr).max(axis=-1), np.abs(yi).max(axis=-1))
        sh = bfp_shift_rule(mx)
        xr = sat16(yr >> sh[..., None])
        xi = sat16(yi >> sh[..., None])
        e += sh
    return xr, xi, e


def fft_pass_bfp(x_re, x_im):
    """silicon_emulator.py:198-205: per-row BFP FFT then global renorm
    y >> (emax - e_row), sat16.  Returns (re, im, e_row)."""
    yr, yi, e = fft_bfp_perrow(x_re, x_im)
    emax = int(e.max()) if e.size else 0
    sh = (emax - e)[..., None]
    return sat16(yr >> sh), sat16(yi >> sh), e


# --------------------------------------------------------------------------- F11
def detect_fixed(z_re, z_im):
    """silicon_emulator.py:208-213: floor(sqrt(re^2+im^2)) of sat16 codes, uint16."""
    re = sat16(z_re); im = sat16(z_im)
    m = np.floor(np.sqrt((re * re + im * im).astype(np.float64)))
    return np.clip(m, 0, 0xFFFF).astype(np.uint16)


# --------------------------------------------------------------------------- F0
def form_image_fixed(sig_re, sig_im, tabs, grid, range_sinc=False, az_sinc=False,
                     hr=None, hc=None, table=None):
    """silicon_emulator.py:216-253 after quantisation: resample -> window ->
    BFP FFT (cross axis, per range row) -> corner-turn -> BFP FFT (range axis)
    -> corner-turn back -> detect.  Returns (uint16 [range][cross], e_a, e_b)."""
    m, n = sig_re.shape
    g_re, g_im = resample_fixed(sig_re, sig_im, tabs, grid, range_sinc, az_sinc, table)
    w_re, w_im = window_fixed(g_re, g_im, m, n, hr, hc)
    del g_re, g_im
    a_re, a_im, ea = fft_pass_bfp(w_re, w_im)
    del w_re, w_im
    b_re, b_im, eb = fft_pass_bfp(np.ascontiguousarray(a_re.T), np.ascontiguousarray(a_im.T))
    del a_re, a_im
    img = detect_fixed(b_re.T, b_im.T)
    return np.ascontiguousarray(img), ea, eb


# --------------------------------------------------------------------------- fixture sibling
def fft_bfp_blockwise(codes_re, codes_im, e_in, nbits=16, nbits_tw=18):
    """fixedpoint.py:49-72 (fft1d_bfp) on one line given its input codes at 2**e_in.

    Block-floating-point after every stage via fit_scale/quant(trunc) over the
    line, twiddles truncated to nbits_tw.  Returns (codes_re, codes_im, exps)
    with exps[0] the input block exponent as in the reference manifest."""
    full = (1 << (nbits - 1)) - 1
    N = codes_re.size
    stages = N.bit_length() - 1
    perm = bitrev_perm(N)
    xr = [int(v) for v in np.asarray(codes_re, np.int64)[perm]]
    xi = [int(v) for v in np.asarray(codes_im, np.int64)[perm]]
    # the input codes are the reference's quant(x, 2**e_in) of the stimulus
    e = int(e_in)
    exps = [e]
    tw = nbits_tw - 1
    for s in range(1, stages + 1):
        m = 1 << s
        h = m >> 1
        wr, wi = twiddle_codes(m, nbits_tw)
        wr = [int(v) for v in wr]; wi = [int(v) for v in wi]
        nr = [0] * N; ni = [0] * N
        for base in range(0, N, m):
            for j in range(h):
                ar, ai = xr[base + j] << tw, xi[base + j] << tw
                br, bi = xr[base + j + h], xi[base + j + h]
                tr = wr[j] * br - wi[j] * bi
                ti = wr[j] * bi + wi[j] * br
                nr[base + j] = ar + tr; ni[base + j] = ai + ti
                nr[base + j + h] = ar - tr; ni[base + j + h] = ai - ti
        e, xr, xi = _refit(nr, ni, e - tw, full)
        exps.append(e)
    return np.array(xr, np.int64), np.array(xi, np.int64), exps


def _refit(vr, vi, scale_exp, full):
    """fit_scale + quant(trunc) (fixedpoint.py:19-41) for integer values v * 2**scale_exp:
    the new exponent is the smallest E with max|v| * 2**scale_exp <= full * 2**E."""
    vmax = max(max(abs(v) for v in vr), max(abs(v) for v in vi))
    if vmax == 0:
        return 0, [0] * len(vr), [0] * len(vi)

    def fits(d):
        return vmax <= (full << d) if d >= 0 else (vmax << -d) <= full
    d = 0
    while not fits(d):
        d += 1
    while fits(d - 1):
        d -= 1

    def q(v):
        c = (v >> d) if d >= 0 else (v << -d)
        return min(max(c, -full - 1), full)
    return scale_exp + d, [q(v) for v in vr], [q(v) for v in vi]
