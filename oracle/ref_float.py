"""numpy restatement of the reference's float Polar Format focus.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``geometry``         -- form_image_pfa.py:177-193, 205-209, 219-221
* ``resample_kspace``  -- form_image_pfa.py:165-222 (LINEAR = np.interp, the
                          reference's interpolator) plus the SINC32 variant
* ``focus``            -- form_image_pfa.py:225-240 with transform2d :144-147
* ``pfa``              -- form_image_pfa.py:243-246
* ``quicklook``        -- form_image_pfa.py:154-162 (hamming2d + transform2d)
* ``prepare_tables``   -- sar_pipeline.py:36-115 (PVP geometry -> focus inputs,
                          KR / KC grids, taper, pixel geometry; O(m n) in numpy)

``store`` selects the intermediate precision: complex64 reproduces the
reference; complex128 gives the float64 oracle that the reduced-precision
forms are scored against.

SINC32 float mode has no reference function.  It is defined here as the
silicon's 32-tap polyphase gather (gen_resample_vectors.py:257-296) carried in
floating point: the np.interp bracket J and fraction f of each output, the Q15
tap row ``table[min(rint(f*32768), 32767) >> 7] / 2**15`` over samples
J-15..J+16 where 15 <= J <= S-17, and np.interp's linear formula elsewhere.
"""
from __future__ import annotations

import math

import numpy as np

from . import ref_fixed

C_LIGHT = 299792458.0


def to_pow2(n):
    """form_image_pfa.py:150-151."""
    return 1 << int(math.ceil(math.log2(n)))


def geometry(freq, ax, ay):
    """Rotated frame, KR/KC grids and pulse order (form_image_pfa.py:177-209)."""
    freq = np.asarray(freq, np.float64)
    ax = np.asarray(ax, np.float64); ay = np.asarray(ay, np.float64)
    m, n = freq.shape
    kmag = 2.0 * freq / C_LIGHT
    dx, dy = ax.mean(), ay.mean()
    dn = math.hypot(dx, dy)
    dx, dy = dx / dn, dy / dn
    cx, cy = -dy, dx
    pr = ax * dx + ay * dy
    pc = ax * cx + ay * cy
    kr = kmag * pr[:, None]
    tan_phi = pc / pr
    lo, hi = kr.max(axis=1).min(), kr.min(axis=1).max()
    if lo > hi:                      # overlapping rows: falls back to the union extent
        lo, hi = kr.min(), kr.max()
    KR = np.linspace(lo, hi, n)
    kc_max = np.abs(KR[:, None] * tan_phi[None, :]).max()
    KC = np.linspace(-kc_max, kc_max, m)
    order = np.argsort(tan_phi)
    dr = 1.0 / (KR[-1] - KR[0]) if KR[-1] != KR[0] else float("nan")
    dc = 1.0 / (KC[-1] - KC[0]) if KC[-1] != KC[0] else float("nan")
    geo = {"dc": dc, "dr": dr, "dhat": (dx, dy), "chat": (cx, cy)}
    return dict(kr=kr, tan_phi=tan_phi, KR=KR, KC=KC, order=order,
                tan_s=tan_phi[order], pr=pr, pc=pc, geo=geo)


def _sinc_line(query, xp, fp, table):
    """SINC32 float gather of one line (module docstring); xp ascending, fp complex."""
    xp = np.asarray(xp, np.float64)
    q = np.asarray(query, np.float64)
    S = xp.size
    J = np.searchsorted(xp, q, side="right") - 1
    out = np.zeros(q.shape, np.complex128)
    inside = (J >= 0) & (q <= xp[-1])
    last = inside & (J == S - 1)
    out[last] = fp[S - 1]
    mid = inside & (J < S - 1)
    Jm = J[mid]
    x0, x1 = xp[Jm], xp[Jm + 1]
    f0, f1 = fp[Jm], fp[Jm + 1]
    frac = (q[mid] - x0) / (x1 - x0)
    interior = (Jm >= 15) & (Jm <= S - 17)
    lin = np.empty(Jm.shape, np.complex128)
    slope_r = (f1.real - f0.real) / (x1 - x0)
    slope_i = (f1.imag - f0.imag) / (x1 - x0)
    lin.real = slope_r * (q[mid] - x0) + f0.real
    lin.imag = slope_i * (q[mid] - x0) + f0.imag
    if interior.any():
        wq = np.minimum(np.rint(frac[interior] * 32768.0), 32767).astype(np.int64)
        c = table[wq >> 7].astype(np.float64) / 32768.0
        cols = Jm[interior][:, None] + np.arange(-15, 17)[None, :]
        lin[interior] = (c * fp[cols]).sum(axis=1)
    out[mid] = lin
    return out


def _interp_line(query, xp, fp, interp, table):
    if interp == "linear":
        return (np.interp(query, xp, fp.real, left=0, right=0)
                + 1j * np.interp(query, xp, fp.imag, left=0, right=0))
    return _sinc_line(query, xp, fp.astype(np.complex128), table)


def resample_kspace(sig, freq, ax, ay, window=True, interp="linear",
                    store=np.complex64, table=None):
    """form_image_pfa.py:165-222.  Returns (g2 (m, n) [cross][range], geo)."""
    sig = np.asarray(sig)
    m, n = sig.shape
    g = geometry(freq, ax, ay)
    if interp == "sinc32" and table is None:
        table = ref_fixed.sinc_table_q15()
    kr, KR, KC = g["kr"], g["KR"], g["KC"]
    g1 = np.empty((m, n), dtype=store)
    for i in range(m):
        xp = kr[i]
        row = sig[i]
        if xp[0] > xp[-1]:
            xp = xp[::-1]; row = row[::-1]
        g1[i] = _interp_line(KR, xp, row, interp, table)
    order = g["order"]
    tan_s = g["tan_s"]
    g1s = g1[order]
    g2 = np.empty((m, n), dtype=store)
    for j in range(n):
        src = KR[j] * tan_s
        g2[:, j] = _interp_line(KC, src, g1s[:, j], interp, table)
    if window:
        w = np.outer(np.hamming(m), np.hamming(n))
        if store == np.complex64:
            w = w.astype(np.float32)
        g2 = g2 * w
    return g2, dict(g["geo"])


def focus(g2, geo, sgn, fft_dtype=None):
    """form_image_pfa.py:225-240: zero-pad to pow2 (top-left), centred 2-D FFT
    (fft2 for sgn < 0, ifft2 otherwise), pixel spacing scaled to the padded grid."""
    m, n = g2.shape
    m2, n2 = to_pow2(m), to_pow2(n)
    dt = g2.dtype if fft_dtype is None else fft_dtype
    pad = np.zeros((m2, n2), dtype=dt)
    pad[:m, :n] = g2
    fwd = np.fft.fft2 if sgn < 0 else np.fft.ifft2
    img = np.fft.fftshift(fwd(np.fft.ifftshift(pad)))
    geo = dict(geo)
    geo["dr"] *= n / n2
    geo["dc"] *= m / m2
    return img, geo


def pfa(sig, freq, ax, ay, sgn, window=True, interp="linear", store=np.complex64):
    """form_image_pfa.py:243-246."""
    g2, geo = resample_kspace(sig, freq, ax, ay, window=window, interp=interp, store=store)
    return focus(g2, geo, sgn)


def quicklook(sig, sgn, window=True):
    """form_image_pfa.py:158-162: the raw history times hamming2d (:154-155, the
    float32 outer product of np.hamming) when window is set, then transform2d
    (:144-147) with no padding."""
    sig = np.asarray(sig)
    if window:
        sig = sig * np.outer(np.hamming(sig.shape[0]), np.hamming(sig.shape[1])).astype(np.float32)
    fwd = np.fft.fft2 if sgn < 0 else np.fft.ifft2
    return np.fft.fftshift(fwd(np.fft.ifftshift(sig)))


def detect(img):
    """accel.py:44: |img| as float32."""
    return np.abs(img).astype(np.float32)


def prepare_tables(reader, meta, deci_pulse=1, deci_sample=1, crop_sample=0):
    """sar_pipeline.py:36-115: read the PVP geometry of a CPHD reader and form
    the focus inputs (ax, ay, the (m, n) freq array, sgn), the resample grids
    KR / KC, tan_phi, the float32 outer Hamming taper and the pixel geometry.
    Every array is built the way numpy builds it there, including the O(m n)
    kr array whose row extrema set KR's extent, and ``np.hypot`` for the mean
    look direction's norm (form_image_pfa.py:181 uses ``math.hypot``)."""
    mu, nu = max(1, deci_pulse), max(1, deci_sample)
    n_vec, n_samp = reader.get_data_size_as_tuple()[0]
    sgn = int(meta.Global.SGN)
    tx = reader.read_pvp_variable("TxPos", 0)[::mu]
    rcv = reader.read_pvp_variable("RcvPos", 0)[::mu]
    srp = reader.read_pvp_variable("SRPPos", 0)[::mu]
    sc0 = reader.read_pvp_variable("SC0", 0)[::mu]
    scss = reader.read_pvp_variable("SCSS", 0)[::mu]
    apc = 0.5 * (tx + rcv)
    planar = meta.SceneCoordinates.ReferenceSurface.Planar
    uiax = np.array([planar.uIAX.X, planar.uIAX.Y, planar.uIAX.Z])
    uiay = np.array([planar.uIAY.X, planar.uIAY.Y, planar.uIAY.Z])
    los = srp - apc
    u = los / np.linalg.norm(los, axis=1, keepdims=True)
    ax, ay = u @ uiax, u @ uiay
    m, n = len(range(0, n_vec, mu)), len(range(0, n_samp, nu))
    d0 = 0
    if crop_sample and crop_sample < n:
        d0 = (n - crop_sample) // 2
        n = crop_sample
    k = d0 + np.arange(n)
    freq = sc0[:, None] + k[None, :] * (scss[:, None] * nu)
    kmag = 2.0 * freq / C_LIGHT
    dx, dy = ax.mean(), ay.mean()
    dn = np.hypot(dx, dy)
    dx, dy = dx / dn, dy / dn
    cx, cy = -dy, dx
    pr = ax * dx + ay * dy
    pc = ax * cx + ay * cy
    kr = kmag * pr[:, None]
    tan_phi = pc / pr
    lo, hi = kr.max(axis=1).min(), kr.min(axis=1).max()
    if lo > hi:
        lo, hi = kr.min(), kr.max()
    KR = np.linspace(lo, hi, n)
    kc_max = np.abs(KR[:, None] * tan_phi[None, :]).max()
    KC = np.linspace(-kc_max, kc_max, m)
    window = np.outer(np.hamming(m), np.hamming(n)).astype(np.float32)
    geo = {"dc": 1.0 / (KC[-1] - KC[0]), "dr": 1.0 / (KR[-1] - KR[0]),
           "dhat": (dx, dy), "chat": (cx, cy)}
    return {"ax": ax, "ay": ay, "freq": freq, "sgn": sgn, "KR": KR, "KC": KC,
            "tan_phi": tan_phi, "window": window, "geo": geo, "center_ecef": srp[0],
            "uiax": uiax, "uiay": uiay, "dims": (m, n), "deci": (mu, nu), "n_vec": n_vec,
            "n_samp": n_samp, "samp0": d0 * nu}
