"""GPU drop-ins for the product passes after detection.

* ``normal_score(img, sigma, exclude_zero)``  mpfs/host/sar_display.py:24-60
  (uint16 magnitudes, the fixed path's output type)
* ``percentile(amp, q)``                       np.percentile of a float32 image
* ``stretch_u8(amp, pct)``                     the uint8 data of
  save_detected_geotiff, src/form_image_pfa.py:286-291

The O(N) passes run in libmi355x_pfa.so: the 65536-bin histogram and the LUT
gather of the normal score, the radix select of the percentile's order
statistics and the uint8 conversion.  The O(bins) steps between them (the
midpoint CDF, the inverse normal CDF, numpy's percentile interpolation) are
the reference's own numpy / scipy operations on the host, so every output
equals the reference's bit for bit.  Inputs are numpy arrays (host entry
points) or CUDA tensors (device entry points, outputs on the device).
"""
from __future__ import annotations

import ctypes as C
from math import sqrt

import numpy as np

from . import _lib
from ._lib import PfaError, context, ptr


def _is_dev(a):
    return hasattr(a, "is_cuda") and a.is_cuda


def _flat(a, dtype):
    if _is_dev(a):
        import torch
        if dtype is np.uint16 and a.dtype == torch.int16:
            a = a.view(torch.uint16)                  # 16-bit magnitudes held as int16
        want = {np.uint16: torch.uint16, np.float32: torch.float32}[dtype]
        if a.dtype != want:
            raise PfaError(-1, f"expected a {want} tensor, got {a.dtype}")
        return a.contiguous().reshape(-1)
    a = np.asarray(a)
    if a.dtype != dtype:
        raise PfaError(-1, f"expected a {np.dtype(dtype)} array, got {a.dtype}")
    return np.ascontiguousarray(a).reshape(-1)


def _sync_torch(a):
    if _is_dev(a):
        import torch
        torch.cuda.synchronize(a.device)


def histogram_u16(img, device=0):
    """np.bincount(img.ravel(), minlength=65536) of a uint16 image, as uint32
    (numpy array; computed on the GPU)."""
    flat = _flat(img, np.uint16)
    ctx = context(device)
    hist = np.zeros(65536, np.uint32)
    if _is_dev(flat):
        import torch
        dh = torch.empty(65536, dtype=torch.int32, device=flat.device)
        _sync_torch(flat)
        ctx.check(_lib.lib().pfa_hist_u16_dev(ctx.handle, ptr(flat), flat.numel(), ptr(dh)),
                  "hist_u16")
        ctx.synchronize()
        hist[:] = dh.cpu().numpy().view(np.uint32)
    else:
        ctx.check(_lib.lib().pfa_hist_u16(ctx.handle, ptr(flat), flat.size, ptr(hist)),
                  "hist_u16")
    return hist


def _erf(x):
    import math
    return math.erf(x)


def normal_score_lut(hist, sigma=3.0, exclude_zero=True):
    """The LUT of sar_display.normal_score (sar_display.py:30-58) from the
    value histogram of a non-empty image: returns (lut uint8[nb], lo, blank)
    with nb = hi - lo + 1; blank marks an image with nothing but zero-fill
    (the reference then returns all zeros)."""
    nz = np.flatnonzero(hist)
    lo, hi = int(nz[0]), int(nz[-1])
    nb = hi - lo + 1
    counts = np.asarray(hist[lo:hi + 1], np.int64).astype(np.float64)
    if exclude_zero and lo <= 0 < lo + nb:
        counts[0 - lo] = 0.0
    tot = counts.sum()
    if tot <= 0:
        return np.zeros(nb, np.uint8), lo, True
    cdf = (np.cumsum(counts) - 0.5 * counts) / tot
    cdf = np.clip(cdf, 1e-9, 1 - 1e-9)
    try:
        from scipy.special import erfinv
        z = sqrt(2.0) * erfinv(2.0 * cdf - 1.0)
    except Exception:
        g = np.linspace(-8.0, 8.0, 200001)
        gc = 0.5 * (1.0 + np.vectorize(_erf)(g / sqrt(2.0)))
        z = np.interp(cdf, gc, g)
    lut = np.clip(128.0 + 127.0 * z / sigma, 0, 255).astype(np.uint8)
    if exclude_zero and lo <= 0 < lo + nb:
        lut[0 - lo] = 0
    return lut, lo, False


def normal_score(img, sigma=3.0, exclude_zero=True, device=0):
    """sar_display.normal_score for a uint16 image: uint8 of the same shape
    whose histogram is Gaussian (each pixel's midpoint rank through the
    inverse normal CDF, sigma standard deviations over 0..255, zero-fill
    excluded from the ranking and kept black)."""
    shape = tuple(img.shape)
    flat = _flat(img, np.uint16)
    if flat.shape[0] == 0:
        raise PfaError(-1, "normal_score of an empty image")   # numpy: min() of an empty array
    hist = histogram_u16(flat, device)
    lut, lo, blank = normal_score_lut(hist, sigma, exclude_zero)
    ctx = context(device)
    if _is_dev(flat):
        import torch
        if blank:
            return torch.zeros(shape, dtype=torch.uint8, device=flat.device)
        out = torch.empty(flat.shape[0], dtype=torch.uint8, device=flat.device)
        dl = torch.from_numpy(np.ascontiguousarray(lut)).to(flat.device)
        _sync_torch(dl)
        ctx.check(_lib.lib().pfa_lut_u16_u8_dev(ctx.handle, ptr(flat), flat.numel(), ptr(dl),
                                                 lo, lut.size, ptr(out)), "lut_u16_u8")
        ctx.synchronize()
        return out.reshape(shape)
    if blank:
        return np.zeros(shape, np.uint8)
    out = np.empty(flat.shape[0], np.uint8)
    ctx.check(_lib.lib().pfa_lut_u16_u8(ctx.handle, ptr(flat), flat.size,
                                        ptr(np.ascontiguousarray(lut)), lo, lut.size, ptr(out)),
              "lut_u16_u8")
    return out.reshape(shape)


def percentile_ranks(n, q):
    """np.percentile's 'linear' method for a float32 array of n values
    (numpy 2.x _quantile / _get_indexes / _get_gamma): q / 100 in float32, the
    virtual index (n - 1) * q in float32, its floor and floor + 1 (rounded in
    float32) clamped to the array, and gamma.  Returns (k0, k1, gamma)."""
    qq = np.asanyarray(np.true_divide(q, np.float32(100)))
    vi = np.asanyarray((n - 1) * qq)
    prev = np.asanyarray(np.floor(vi))
    nxt = np.asanyarray(prev + 1)
    if vi >= n - 1:
        prev, nxt = np.asanyarray(-1.0), np.asanyarray(-1.0)
    if vi < 0:
        prev, nxt = np.asanyarray(0.0), np.asanyarray(0.0)
    gamma = np.asanyarray(vi - prev, dtype=vi.dtype)
    k0, k1 = int(prev) % n, int(nxt) % n
    return k0, min(k1, n - 1), gamma


def lerp(a, b, t):
    """numpy's _lerp for 0-d float32 operands: a + (b - a) t, or
    b - (b - a)(1 - t) where t >= 0.5, each operation rounded in float32."""
    a, b = np.float32(a), np.float32(b)
    diff = np.subtract(b, a)
    r = np.asanyarray(np.add(a, diff * t))
    if t >= 0.5:
        r = np.asanyarray(np.subtract(b, diff * (1 - t)), dtype=r.dtype)
    return r[()]


def percentile(amp, q, device=0):
    """np.percentile(amp, q) of a float32 image (any shape, flattened), exact:
    the two order statistics come from the GPU radix select, the
    interpolation is numpy's, in float32.  NaN anywhere gives NaN."""
    flat = _flat(amp, np.float32)
    n = flat.shape[0]
    if n == 0:
        raise PfaError(-1, "percentile of an empty array")
    k0, k1, gamma = percentile_ranks(n, q)
    ctx = context(device)
    v0, v1, nn = C.c_float(), C.c_float(), _lib._i64()
    if _is_dev(flat):
        _sync_torch(flat)
        ctx.check(_lib.lib().pfa_select_f32_dev(ctx.handle, ptr(flat), n, k0, k1, C.byref(v0),
                                                 C.byref(v1), C.byref(nn)), "select_f32")
    else:
        ctx.check(_lib.lib().pfa_select_f32(ctx.handle, ptr(flat), n, k0, k1, C.byref(v0),
                                             C.byref(v1), C.byref(nn)), "select_f32")
    if nn.value:
        return np.float32(np.nan)
    return np.float32(lerp(v0.value, v1.value, gamma))


def stretch_u8(amp, pct=99.7, device=0):
    """The uint8 product of save_detected_geotiff (form_image_pfa.py:286-291):
    hi = np.percentile(amp, pct), data = clip(amp / (hi + 1e-12) * 255, 0,
    255).astype(uint8).  Returns (data, hi); data is on the input's side
    (numpy array or CUDA tensor)."""
    shape = tuple(amp.shape)
    flat = _flat(amp, np.float32)
    hi = percentile(flat, pct, device)
    d = hi + 1e-12                                 # float32, as numpy forms it
    ctx = context(device)
    if _is_dev(flat):
        import torch
        out = torch.empty(flat.shape[0], dtype=torch.uint8, device=flat.device)
        ctx.check(_lib.lib().pfa_stretch_u8_dev(ctx.handle, ptr(flat), flat.numel(), float(d),
                                                 ptr(out)), "stretch_u8")
        ctx.synchronize()
        return out.reshape(shape), hi
    out = np.empty(flat.shape[0], np.uint8)
    ctx.check(_lib.lib().pfa_stretch_u8(ctx.handle, ptr(flat), flat.size, float(d), ptr(out)),
              "stretch_u8")
    return out.reshape(shape), hi
