"""GPU drop-ins for the fixed-point datapath of mpfs/host/silicon_emulator.py.

Same names, arguments, array layouts and dtypes as the reference functions:
complex arrays holding integer codes in, complex128 (or uint16) out.  All
arithmetic runs in libmi355x_pfa.so on the GPU; this module only converts
between the reference's complex arrays and the library's int16 I/Q layout.

* ``resample_fixed``   silicon_emulator.py:125-156
* ``window_fixed``     silicon_emulator.py:165-195
* ``fft_pass_bfp``     silicon_emulator.py:198-205
* ``detect_fixed``     silicon_emulator.py:208-213
* ``form_image``       silicon_emulator.py:216-253 (from an in-memory signal and
                       prepare_tables() dict instead of a CPHD path)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import PfaError, context, ptr
from .tables import fixed_tables, hamming_q15, sinc_table_q15


def _to_iq(z) -> np.ndarray:
    """complex integer-valued array -> int16 [..., 2] (I, Q)."""
    z = np.asarray(z)
    out = np.empty(z.shape + (2,), np.int16)
    re, im = np.real(z), np.imag(z)
    out[..., 0] = np.clip(np.floor(re + 0.0), -32768, 32767)
    out[..., 1] = np.clip(np.floor(im + 0.0), -32768, 32767)
    return out


def _from_iq(iq) -> np.ndarray:
    return iq[..., 0].astype(np.float64) + 1j * iq[..., 1].astype(np.float64)


class FixedPlan:
    """A fixed-mode plan: scene tables resident on the device."""

    def __init__(self, tabs: dict, m: int, n: int, grid: int, range_sinc=False,
                 az_sinc=False, lerp_exact=False, device: int = 0):
        self.ctx = context(device)
        self.m, self.n, self.grid = m, n, grid
        flags = ((_lib.FIXED_RANGE_SINC if range_sinc else 0)
                 | (_lib.FIXED_AZ_SINC if az_sinc else 0)
                 | (_lib.FIXED_LERP_EXACT if lerp_exact else 0))
        types = {"inv_order": np.int32, "hamr": np.int32, "hamc": np.int32}
        # keep every table alive (and C-contiguous, right dtype) for the call
        self._keep = {k: np.ascontiguousarray(v, types.get(k, np.float32))
                      for k, v in tabs.items()}
        sinc = sinc_table_q15() if (range_sinc or az_sinc) else None
        self._sinc = None if sinc is None else np.ascontiguousarray(sinc)
        k = self._keep
        self._h = C.c_void_p()
        self.ctx.call("pfa_plan_fixed", C.byref(self._h), m, n, grid,
                      ptr(k["f0"]), ptr(k["df"]), ptr(k["pr"]), ptr(k["tan_s"]),
                      ptr(k["inv_order"]), ptr(k["krp"]), ptr(k["kcp"]),
                      ptr(k["hamr"]), ptr(k["hamc"]),
                      ptr(self._sinc), flags)

    def resample(self, sig_iq: np.ndarray) -> np.ndarray:
        sig_iq = np.ascontiguousarray(sig_iq, np.int16)
        if sig_iq.shape != (self.m, self.n, 2):
            raise PfaError(-1, f"signal shape {sig_iq.shape} != {(self.m, self.n, 2)}")
        g2 = np.empty((self.grid, self.grid, 2), np.int16)
        self.ctx.check(_lib.lib().pfa_resample_fixed(self._h, ptr(sig_iq), ptr(g2)), "resample")
        return g2

    def form_image(self, sig_iq: np.ndarray):
        sig_iq = np.ascontiguousarray(sig_iq, np.int16)
        if sig_iq.shape != (self.m, self.n, 2):
            raise PfaError(-1, f"signal shape {sig_iq.shape} != {(self.m, self.n, 2)}")
        G = self.grid
        img = np.empty((G, G), np.uint16)
        ea = np.empty(G, np.int32)
        eb = np.empty(G, np.int32)
        self.ctx.check(_lib.lib().pfa_form_image_fixed(self._h, ptr(sig_iq), ptr(img), ptr(ea),
                                                       ptr(eb)), "form_image")
        return img, ea, eb

    def form_image_dev(self, sig_dev, out_dev, ea_dev=None, eb_dev=None):
        """Device pointers / tensors in, device out, enqueued on the context stream."""
        self.ctx.check(_lib.lib().pfa_form_image_fixed_dev(self._h, ptr(sig_dev), ptr(out_dev),
                                                           ptr(ea_dev), ptr(eb_dev)),
                       "form_image_dev")

    def close(self):
        if self._h:
            _lib.lib().pfa_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def quantize_iq16(signal, device: int = 0):
    """fit_scale + floor + sat16 of a complex64 signal (silicon_emulator.py:226-229).
    Returns (int16 [m, n, 2], exponent)."""
    sig = np.ascontiguousarray(signal, np.complex64)
    out = np.empty(sig.shape + (2,), np.int16)
    e = C.c_int(0)
    context(device).call("pfa_quantize_iq16", ptr(sig), sig.size, ptr(out), C.byref(e))
    return out, e.value


def resample_fixed(signal_i16, tables, grid, range_sinc=False, az_sinc=False, device: int = 0):
    """silicon_emulator.resample_fixed.  A complex64 signal gets form_image's
    float32 pass-1 lerp product, complex128 the exact one (as the reference)."""
    sig = np.asarray(signal_i16)
    m, n = sig.shape
    tabs = fixed_tables(tables, m, n, grid)
    plan = FixedPlan(tabs, m, n, grid, range_sinc, az_sinc,
                     lerp_exact=(sig.dtype != np.complex64), device=device)
    g2 = plan.resample(_to_iq(sig))
    plan.close()
    return _from_iq(g2), (m, n)


def window_fixed(g2, m, n, device: int = 0):
    """silicon_emulator.window_fixed on g2 [range][cross]."""
    iq = _to_iq(g2)
    Np, Mp = iq.shape[:2]
    hr = hamming_q15(n, Np)
    hc = hamming_q15(m, Mp)
    out = np.empty_like(iq)
    context(device).call("pfa_window_fixed", ptr(iq), Np, Mp, ptr(hr), ptr(hc), ptr(out))
    return _from_iq(out)


def fft_pass_bfp(x, device: int = 0):
    """silicon_emulator.fft_pass_bfp over the last axis; returns (y, e)."""
    iq = _to_iq(x)
    n = iq.shape[-2]
    lead = iq.shape[:-2]
    rows = int(np.prod(lead)) if lead else 1
    src = np.ascontiguousarray(iq.reshape(rows, n, 2))
    y = np.empty_like(src)
    e = np.empty(rows, np.int32)
    context(device).call("pfa_fft_pass_bfp", ptr(src), rows, n, ptr(y), ptr(e))
    return _from_iq(y.reshape(iq.shape)), e.reshape(lead)


def fft_bfp_perrow(x, device: int = 0):
    """fixedpoint.fft1d_bfp_hw_perrow(x, 16, 16, perm) over the last axis."""
    iq = _to_iq(x)
    n = iq.shape[-2]
    lead = iq.shape[:-2]
    rows = int(np.prod(lead)) if lead else 1
    src = np.ascontiguousarray(iq.reshape(rows, n, 2))
    y = np.empty_like(src)
    e = np.empty(rows, np.int32)
    context(device).call("pfa_fft_bfp_perrow", ptr(src), rows, n, ptr(y), ptr(e))
    return _from_iq(y.reshape(iq.shape)), e.reshape(lead)


def detect_fixed(z, device: int = 0):
    """silicon_emulator.detect_fixed -> uint16."""
    iq = np.ascontiguousarray(_to_iq(z))
    out = np.empty(iq.shape[:-1], np.uint16)
    context(device).call("pfa_detect_fixed", ptr(iq), out.size, ptr(out))
    return out


def form_image(signal, tables, grid, range_sinc=False, az_sinc=False, device: int = 0):
    """silicon_emulator.form_image from an in-memory complex64 signal and a
    prepare_tables() dict: quantise, resample, window, two BFP FFT passes with
    the corner-turn, detect.  Returns the uint16 image [range][cross]."""
    sig = np.ascontiguousarray(signal, np.complex64)
    m, n = sig.shape
    iq, _ = quantize_iq16(sig, device)
    plan = FixedPlan(fixed_tables(tables, m, n, grid), m, n, grid, range_sinc, az_sinc,
                     device=device)
    img, _, _ = plan.form_image(iq)
    plan.close()
    return img
