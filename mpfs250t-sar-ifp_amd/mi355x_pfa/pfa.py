"""GPU drop-ins for the float Polar Format focus of src/form_image_pfa.py and the
accelerator seam of mpfs/host/accel.py.

* ``resample_kspace(sig, freq, ax, ay)``  form_image_pfa.py:165-222
* ``focus(g2, geo, sgn)``                 form_image_pfa.py:225-240
* ``pfa(sig, freq, ax, ay, sgn)``         form_image_pfa.py:243-246
* ``GpuBackend`` / ``make_backend``       accel.py:23-78 (``backend.focus(signal, tables)``)

Only the O(m + n) scene geometry is formed here on the host, with the same
numpy operations the reference uses (so KR, KC, tan_phi and the pulse order
are identical); the O(m n) k-space positions, both resample passes, window,
2-D FFT and detection run in libmi355x_pfa.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from ._lib import PfaError, context, ptr
from .tables import sinc_table_q15

C_LIGHT = 299792458.0
FLOAT_SINC32 = 0x1
FLOAT_FFT_BF16, FLOAT_FFT_FP16, FLOAT_FFT_SPLIT = 0x2, 0x4, 0x8
OUT_C64, OUT_ABS_F32 = 0, 1
# FFT forms of the focus: the FP32 vector FFT (the reference's arithmetic) or
# the matrix-core DFT stages with 16-bit operands (precision form)
FFT_FORMS = {"fp32": 0, "bf16": FLOAT_FFT_BF16, "fp16": FLOAT_FFT_FP16,
             "bf16-split": FLOAT_FFT_BF16 | FLOAT_FFT_SPLIT}


def to_pow2(n: int) -> int:
    return 1 << int(math.ceil(math.log2(n)))


def _kr_row_extrema(freq, pr, device=0):
    """Per-pulse max / min of kr = (2 freq / c) pr[:, None] on the GPU in FP64
    (numpy's operation order and rounding): the O(m n) part of the geometry.
    freq: numpy (m, n) float64, or a CUDA float64 tensor."""
    m, n = freq.shape
    ctx = context(device)
    if isinstance(freq, np.ndarray):
        f = np.ascontiguousarray(freq, np.float64)
        p = np.ascontiguousarray(pr, np.float64)
        rmax, rmin = np.empty(m), np.empty(m)
        ctx.check(_lib.lib().pfa_kr_extrema(ctx.handle, ptr(f), ptr(p), m, n, ptr(rmax),
                                             ptr(rmin)), "kr_extrema")
        return rmax, rmin
    import torch
    p = torch.as_tensor(np.asarray(pr, np.float64), device=freq.device)
    rmax = torch.empty(m, dtype=torch.float64, device=freq.device)
    rmin = torch.empty_like(rmax)
    torch.cuda.synchronize()
    ctx.check(_lib.lib().pfa_kr_extrema_dev(ctx.handle, ptr(freq.contiguous()), ptr(p), m, n,
                                            ptr(rmax), ptr(rmin)), "kr_extrema")
    ctx.synchronize()
    return rmax.cpu().numpy(), rmin.cpu().numpy()


def float_geometry(ax, ay, freq=None, f0=None, df=None, n=None, device=0, k0=0,
                   _hypot=math.hypot):
    """form_image_pfa.py:178-193, 205-209, 219-221.

    The O(m + n) tables are formed on the host with the reference's numpy
    operations.  With the full freq (m, n) array (numpy or a CUDA tensor) the
    O(m n) per-pulse extrema of kr are reduced on the GPU in FP64 (same values);
    with compact per-pulse f0/df (freq[i, k] = f0[i] + (k0 + k) * df[i], rows
    monotone) each row's extrema are its end points.  ``_hypot`` is the norm of
    the mean look direction: math.hypot in form_image_pfa.py:181, np.hypot in
    sar_pipeline.py:95."""
    ax = np.asarray(ax, np.float64)
    ay = np.asarray(ay, np.float64)
    dx, dy = ax.mean(), ay.mean()
    dn = _hypot(dx, dy)
    dx, dy = dx / dn, dy / dn
    cx, cy = -dy, dx
    pr = ax * dx + ay * dy
    pc = ax * cx + ay * cy
    tan_phi = pc / pr
    if freq is not None:
        if isinstance(freq, np.ndarray) or not hasattr(freq, "is_cuda"):
            freq = np.asarray(freq, np.float64)
        n = freq.shape[1]
        row_max, row_min = _kr_row_extrema(freq, pr, device)
        gmin, gmax = row_min.min(), row_max.max()
    else:
        f0 = np.asarray(f0, np.float64)
        df = np.asarray(df, np.float64)
        kk = np.array([k0, k0 + n - 1])
        ends = f0[:, None] + kk[None, :] * df[:, None]
        kr = (2.0 * ends / C_LIGHT) * pr[:, None]
        row_max, row_min = kr.max(axis=1), kr.min(axis=1)
        gmin, gmax = kr.min(), kr.max()
    lo, hi = row_max.min(), row_min.max()
    if lo > hi:
        lo, hi = gmin, gmax
    KR = np.linspace(lo, hi, n)
    kc_max = float(np.max(np.abs(KR[[0, -1]])) * np.max(np.abs(tan_phi)))
    m = ax.size
    KC = np.linspace(-kc_max, kc_max, m)
    order = np.argsort(tan_phi)
    dr = 1.0 / (KR[-1] - KR[0]) if KR[-1] != KR[0] else float("nan")
    dc = 1.0 / (KC[-1] - KC[0]) if KC[-1] != KC[0] else float("nan")
    return {"pr": pr, "pc": pc, "tan_phi": tan_phi, "order": order.astype(np.int32),
            "tan_s": np.ascontiguousarray(tan_phi[order]), "KR": KR, "KC": KC,
            "geo": {"dc": dc, "dr": dr, "dhat": (dx, dy), "chat": (cx, cy)}}


class FloatPlan:
    """A float-mode plan: geometry tables resident on the device."""

    def __init__(self, m, n, sgn, geom, f0=None, df=None, window=True, interp="linear",
                 device=0, fft="fp32"):
        if interp not in ("linear", "sinc32"):
            raise PfaError(-1, f"unknown interpolator {interp!r}")
        if fft not in FFT_FORMS:
            raise PfaError(-1, f"unknown FFT form {fft!r} (one of {sorted(FFT_FORMS)})")
        self.fft = fft
        self.ctx = context(device)
        self.m, self.n, self.sgn = m, n, sgn
        self.m2, self.n2 = max(64, to_pow2(m)), max(64, to_pow2(n))
        keep = {"pr": geom["pr"], "order": geom["order"], "tan_s": geom["tan_s"],
                "KR": geom["KR"], "KC": geom["KC"]}
        if f0 is not None:
            keep["f0"] = np.asarray(f0, np.float64)
            keep["df"] = np.asarray(df, np.float64)
        if window:
            keep["hm"] = np.hamming(m)
            keep["hn"] = np.hamming(n)
        self._keep = {k: np.ascontiguousarray(v, np.int32 if k == "order" else np.float64)
                      for k, v in keep.items()}
        k = self._keep
        sinc = np.ascontiguousarray(sinc_table_q15()) if interp == "sinc32" else None
        self._sinc = sinc
        self._h = C.c_void_p()
        self.ctx.call("pfa_plan_float", C.byref(self._h), m, n, int(sgn), ptr(k["pr"]),
                      ptr(k["order"]), ptr(k["tan_s"]), ptr(k["KR"]), ptr(k["KC"]),
                      ptr(k.get("f0")), ptr(k.get("df")), ptr(k.get("hm")), ptr(k.get("hn")),
                      ptr(sinc), (FLOAT_SINC32 if interp == "sinc32" else 0) | FFT_FORMS[fft])
        self.geo = dict(geom["geo"])

    def _sig(self, sig):
        sig = np.ascontiguousarray(sig, np.complex64)
        if sig.shape != (self.m, self.n):
            raise PfaError(-1, f"signal shape {sig.shape} != {(self.m, self.n)}")
        return sig

    def _freq(self, freq):
        if freq is None:
            return None
        f = np.ascontiguousarray(freq, np.float64)
        if f.shape != (self.m, self.n):
            raise PfaError(-1, f"freq shape {f.shape} != {(self.m, self.n)}")
        return f

    def focus(self, sig, freq=None, detect=False):
        sig, fr = self._sig(sig), self._freq(freq)
        out = np.empty((self.m2, self.n2), np.float32 if detect else np.complex64)
        self.ctx.check(_lib.lib().pfa_focus_f32(self._h, ptr(sig), ptr(fr), ptr(out),
                                                OUT_ABS_F32 if detect else OUT_C64), "focus")
        return out

    def focus_dev(self, sig_dev, out_dev, freq_dev=None, detect=False):
        self.ctx.check(_lib.lib().pfa_focus_f32_dev(self._h, ptr(sig_dev), ptr(freq_dev),
                                                    ptr(out_dev),
                                                    OUT_ABS_F32 if detect else OUT_C64),
                       "focus_dev")

    def resample_kspace(self, sig, freq=None):
        sig, fr = self._sig(sig), self._freq(freq)
        g2 = np.empty((self.m, self.n), np.complex64)
        self.ctx.check(_lib.lib().pfa_resample_kspace_f32(self._h, ptr(sig), ptr(fr), ptr(g2)),
                       "resample_kspace")
        return g2

    def transform(self, g2, detect=False):
        g2 = self._sig(g2)
        out = np.empty((self.m2, self.n2), np.float32 if detect else np.complex64)
        self.ctx.check(_lib.lib().pfa_transform_f32(self._h, ptr(g2), ptr(out),
                                                    OUT_ABS_F32 if detect else OUT_C64),
                       "transform")
        return out

    def padded_geo(self):
        geo = dict(self.geo)
        geo["dr"] *= self.n / self.n2
        geo["dc"] *= self.m / self.m2
        return geo

    def close(self):
        if self._h:
            _lib.lib().pfa_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _plan_for(sig, freq, ax, ay, sgn, window=True, interp="linear", device=0, fft="fp32"):
    m, n = np.shape(sig)
    geom = float_geometry(ax, ay, freq=freq)
    return FloatPlan(m, n, sgn, geom, window=window, interp=interp, device=device, fft=fft)


def resample_kspace(sig, freq, ax, ay, window=True, interp="linear", device=0):
    """form_image_pfa.resample_kspace -> (g2 complex64 (m, n), geo)."""
    plan = _plan_for(sig, freq, ax, ay, -1, window, interp, device)
    g2 = plan.resample_kspace(sig, freq)
    geo = dict(plan.geo)
    plan.close()
    return g2, geo


def focus(g2, geo, sgn, device=0):
    """form_image_pfa.focus -> (img complex64 (m2, n2), geo scaled to the padded grid)."""
    g2 = np.ascontiguousarray(g2, np.complex64)
    m, n = g2.shape
    dummy = {"pr": np.ones(m), "order": np.arange(m, dtype=np.int32), "tan_s": np.zeros(m),
             "KR": np.ones(n), "KC": np.zeros(m), "geo": dict(geo)}
    plan = FloatPlan(m, n, sgn, dummy, window=False, device=device)
    img = plan.transform(g2)
    out_geo = plan.padded_geo()
    plan.close()
    return img, out_geo


def quicklook(sig, sgn, window=True, device=0):
    """form_image_pfa.quicklook -> complex64 (m, n): window (hamming2d) + one
    centred 2-D FFT of the raw history (the polar grid taken as Cartesian), no
    padding, any shape: lengths up to 8192 on either axis, or powers of two up
    to 16384 (the reference's numpy FFT takes any size).  sig: numpy array or
    a CUDA complex64 tensor (the result stays on its side)."""
    ctx = context(device)
    if isinstance(sig, np.ndarray) or not hasattr(sig, "is_cuda"):
        x = np.ascontiguousarray(sig, np.complex64)
        m, n = x.shape
        out = np.empty((m, n), np.complex64)
        ctx.check(_lib.lib().pfa_quicklook_f32(ctx.handle, ptr(x), m, n, int(sgn), int(window),
                                               ptr(out), OUT_C64), "quicklook")
        return out
    import torch
    x = sig.contiguous()
    m, n = x.shape
    out = torch.empty((m, n), dtype=torch.complex64, device=x.device)
    torch.cuda.synchronize()
    ctx.check(_lib.lib().pfa_quicklook_f32_dev(ctx.handle, ptr(x), m, n, int(sgn), int(window),
                                               ptr(out), OUT_C64), "quicklook")
    ctx.synchronize()
    return out


def pfa(sig, freq, ax, ay, sgn, window=True, interp="linear", device=0, fft="fp32"):
    """form_image_pfa.pfa -> (img complex64 (m2, n2) centred, geo).  fft selects
    the FP32 FFT (default, the reference's arithmetic) or a matrix-core form
    ("bf16", "fp16", "bf16-split")."""
    plan = _plan_for(sig, freq, ax, ay, sgn, window, interp, device, fft)
    img = plan.focus(sig, freq)
    geo = plan.padded_geo()
    plan.close()
    return img, geo


class GpuBackend:
    """accel.AccelBackend on the MI355X: focus(signal, tables) -> |img| float32
    (the pow2-padded centred image, as NumpyBackend returns)."""
    name = "gpu"

    def __init__(self, interp="linear", device=0, fft="fp32"):
        self.interp = interp
        self.device = device
        self.fft = fft

    def focus(self, signal, tables):
        """tables from sar_pipeline.prepare_tables (freq an (m, n) array) or
        from pipeline.prepare_tables (freq a LinearFreq: only the per-pulse
        sc0 / step reach the GPU when no samples were cropped)."""
        from .pipeline import LinearFreq
        fr = tables["freq"]
        if isinstance(fr, LinearFreq) and fr.k0 == 0:
            m, n = np.shape(signal)
            if fr.shape != (m, n):
                raise PfaError(-1, f"freq shape {fr.shape} != signal shape {(m, n)}")
            geom = float_geometry(tables["ax"], tables["ay"], f0=fr.sc0, df=fr.step, n=n)
            plan = FloatPlan(m, n, tables["sgn"], geom, f0=fr.sc0, df=fr.step, window=True,
                             interp=self.interp, device=self.device, fft=self.fft)
            out = plan.focus(signal, None, detect=True)
        else:
            if isinstance(fr, LinearFreq):            # cropped: freq starts at sample k0
                fr = np.asarray(fr)
            plan = _plan_for(signal, fr, tables["ax"], tables["ay"], tables["sgn"],
                             True, self.interp, self.device, self.fft)
            out = plan.focus(signal, fr, detect=True)
        plan.close()
        return out


def make_backend(name, ref_module=None):
    """accel.make_backend with the GPU backend added ("gpu", "gpu-sinc32")."""
    if name == "gpu":
        return GpuBackend("linear")
    if name == "gpu-sinc32":
        return GpuBackend("sinc32")
    if name in ("gpu-bf16", "gpu-fp16"):
        return GpuBackend("sinc32", fft=name[4:])
    if name == "numpy":
        if ref_module is None:
            raise ValueError("the numpy backend needs the reference module")

        class _Numpy:
            name = "numpy"

            def focus(self, signal, tables):
                img, _ = ref_module.pfa(signal, tables["freq"], tables["ax"], tables["ay"],
                                        tables["sgn"])
                return np.abs(img).astype(np.float32)
        return _Numpy()
    raise ValueError(f"unknown backend: {name}")
