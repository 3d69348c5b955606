"""ctypes binding of libmi355x_pfa.so (include/mi355x_pfa.h).

The library is the only compute path: if it cannot be loaded, or no HIP device
is present, calls raise ``PfaError`` -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libmi355x_pfa.so"

PFA_OK = 0
STATUS = {-1: "PFA_ERR_ARG", -2: "PFA_ERR_DIMS", -3: "PFA_ERR_GEOMETRY", -4: "PFA_ERR_HIP",
          -5: "PFA_ERR_NOMEM", -6: "PFA_ERR_STATE"}

FIXED_RANGE_SINC = 0x1
FIXED_AZ_SINC = 0x2
FIXED_LERP_EXACT = 0x4

TIMER_RANGE, TIMER_AZIMUTH, TIMER_FFT_B, TIMER_DETECT, TIMER_OTHER = range(5)


class PfaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_vp, _i, _i64 = C.c_void_p, C.c_int, C.c_int64

# name -> argtypes; every function returns int status unless listed in _VOID/_STR
PROTOS = {
    "pfa_abi_version": [],
    "pfa_create": [C.POINTER(_vp), _i],
    "pfa_destroy": [_vp],
    "pfa_last_error": [_vp],
    "pfa_set_stream": [_vp, _vp],
    "pfa_synchronize": [_vp],
    "pfa_timer_enable": [_vp, _i],
    "pfa_timer_read": [_vp, _i, C.POINTER(C.c_float), C.POINTER(_i)],
    "pfa_timer_reset": [_vp],
    "pfa_fallback_counts": [_vp, _vp, _i],
    "pfa_twiddle_codes": [_i, _i, _vp],
    "pfa_quantize_iq16": [_vp, _vp, _i64, _vp, C.POINTER(_i)],
    "pfa_quantize_iq16_dev": [_vp, _vp, _i64, _vp, C.POINTER(_i)],
    "pfa_fft_pass_bfp": [_vp, _vp, _i, _i, _vp, _vp],
    "pfa_fft_pass_bfp_dev": [_vp, _vp, _i, _i, _vp, _vp],
    "pfa_fft_bfp_perrow": [_vp, _vp, _i, _i, _vp, _vp],
    "pfa_window_fixed": [_vp, _vp, _i, _i, _vp, _vp, _vp],
    "pfa_detect_fixed": [_vp, _vp, _i64, _vp],
    "pfa_detect_fixed_dev": [_vp, _vp, _i64, _vp],
    "pfa_plan_fixed": [_vp, C.POINTER(_vp), _i, _i, _i] + [_vp] * 10 + [_i],
    "pfa_resample_fixed": [_vp, _vp, _vp],
    "pfa_form_image_fixed": [_vp, _vp, _vp, _vp, _vp],
    "pfa_form_image_fixed_dev": [_vp, _vp, _vp, _vp, _vp],
    "pfa_form_image_fixed_dev_noalloc": [_vp, _vp, _vp, _vp, _vp],
    "pfa_plan_float": [_vp, C.POINTER(_vp), _i, _i, _i] + [_vp] * 10 + [_i],
    "pfa_focus_f32": [_vp, _vp, _vp, _vp, _i],
    "pfa_focus_f32_dev": [_vp, _vp, _vp, _vp, _i],
    "pfa_resample_kspace_f32": [_vp, _vp, _vp, _vp],
    "pfa_transform_f32": [_vp, _vp, _vp, _i],
    "pfa_float_geometry": [_i, _i] + [_vp] * 13,
    "pfa_plan_float_pfa": [_vp, C.POINTER(_vp), _i, _i, _i] + [_vp] * 5 + [_i, _vp, _i, _vp],
    "pfa_pfa_f32": [_vp, _vp, _i, _i, _vp, _vp, _vp, _i, _i, _vp, _i, _vp],
    "pfa_quicklook_f32": [_vp, _vp, _i, _i, _i, _i, _vp, _i],
    "pfa_quicklook_f32_dev": [_vp, _vp, _i, _i, _i, _i, _vp, _i],
    "pfa_kr_extrema": [_vp, _vp, _vp, _i, _i, _vp, _vp],
    "pfa_kr_extrema_dev": [_vp, _vp, _vp, _i, _i, _vp, _vp],
    "pfa_hist_u16": [_vp, _vp, _i64, _vp],
    "pfa_hist_u16_dev": [_vp, _vp, _i64, _vp],
    "pfa_lut_u16_u8": [_vp, _vp, _i64, _vp, _i, _i, _vp],
    "pfa_lut_u16_u8_dev": [_vp, _vp, _i64, _vp, _i, _i, _vp],
    "pfa_select_f32": [_vp, _vp, _i64, _i64, _i64, C.POINTER(C.c_float),
                       C.POINTER(C.c_float), C.POINTER(_i64)],
    "pfa_select_f32_dev": [_vp, _vp, _i64, _i64, _i64, C.POINTER(C.c_float),
                           C.POINTER(C.c_float), C.POINTER(_i64)],
    "pfa_stretch_u8": [_vp, _vp, _i64, C.c_float, _vp],
    "pfa_stretch_u8_dev": [_vp, _vp, _i64, C.c_float, _vp],
    "pfa_af_checkpoint": [_vp, _vp, _vp],
    "pfa_af_forward": [_vp, _vp, _vp, _vp],
    "pfa_af_backward": [_vp, _vp, _vp, _vp, _vp],
    "pfa_plan_destroy": [_vp],
}
_VOID = {"pfa_destroy", "pfa_plan_destroy"}
_STR = {"pfa_last_error"}

_lib = None
_lock = threading.Lock()


def _share_torch_runtime():
    """PyTorch-ROCm wheels carry their own HIP runtime (libamdhip64.so,
    libhsa-runtime64.so under torch/lib).  If the library bound the system
    runtime first, a later ``import torch`` in the same process would load the
    second copy, whose device initialisation then fails ("No HIP GPUs are
    available").  When torch is installed, its runtime is loaded globally
    before the library, so the library's sonames resolve to it and the process
    holds one runtime whichever is imported first."""
    import importlib.util
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    if spec is None or not spec.submodule_search_locations:
        return
    d = Path(list(spec.submodule_search_locations)[0]) / "lib"
    for name in ("libhsa-runtime64.so", "libamdhip64.so"):
        p = d / name
        if p.exists():
            C.CDLL(str(p), mode=C.RTLD_GLOBAL)


def lib():
    """Load the shared library once; raise loudly if it is missing."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise PfaError(-4, f"{LIB_PATH} is not built (run mpfs250t-sar-ifp_amd/build.py)")
            _share_torch_runtime()
            L = C.CDLL(str(LIB_PATH))
            for name, args in PROTOS.items():
                f = getattr(L, name)
                f.argtypes = args
                f.restype = None if name in _VOID else (C.c_char_p if name in _STR else C.c_int)
            _lib = L
        return _lib


def ptr(a):
    """Address of a numpy array (host) or a torch tensor (device)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags.c_contiguous:
            raise PfaError(-1, "arrays passed to the library must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        if hasattr(a, "is_contiguous") and not a.is_contiguous():
            raise PfaError(-1, "tensors passed to the library must be contiguous")
        return a.data_ptr()
    return int(a)


class Context:
    """One library context (device, stream, workspace)."""

    def __init__(self, device: int = 0):
        self._h = _vp()
        st = lib().pfa_create(C.byref(self._h), int(device))
        if st != PFA_OK:
            raise PfaError(st, f"pfa_create(device={device}) failed (is a HIP device present?)")
        self.device = device

    @property
    def handle(self):
        return self._h

    def check(self, st, what=""):
        if st != PFA_OK:
            msg = lib().pfa_last_error(self._h)
            raise PfaError(st, f"{what}: {msg.decode() if msg else ''}")

    def call(self, name, *args):
        self.check(getattr(lib(), name)(self._h, *args), name)

    def set_stream(self, stream_ptr):
        self.call("pfa_set_stream", stream_ptr)

    def synchronize(self):
        self.call("pfa_synchronize")

    def timer_enable(self, on=True):
        self.call("pfa_timer_enable", int(on))

    def timer_reset(self):
        self.call("pfa_timer_reset")

    def timer_read(self, kind):
        ms = C.c_float(0)
        n = _i(0)
        self.call("pfa_timer_read", int(kind), C.byref(ms), C.byref(n))
        return ms.value, n.value

    def fallback_counts(self, reset=True):
        out = np.zeros(2, np.uint32)
        self.call("pfa_fallback_counts", out.ctypes.data, int(reset))
        return int(out[0]), int(out[1])

    def close(self):
        if self._h:
            lib().pfa_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts = {}


def context(device: int = 0) -> Context:
    with _lock:
        ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        with _lock:
            _contexts[device] = ctx
    return ctx


def exported_symbols():
    return list(PROTOS)
