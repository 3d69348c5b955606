"""The precision form of the float focus (matrix-core DFT stages with BF16 /
FP16 operands, FP32 accumulation, per-line block scaling, 16-bit intermediate)
against the FP64 oracle (oracle/ref_float with complex128 intermediates and
numpy's double-precision FFT).

Bar: peak-normalised RMS <= 1e-3 against FP64, each of the 8 strongest point
targets within 1 pixel of its FP64 position (16-bit operands can flip the
maximum between two nearly equal neighbouring pixels; the shift is stated),
PSLR / ISLR of the brightest target within 0.5 dB of the FP64 image.  The shapes cover every matrix-core FFT length (256, 512, 1024,
4096 and 8192) on both FFT passes."""
import numpy as np
import pytest

from oracle import ref_float as rfl

pytestmark = pytest.mark.gpu
TOL = 1e-3


@pytest.fixture(scope="module")
def gp():
    from mi355x_pfa import pfa
    return pfa


@pytest.fixture(scope="module")
def met():
    from mi355x_pfa import metrics
    return metrics


def _scene(m, n, seed, targets=16):
    from mi355x_pfa import synth
    sc = synth.make_scene(synth.SceneSpec(m, n, targets=targets, seed=seed,
                                          scene_half=0.2 * min(m, n)))
    ax, ay = sc.look()
    return sc, sc.freq(), ax, ay


_CACHE = {}


def _ref64(m, n, seed, sgn, interp):
    key = (m, n, seed, sgn, interp)
    if key not in _CACHE:
        sc, freq, ax, ay = _scene(m, n, seed)
        ref, _ = rfl.pfa(sc.sig, freq, ax, ay, sgn, interp=interp, store=np.complex128)
        _CACHE[key] = (sc, freq, ax, ay, ref)
    return _CACHE[key]


CASES = [(1024, 1024, -1, "sinc32"),      # 1024 x 1024
         (300, 4100, -1, "linear"),       # 512 cross x 8192 range
         (2100, 200, 1, "linear"),        # 4096 cross x 256 range, inverse transform
         (600, 500, 1, "linear")]         # 1024 cross x 512 range, inverse, radix-32 stages


@pytest.mark.parametrize("m,n,sgn,interp", CASES)
@pytest.mark.parametrize("form", ["bf16", "fp16", "bf16-split"])
def test_precision_form_against_fp64(gp, met, m, n, sgn, interp, form):
    sc, freq, ax, ay, ref = _ref64(m, n, 7 * m + n, sgn, interp)
    print(f"precision form {form} {m}x{n} sgn={sgn} {interp}", flush=True)
    img, _ = gp.pfa(sc.sig, freq, ax, ay, sgn, interp=interp, fft=form)
    rep = met.error_report(img, ref)
    print(rep, flush=True)
    assert rep["pn_rms"] <= TOL
    assert rep["peak_shift_px"] <= 1.0
    assert abs(rep["pslr_db"] - rep["pslr_ref_db"]) < 0.5
    assert abs(rep["islr_db"] - rep["islr_ref_db"]) < 0.5


def test_precision_ordering_and_fp32_baseline(gp, met):
    """FP16 and hi/lo BF16 are closer to FP64 than plain BF16; FP32 closer still."""
    sc, freq, ax, ay, ref = _ref64(1024, 1024, 7 * 1024 + 1024, -1, "sinc32")
    err = {}
    for form in ("fp32", "bf16", "bf16-split", "fp16"):
        img, _ = gp.pfa(sc.sig, freq, ax, ay, -1, interp="sinc32", fft=form)
        err[form] = met.pn_rms(img, ref)
    print(err, flush=True)
    assert err["fp32"] < err["fp16"] < err["bf16"]
    assert err["bf16-split"] < err["bf16"]


def test_detected_output_matches_complex(gp):
    sc, freq, ax, ay = _scene(300, 4100, seed=5)
    geom = gp.float_geometry(ax, ay, freq=freq)
    plan = gp.FloatPlan(300, 4100, -1, geom, fft="bf16")
    c = plan.focus(sc.sig, freq)
    d = plan.focus(sc.sig, freq, detect=True)
    np.testing.assert_allclose(d, np.abs(c), rtol=2e-6, atol=1e-6 * np.abs(c).max())


def test_unsupported_size_is_refused(gp):
    from mi355x_pfa._lib import PfaError
    sc, freq, ax, ay = _scene(1500, 300, seed=6)        # cross padded to 2048
    geom = gp.float_geometry(ax, ay, freq=freq)
    with pytest.raises(PfaError, match="matrix-core FFT sizes"):
        gp.FloatPlan(1500, 300, -1, geom, fft="fp16")
