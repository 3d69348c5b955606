"""The float focus at full size, 8167 x 8192 -> 8192^2 (the benchmark frame) and
5634 x 4319 -> 8192 x 8192 (the board scene's shape), checked
through properties that do not need an oracle run of that size (the numpy
oracle takes minutes per frame): linearity of the whole chain, point targets at
the pixels the geometry predicts, the matrix-core FFT forms within their bound
of the FP32 form, and the detected output equal to |complex output|.

Tolerances: linearity to 1e-6 of the peak (FP32 rounding of three frames);
matrix-core forms to 1e-3 peak-normalised RMS of the FP32 image (the C4 bound);
target peaks within 1 pixel of the prediction."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=["C3", "C2"])
def frame(request):
    import torch
    from mi355x_pfa import synth
    from mi355x_pfa.pfa import float_geometry
    # targets within 40 m of the scene centre: the polar format's own geometric
    # distortion, about r^2 / (2 R) at 10 km standoff, stays below 0.2 pixel
    base = {"C3": synth.C3, "C2": synth.C2}[request.param]
    spec = dataclasses.replace(base, targets=8, clutter=0, scene_half=40.0, seed=11)
    print(f"full-size frame {spec.pulses}x{spec.samples}", flush=True)
    sc = synth.make_scene(spec, device="cuda")
    ax, ay = sc.look()
    geom = float_geometry(ax, ay, f0=sc.sc0, df=sc.scss, n=spec.samples)
    torch.cuda.synchronize()
    return spec, sc, geom


def _plan(frame, fft="fp32", interp="sinc32"):
    from mi355x_pfa.pfa import FloatPlan
    spec, sc, geom = frame
    return FloatPlan(spec.pulses, spec.samples, -1, geom, f0=sc.sc0, df=sc.scss, window=True,
                     interp=interp, fft=fft)


def _focus(plan, sig, detect=False):
    import torch
    out = torch.empty((plan.m2, plan.n2), dtype=torch.float32 if detect else torch.complex64,
                      device="cuda")
    torch.cuda.synchronize()                      # inputs written on torch's stream
    plan.focus_dev(sig, out, detect=detect)
    plan.ctx.synchronize()
    return out


def test_fullsize_linearity(frame):
    import torch
    spec, sc, geom = frame
    plan = _plan(frame)
    g = torch.Generator(device="cuda").manual_seed(5)
    b = torch.complex(torch.randn(sc.sig.shape, device="cuda", generator=g),
                      torch.randn(sc.sig.shape, device="cuda", generator=g)) * 0.1
    ia = _focus(plan, sc.sig)
    ib = _focus(plan, b)
    ic = _focus(plan, (sc.sig + 2 * b).contiguous())
    peak = float(ic.abs().max())
    err = float((ic - (ia + 2 * ib)).abs().max()) / peak
    print(f"linearity: max |I(a + 2b) - I(a) - 2 I(b)| / peak = {err:.2e}", flush=True)
    assert err <= 1e-6
    plan.close()


def test_fullsize_point_targets_at_predicted_pixels(frame):
    spec, sc, geom = frame
    plan = _plan(frame)
    a = _focus(plan, sc.sig, detect=True).cpu().numpy()
    geo = plan.padded_geo()
    dh, ch = np.array(geo["dhat"]), np.array(geo["chat"])
    for p in sc.targets:
        col = plan.n2 // 2 - (p[:2] @ dh) / geo["dr"]
        row = plan.m2 // 2 - (p[:2] @ ch) / geo["dc"]
        r0, c0 = int(round(row)), int(round(col))
        win = a[r0 - 2:r0 + 3, c0 - 2:c0 + 3]
        rr, cc = np.unravel_index(np.argmax(win), win.shape)
        print(f"target at predicted ({row:.2f}, {col:.2f}): peak ({r0 - 2 + rr}, {c0 - 2 + cc})",
              flush=True)
        assert abs(r0 - 2 + rr - row) <= 1.0 and abs(c0 - 2 + cc - col) <= 1.0
    plan.close()


@pytest.mark.parametrize("form", ["bf16", "fp16"])
def test_fullsize_matrix_core_forms_within_bound(frame, form):
    import torch
    spec, sc, geom = frame
    p32, pf = _plan(frame), _plan(frame, fft=form)
    i32 = _focus(p32, sc.sig)
    i16 = _focus(pf, sc.sig)
    peak = float(i32.abs().max())
    rms = float(torch.sqrt(torch.mean((i16 - i32).abs().double() ** 2))) / peak
    print(f"{form}: pn-RMS vs FP32 at 8192^2 = {rms:.2e}", flush=True)
    assert rms <= 1e-3
    ka = int(torch.argmax(i32.abs()))
    kb = int(torch.argmax(i16.abs()))
    n2 = p32.n2
    assert abs(ka // n2 - kb // n2) <= 1 and abs(ka % n2 - kb % n2) <= 1
    p32.close()
    pf.close()


def test_fullsize_detect_equals_abs_of_complex(frame):
    import torch
    spec, sc, geom = frame
    plan = _plan(frame, interp="linear")
    c = _focus(plan, sc.sig)
    d = _focus(plan, sc.sig, detect=True)
    assert torch.allclose(d, c.abs(), rtol=2e-6, atol=1e-6 * float(d.max()))
    plan.close()


def test_fullsize_fixed_targets_match_float(frame):
    """The fixed datapath (quantised int16 history, Q15 tapers, BFP FFTs) puts
    every target where the float focus puts it.  The fixed image is
    [range][cross] with no centring: its pixel (j, k) is the float image's
    ((k - m2/2) mod m2, (j - n2/2) mod n2).  Peaks agree within one pixel."""
    import torch
    from mi355x_pfa import _lib
    from mi355x_pfa.emulator import FixedPlan
    from mi355x_pfa.tables import fixed_tables
    import ctypes as C
    spec, sc, geom = frame
    m, n, G = spec.pulses, spec.samples, 8192
    plan_f = _plan(frame, interp="linear")
    a = _focus(plan_f, sc.sig, detect=True).cpu().numpy()
    freq2 = np.zeros((m, n))
    freq2[:, 0], freq2[:, 1] = sc.sc0, sc.sc0 + sc.scss
    ax, ay = sc.look()
    tabs = fixed_tables({"ax": ax, "ay": ay, "freq": freq2, "KR": geom["KR"], "KC": geom["KC"],
                         "tan_phi": geom["tan_phi"]}, m, n, G)
    plan_x = FixedPlan(tabs, m, n, G)
    sig_iq = torch.empty((m, n, 2), dtype=torch.int16, device="cuda")
    e = C.c_int(0)
    torch.cuda.synchronize()
    plan_x.ctx.check(_lib.lib().pfa_quantize_iq16_dev(plan_x.ctx.handle, sc.sig.data_ptr(), m * n,
                                                      sig_iq.data_ptr(), C.byref(e)), "quantize")
    out = torch.empty((G, G), dtype=torch.int16, device="cuda")
    plan_x.form_image_dev(sig_iq, out)
    plan_x.ctx.synchronize()
    fx = out.view(torch.uint16).to(torch.int32).cpu().numpy()
    m2, n2 = plan_f.m2, plan_f.n2
    # fixed [range j][cross k] -> float layout [cross r][range c]
    fx_c = np.roll(np.roll(fx.T, -m2 // 2, axis=0), -n2 // 2, axis=1)
    geo = plan_f.padded_geo()
    dh, ch = np.array(geo["dhat"]), np.array(geo["chat"])
    for p in sc.targets:
        r0 = int(round(m2 // 2 - (p[:2] @ ch) / geo["dc"]))
        c0 = int(round(n2 // 2 - (p[:2] @ dh) / geo["dr"]))
        wa = a[r0 - 2:r0 + 3, c0 - 2:c0 + 3]
        wf = fx_c[r0 - 2:r0 + 3, c0 - 2:c0 + 3]
        pa = np.unravel_index(np.argmax(wa), wa.shape)
        pf = np.unravel_index(np.argmax(wf), wf.shape)
        print(f"target near ({r0}, {c0}): float peak {pa}, fixed peak {pf}, fixed max {wf.max()}",
              flush=True)
        assert wf.max() > 0
        assert abs(pa[0] - pf[0]) <= 1 and abs(pa[1] - pf[1]) <= 1
    plan_f.close()
    plan_x.close()
