"""Gradient autofocus on the GPU: the library's hand-written adjoint against
the FP64 torch-autograd reference (tests/af_reference.py, itself pinned to the
oracle and to finite differences), against FP32 finite differences of the
library's own forward, and a short autofocus run that removes an injected
phase error."""
import numpy as np
import pytest

from af_reference import Reference

pytestmark = pytest.mark.gpu


def _scene(m, n, seed, targets=6, phase=None):
    from mi355x_pfa import synth
    sc = synth.make_scene(synth.SceneSpec(m, n, targets=targets, seed=seed,
                                          scene_half=0.2 * min(m, n), phase_error=phase))
    ax, ay = sc.look()
    return sc, sc.freq(), ax, ay


def _plan(sc, freq, ax, ay, sgn, interp):
    import torch
    from mi355x_pfa import pfa
    from mi355x_pfa.autofocus import AutofocusPlan
    geom = pfa.float_geometry(ax, ay, freq=freq)
    sig = torch.from_numpy(np.ascontiguousarray(sc.sig, np.complex64)).cuda()
    fr = torch.from_numpy(np.ascontiguousarray(freq, np.float64)).cuda()
    return AutofocusPlan(sig, geom, freq=fr, sgn=sgn, interp=interp), sig, fr


@pytest.mark.parametrize("m,n,sgn,interp", [(120, 128, -1, "sinc32"), (100, 90, 1, "linear"),
                                            (250, 64, -1, "sinc32")])
def test_adjoint_gradient_matches_fp64_autograd(m, n, sgn, interp):
    import torch
    sc, freq, ax, ay = _scene(m, n, seed=m + n)
    print(f"af gradient {m}x{n} sgn={sgn} {interp}", flush=True)
    phi = np.random.RandomState(m).normal(0.0, 0.4, m)
    s_ref, g_ref = Reference(sc.sig, freq, ax, ay, sgn, interp).gradient(phi)
    af, _, _ = _plan(sc, freq, ax, ay, sgn, interp)
    pt = torch.tensor(phi, dtype=torch.float32, device="cuda")
    S, img, stats = af.forward(pt)
    g = af.backward(pt, img, stats).cpu().numpy()
    torch.cuda.synchronize()
    assert abs(float(S) - s_ref) <= 1e-4 * s_ref
    err = np.max(np.abs(g - g_ref)) / np.max(np.abs(g_ref))
    print(f"S={float(S):.6e} ref={s_ref:.6e} grad rel err={err:.2e}", flush=True)
    assert err <= 2e-3
    af.close()


def test_forward_at_zero_phase_is_the_focus():
    import torch
    from mi355x_pfa import pfa
    sc, freq, ax, ay = _scene(120, 128, seed=4)
    af, sig, fr = _plan(sc, freq, ax, ay, -1, "sinc32")
    S, img, stats = af.forward(torch.zeros(120, dtype=torch.float32, device="cuda"))
    ref, _ = pfa.pfa(sc.sig, freq, ax, ay, -1, interp="sinc32")
    assert np.array_equal(img.cpu().numpy(), ref)
    a = np.abs(ref.astype(np.complex128)) ** 2
    assert abs(float(S) - (a * a).sum() / a.sum() ** 2) <= 1e-5 * float(S)
    af.close()


def test_gradient_against_fp32_finite_differences():
    import torch
    sc, freq, ax, ay = _scene(120, 128, seed=8)
    af, _, _ = _plan(sc, freq, ax, ay, -1, "sinc32")
    phi = torch.tensor(np.random.RandomState(2).normal(0, 0.4, 120), dtype=torch.float32,
                       device="cuda")
    S, img, stats = af.forward(phi)
    g = af.backward(phi, img, stats).cpu().numpy()
    h = 2e-2
    for i in (3, 60, 119):
        e = torch.zeros_like(phi)
        e[i] = h
        sp = float(af.forward(phi + e)[0])
        sm = float(af.forward(phi - e)[0])
        fd = (sp - sm) / (2 * h)
        assert abs(fd - g[i]) <= 3e-2 * np.max(np.abs(g)), (i, fd, g[i])
    af.close()


def test_autofocus_removes_injected_phase_error():
    import torch
    from mi355x_pfa.autofocus import autofocus
    m = n = 256
    sc, freq, ax, ay = _scene(m, n, seed=5, targets=16,
                              phase={"quad_rad": 3.0, "walk_sigma": 0.05})
    af, _, _ = _plan(sc, freq, ax, ay, -1, "sinc32")
    s0 = float(af.forward(torch.zeros(m, dtype=torch.float32, device="cuda"))[0])
    phi, hist = autofocus(af, steps=60, lr=0.1)
    s1 = float(af.forward(phi)[0])
    # residual error modulo the phases S cannot see (constant and linear in pulse index)
    resid = phi.cpu().numpy() + sc.phase_error
    t = np.arange(m)
    resid -= np.polyval(np.polyfit(t, resid, 1), t)
    inj = sc.phase_error - np.polyval(np.polyfit(t, sc.phase_error, 1), t)
    print(f"S {s0:.4e} -> {s1:.4e}; phase error rms {inj.std():.3f} -> {resid.std():.3f} rad",
          flush=True)
    assert s1 > 1.2 * s0
    assert resid.std() < 0.2 * inj.std()
    af.close()
