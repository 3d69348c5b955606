"""CPU checks of the FP64 autofocus reference (tests/af_reference.py): its
pass-2 matrices reproduce the oracle's gathers, its image at phi = 0 is the
oracle's FP64 focus, and its autograd gradient agrees with central finite
differences."""
import numpy as np
import pytest

from af_reference import Reference
from oracle import ref_float as rfl


def _scene(m, n, seed, phase=None):
    from mi355x_pfa import synth
    sc = synth.make_scene(synth.SceneSpec(m, n, targets=6, seed=seed, scene_half=0.2 * min(m, n),
                                          phase_error=phase))
    ax, ay = sc.look()
    return sc, sc.freq(), ax, ay


@pytest.fixture(scope="module")
def ref():
    sc, freq, ax, ay = _scene(60, 64, seed=11)
    return Reference(sc.sig, freq, ax, ay, -1, interp="sinc32"), (sc, freq, ax, ay)


def test_matrices_reproduce_oracle_gather(ref):
    r, _ = ref
    Z, KC, xp, interp, table = r.check_line
    f = np.random.RandomState(0).standard_normal(xp.size) + 1j * np.random.RandomState(1).standard_normal(xp.size)
    np.testing.assert_allclose(Z @ f, rfl._interp_line(KC, xp, f, interp, table), rtol=0, atol=1e-12)


def test_zero_phase_image_is_oracle_fp64_focus(ref):
    r, (sc, freq, ax, ay) = ref
    import torch
    img = r.image(torch.zeros(r.m, dtype=torch.float64)).numpy()
    ora, _ = rfl.pfa(sc.sig, freq, ax, ay, -1, interp="sinc32", store=np.complex128)
    assert np.max(np.abs(img - ora)) <= 1e-9 * np.max(np.abs(ora))


def test_autograd_gradient_matches_finite_differences(ref):
    r, _ = ref
    import torch
    phi = np.random.RandomState(3).normal(0, 0.3, r.m)
    s, g = r.gradient(phi)
    h = 1e-5
    for i in (0, 7, 31, r.m - 1):
        e = np.zeros(r.m); e[i] = h
        sp = float(r.sharpness(torch.tensor(phi + e)))
        sm = float(r.sharpness(torch.tensor(phi - e)))
        assert abs((sp - sm) / (2 * h) - g[i]) <= 1e-6 * np.max(np.abs(g)) + 1e-12 * s
