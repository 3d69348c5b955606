"""The allocation-free fixed focus (pfa_form_image_fixed_dev_noalloc) against
the allocating entry point (pfa_form_image_fixed_dev), through the C ABI,
bit for bit: two calls in a row on the same device, the second on a smaller
scene, and a grid above the scratch limit refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fn():
    from mi355x_pfa import _lib
    f = _lib.lib().pfa_form_image_fixed_dev_noalloc
    f.argtypes = [C.c_void_p] * 5
    f.restype = C.c_int
    return f


def _case(m, n, grid, seed, rs, az):
    import torch
    from mi355x_pfa import synth
    from mi355x_pfa.emulator import FixedPlan, quantize_iq16
    from mi355x_pfa.pfa import float_geometry
    from mi355x_pfa.tables import fixed_tables
    sc = synth.make_scene(synth.SceneSpec(m, n, targets=6, seed=seed, scene_half=0.1 * m))
    ax, ay = sc.look()
    freq = sc.freq()
    g = float_geometry(ax, ay, freq=freq)
    tabs = fixed_tables({"ax": ax, "ay": ay, "freq": freq, "KR": g["KR"], "KC": g["KC"],
                         "tan_phi": g["tan_phi"]}, m, n, grid)
    iq, _ = quantize_iq16(sc.sig)
    plan = FixedPlan(tabs, m, n, grid, rs, az)
    sig = torch.from_numpy(iq).cuda()
    return plan, sig


def _run(plan, sig, noalloc):
    import torch
    G = plan.grid
    out = torch.empty((G, G), dtype=torch.int16, device="cuda")
    ea = torch.empty(G, dtype=torch.int32, device="cuda")
    eb = torch.empty(G, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    if noalloc:
        st = _fn()(plan._h, sig.data_ptr(), out.data_ptr(), ea.data_ptr(), eb.data_ptr())
        plan.ctx.check(st, "form_image_fixed_dev_noalloc")
    else:
        plan.form_image_dev(sig, out, ea, eb)
    plan.ctx.synchronize()
    return out.cpu().numpy(), ea.cpu().numpy(), eb.cpu().numpy()


def test_noalloc_equals_allocating_entry_point_twice():
    cases = [(1500, 1400, 2048, 3, True, True), (300, 250, 512, 4, True, False)]
    for m, n, grid, seed, rs, az in cases:
        print(f"noalloc fixed focus {m}x{n} -> {grid}^2 sinc=({rs},{az})", flush=True)
        plan, sig = _case(m, n, grid, seed, rs, az)
        a = _run(plan, sig, noalloc=True)
        b = _run(plan, sig, noalloc=False)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert a[0].max() > 0
        plan.close()


def test_noalloc_without_exponent_outputs_and_limit():
    import torch
    plan, sig = _case(200, 180, 256, 5, False, False)
    G = plan.grid
    out = torch.empty((G, G), dtype=torch.int16, device="cuda")
    torch.cuda.synchronize()
    plan.ctx.check(_fn()(plan._h, sig.data_ptr(), out.data_ptr(), None, None), "noalloc")
    plan.ctx.synchronize()
    ref = _run(plan, sig, noalloc=False)[0]
    assert np.array_equal(out.cpu().numpy(), ref)
    plan.close()
    big, sig2 = _case(100, 100, 4096, 6, False, False)
    out2 = torch.empty((4096, 4096), dtype=torch.int16, device="cuda")
    assert _fn()(big._h, sig2.data_ptr(), out2.data_ptr(), None, None) == -2
    big.close()
