"""GPU checks at the production shapes and of the C ABI's lifetime rules.

* fixed mode at 8167 x 8192 and 5634 x 4319 -> 8192^2, bit-exact: the image
  and both exponent vectors equal the oracle's (oracle/ref_fixed
  .form_image_fixed, the restatement of silicon_emulator.form_image) through
  SHA-256 digests computed in the build container by
  tests/golden/make_fixed_digests.py from inputs regenerated here bit for bit
  (their own digests are checked first);
* float mode at 8167 x 8192, np.interp in both passes, against the oracle's
  numpy pfa run on this host: peak-normalised RMS <= 1e-5, same top-8 peaks;
* gradient autofocus at the 4096-pulse size (the 4096-point cross line):
  forward at zero phase = the focus, adjoint gradient = central FP32 finite
  differences of the library's own forward on pulses across the aperture;
* two autofocus plans interleaved on one context keep their own checkpoints;
* a plan outlives its context: later calls fail with PFA_ERR_STATE;
* form_image_pfa.pfa(sig, freq, ax, ay, sgn) in one C call;
* detect_fixed over all 2^32 int16 pairs against the integer square root.
"""
import ctypes as C
import json

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

DIGESTS = json.loads((GOLDEN / "fixed_digests.json").read_text())


# ------------------------------------------------------------------ fixed, bit-exact at full size
@pytest.fixture(scope="module")
def fixed_inputs_cache():
    return {}


@pytest.mark.parametrize("case", sorted(DIGESTS))
def test_fixed_fullsize_bit_exact(case, fixed_inputs_cache):
    import torch
    import fixed_inputs as fi
    from mi355x_pfa.emulator import FixedPlan
    d = DIGESTS[case]
    config = d["config"]
    print(f"fixed {case}: {d['m']}x{d['n']} -> {d['grid']}^2 range_sinc={d['range_sinc']} "
          f"az_sinc={d['az_sinc']}", flush=True)
    if config not in fixed_inputs_cache:
        m, n, tabs = fi.scene_tables(config)
        iq = fi.phase_history_iq(m, n, d["seed"], device="cuda")
        torch.cuda.synchronize()
        assert fi.sha(iq.cpu().numpy()) == d["input_sha256"], "regenerated input differs"
        fixed_inputs_cache.clear()
        fixed_inputs_cache[config] = (m, n, tabs, iq)
    m, n, tabs, iq = fixed_inputs_cache[config]
    for k, h in d["tables_sha256"].items():
        assert fi.sha(tabs[k]) == h, f"regenerated table {k} differs"
    G = d["grid"]
    plan = FixedPlan(tabs, m, n, G, d["range_sinc"], d["az_sinc"])
    out = torch.empty((G, G), dtype=torch.int16, device="cuda")
    ea = torch.empty(G, dtype=torch.int32, device="cuda")
    eb = torch.empty(G, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    plan.form_image_dev(iq, out, ea, eb)
    plan.ctx.synchronize()
    img = out.view(torch.uint16).cpu().numpy()
    print(f"image max {int(img.max())} (oracle {d['image_max']}), emax_a "
          f"{int(ea.max())} ({d['emax_a']}), emax_b {int(eb.max())} ({d['emax_b']})", flush=True)
    assert fi.sha(ea.cpu().numpy()) == d["exp_a_sha256"]
    assert fi.sha(eb.cpu().numpy()) == d["exp_b_sha256"]
    assert fi.sha(img) == d["image_sha256"]
    plan.close()


# ------------------------------------------------------------------ float at full size vs the oracle
def test_float_fullsize_linear_matches_numpy_oracle():
    """C3 frame (8167 x 8192 -> 8192^2) with the reference's np.interp in both
    passes against the oracle's numpy pfa of the same complex64 history on
    this host (tolerance: peak-normalised RMS 1e-5, identical top-8 peaks)."""
    import dataclasses
    import time
    import torch
    from mi355x_pfa import metrics, synth
    from mi355x_pfa.pfa import FloatPlan, float_geometry
    from oracle import ref_float as rfl
    spec = dataclasses.replace(synth.C3, targets=32, clutter=64, seed=21)
    print(f"float full-size oracle check {spec.pulses}x{spec.samples}", flush=True)
    sc = synth.make_scene(spec, device="cuda")
    ax, ay = sc.look()
    geom = float_geometry(ax, ay, f0=sc.sc0, df=sc.scss, n=spec.samples)
    plan = FloatPlan(spec.pulses, spec.samples, -1, geom, f0=sc.sc0, df=sc.scss, window=True,
                     interp="linear")
    img = torch.empty((plan.m2, plan.n2), dtype=torch.complex64, device="cuda")
    torch.cuda.synchronize()
    plan.focus_dev(sc.sig, img, detect=False)
    plan.ctx.synchronize()
    mine = img.cpu().numpy()
    sig = sc.sig.cpu().numpy()
    del img, sc.sig
    plan.close()
    torch.cuda.empty_cache()
    t0 = time.time()
    ref, _ = rfl.pfa(sig, sc.freq(), ax, ay, -1)
    print(f"numpy oracle {time.time() - t0:.1f} s", flush=True)
    peak = np.abs(ref).max()
    rms = float(np.sqrt(np.mean(np.abs(mine.astype(np.complex128) - ref) ** 2)) / peak)
    print(f"pn-RMS vs oracle {rms:.2e}", flush=True)
    assert rms <= 1e-5
    assert metrics.peak_shift(np.abs(mine), np.abs(ref), 8) == 0


# ------------------------------------------------------------------ autofocus at 4096 pulses
def _af_scene(m, n, seed):
    from mi355x_pfa import synth
    from mi355x_pfa.pfa import float_geometry
    spec = synth.SceneSpec(m, n, targets=64, seed=seed, phase_error={"quad_rad": 3.0,
                                                                     "walk_sigma": 0.05})
    sc = synth.make_scene(spec, device="cuda")
    ax, ay = sc.look()
    geom = float_geometry(ax, ay, f0=sc.sc0, df=sc.scss, n=n)
    return sc, geom


@pytest.mark.parametrize("m,n", [(4096, 4096), (4096, 512), (2000, 300), (1500, 1100)])
def test_autofocus_large_cross_sizes(m, n):
    """The 2048- and 4096-point cross lines (the 4096 x 4096 autofocus frame
    among them): forward at zero phase is the focus bit for bit; the adjoint
    gradient equals central finite differences of the forward (FP32 image,
    S in FP64) on 8 pulses spread over the aperture to 5 % of max |grad|."""
    import torch
    from mi355x_pfa.autofocus import AutofocusPlan
    from mi355x_pfa.pfa import FloatPlan
    print(f"autofocus {m}x{n}", flush=True)
    sc, geom = _af_scene(m, n, seed=m + n)
    af = AutofocusPlan(sc.sig, geom, f0=sc.sc0, df=sc.scss, sgn=-1, interp="sinc32")
    full = FloatPlan(m, n, -1, geom, f0=sc.sc0, df=sc.scss, window=True, interp="sinc32")
    ref = torch.empty((full.m2, full.n2), dtype=torch.complex64, device="cuda")
    torch.cuda.synchronize()
    full.focus_dev(sc.sig, ref, detect=False)
    full.ctx.synchronize()
    zero = torch.zeros(m, dtype=torch.float32, device="cuda")
    S0, img0, _ = af.forward(zero)
    torch.cuda.synchronize()
    assert torch.equal(img0, ref)
    full.close()
    del ref, img0
    phi = torch.tensor(np.random.RandomState(m).normal(0.0, 0.3, m), dtype=torch.float32,
                       device="cuda")
    S, img, stats = af.forward(phi)
    g = af.backward(phi, img, stats).cpu().numpy()
    gmax = float(np.abs(g).max())
    h = 0.05
    for i in np.linspace(0, m - 1, 8).astype(int):
        e = torch.zeros_like(phi)
        e[i] = h
        sp = float(af.forward(phi + e)[0])
        sm = float(af.forward(phi - e)[0])
        fd = (sp - sm) / (2 * h)
        print(f"pulse {i}: grad {g[i]:.4e} fd {fd:.4e} (max |grad| {gmax:.3e})", flush=True)
        assert abs(fd - g[i]) <= 5e-2 * gmax, (i, fd, g[i])
    af.close()


def test_two_autofocus_plans_keep_their_own_checkpoints():
    import torch
    from mi355x_pfa.autofocus import AutofocusPlan
    sa, ga = _af_scene(300, 200, seed=1)
    sb, gb = _af_scene(700, 260, seed=2)
    pa = AutofocusPlan(sa.sig, ga, f0=sa.sc0, df=sa.scss, sgn=-1, interp="sinc32")
    phi_a = torch.zeros(300, dtype=torch.float32, device="cuda")
    phi_b = torch.full((700,), 0.1, dtype=torch.float32, device="cuda")
    Sa, ia, _ = pa.forward(phi_a)
    ia = ia.clone()
    pb = AutofocusPlan(sb.sig, gb, f0=sb.sc0, df=sb.scss, sgn=-1, interp="sinc32")
    Sb, ib, _ = pb.forward(phi_b)
    ib = ib.clone()
    Sa2, ia2, sta = pa.forward(phi_a)            # after the second plan's checkpoint
    ga2 = pa.backward(phi_a, ia2, sta)
    Sb2, ib2, _ = pb.forward(phi_b)
    torch.cuda.synchronize()
    assert torch.equal(ia, ia2) and torch.equal(ib, ib2)
    assert float(Sa) == float(Sa2) and float(Sb) == float(Sb2)
    pa1 = AutofocusPlan(sa.sig, ga, f0=sa.sc0, df=sa.scss, sgn=-1, interp="sinc32")
    S1, i1, st1 = pa1.forward(phi_a)
    g1 = pa1.backward(phi_a, i1, st1)
    torch.cuda.synchronize()
    assert torch.equal(ga2, g1)
    for p in (pa, pb, pa1):
        p.close()


def test_plan_outliving_its_context_fails_cleanly():
    from mi355x_pfa import _lib
    from mi355x_pfa.pfa import float_geometry
    from mi355x_pfa import synth
    L = _lib.lib()
    sc = synth.make_scene(synth.SceneSpec(64, 64, targets=1, seed=1, scene_half=5.0),
                          signal=False)
    ax, ay = sc.look()
    g = float_geometry(ax, ay, f0=sc.sc0, df=sc.scss, n=64)
    ctx = C.c_void_p()
    assert L.pfa_create(C.byref(ctx), 0) == 0
    plan = C.c_void_p()
    keep = [np.ascontiguousarray(g[k]) for k in ("pr", "order", "tan_s", "KR", "KC")]
    keep += [np.ascontiguousarray(sc.sc0), np.ascontiguousarray(sc.scss)]
    st = L.pfa_plan_float(ctx, C.byref(plan), 64, 64, -1, *[a.ctypes.data for a in keep],
                          None, None, None, 0)
    assert st == 0
    L.pfa_destroy(ctx)
    sig = np.zeros((64, 64), np.complex64)
    out = np.zeros((64, 64), np.complex64)
    assert L.pfa_focus_f32(plan, sig.ctypes.data, None, out.ctypes.data, 0) == -6
    L.pfa_plan_destroy(plan)


@pytest.mark.parametrize("m,n,sgn", [(96, 80, -1), (130, 72, 1)])
def test_pfa_one_call_matches_oracle(m, n, sgn):
    """pfa_pfa_f32 from (sig, freq, ax, ay, sgn): image within 1e-5 of the
    oracle's numpy pfa with the same top-8 peaks, and pfa's geo exactly."""
    from mi355x_pfa import _lib, metrics, synth
    from mi355x_pfa._lib import context
    from oracle import ref_float as rfl
    sc = synth.make_scene(synth.SceneSpec(m, n, targets=5, seed=m, scene_half=15.0))
    ax, ay = sc.look()
    freq = np.ascontiguousarray(sc.freq())
    ref, geo_r = rfl.pfa(sc.sig, freq, ax, ay, sgn)
    ctx = context(0)
    out = np.empty(ref.shape, np.complex64)
    geo = np.empty(6)
    sig = np.ascontiguousarray(sc.sig, np.complex64)
    ctx.check(_lib.lib().pfa_pfa_f32(ctx.handle, sig.ctypes.data, m, n, freq.ctypes.data,
                                     np.ascontiguousarray(ax).ctypes.data,
                                     np.ascontiguousarray(ay).ctypes.data, sgn, 0,
                                     out.ctypes.data, 0, geo.ctypes.data), "pfa_pfa_f32")
    peak = np.abs(ref).max()
    assert np.sqrt(np.mean(np.abs(out - ref) ** 2)) / peak <= 1e-5
    assert metrics.peak_shift(np.abs(out), np.abs(ref), 8) == 0
    assert np.array_equal(geo, [geo_r["dc"], geo_r["dr"], *geo_r["dhat"], *geo_r["chat"]])


# ------------------------------------------------------------------ detect, exhaustive
def test_detect_fixed_all_int16_pairs():
    """floor(sqrt(I^2 + Q^2)) for every (I, Q) in int16^2, in 16 chunks of 2^28
    pairs made on the device; checked against s^2 <= I^2 + Q^2 < (s + 1)^2."""
    import torch
    from mi355x_pfa import _lib
    from mi355x_pfa._lib import context
    ctx = context(0)
    torch.cuda.synchronize()
    q = torch.arange(-32768, 32768, dtype=torch.int32, device="cuda")
    rows = 4096
    out = torch.empty(rows * 65536, dtype=torch.int16, device="cuda")
    for c in range(16):
        re = torch.arange(-32768 + c * rows, -32768 + (c + 1) * rows, dtype=torch.int32,
                          device="cuda")
        z = torch.stack([re[:, None].expand(rows, 65536), q[None, :].expand(rows, 65536)],
                        dim=-1).to(torch.int16).contiguous()
        torch.cuda.synchronize()
        ctx.check(_lib.lib().pfa_detect_fixed_dev(ctx.handle, z.data_ptr(), rows * 65536,
                                                  out.data_ptr()), "detect")
        ctx.synchronize()
        s = out.view(torch.uint16).to(torch.int64).view(rows, 65536)
        v = (re.to(torch.int64)[:, None] ** 2 + q.to(torch.int64)[None, :] ** 2)
        ok = (s * s <= v) & ((s + 1) * (s + 1) > v)
        assert bool(ok.all()), f"chunk {c}"
        del z, s, v, ok
    print("detect: all 2^32 pairs exact", flush=True)
