"""Benchmark: ms per 8192^2 Polar Format frame on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config C3] [--interp sinc32]
                    [--no-fixed] [--no-cpu] [--copy-calibration]

Workload (default): 8167 pulses x 8192 samples -> 8192^2 grid, 32-tap sinc in
both resample passes, FP32, synthetic spotlight phase history (256 point
targets + 512 exponential-power clutter scatterers, 20 dB SNR) synthesised on
the GPU.  A step is one full focus of the frame (range resample + corner-turn,
azimuth resample + window + FFT along cross, FFT along range + detect) with the
phase history already resident in HBM.  The fixed-point (bit-exact) frame on the
same scene is timed as well.  Per-kernel device times come from hipEvents the
library records on the stream its kernels run on.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO))
sys.path.insert(0, str(REPO / "mpfs250t-sar-ifp_amd"))

import numpy as np  # noqa: E402

METRIC = "ms per 8192² PFA frame (FP32 vs BF16-MFMA FFT @ stated error); HBM GB/s"
PEAK_HBM = 8.0e12            # MI355X published HBM3E bandwidth, bytes/s
FPGA_FRAME_MS = 15738.0      # reference silicon, 8167x8192 -> 8192^2 fixed, sinc both passes


def cpu_info():
    model = platform.processor() or ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "cpu_count": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0))}


def make_inputs(spec, device):
    from mi355x_pfa import synth
    from mi355x_pfa.pfa import float_geometry
    sc = synth.make_scene(spec, device=device)
    ax, ay = sc.look()
    geom = float_geometry(ax, ay, f0=sc.sc0, df=sc.scss, n=spec.samples)
    return sc, ax, ay, geom


def algorithmic_bytes_float(m, n, m2, n2, s_in=8, s_mid=8, s_a=8, s_out=4):
    """Compulsory HBM bytes per kernel of the 3-kernel float staging (s_a: bytes
    per element of A, 4 in the matrix-core forms that store it as 16-bit pairs)."""
    return {"range": s_in * m * n + s_mid * n * m,          # read phase history, write g1t
            "azimuth_fft": s_mid * n * m + s_a * n * m2,    # read g1t, write A
            "fft_b_detect": s_a * n * m2 + s_out * m2 * n2}


def algorithmic_bytes_fixed(m, n, G):
    """Compulsory HBM bytes per kernel of the fixed frame (int16 I/Q = 4 bytes,
    uint16 image): the range pass reads the codes and writes the G corner-turned
    rows, the azimuth pass + FFT-A reads them and writes A, FFT-B reads A and
    writes B, detect reads B and writes the image (the B4 staging)."""
    return {"range": 4 * m * n + 4 * m * G, "azimuth_fft": 4 * m * G + 4 * G * G,
            "fft_b": 8 * G * G, "detect": 4 * G * G + 2 * G * G}


def fixed_b3(m, n, G):
    """The B3 staging of the same frame (FFT-B with detect fused), the
    cross-design figure."""
    return 4 * m * n + 8 * m * G + 8 * G * G + 2 * G * G


SETTLE_S = 0.25


def settle(frame, torch, seconds=SETTLE_S):
    """Untimed frames for `seconds` of GPU work before the warm-up frames: a GPU
    that sat idle runs its first ~20 frames about 3 % slower while its clocks
    come up (1.46 against 1.417 ms/frame measured at C3), more than the few
    warm-up frames cover.  Returns the frames run."""
    torch.cuda.synchronize()
    t0, k = time.perf_counter(), 0
    while time.perf_counter() - t0 < seconds:
        for _ in range(10):
            frame()
        k += 10
        torch.cuda.synchronize()
    return k


def timed_frames(ctx, frame, steps, timers, torch):
    """ms per frame of `steps` back-to-back frames (the library's event timers off),
    then the mean event time of each timer over a second, instrumented pass of
    `steps` frames."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        frame()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps * 1e3
    ctx.timer_reset()
    ctx.timer_enable(True)
    for _ in range(steps):
        frame()
    torch.cuda.synchronize()
    kern = {}
    for name, kind in timers:
        ms, cnt = ctx.timer_read(kind)
        kern[name] = ms / max(cnt, 1)
    ctx.timer_enable(False)
    ctx.timer_reset()
    return wall, kern


def time_float(plan, sig, out, steps, warmup, torch):
    from mi355x_pfa import _lib
    settle(lambda: plan.focus_dev(sig, out, detect=True), torch)
    for _ in range(warmup):
        plan.focus_dev(sig, out, detect=True)
    return timed_frames(plan.ctx, lambda: plan.focus_dev(sig, out, detect=True), steps,
                        (("range", _lib.TIMER_RANGE), ("azimuth_fft", _lib.TIMER_AZIMUTH),
                         ("fft_b_detect", _lib.TIMER_FFT_B)), torch)


def time_fixed(sc, ax, ay, geom, steps, warmup, torch, product=None):
    from mi355x_pfa import _lib
    from mi355x_pfa.emulator import FixedPlan
    from mi355x_pfa.tables import fixed_tables
    import ctypes as C
    spec = sc.spec
    m, n, G = spec.pulses, spec.samples, 8192
    freq2 = np.stack([sc.sc0, sc.sc0 + sc.scss], axis=1)
    tabs = fixed_tables({"ax": ax, "ay": ay, "freq": np.pad(freq2, ((0, 0), (0, n - 2))),
                         "KR": geom["KR"], "KC": geom["KC"], "tan_phi": geom["tan_phi"]}, m, n, G)
    plan = FixedPlan(tabs, m, n, G, True, True)
    ctx = plan.ctx
    sig_iq = torch.empty((m, n, 2), dtype=torch.int16, device="cuda")
    e = C.c_int(0)
    ctx.check(_lib.lib().pfa_quantize_iq16_dev(ctx.handle, sc.sig.data_ptr(), m * n,
                                               sig_iq.data_ptr(), C.byref(e)), "quantize")
    out = torch.empty((G, G), dtype=torch.int16, device="cuda")
    settle(lambda: plan.form_image_dev(sig_iq, out), torch)
    for _ in range(warmup):
        plan.form_image_dev(sig_iq, out)
    wall, kern = timed_frames(ctx, lambda: plan.form_image_dev(sig_iq, out), steps,
                              (("range", _lib.TIMER_RANGE), ("azimuth_fft", _lib.TIMER_AZIMUTH),
                               ("fft_b", _lib.TIMER_FFT_B), ("detect", _lib.TIMER_DETECT)), torch)
    plan.close()
    if product is not None:                       # normal score of the uint16 [range][cross] image
        product["normal_score"] = time_product("normal_score", out.view(torch.uint16), torch,
                                               product.pop("_with_cpu"))
    del sig_iq, out
    return wall, kern


def time_product(kind, img, torch, with_cpu, reps=5):
    """A product pass on a device-resident detected frame: "geotiff_u8" (the
    99.7-percentile uint8 data of save_detected_geotiff) on a float32 image, or
    "normal_score" (sar_display.normal_score) on a uint16 image.  Wall time per
    call (host syncs, the O(bins) host steps and the small copies included) and
    the GPU kernels' event time; with_cpu also times the numpy oracle on the
    same frame and checks the outputs are equal."""
    from mi355x_pfa import _lib, display
    from mi355x_pfa._lib import context
    ctx = context(0)
    fn = (lambda: display.stretch_u8(img)[0]) if kind == "geotiff_u8" else \
        (lambda: display.normal_score(img))
    dev_out = fn()
    torch.cuda.synchronize()
    ctx.timer_reset()
    ctx.timer_enable(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps * 1e3
    ms, _ = ctx.timer_read(_lib.TIMER_OTHER)
    ctx.timer_enable(False)
    ctx.timer_reset()
    n = img.numel()
    res = {"ms": round(wall, 3), "kernels_ms": round(ms / reps, 4), "pixels": n,
           "dtype_in": str(img.dtype).replace("torch.", "")}
    if with_cpu:
        from oracle import ref_display as rd
        host = img.cpu().numpy()
        t0 = time.perf_counter()
        ref = rd.geotiff_u8(host)[0] if kind == "geotiff_u8" else rd.normal_score(host)
        res["numpy_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        res["equal_to_numpy"] = bool(np.array_equal(dev_out.cpu().numpy(), ref))
    return res


def _fp64_reference(spec, sc, geom, interp, torch):
    """Centred 2-D FFT in double precision (on the GPU) of the frame's FP32
    k-space (the resample stage every FFT form shares)."""
    from mi355x_pfa.pfa import FloatPlan
    m, n = spec.pulses, spec.samples
    base = FloatPlan(m, n, -1, geom, f0=sc.sc0, df=sc.scss, window=True, interp=interp)
    m2, n2 = base.m2, base.n2
    g2 = base.resample_kspace(sc.sig.cpu().numpy())
    base.close()
    pad = torch.zeros((m2, n2), dtype=torch.complex128, device="cuda")
    pad[:m, :n] = torch.from_numpy(g2).to("cuda")
    del g2
    ref = torch.fft.fftshift(torch.fft.fft2(torch.fft.ifftshift(pad)))
    del pad
    return ref


def precision_forms(spec, sc, geom, interp, forms, steps, warmup, torch):
    """Time each FFT form on the frame and state its error against an FP64
    reference of the same FP32 k-space: RMS and peak positions on this frame,
    PSLR / ISLR on a one-target frame of the same geometry without clutter (in
    the cluttered frame the side-lobe region holds other scatterers)."""
    import dataclasses
    from mi355x_pfa import metrics
    from mi355x_pfa.pfa import FloatPlan
    m, n = spec.pulses, spec.samples
    ref = _fp64_reference(spec, sc, geom, interp, torch)
    m2, n2 = ref.shape
    ref_abs = ref.abs()
    peak = float(ref_abs.max())
    ref_host = ref_abs.float().cpu().numpy()
    out = {}
    for form in forms:
        plan = FloatPlan(m, n, -1, geom, f0=sc.sc0, df=sc.scss, window=True, interp=interp,
                         fft=form)
        det = torch.empty((m2, n2), dtype=torch.float32, device="cuda")
        wall, kern = time_float(plan, sc.sig, det, steps, warmup, torch)
        img = torch.empty((m2, n2), dtype=torch.complex64, device="cuda")
        plan.focus_dev(sc.sig, img, detect=False)
        torch.cuda.synchronize()
        err = float(torch.sqrt(torch.mean((img.to(torch.complex128) - ref).abs() ** 2)) / peak)
        mag = img.abs().cpu().numpy()
        bk = algorithmic_bytes_float(m, n, m2, n2, s_a=8 if form == "fp32" else 4)
        out[form] = {"ms_per_frame": round(wall, 4),
                     "kernels_ms": {k: round(v, 4) for k, v in kern.items()},
                     "kernels_gbps": {k: round(bk[k] / (kern[k] * 1e-3) / 1e9, 1) for k in kern},
                     "error_vs_fp64": {"pn_rms": float(f"{err:.3e}"),
                                       "peak_shift_px": metrics.peak_shift(mag, ref_host, 8)}}
        del det, img, mag
        plan.close()
    del ref, ref_abs, ref_host
    torch.cuda.empty_cache()
    # point-target frame for the side-lobe figures
    spec1 = dataclasses.replace(spec, targets=1, clutter=0, scene_half=50.0)
    sc1, _, _, geom1 = make_inputs(spec1, "cuda")
    ref1 = _fp64_reference(spec1, sc1, geom1, interp, torch).abs().float().cpu().numpy()
    pk = metrics.top_peaks(ref1, 1)[0]
    ps_r, isl_r = metrics.pslr_islr(ref1, pk)
    for form in forms:
        plan = FloatPlan(m, n, -1, geom1, f0=sc1.sc0, df=sc1.scss, window=True, interp=interp,
                         fft=form)
        img = torch.empty((m2, n2), dtype=torch.float32, device="cuda")
        plan.focus_dev(sc1.sig, img, detect=True)
        torch.cuda.synchronize()
        ps, isl = metrics.pslr_islr(img.cpu().numpy(), pk)
        out[form]["error_vs_fp64"].update({
            "pslr_db": round(ps, 3), "islr_db": round(isl, 3),
            "pslr_fp64_db": round(ps_r, 3), "islr_fp64_db": round(isl_r, 3),
            "sidelobe_frame": "1 point target, no clutter, same geometry and shape"})
        del img
        plan.close()
    del sc1, ref1
    torch.cuda.empty_cache()
    return out


def autofocus_c5(interp, steps_af, torch):
    """Gradient autofocus at 4096 x 4096 (64 targets, quadratic 3 rad + random-walk
    phase error): ms of one sharpness forward, of one gradient step (forward +
    adjoint), of the full focus of the same frame, and a steps_af-step Adam run."""
    from mi355x_pfa import synth
    from mi355x_pfa.autofocus import AutofocusPlan, autofocus
    from mi355x_pfa.pfa import FloatPlan
    spec = synth.C5
    m, n = spec.pulses, spec.samples
    sc, ax, ay, geom = make_inputs(spec, "cuda")
    af = AutofocusPlan(sc.sig, geom, f0=sc.sc0, df=sc.scss, sgn=-1, interp=interp)
    phi = torch.zeros(m, dtype=torch.float32, device="cuda")

    def timed(fn, reps=10):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3

    def fwd():
        return af.forward(phi)

    def step():
        S, img, st = af.forward(phi)
        return af.backward(phi, img, st)

    full = FloatPlan(m, n, -1, geom, f0=sc.sc0, df=sc.scss, window=True, interp=interp)
    out = torch.empty((full.m2, full.n2), dtype=torch.float32, device="cuda")
    t_fwd, t_step = timed(fwd), timed(step)
    t_full = timed(lambda: full.focus_dev(sc.sig, out, detect=True))
    full.close()
    s0, img0, _ = af.forward(phi)
    s0, pk0 = float(s0), float(img0.abs().max())
    truth = torch.tensor(-sc.phase_error, dtype=torch.float32, device="cuda")
    st, imgt, _ = af.forward(truth)
    st, pkt = float(st), float(imgt.abs().max())
    del img0, imgt
    t0 = time.perf_counter()
    phi_hat, hist = autofocus(af, steps=steps_af, lr=0.1)
    torch.cuda.synchronize()
    t_af = time.perf_counter() - t0
    s1, img1, _ = af.forward(phi_hat)
    s1, pk1 = float(s1), float(img1.abs().max())
    del img1
    t = np.arange(m)

    def detrend(v):
        return v - np.polyval(np.polyfit(t, v, 1), t)

    def lowpass(v, w=64):                       # components S resolves with few targets
        return np.convolve(v, np.ones(w) / w, mode="same")

    inj = detrend(sc.phase_error)
    res = detrend(phi_hat.cpu().numpy().astype(np.float64) + sc.phase_error)
    w_inj = wrapped_residual(sc.phase_error)
    w_res = wrapped_residual(phi_hat.cpu().numpy().astype(np.float64) + sc.phase_error)
    af.close()
    return {"workload": f"C5: {m} pulses x {n} samples -> 4096^2, {interp}, FP32, "
                        "quadratic 3 rad + random walk 0.05 rad/pulse phase error",
            "ms_sharpness_forward": round(t_fwd, 3),
            "ms_gradient_step": round(t_step, 3),
            "ms_focus_full": round(t_full, 3),
            "gradient_step_over_forward": round(t_step / t_fwd, 2),
            "gradient_step_over_full_focus": round(t_step / t_full, 2),
            "steps": steps_af, "s_autofocus_total": round(t_af, 2),
            "sharpness_before_after_truth": [float(f"{s0:.5e}"), float(f"{s1:.5e}"),
                                             float(f"{st:.5e}")],
            "peak_over_truth_before_after": [round(pk0 / pkt, 4), round(pk1 / pkt, 4)],
            "phase_error_rms_rad_before_after": [round(float(inj.std()), 4),
                                                 round(float(res.std()), 4)],
            "phase_error_rms_rad_lowpass64_before_after": [
                round(float(lowpass(inj)[64:-64].std()), 4),
                round(float(lowpass(res)[64:-64].std()), 4)],
            # modulo 2 pi per pulse (S cannot tell phi from phi + 2 pi k) and
            # modulo the constant + linear phase (an image shift)
            "phase_residual_wrapped_rms_rad_before_after": [round(w_inj[0], 4),
                                                            round(w_res[0], 4)],
            "phase_coherence_before_after": [round(w_inj[1], 4), round(w_res[1], 4)]}


def wrapped_residual(res, pad=64):
    """A per-pulse phase residual modulo 2 pi and modulo its best constant +
    linear phase: the ramp b maximising |sum_i exp(j (res_i - b i))| (a
    zero-padded FFT, then a parabolic refinement), the coherence
    |sum_i exp(j (res_i - b i))| / m, and the RMS of the wrapped remainder."""
    m = res.size
    z = np.exp(1j * res)
    spec = np.fft.fft(z, pad * m)
    k = int(np.argmax(np.abs(spec)))
    a0, a1, a2 = (np.abs(spec[(k + d) % (pad * m)]) for d in (-1, 0, 1))
    den = a0 - 2 * a1 + a2
    kk = k + (0.5 * (a0 - a2) / den if den != 0 else 0.0)
    b = 2 * np.pi * kk / (pad * m)
    c = np.sum(z * np.exp(-1j * b * np.arange(m)))
    r = np.angle(z * np.exp(-1j * b * np.arange(m)) * np.exp(-1j * np.angle(c)))
    return float(np.sqrt(np.mean(r * r))), float(np.abs(c) / m)


def cpu_baseline_full(sc, ax, ay):
    """The reference's float focus restated in numpy (oracle/ref_float.pfa:
    np.interp in both passes, hamming2d, pad, centred fft2; form_image_pfa.py
    :165-246) on the whole frame on this host, then |img| (accel.py:44): one
    run, numpy's interp and FFT single-threaded."""
    from oracle import ref_float as rfl
    sig = sc.sig.cpu().numpy()
    freq = sc.freq()
    t0 = time.perf_counter()
    img, _ = rfl.pfa(sig, freq, ax, ay, -1)
    np.abs(img).astype(np.float32)
    t = time.perf_counter() - t0
    del img, sig, freq
    return t


def cpu_baseline(spec, interp, geom, budget_rows=800):
    """The numpy oracle on the host, on a bounded sample of the frame: pass 1 on
    `budget_rows` pulses and pass 2 on `budget_rows` range bins of the real
    geometry (each scaled to the full frame), plus the full 2-D FFT + detect."""
    from oracle import ref_float as rfl
    from oracle import ref_fixed as rf
    m, n = spec.pulses, spec.samples
    rs = np.random.RandomState(0)
    table = rf.sinc_table_q15()
    row = (rs.standard_normal(n) + 1j * rs.standard_normal(n)).astype(np.complex64)
    col = (rs.standard_normal(m) + 1j * rs.standard_normal(m)).astype(np.complex64)
    f = (spec.f0 - spec.bandwidth / 2) + np.arange(n) * (spec.bandwidth / (n - 1))
    KR, KC, pr, tan_s = geom["KR"], geom["KC"], geom["pr"], geom["tan_s"]
    t0 = time.perf_counter()
    for i in range(budget_rows):
        kr_i = (2.0 * f / rfl.C_LIGHT) * pr[(i * 7) % m]
        rfl._interp_line(KR, kr_i, row, interp, table)
    t1 = time.perf_counter()
    for j in range(budget_rows):
        rfl._interp_line(KC, KR[(j * 11) % n] * tan_s, col, interp, table)
    t2 = time.perf_counter()
    pad = np.zeros((8192, 8192), np.complex64)
    pad[:256, :n] = row
    t3 = time.perf_counter()
    img = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(pad)))
    np.abs(img).astype(np.float32)
    t4 = time.perf_counter()
    p1 = (t1 - t0) * m / budget_rows
    p2 = (t2 - t1) * n / budget_rows
    total = p1 + p2 + (t4 - t3)
    return {"value": round(total * 1e3, 1), "unit": "ms/frame", "cores": 1, "kind": "port",
            "sample": (f"oracle/ref_float {interp}: pass 1 on {budget_rows} of {m} pulses and "
                       f"pass 2 on {budget_rows} of {n} range bins of this frame's geometry, "
                       f"each scaled to the frame, plus the full 8192^2 numpy FFT + detect "
                       f"(numpy interp/fft run single-threaded)"),
            "measured_s": round(t2 - t0 + t4 - t3, 2),
            "split_ms": {"pass1": round(p1 * 1e3), "pass2": round(p2 * 1e3),
                         "fft_detect": round((t4 - t3) * 1e3)}, **cpu_info()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="C3", choices=["C1", "C2", "C3"])
    ap.add_argument("--interp", default="sinc32", choices=["sinc32", "linear"])
    ap.add_argument("--no-fixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quick-scene", action="store_true",
                    help="16 targets, no clutter (same shapes; for profiler runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 512 MiB device copy (counter calibration)")
    ap.add_argument("--fft", default="fp32", choices=["fp32", "bf16", "fp16", "bf16-split"],
                    help="FFT form of the timed frames (value)")
    ap.add_argument("--forms", default="fp32,bf16,fp16",
                    help="FFT forms timed beside, with their error against FP64 ('' = none)")
    ap.add_argument("--autofocus-steps", type=int, default=50,
                    help="gradient autofocus steps at 4096^2 (0 = skip)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X: the focus path is built and timed on a single "
              "GPU (multi-GPU sharding is a design note only); run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from mi355x_pfa import synth
    from mi355x_pfa.pfa import FloatPlan
    from mi355x_pfa._lib import context

    spec = {"C1": synth.C1, "C2": synth.C2, "C3": synth.C3}[a.config]
    if a.quick_scene:
        import dataclasses
        spec = dataclasses.replace(spec, targets=16, clutter=0)
    m, n = spec.pulses, spec.samples
    print(f"[bench] synthesising {a.config} {m}x{n} on the GPU", file=sys.stderr, flush=True)
    sc, ax, ay, geom = make_inputs(spec, "cuda")
    ctx = context(0)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    plan = FloatPlan(m, n, -1, geom, f0=sc.sc0, df=sc.scss, window=True, interp=a.interp,
                     fft=a.fft)
    out = torch.empty((plan.m2, plan.n2), dtype=torch.float32, device="cuda")
    print("[bench] float frames", file=sys.stderr, flush=True)
    wall, kern = time_float(plan, sc.sig, out, a.steps, a.warmup, torch)
    ctx.fallback_counts(reset=True)                  # one more frame, counted on its own
    plan.focus_dev(sc.sig, out, detect=True)
    torch.cuda.synchronize()
    fb = ctx.fallback_counts(reset=True)
    bytes_k = algorithmic_bytes_float(m, n, plan.m2, plan.n2, s_a=8 if a.fft == "fp32" else 4)
    frame_dev = sum(kern.values())
    B3 = sum(bytes_k.values())
    dom = max(kern, key=kern.get)
    achieved = bytes_k[dom] / (kern[dom] * 1e-3)
    traffic = None
    tfile = REPO / "profiles" / f"traffic_{a.config.lower()}_{a.interp}.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get(dom)
        except Exception:
            traffic = None
    prof_us = None                    # the same kernel's average under rocprofv3 --kernel-trace
    pfile = REPO / "profiles" / f"kernel_trace_{a.config.lower()}.json"
    if pfile.exists() and a.fft == "fp32" and a.interp == "sinc32":
        try:
            prof_us = json.loads(pfile.read_text())["float_fp32"].get(dom)
        except Exception:
            prof_us = None
    res = {
        "metric": METRIC, "value": round(wall, 4), "unit": "ms/frame",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(wall, 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if a.fft == "fp32" else f"f32 resample, {a.fft} matrix-core FFT",
        "data": "synthetic (GPU-synthesised spotlight phase history)",
        "config": {"workload": f"{a.config}: {m} pulses x {n} samples -> "
                               f"{plan.m2}x{plan.n2} PFA frame, {a.interp} both passes, FP32",
                   "pulses": m, "samples": n, "grid": plan.m2, "interp": a.interp,
                   "staging": "B3 (3 kernels)"},
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 1),
                     "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM, 4), "traffic": traffic,
                     "kernel": dom, "bytes_per_launch": bytes_k[dom],
                     "time_source": "hipEvents on the library's stream, mean over a second pass of as many frames (the timed frames run without events)",
                     "rocprof_avg_us": prof_us,
                     "frac_at_rocprof_avg": (round(bytes_k[dom] / (prof_us * 1e-6) / PEAK_HBM, 4)
                                             if prof_us else None)},
        "frame_device_ms": round(frame_dev, 4),
        "frame_hbm_gbps_B3": round(B3 / (frame_dev * 1e-3) / 1e9, 1),
        "frame_roofline_frac_B3": round(B3 / (frame_dev * 1e-3) / PEAK_HBM, 4),
        "kernels_ms": {k: round(v, 4) for k, v in kern.items()},
        "kernels_gbps": {k: round(bytes_k[k] / (kern[k] * 1e-3) / 1e9, 1) for k in kern},
        "fallback_outputs_per_frame": list(fb),
        "settle_s": SETTLE_S,
    }
    print("[bench] product passes", file=sys.stderr, flush=True)
    res["products"] = {"geotiff_u8": time_product("geotiff_u8", out, torch, not a.no_cpu)}
    # end to end through the host-buffer entry point (pageable numpy in and
    # out, as the reference's ctypes caller would pass them): never `value`
    sig_host = sc.sig.cpu().numpy()
    plan.focus(sig_host, detect=True)
    t0 = time.perf_counter()
    for _ in range(3):
        plan.focus(sig_host, detect=True)
    res["end_to_end_host_buffers_ms"] = round((time.perf_counter() - t0) / 3 * 1e3, 2)
    del sig_host
    # scene geometry from the full freq (m, n) array: the O(m n) kr extrema on
    # the GPU, from a host array (staged) and from a device tensor
    from mi355x_pfa.pfa import float_geometry
    freq_host = sc.freq()
    freq_dev = torch.from_numpy(freq_host).to("cuda")
    g = {}
    for name, f in (("host_array", freq_host), ("device_tensor", freq_dev)):
        float_geometry(ax, ay, freq=f)
        t0 = time.perf_counter()
        for _ in range(3):
            float_geometry(ax, ay, freq=f)
        g[name] = round((time.perf_counter() - t0) / 3 * 1e3, 2)
    res["geometry_full_freq_ms"] = g
    del freq_host, freq_dev
    # the pipeline's table preparation (sar_pipeline.prepare_tables) from the
    # scene's PVP arrays, with no (m, n) host array
    from mi355x_pfa.pipeline import prepare_tables
    reader = synth.SceneReader(sc)
    prepare_tables(reader, reader.cphd_meta)
    t0 = time.perf_counter()
    for _ in range(3):
        prepare_tables(reader, reader.cphd_meta)
    res["prepare_tables_ms"] = round((time.perf_counter() - t0) / 3 * 1e3, 2)
    del out
    plan.close()
    if a.copy_calibration:
        # plain 512 MiB device copies of known size for the HBM byte counters:
        # torch's elementwise copy and the HIP runtime's copy kernel
        import ctypes
        x = torch.empty(512 << 20, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_int, ctypes.c_void_p]
        for _ in range(3):
            y.copy_(x)
            hip.hipMemcpyAsync(y.data_ptr(), x.data_ptr(), 512 << 20, 3,
                               torch.cuda.current_stream().cuda_stream)
        # read-mostly kernels of known size: torch's sum of the 512 MiB, and the
        # library's uint8 stretch of 128 Mi float32 values (512 MiB read, 128 MiB
        # written, 16-byte loads like the focus kernels)
        from mi355x_pfa import _lib as L
        xf = x.view(torch.float32)
        u8 = torch.empty(xf.numel(), dtype=torch.uint8, device="cuda")
        for _ in range(3):
            x.sum()
            ctx.check(L.lib().pfa_stretch_u8_dev(ctx.handle, xf.data_ptr(), xf.numel(), 1.0,
                                                 u8.data_ptr()), "stretch")
        torch.cuda.synchronize()
        res["copy_calibration_bytes"] = 512 << 20
        del x, y, xf, u8
    forms = [f for f in a.forms.split(",") if f]
    if forms:
        print("[bench] FFT forms vs FP64", file=sys.stderr, flush=True)
        res["fft_forms"] = precision_forms(spec, sc, geom, a.interp, forms, a.steps, a.warmup,
                                           torch)
    if a.autofocus_steps > 0:
        print("[bench] gradient autofocus 4096^2", file=sys.stderr, flush=True)
        res["autofocus"] = autofocus_c5(a.interp, a.autofocus_steps, torch)
    if not a.no_fixed and a.config == "C3":
        print("[bench] fixed-point frames", file=sys.stderr, flush=True)
        res["products"]["_with_cpu"] = not a.no_cpu
        fwall, fk = time_fixed(sc, ax, ay, geom, max(3, a.steps // 2), 2, torch, res["products"])
        fb = algorithmic_bytes_fixed(m, n, 8192)
        fdev = sum(fk.values())
        fdom = max(fk, key=fk.get)
        res["fixed"] = {"ms_per_frame": round(fwall, 3),
                        "kernels_ms": {k: round(v, 4) for k, v in fk.items()},
                        "kernels_gbps": {k: round(fb[k] / (fk[k] * 1e-3) / 1e9, 1) for k in fk},
                        "roofline": {"bound": "hbm", "kernel": fdom,
                                     "achieved": round(fb[fdom] / (fk[fdom] * 1e-3) / 1e9, 1),
                                     "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                                     "frac": round(fb[fdom] / (fk[fdom] * 1e-3) / PEAK_HBM, 4),
                                     "bytes_per_launch": fb[fdom]},
                        "frame_device_ms": round(fdev, 4),
                        "frame_hbm_gbps_B4": round(sum(fb.values()) / (fdev * 1e-3) / 1e9, 1),
                        "frame_roofline_frac_B4": round(sum(fb.values()) / (fdev * 1e-3) / PEAK_HBM, 4),
                        "frame_roofline_frac_B3": round(fixed_b3(m, n, 8192) / (fdev * 1e-3) / PEAK_HBM, 4),
                        "speedup_vs_fpga_silicon": round(FPGA_FRAME_MS / fwall, 1)}
    if not a.no_cpu:
        print("[bench] cpu baseline (numpy oracle: the whole frame, np.interp)", file=sys.stderr,
              flush=True)
        t_full = cpu_baseline_full(sc, ax, ay)
        sample = cpu_baseline(spec, a.interp, geom)
        res["cpu_baseline"] = {
            "value": round(t_full * 1e3, 1), "unit": "ms/frame", "cores": 1, "kind": "port",
            "sample": (f"the whole {m} x {n} -> {plan.m2}^2 frame once: oracle/ref_float.pfa "
                       "(the reference's np.interp focus restated in numpy) + |img|, "
                       "numpy interp/fft single-threaded"),
            "sinc32_estimate": {k: sample[k] for k in ("value", "unit", "sample", "measured_s",
                                                       "split_ms")},
            **cpu_info()}
        # the reference's numpy table preparation on the same scene, O(m n)
        from oracle import ref_float as rfl
        t0 = time.perf_counter()
        rfl.prepare_tables(reader, reader.cphd_meta)
        res["cpu_baseline"]["prepare_tables_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
