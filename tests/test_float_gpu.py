"""Float mode on the GPU against form_image_pfa.py.

Tolerance (the float contract): peak-normalised RMS error
    rms(img_gpu - img_ref) / max|img_ref| <= 1e-5
on the complex image, with identical point-target peak positions.  The
reference fixtures come from the reference itself; larger and SINC32 cases use
the oracle (oracle/ref_float.py, pinned to those fixtures)."""
import numpy as np
import pytest

from conftest import load_golden
from oracle import ref_float as rfl

pytestmark = pytest.mark.gpu
TOL = 1e-5


def pn_rms(a, b):
    return float(np.sqrt(np.mean(np.abs(a.astype(np.complex128) - b) ** 2)) / np.max(np.abs(b)))


def top_peaks(img, k, guard=3):
    a = np.abs(img).astype(np.float64).copy()
    out = []
    for _ in range(k):
        r, c = np.unravel_index(np.argmax(a), a.shape)
        out.append((int(r), int(c)))
        a[max(r - guard, 0):r + guard + 1, max(c - guard, 0):c + guard + 1] = 0
    return out


@pytest.fixture(scope="module")
def gp():
    from mi355x_pfa import pfa
    return pfa


@pytest.mark.parametrize("name", ["small", "rect"])
def test_resample_kspace_matches_reference(gp, name):
    g = load_golden(f"float_{name}.npz")
    print(f"resample_kspace {name}", flush=True)
    g2, geo = gp.resample_kspace(g["sig"], g["freq"], g["ax"], g["ay"])
    ref = g["g2"]
    err = np.max(np.abs(g2.astype(np.complex128) - ref)) / np.max(np.abs(ref))
    assert err <= 1e-6, err
    # resample_kspace's geo is the unpadded one; focus scales it by n / n2 and
    # m / m2 (form_image_pfa.py:237-238) into the fixture's (pfa's) geo
    m, n = ref.shape
    m2, n2 = 1 << (m - 1).bit_length(), 1 << (n - 1).bit_length()
    assert geo["dr"] * (n / n2) == g["geo"][1] and geo["dc"] * (m / m2) == g["geo"][0]
    assert [*geo["dhat"], *geo["chat"]] == list(g["geo"][2:])


@pytest.mark.parametrize("name", ["small", "rect"])
def test_focus_and_pfa_match_reference(gp, name):
    g = load_golden(f"float_{name}.npz")
    sgn = int(g["sgn"])
    print(f"focus/pfa {name}", flush=True)
    img, geo = gp.focus(g["g2"], {"dc": 1.0, "dr": 1.0, "dhat": (1, 0), "chat": (0, 1)}, sgn)
    assert pn_rms(img, g["img"]) <= TOL
    img2, geo2 = gp.pfa(g["sig"], g["freq"], g["ax"], g["ay"], sgn)
    assert pn_rms(img2, g["img"]) <= TOL
    assert np.allclose([geo2["dc"], geo2["dr"], *geo2["dhat"], *geo2["chat"]], g["geo"],
                       rtol=1e-15, atol=0)
    det = gp.GpuBackend().focus(g["sig"], {"freq": g["freq"], "ax": g["ax"], "ay": g["ay"],
                                           "sgn": sgn})
    assert det.dtype == np.float32 and det.shape == g["det"].shape
    assert float(np.sqrt(np.mean((det - g["det"]) ** 2)) / g["det"].max()) <= TOL
    assert top_peaks(img2, 3) == top_peaks(g["img"], 3)


def _scene(m, n, seed, targets=16, half=None, desc=False):
    from mi355x_pfa import synth
    sc = synth.make_scene(synth.SceneSpec(m, n, targets=targets, seed=seed,
                                          scene_half=half or 0.2 * min(m, n)))
    freq = sc.freq()
    if desc:
        freq = freq[:, ::-1].copy()
    ax, ay = sc.look()
    return sc, freq, ax, ay


@pytest.mark.parametrize("m,n,sgn,desc", [(1024, 1024, -1, False), (700, 1100, 1, False),
                                          (300, 257, -1, True), (600, 9000, -1, False),
                                          (300, 8300, 1, True)])
def test_linear_pfa_matches_oracle(gp, m, n, sgn, desc):
    sc, freq, ax, ay = _scene(m, n, seed=m + n, desc=desc)
    print(f"linear pfa {m}x{n} sgn={sgn} desc={desc}", flush=True)
    img, _ = gp.pfa(sc.sig, freq, ax, ay, sgn)
    ref, _ = rfl.pfa(sc.sig, freq, ax, ay, sgn)
    assert pn_rms(img, ref) <= TOL
    assert top_peaks(img, 8) == top_peaks(ref, 8)


@pytest.mark.parametrize("m,n", [(256, 300), (1024, 1024), (256, 8200)])
def test_sinc32_pfa_matches_fp64_oracle(gp, m, n):
    sc, freq, ax, ay = _scene(m, n, seed=3 * m + n)
    print(f"sinc32 pfa {m}x{n}", flush=True)
    img, _ = gp.pfa(sc.sig, freq, ax, ay, -1, interp="sinc32")
    ref, _ = rfl.pfa(sc.sig, freq, ax, ay, -1, interp="sinc32", store=np.complex128)
    assert pn_rms(img, ref) <= TOL
    assert top_peaks(img, 8) == top_peaks(ref, 8)


def test_compact_frequencies_equal_full_array(gp):
    sc, freq, ax, ay = _scene(200, 180, seed=9)
    m, n = freq.shape
    geom_full = gp.float_geometry(ax, ay, freq=freq)
    geom_c = gp.float_geometry(ax, ay, f0=sc.sc0, df=sc.scss, n=n)
    for k in ("KR", "KC", "tan_s", "order", "pr"):
        assert np.array_equal(geom_full[k], geom_c[k]), k
    p_full = gp.FloatPlan(m, n, -1, geom_full)
    p_c = gp.FloatPlan(m, n, -1, geom_c, f0=sc.sc0, df=sc.scss)
    a = p_full.focus(sc.sig, freq)
    b = p_c.focus(sc.sig)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("m,n,sgn,window", [(256, 512, -1, True), (1024, 1024, 1, True),
                                            (512, 256, -1, False), (128, 16384, -1, True),
                                            (5634, 4319, -1, True), (4319, 5634, 1, True),
                                            (33, 47, -1, True), (1, 7, 1, False), (7, 1, -1, True),
                                            (4099, 64, -1, False), (100, 8192, 1, True),
                                            (8191, 300, -1, True), (8192, 257, 1, False)])
def test_quicklook_matches_oracle(gp, m, n, sgn, window):
    """form_image_pfa.quicklook (window + one centred 2-D FFT of the raw history,
    no padding) at any shape, the board scene's 5634 x 4319 among them; odd
    lengths take both ifftshift / fftshift conventions.  Against the oracle's
    numpy (complex64 FFT) within 1e-5 of the peak with the same top peaks, and
    against a complex128 transform of the same windowed history."""
    import torch
    from mi355x_pfa import synth
    if min(m, n) >= 16:
        sc = synth.make_scene(synth.SceneSpec(m, n, targets=4, seed=m + n,
                                              scene_half=0.2 * min(m, n)), device="cuda")
        sig = sc.sig.cpu().numpy()
    else:
        rs = np.random.RandomState(m * n)
        sig = (rs.standard_normal((m, n)) + 1j * rs.standard_normal((m, n))).astype(np.complex64)
    print(f"quicklook {m}x{n} sgn={sgn} window={window}", flush=True)
    img = gp.quicklook(sig, sgn, window=window)
    ref = rfl.quicklook(sig, sgn, window=window)
    assert img.shape == (m, n)
    assert pn_rms(img, ref) <= TOL
    x = sig.astype(np.complex128)
    if window:
        x = x * np.outer(np.hamming(m), np.hamming(n)).astype(np.float32)
    fwd = np.fft.fft2 if sgn < 0 else np.fft.ifft2
    ref64 = np.fft.fftshift(fwd(np.fft.ifftshift(x)))
    err = pn_rms(img, ref64)
    print(f"pn-RMS vs complex128 {err:.2e}", flush=True)
    assert err <= TOL
    if min(m, n) >= 16:
        assert top_peaks(img, 4) == top_peaks(ref, 4)
    dev = gp.quicklook(torch.from_numpy(sig).cuda(), sgn, window=window).cpu().numpy()
    assert np.array_equal(dev, img)


def test_quicklook_refuses_other_sizes(gp):
    from mi355x_pfa._lib import PfaError
    with pytest.raises(PfaError, match="1..8192"):
        gp.quicklook(np.zeros((8193, 64), np.complex64), -1)
    with pytest.raises(PfaError, match="1..8192"):
        gp.quicklook(np.zeros((64, 32768), np.complex64), -1)


def test_padded_size_limits(gp):
    """Padded range (sample) and cross (pulse) sizes up to 16384 in the FP32
    form; beyond either the plan is refused with a dimension error, and the
    matrix-core forms stop at 8192."""
    from mi355x_pfa._lib import PfaError

    def plan(m, n, fft="fp32"):
        geom = {"pr": np.ones(m), "order": np.arange(m, dtype=np.int32), "tan_s": np.zeros(m),
                "KR": np.linspace(1.0, 2.0, n), "KC": np.linspace(-1.0, 1.0, m), "geo": {}}
        return gp.FloatPlan(m, n, -1, geom, f0=np.ones(m), df=np.ones(m), fft=fft)
    p = plan(100, 16384)
    assert (p.m2, p.n2) == (128, 16384)
    p.close()
    p = plan(16384, 100)
    assert (p.m2, p.n2) == (16384, 128)
    p.close()
    with pytest.raises(PfaError, match="range"):
        plan(100, 16385)
    with pytest.raises(PfaError, match="cross"):
        plan(16385, 100)
    with pytest.raises(PfaError, match="matrix-core"):
        plan(9000, 256, fft="bf16")


@pytest.mark.parametrize("m,n,sgn,interp", [(9000, 300, -1, "linear"), (12000, 200, 1, "linear"),
                                            (8200, 256, -1, "sinc32"), (16384, 128, -1, "sinc32")])
def test_cross_sizes_above_8192(gp, m, n, sgn, interp):
    """Scenes of more than 8192 pulses (padded cross size 16384,
    form_image_pfa.py:225-240): the np.interp focus against the oracle's numpy
    pfa, SINC32 against the FP64 oracle, within 1e-5 with the same peaks; the
    windowed k-space of resample_kspace elementwise as well."""
    print(f"cross size above 8192: {m}x{n} sgn={sgn} {interp}", flush=True)
    sc, freq, ax, ay = _scene(m, n, seed=m % 1000 + n, targets=8)
    img, geo = gp.pfa(sc.sig, freq, ax, ay, sgn, interp=interp)
    store = np.complex64 if interp == "linear" else np.complex128
    ref, geo_r = rfl.pfa(sc.sig, freq, ax, ay, sgn, interp=interp, store=store)
    assert img.shape == (16384, 512 if n > 256 else 256 if n > 128 else 128)
    assert pn_rms(img, ref) <= TOL
    assert top_peaks(img, 4) == top_peaks(ref, 4)
    assert [geo["dc"], geo["dr"]] == [geo_r["dc"], geo_r["dr"]]


@pytest.mark.parametrize("desc", [False, True])
def test_geometry_extrema_on_gpu_equal_numpy(gp, desc):
    """The O(m n) part of the geometry (per-pulse extrema of kr) reduced on the
    GPU gives the oracle's KR / KC grids bit for bit, from a host array and from
    a device tensor."""
    import torch
    sc, freq, ax, ay = _scene(700, 900, seed=21, desc=desc)
    ref = rfl.geometry(freq, ax, ay)
    for f in (freq, torch.from_numpy(np.ascontiguousarray(freq)).cuda()):
        g = gp.float_geometry(ax, ay, freq=f)
        assert np.array_equal(g["KR"], ref["KR"]) and np.array_equal(g["KC"], ref["KC"])


def test_host_buffer_path_equals_device_path(gp):
    """The host-buffer entry points stage large copies through a ring of pinned
    buffers on a copy stream: 18.5 MB of signal and of frequencies in (two chunks,
    the last ragged), a 2048 x 4096 image out (3 or 5 chunks).  The images equal
    the device-resident entry point's bit for bit, on repeated calls, and so does
    resample_kspace through the same staging."""
    import torch
    sc, freq, ax, ay = _scene(1100, 2100, seed=12)
    m, n = freq.shape
    geom = gp.float_geometry(ax, ay, freq=freq)
    plan = gp.FloatPlan(m, n, -1, geom, interp="sinc32")
    print(f"host-buffer staging {m}x{n} -> {plan.m2}x{plan.n2}", flush=True)
    sig_d = torch.from_numpy(np.ascontiguousarray(sc.sig)).cuda()
    freq_d = torch.from_numpy(np.ascontiguousarray(freq)).cuda()
    torch.cuda.synchronize()
    for detect in (True, False):
        out_d = torch.empty((plan.m2, plan.n2), dtype=torch.float32 if detect else torch.complex64,
                            device="cuda")
        plan.focus_dev(sig_d, out_d, freq_d, detect=detect)
        plan.ctx.synchronize()
        ref = out_d.cpu().numpy()
        for _ in range(2):
            assert np.array_equal(plan.focus(sc.sig, freq, detect=detect), ref)
    g2a = plan.resample_kspace(sc.sig, freq)
    g2b = plan.resample_kspace(sc.sig, freq)
    assert np.array_equal(g2a, g2b)
    plan.close()


@pytest.mark.parametrize("m,n,interp", [(100, 90, "linear"), (64, 90, "linear"), (100, 100, "sinc32"),
                                        (128, 90, "linear")])
def test_resample_kspace_elementwise_against_oracle(gp, m, n, interp):
    """Every k-space sample, not only the image RMS: the range grid's end points
    coincide with some pulse's own k-space samples, and np.interp's decision
    there (last sample vs zero) follows numpy's rounding of kr = (2 f / c) pr."""
    sc, freq, ax, ay = _scene(m, n, seed=m + n)
    print(f"resample_kspace elementwise {m}x{n} {interp}", flush=True)
    g2o, _ = rfl.resample_kspace(sc.sig, freq, ax, ay, interp=interp)
    g2g, _ = gp.resample_kspace(sc.sig, freq, ax, ay, interp=interp)
    assert np.abs(g2g.astype(np.complex128) - g2o).max() <= 1e-5 * np.abs(g2o).max()
