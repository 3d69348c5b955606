"""The oracle pinned against the reference's committed vectors and fixtures.

Fixtures: tests/golden/*.npz were written by tests/golden/make_golden.py from
the reference's own functions (and its committed CoreFFT vectors).  Closed
forms and invariants pin what the fixtures leave open."""
import numpy as np
import pytest

from oracle import ref_fixed as rf
from oracle import ref_float as rfl
from conftest import load_golden


def _c(iq):
    return rf.from_iq(iq)


# ------------------------------------------------------------------ fixed datapath
def test_fft1d_bfp_matches_committed_corefft_vectors():
    g = load_golden("corefft_vectors.npz")
    names = sorted({k.rsplit("_", 1)[0] for k in g.files})
    assert len(names) == 8
    for name in names:
        i = g[f"{name}_in"].astype(np.int64)
        o = g[f"{name}_out"].astype(np.int64)
        exps = list(g[f"{name}_exps"])
        yr, yi, ex = rf.fft_bfp_blockwise(i[:, 0], i[:, 1], exps[0])
        assert np.array_equal(yr, o[:, 0]) and np.array_equal(yi, o[:, 1]), name
        assert ex == exps, name


@pytest.mark.parametrize("n", [64, 1024, 8192])
def test_fft_bfp_perrow_matches_reference(n):
    g = load_golden("fft_hw_perrow.npz")
    xr, xi = _c(g[f"x{n}"])
    yr, yi, e = rf.fft_bfp_perrow(xr, xi)
    er, ei = _c(g[f"y{n}"])
    assert np.array_equal(yr, er) and np.array_equal(yi, ei)
    assert np.array_equal(e, g[f"e{n}"])
    pr, pi, pe = rf.fft_pass_bfp(xr, xi)
    qr, qi = _c(g[f"pass_y{n}"])
    assert np.array_equal(pr, qr) and np.array_equal(pi, qi)
    assert np.array_equal(pe, g[f"pass_e{n}"])


def test_fft_bfp_closed_forms():
    n = 1024
    # zero row: no growth, exponent 0
    z = np.zeros((1, n), np.int64)
    yr, yi, e = rf.fft_bfp_perrow(z, z)
    assert e[0] == 0 and not yr.any() and not yi.any()
    # impulse at 0 (below the shift threshold): flat spectrum, no shifts
    x = np.zeros((1, n), np.int64); x[0, 0] = 1000
    yr, yi, e = rf.fft_bfp_perrow(x, np.zeros_like(x))
    assert e[0] == 0 and np.all(yr == 1000) and not yi.any()
    # exponents are bounded by the 13-stage, 2-bit-per-stage rule
    g = load_golden("fft_hw_perrow.npz")
    assert g["e8192"].max() <= 26 and g["e8192"].min() >= 0


def test_twiddles_match_c_library_formula():
    # the quarter-wave points are exact; floor keeps the cos(pi/2)=6e-17 at 0
    wr, wi = rf.twiddle_codes(8192, 16)
    assert wr[0] == 32767 and wi[0] == 0
    assert wr[2048] == 0 and wi[2048] == -32768


def test_interp_coeffs_matches_reference():
    g = load_golden("interp_coeffs.npz")
    for c in g["cases"]:
        idx, wq = rf.interp_coeffs(g[f"q{c}"], g[f"xp{c}"])
        assert np.array_equal(idx, g[f"idx{c}"]) and np.array_equal(wq, g[f"wq{c}"])


def test_sinc_table_matches_reference_and_invariants():
    tab = rf.sinc_table_q15()
    ref = np.load(load_golden.__globals__["GOLDEN"] / "sinc_table_q15.npy")
    assert np.array_equal(tab, ref.astype(np.int64))
    assert np.all(tab.sum(axis=1) == 1 << 15)
    assert tab.min() >= -32768 and tab.max() <= 32767


def _fixed_scene(name):
    g = load_golden(f"fixed_{name}.npz")
    grid = int(g["grid"])
    sr, si = _c(g["sig_i16"])
    m, n = sr.shape
    tabs = rf.fixed_tables(g["tab_ax"], g["tab_ay"], g["tab_freq"], g["tab_KR"], g["tab_KC"],
                           g["tab_tan_phi"], m, n, grid)
    return g, grid, sr, si, tabs


@pytest.mark.parametrize("name,tags", [("tiny", ["00", "11", "10"]), ("small", ["00", "11"])])
def test_fixed_stages_match_emulator(name, tags):
    g, grid, sr, si, tabs = _fixed_scene(name)
    m, n = sr.shape
    qr, qi, e = rf.quantize_iq16(g["sig"])
    assert e == int(g["sig_exp"])
    assert np.array_equal(qr, sr) and np.array_equal(qi, si)
    for tag in tags:
        rs, az = tag[0] == "1", tag[1] == "1"
        gr, gi = rf.resample_fixed(sr, si, tabs, grid, rs, az)
        er, ei = _c(g[f"g2_{tag}"])
        assert np.array_equal(gr, er) and np.array_equal(gi, ei), tag
        wr, wi = rf.window_fixed(gr, gi, m, n)
        er, ei = _c(g[f"win_{tag}"])
        assert np.array_equal(wr, er) and np.array_equal(wi, ei), tag
        img, ea, eb = rf.form_image_fixed(sr, si, tabs, grid, rs, az)
        assert np.array_equal(ea, g[f"ea_{tag}"]) and np.array_equal(eb, g[f"eb_{tag}"])
        assert np.array_equal(img, g[f"img_{tag}"]), tag


def test_detect_isqrt_closed_form():
    rs = np.random.RandomState(0)
    re = rs.randint(-32768, 32768, 200000); im = rs.randint(-32768, 32768, 200000)
    re[:4] = [-32768, 32767, 0, -32768]; im[:4] = [-32768, 32767, 0, 0]
    d = rf.detect_fixed(re, im).astype(np.int64)
    s = re * re + im * im
    assert np.all(d * d <= s) and np.all((d + 1) * (d + 1) > s)


# ------------------------------------------------------------------ float path
@pytest.mark.parametrize("name", ["small", "rect"])
def test_float_pfa_matches_reference(name):
    g = load_golden(f"float_{name}.npz")
    g2, geo = rfl.resample_kspace(g["sig"], g["freq"], g["ax"], g["ay"])
    assert np.array_equal(g2, g["g2"])
    img, geo2 = rfl.focus(g2, geo, int(g["sgn"]))
    assert np.array_equal(img, g["img"])
    assert np.array_equal(rfl.detect(img), g["det"])
    assert np.allclose([geo2["dc"], geo2["dr"], *geo2["dhat"], *geo2["chat"]], g["geo"], rtol=0, atol=0)


def test_float_point_targets_at_predicted_pixels():
    from mi355x_pfa import synth
    sc = synth.make_scene(synth.SceneSpec(256, 256, targets=4, seed=1, scene_half=50))
    ax, ay = sc.look()
    img, geo = rfl.pfa(sc.sig, sc.freq(), ax, ay, sc.sgn)
    a = np.abs(img)
    dh, ch = np.array(geo["dhat"]), np.array(geo["chat"])
    for p in sc.targets:
        col = 128 - (p[:2] @ dh) / geo["dr"]
        row = 128 - (p[:2] @ ch) / geo["dc"]
        r0, c0 = int(round(row)), int(round(col))
        win = a[r0 - 2:r0 + 3, c0 - 2:c0 + 3]
        rr, cc = np.unravel_index(np.argmax(win), win.shape)
        assert abs(r0 - 2 + rr - row) <= 1.0 and abs(c0 - 2 + cc - col) <= 1.0


def test_float_plain_fft_special_case_and_parseval():
    # uniform polar grid that is already Cartesian (pr = 1, tan_phi = 0 for all but
    # tiny spread) and no window reduces to a centred 2-D FFT of the input
    m, n = 32, 64
    rs = np.random.RandomState(3)
    sig = (rs.standard_normal((m, n)) + 1j * rs.standard_normal((m, n))).astype(np.complex64)
    g2 = sig
    img, _ = rfl.focus(g2, {"dc": 1.0, "dr": 1.0, "dhat": (1, 0), "chat": (0, 1)}, -1)
    ref = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(g2.astype(np.complex128))))
    assert np.max(np.abs(img - ref)) / np.max(np.abs(ref)) < 1e-5
    e_in = np.sum(np.abs(g2.astype(np.complex128)) ** 2) * m * n
    assert abs(np.sum(np.abs(img.astype(np.complex128)) ** 2) - e_in) / e_in < 1e-5


def test_quicklook_oracle_pinned():
    """quicklook = hamming2d x history, then transform2d without padding.  On
    power-of-two shapes it equals focus() (pinned to the reference fixtures) of
    the same array, windowed or not; a tone at DFT bin (a, b) becomes a single
    pixel at the centred bin, with sgn selecting fft2 / ifft2."""
    rs = np.random.RandomState(3)
    x = (rs.randn(64, 128) + 1j * rs.randn(64, 128)).astype(np.complex64)
    for sgn in (-1, 1):
        assert np.array_equal(rfl.quicklook(x, sgn, window=False), rfl.focus(x, {"dc": 1, "dr": 1}, sgn)[0])
        w = np.outer(np.hamming(64), np.hamming(128)).astype(np.float32)
        assert np.array_equal(rfl.quicklook(x, sgn), rfl.focus(x * w, {"dc": 1, "dr": 1}, sgn)[0])
    m, n, a, b = 64, 128, 5, 17
    k = np.arange(m)[:, None] - m // 2
    c = np.arange(n)[None, :] - n // 2
    tone = np.exp(2j * np.pi * (a * k / m + b * c / n))
    img = np.abs(rfl.quicklook(tone, 1, window=False))       # ifft2 (1/(m n)): peak at -(a, b)
    assert np.unravel_index(np.argmax(img), img.shape) == (m // 2 - a, n // 2 - b)
    assert np.isclose(img.max(), 1.0) and np.sort(img.ravel())[-2] < 1e-9
    img = np.abs(rfl.quicklook(tone, -1, window=False))      # fft2: peak m n at +(a, b)
    assert np.unravel_index(np.argmax(img), img.shape) == (m // 2 + a, n // 2 + b)
    assert np.isclose(img.max(), m * n)


def test_sinc32_float_consistent_with_fixed_sinc():
    """Float SINC32 on integer data equals the reference-pinned fixed gather
    rounded half-up, wherever the two brackets agree (no exact hits)."""
    rs = np.random.RandomState(5)
    S = 120
    xp = np.cumsum(rs.uniform(0.5, 1.5, S))
    xr = rs.randint(-20000, 20000, S).astype(np.int64)
    xi = rs.randint(-20000, 20000, S).astype(np.int64)
    q = np.sort(rs.uniform(xp[0] - 2, xp[-1] + 2, 400))
    tab = rf.sinc_table_q15()
    out = rfl._sinc_line(q, xp, xr + 1j * xi, tab)
    idx, wq = rf.interp_coeffs(q, xp)
    fr, fi = rf.sinc_fixed(xr, xi, idx, wq, S, tab)
    interior = (idx >= 15) & (idx <= S - 17)
    assert interior.sum() > 200
    assert np.array_equal(np.floor(out.real[interior] + 0.5).astype(np.int64), fr[interior])
    assert np.array_equal(np.floor(out.imag[interior] + 0.5).astype(np.int64), fi[interior])


def test_sinc32_float_reconstructs_bandlimited_tone():
    S = 400
    xp = np.arange(S, dtype=np.float64) * 0.75 + 3.0
    f = 0.23                                     # cycles per sample, well inside the band
    fp = np.exp(2j * np.pi * f * np.arange(S))
    q = np.linspace(xp[20], xp[-20], 777)
    out = rfl._sinc_line(q, xp, fp, rf.sinc_table_q15())
    ideal = np.exp(2j * np.pi * f * (q - 3.0) / 0.75)
    err = np.abs(out - ideal)
    assert err.max() < 5e-2          # truncated-sinc ripple + 256-phase delay quantisation
    lin = np.interp(q, xp, fp.real) + 1j * np.interp(q, xp, fp.imag)
    assert np.sqrt(np.mean(err ** 2)) < 0.2 * np.sqrt(np.mean(np.abs(lin - ideal) ** 2))


def test_reference_cpp_fft_model_agrees(tmp_path):
    """oracle/_ref/fft_tb (the reference's fft_tb.cpp, built by oracle/build_ref.py)
    reproduces the committed vectors from the committed inputs and our twiddles."""
    from conftest import REPO
    exe = REPO / "oracle" / "_ref" / "fft_tb"
    if not exe.exists():
        pytest.skip("oracle/_ref not built (no reference checkout here)")
    import subprocess
    g = load_golden("corefft_vectors.npz")
    wr, wi = rf.twiddle_codes(8192, 18)
    (tmp_path / "tw.csv").write_text("\n".join(f"{a},{b}" for a, b in zip(wr, wi)) + "\n")
    for name in ("impulse_k", "random", "twotone_hidr"):
        i = g[f"{name}_in"].astype(np.int64)
        words = ((i[:, 0] & 0xFFFF) << 16) | (i[:, 1] & 0xFFFF)
        (tmp_path / "in.hex").write_text("\n".join(f"{w:08x}" for w in words) + "\n")
        subprocess.run([str(exe), "8192", str(tmp_path / "in.hex"), str(tmp_path / "tw.csv"),
                        str(tmp_path / "out.hex")], check=True, capture_output=True)
        w = np.array([int(s, 16) for s in (tmp_path / "out.hex").read_text().split()], np.int64)
        re = (w >> 16) & 0xFFFF; im = w & 0xFFFF
        re = np.where(re >= 0x8000, re - 0x10000, re); im = np.where(im >= 0x8000, im - 0x10000, im)
        o = g[f"{name}_out"].astype(np.int64)
        assert np.array_equal(re, o[:, 0]) and np.array_equal(im, o[:, 1]), name
