"""Fixed-point mode on the GPU, bit-exact against silicon_emulator.py.

Every call goes through libmi355x_pfa.so (C ABI) via mi355x_pfa.emulator.
Fixtures come from the reference itself (tests/golden/make_golden.py); the
larger cases compare with the oracle (oracle/ref_fixed.py, pinned to those
fixtures) on the same seeded inputs."""
import numpy as np
import pytest

from conftest import load_golden
from oracle import ref_fixed as rf

pytestmark = pytest.mark.gpu


def _c(iq):
    return iq[..., 0].astype(np.float64) + 1j * iq[..., 1].astype(np.float64)


@pytest.fixture(scope="module")
def emu():
    from mi355x_pfa import emulator
    return emulator


@pytest.mark.parametrize("n", [64, 1024, 8192])
def test_fft_pass_bfp_matches_reference(emu, n):
    g = load_golden("fft_hw_perrow.npz")
    print(f"fft_pass_bfp n={n}", flush=True)
    y, e = emu.fft_pass_bfp(_c(g[f"x{n}"]))
    assert np.array_equal(y, _c(g[f"pass_y{n}"]))
    assert np.array_equal(e, g[f"pass_e{n}"])
    y2, e2 = emu.fft_bfp_perrow(_c(g[f"x{n}"]))
    assert np.array_equal(y2, _c(g[f"y{n}"]))
    assert np.array_equal(e2, g[f"e{n}"])


@pytest.mark.parametrize("n", [128, 256, 512, 2048, 4096])
def test_fft_pass_bfp_matches_oracle_all_lengths(emu, n):
    rs = np.random.RandomState(n)
    rows = 24
    x = (rs.randint(-32768, 32768, (rows, n)) >> rs.randint(0, 15, (rows, 1))) \
        + 1j * (rs.randint(-32768, 32768, (rows, n)) >> rs.randint(0, 15, (rows, 1)))
    print(f"fft oracle n={n}", flush=True)
    y, e = emu.fft_pass_bfp(x)
    ore, oim, oe = rf.fft_pass_bfp(x.real.astype(np.int64), x.imag.astype(np.int64))
    assert np.array_equal(y.real, ore) and np.array_equal(y.imag, oim)
    assert np.array_equal(e, oe)


def test_detect_closed_form_wide_sweep(emu):
    re = np.repeat(np.arange(-32768, 32768, 7, dtype=np.int64), 64)
    im = np.tile(np.arange(-32768, 32768, 1024, dtype=np.int64), re.size // 64)
    re = np.concatenate([re, [-32768, -32768, 32767, 0]])
    im = np.concatenate([im, [-32768, 32767, 32767, 0]])
    print(f"detect sweep {re.size}", flush=True)
    d = emu.detect_fixed(re + 1j * im).astype(np.int64)
    s = re * re + im * im
    assert np.all(d * d <= s) and np.all((d + 1) * (d + 1) > s)


def _fixed_case(name):
    g = load_golden(f"fixed_{name}.npz")
    tables = {k: g["tab_" + k] for k in ("ax", "ay", "freq", "KR", "KC", "tan_phi")}
    return g, tables, int(g["grid"])


@pytest.mark.parametrize("name,tags", [("tiny", ["00", "11", "10"]), ("small", ["00", "11"])])
def test_resample_window_match_reference(emu, name, tags):
    g, tables, grid = _fixed_case(name)
    sig64 = _c(g["sig_i16"]).astype(np.complex64)        # form_image's complex64 line
    m, n = sig64.shape
    for tag in tags:
        print(f"resample {name} {tag}", flush=True)
        g2, dims = emu.resample_fixed(sig64, tables, grid, tag[0] == "1", tag[1] == "1")
        assert dims == (m, n)
        assert np.array_equal(g2, _c(g[f"g2_{tag}"])), tag
        w = emu.window_fixed(g2, m, n)
        assert np.array_equal(w, _c(g[f"win_{tag}"])), tag


def test_resample_exact_product_for_complex128_line(emu):
    g, tables, grid = _fixed_case("tiny")
    sig = g["sig_i16"]
    m, n = sig.shape[:2]
    g2, _ = emu.resample_fixed(_c(sig), tables, grid)            # complex128 -> exact lerp
    tabs = rf.fixed_tables(*(tables[k] for k in ("ax", "ay", "freq", "KR", "KC", "tan_phi")),
                           m, n, grid)
    ore, oim = rf.resample_fixed(sig[..., 0].astype(np.int64), sig[..., 1].astype(np.int64),
                                 tabs, grid, pass1_f32=False)
    assert np.array_equal(g2.real, ore) and np.array_equal(g2.imag, oim)


@pytest.mark.parametrize("name,tags", [("tiny", ["00", "11", "10"]), ("small", ["00", "11"])])
def test_form_image_matches_reference(emu, name, tags):
    g, tables, grid = _fixed_case(name)
    for tag in tags:
        print(f"form_image {name} {tag}", flush=True)
        img = emu.form_image(g["sig"], tables, grid, tag[0] == "1", tag[1] == "1")
        assert np.array_equal(img, g[f"img_{tag}"]), tag


@pytest.mark.parametrize("m,n,grid,rs,az", [(1000, 900, 1024, True, True),
                                            (1024, 1024, 1024, False, False),
                                            (1500, 2048, 2048, True, False)])
def test_form_image_matches_oracle_midsize(emu, m, n, grid, rs, az):
    from mi355x_pfa import synth
    from mi355x_pfa.emulator import FixedPlan
    from mi355x_pfa.tables import fixed_tables
    from oracle import ref_float as rfl
    sc = synth.make_scene(synth.SceneSpec(m, n, targets=12, clutter=64, seed=m + n,
                                          scene_half=m / 8, clutter_half=m / 6))
    ax, ay = sc.look()
    freq = sc.freq()
    geo = rfl.geometry(freq, ax, ay)
    tables = {"ax": ax, "ay": ay, "freq": freq, "KR": geo["KR"], "KC": geo["KC"],
              "tan_phi": geo["tan_phi"]}
    qr, qi, e = rf.quantize_iq16(sc.sig)
    print(f"midsize {m}x{n}->{grid} sinc={rs},{az}", flush=True)
    iq, e_gpu = emu.quantize_iq16(sc.sig)
    assert e_gpu == e
    assert np.array_equal(iq[..., 0], qr) and np.array_equal(iq[..., 1], qi)
    plan = FixedPlan(fixed_tables(tables, m, n, grid), m, n, grid, rs, az)
    img, ea, eb = plan.form_image(iq)
    tabs = rf.fixed_tables(ax, ay, freq, geo["KR"], geo["KC"], geo["tan_phi"], m, n, grid)
    oimg, oea, oeb = rf.form_image_fixed(qr, qi, tabs, grid, rs, az)
    assert np.array_equal(ea, oea) and np.array_equal(eb, oeb)
    assert np.array_equal(img, oimg)
