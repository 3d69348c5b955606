"""CPU checks of the float focus's scene geometry built in the C library from the
reference's own arguments (pfa_float_geometry, form_image_pfa.py:178-221),
against the oracle's numpy restatement and the reference's fixtures, plus the
table preparation pinned to the reference's own prepare_tables output
(tests/golden/fixed_*.npz tab_*, written by sar_pipeline.prepare_tables)."""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import load_golden
from oracle import ref_float as rfl


def _c_geometry(ax, ay, freq=None, f0=None, df=None, n=None, window=True):
    from mi355x_pfa import _lib
    ax = np.ascontiguousarray(ax, np.float64)
    ay = np.ascontiguousarray(ay, np.float64)
    m = ax.size
    if freq is not None:
        freq = np.ascontiguousarray(freq, np.float64)
        n = freq.shape[1]
    else:
        f0 = np.ascontiguousarray(f0, np.float64)
        df = np.ascontiguousarray(df, np.float64)
    out = {"pr": np.empty(m), "order": np.empty(m, np.int32), "tan_s": np.empty(m),
           "KR": np.empty(n), "KC": np.empty(m), "ham_c": np.empty(m), "ham_r": np.empty(n),
           "geo": np.empty(6)}
    p = lambda a: None if a is None else a.ctypes.data  # noqa: E731
    st = _lib.lib().pfa_float_geometry(
        m, n, p(freq), p(f0), p(df), p(ax), p(ay), p(out["pr"]), p(out["order"]),
        p(out["tan_s"]), p(out["KR"]), p(out["KC"]),
        p(out["ham_c"]) if window else None, p(out["ham_r"]) if window else None, p(out["geo"]))
    return st, out


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)


def _check(out, ref, m, n):
    assert _same(out["pr"], ref["pr"])
    assert _same(out["order"], ref["order"])
    assert _same(out["tan_s"], ref["tan_s"])
    assert _same(out["KR"], ref["KR"])
    assert _same(out["KC"], ref["KC"])
    g = ref["geo"]
    assert _same(out["geo"], [g["dc"], g["dr"], *g["dhat"], *g["chat"]])
    assert _same(out["ham_c"], np.hamming(m)) and _same(out["ham_r"], np.hamming(n))


@pytest.mark.parametrize("m,n,seed", [(40, 48, 1), (96, 80, 2), (333, 257, 3), (1000, 64, 4),
                                      (9001, 40, 5), (20000, 33, 6)])
def test_c_geometry_equals_numpy_geometry_full_freq(m, n, seed):
    from mi355x_pfa import synth
    sc = synth.make_scene(synth.SceneSpec(m, n, targets=1, seed=seed, scene_half=5.0),
                          signal=False)
    ax, ay = sc.look()
    freq = sc.freq()
    st, out = _c_geometry(ax, ay, freq=freq)
    assert st == 0
    _check(out, rfl.geometry(freq, ax, ay), m, n)


@pytest.mark.parametrize("m,n,seed", [(120, 100, 7), (4096, 4096, 8), (8167, 8192, 9)])
def test_c_geometry_compact_frequencies_equal_the_full_array(m, n, seed):
    from mi355x_pfa import synth
    sc = synth.make_scene(synth.SceneSpec(m, n, targets=1, seed=seed, scene_half=5.0),
                          signal=False)
    ax, ay = sc.look()
    st, out = _c_geometry(ax, ay, f0=sc.sc0, df=sc.scss, n=n)
    assert st == 0
    _check(out, rfl.geometry(sc.freq(), ax, ay), m, n)


def test_c_geometry_on_random_look_angles_and_descending_rows():
    """Look directions off the synthetic straight line, unsorted pulses, rows
    whose frequencies descend, and rows that do not all overlap."""
    rs = np.random.RandomState(11)
    for trial in range(20):
        m, n = int(rs.randint(2, 300)), int(rs.randint(2, 200))
        th = rs.uniform(-0.3, 0.3, m) + rs.uniform(-3, 3)
        r = rs.uniform(0.5, 1.0, m)
        ax, ay = r * np.cos(th), r * np.sin(th)
        f0 = rs.uniform(9.0e9, 10e9, m)
        step = rs.uniform(1e5, 1e6, m) * np.where(rs.rand(m) < 0.3, -1.0, 1.0)
        freq = f0[:, None] + np.arange(n)[None, :] * step[:, None]
        st, out = _c_geometry(ax, ay, freq=freq)
        assert st == 0, trial
        _check(out, rfl.geometry(freq, ax, ay), m, n)


def test_c_geometry_equals_the_reference_fixture():
    """form_image_pfa.pfa's geo of the float fixture (reference import): the C
    plan constructor's padded geo equals it; the unpadded geo equals the
    oracle's resample_kspace geo."""
    g = load_golden("float_small.npz")
    st, out = _c_geometry(g["ax"], g["ay"], freq=g["freq"])
    assert st == 0
    m, n = g["freq"].shape
    m2, n2 = rfl.to_pow2(m), rfl.to_pow2(n)
    geo = out["geo"].copy()
    geo[1] *= n / n2
    geo[0] *= m / m2
    assert _same(geo, g["geo"])


def test_c_hypot_is_correctly_rounded_like_math_hypot():
    """The rotated frame's norm is math.hypot (form_image_pfa.py:181): checked
    through the mean look direction of two-pulse scenes (mean of one value is
    the value itself when both pulses are equal)."""
    rs = np.random.RandomState(3)
    for _ in range(3000):
        x, y = rs.standard_normal(2) * 10.0 ** rs.uniform(-3, 3, 2)
        ax = np.array([x, x])
        ay = np.array([y, y])
        freq = np.array([[1e10, 1.1e10], [1e10, 1.1e10]])
        st, out = _c_geometry(ax, ay, freq=freq, window=False)
        assert st == 0
        dn = math.hypot(x, y)
        assert out["geo"][2] == x / dn and out["geo"][3] == y / dn


def test_c_geometry_mean_is_numpys_blocked_pairwise_sum():
    """ndarray.mean of more than 8192 values (the ufunc buffer) and of every
    length up to 300: dhat equals numpy's."""
    rs = np.random.RandomState(5)
    for m in list(range(2, 300, 7)) + [8192, 8193, 9000, 16400, 30001]:
        ax = rs.standard_normal(m) * 10.0 ** rs.uniform(-2, 2) + 3.0
        ay = rs.standard_normal(m) + 1.0
        freq = np.tile([1e10, 1.1e10], (m, 1))
        st, out = _c_geometry(ax, ay, freq=freq, window=False)
        assert st == 0
        dx, dy = ax.mean(), ay.mean()
        dn = math.hypot(dx, dy)
        assert out["geo"][2] == dx / dn and out["geo"][3] == dy / dn, m


def test_c_geometry_rejects_bad_input():
    ax = np.array([1.0, 1.0, np.nan])
    ay = np.zeros(3)
    freq = np.tile([1e10, 1.1e10], (3, 1))
    assert _c_geometry(ax, ay, freq=freq)[0] == -3
    assert _c_geometry(np.zeros(3), np.zeros(3), freq=freq)[0] == -3
    assert _c_geometry(np.ones(1), np.ones(1), freq=np.ones((1, 2)))[0] == -2


@pytest.mark.parametrize("name,spec", [
    ("tiny", dict(pulses=40, samples=48, targets=3, scene_half=8.0, seed=11)),
    ("small", dict(pulses=200, samples=180, targets=8, scene_half=30.0, clutter=16,
                   clutter_half=40.0, seed=12))])
def test_prepare_tables_equal_the_reference_prepare_tables(name, spec):
    """tab_* of the fixed fixtures are sar_pipeline.prepare_tables' output
    (reference import) on these synthetic scenes: the oracle's and the
    product's table preparation reproduce them bit for bit."""
    from mi355x_pfa import synth
    from mi355x_pfa.pipeline import prepare_tables
    g = load_golden(f"fixed_{name}.npz")
    sc = synth.make_scene(synth.SceneSpec(**spec))
    rd = synth.SceneReader(sc)
    ref = rfl.prepare_tables(rd, rd.cphd_meta, 1, 1)
    mine = prepare_tables(rd, rd.cphd_meta, 1, 1)
    for key in ("ax", "ay", "freq", "KR", "KC", "tan_phi"):
        assert _same(np.asarray(ref[key]), g[f"tab_{key}"]), key
        assert _same(np.asarray(mine[key]), g[f"tab_{key}"]), key


def test_tied_look_angles_order_like_the_staging_tables():
    """Duplicated look angles: the in-memory fixed tables and the staged role
    arrays use the reference's np.argsort(tan_phi) and agree."""
    from mi355x_pfa import synth
    from mi355x_pfa.staging import staging_tables
    from mi355x_pfa.tables import fixed_tables
    sc = synth.make_scene(synth.SceneSpec(64, 40, targets=1, seed=3, scene_half=5.0),
                          signal=False)
    ax, ay = sc.look()
    ax = np.repeat(ax[::2], 2)
    ay = np.repeat(ay[::2], 2)
    freq = sc.freq()
    geo = rfl.geometry(freq, ax, ay)
    tables = {"ax": ax, "ay": ay, "freq": freq, "KR": geo["KR"], "KC": geo["KC"],
              "tan_phi": geo["tan_phi"], "dims": (64, 40)}
    ft = fixed_tables(tables, 64, 40, 64)
    st = staging_tables(tables, 64)
    assert np.array_equal(ft["inv_order"], st["invorder"])
    assert np.array_equal(ft["tan_s"], st["tans"])
