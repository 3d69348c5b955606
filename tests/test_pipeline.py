"""Table preparation (sar_pipeline.prepare_tables, sar_pipeline.py:36-115)
without the O(m n) host arrays, against the oracle's numpy restatement that
builds them.

The oracle is pinned here by what the geometry fixes independently of its
code: the closed-form look direction of the synthetic straight-line flight,
pulse / sample decimation equal to slicing the undecimated tables, a crop
equal to the centred column slice, the form_image_pfa grids it shares, and
the Hamming taper's closed-form end and centre values."""
import dataclasses
import math

import numpy as np
import pytest

from oracle import ref_float as rfl


@pytest.fixture(scope="module")
def reader():
    from mi355x_pfa import synth
    spec = dataclasses.replace(synth.C1, pulses=301, samples=256, targets=4, seed=7)
    return synth.SceneReader(synth.make_scene(spec))


def _same(a, b):
    if isinstance(b, dict):
        return all(_same(a[k], b[k]) for k in b) and a.keys() == b.keys()
    if isinstance(b, (np.ndarray, tuple)):
        a2, b2 = np.asarray(a), np.asarray(b)
        return a2.shape == b2.shape and a2.dtype == b2.dtype and np.array_equal(a2, b2)
    return a == b


@pytest.mark.parametrize("deci_pulse,deci_sample,crop",
                         [(1, 1, 0), (2, 1, 0), (1, 2, 0), (3, 2, 0), (1, 1, 200), (2, 2, 100),
                          (1, 1, 256), (1, 1, 999)])
def test_prepare_tables_equal_oracle(reader, deci_pulse, deci_sample, crop):
    from mi355x_pfa.pipeline import prepare_tables
    a = prepare_tables(reader, reader.cphd_meta, deci_pulse, deci_sample, crop)
    b = rfl.prepare_tables(reader, reader.cphd_meta, deci_pulse, deci_sample, crop)
    assert a.keys() == b.keys()
    bad = [k for k in b if not _same(a[k], b[k])]
    assert not bad, bad


def test_compact_tables_are_o_m_plus_n(reader):
    from mi355x_pfa.pipeline import LinearFreq, OuterWindow, prepare_tables
    t = prepare_tables(reader, reader.cphd_meta)
    assert isinstance(t["freq"], LinearFreq) and isinstance(t["window"], OuterWindow)
    assert t["freq"].shape == t["window"].shape == t["dims"] == (301, 256)
    held = [t["freq"].sc0, t["freq"].step, t["window"].hm, t["window"].hn]
    assert sum(v.size for v in held) == 2 * 301 + 301 + 256
    assert np.asarray(t["freq"], np.float32).dtype == np.float32


def test_oracle_look_direction_closed_form(reader):
    """Flight along y at x = -standoff, height h, SRP at the origin, image plane
    = ground: the unit line of sight is (s, -y, -h) / sqrt(s^2 + y^2 + h^2)."""
    sc = reader.scene
    t = rfl.prepare_tables(reader, reader.cphd_meta)
    s, h = sc.spec.standoff, sc.spec.altitude
    y = sc.tx[:, 1]
    r = np.sqrt(s * s + y * y + h * h)
    np.testing.assert_allclose(t["ax"], s / r, rtol=1e-15, atol=0)
    np.testing.assert_allclose(t["ay"], -y / r, rtol=1e-14, atol=1e-18)
    np.testing.assert_array_equal(t["center_ecef"], sc.srp[0])
    assert t["sgn"] == -1 and t["samp0"] == 0


def test_oracle_decimation_is_slicing(reader):
    full = rfl.prepare_tables(reader, reader.cphd_meta)
    dec = rfl.prepare_tables(reader, reader.cphd_meta, 2, 2)
    assert dec["dims"] == (151, 128) and dec["deci"] == (2, 2)
    np.testing.assert_array_equal(dec["freq"], full["freq"][::2, ::2])
    np.testing.assert_array_equal(dec["ax"], full["ax"][::2])
    np.testing.assert_array_equal(dec["ay"], full["ay"][::2])


def test_oracle_crop_is_centred_slice(reader):
    full = rfl.prepare_tables(reader, reader.cphd_meta)
    crop = rfl.prepare_tables(reader, reader.cphd_meta, 1, 1, 200)
    d0 = (256 - 200) // 2
    assert crop["samp0"] == d0 and crop["dims"] == (301, 200)
    np.testing.assert_array_equal(crop["freq"], full["freq"][:, d0:d0 + 200])
    # KR spans the cropped band, which lies inside the full band
    assert full["KR"][0] < crop["KR"][0] < crop["KR"][-1] < full["KR"][-1]
    dec = rfl.prepare_tables(reader, reader.cphd_meta, 1, 2, 100)
    assert dec["samp0"] == ((128 - 100) // 2) * 2


def test_oracle_grids_agree_with_form_image_pfa(reader):
    """prepare_tables mirrors form_image_pfa's geometry (np.hypot for
    math.hypot): the grids agree to the last few ulp."""
    t = rfl.prepare_tables(reader, reader.cphd_meta)
    g = rfl.geometry(t["freq"], t["ax"], t["ay"])
    for k in ("KR", "KC", "tan_phi"):
        np.testing.assert_allclose(t[k], g[k], rtol=4e-16, atol=0)
    assert math.isclose(t["geo"]["dr"], g["geo"]["dr"], rel_tol=1e-15)


def test_oracle_window_closed_form(reader):
    w = rfl.prepare_tables(reader, reader.cphd_meta)["window"]
    assert w.dtype == np.float32 and w.shape == (301, 256)
    assert abs(float(w[0, 0]) - 0.08 * 0.08) < 1e-9    # Hamming ends: 0.54 - 0.46
    assert abs(float(w[150, 0]) - 0.08) < 1e-7          # odd m: exact centre 1.0
    np.testing.assert_array_equal(w, w[::-1, ::-1])
