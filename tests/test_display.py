"""Product passes after detection (CPU side): numpy's percentile restated as
order statistics plus its float32 interpolation, checked against
np.percentile itself, and the oracle of the normal-score stretch
(sar_display.py:24-60) and of the uint8 GeoTIFF data (form_image_pfa.py:
286-291) pinned by closed forms."""
import numpy as np
import pytest

from oracle import ref_display as rd


def _ordered(x, k0, k1):
    p = np.partition(x, sorted({k0, k1}))
    return p[k0], p[k1]


@pytest.mark.parametrize("n", [1, 2, 3, 5, 10, 101, 1000, 12345, 99991, 3_000_017,
                               16_777_219, 20_000_001])
def test_percentile_restatement_equals_numpy(n):
    from mi355x_pfa.display import lerp, percentile_ranks
    rs = np.random.RandomState(n % 1000)
    x = rs.rayleigh(size=n).astype(np.float32)
    if n > 100:
        x[rs.randint(0, n, size=n // 10)] = np.float32(1.5)      # ties
    qs = [0.0, 12.5, 33.3, 50.0, 99.7, 100.0] if n < 10**6 else [50.0, 99.7]
    for q in qs:
        k0, k1, g = percentile_ranks(n, q)
        a, b = _ordered(x, k0, k1)
        ours = np.float32(lerp(a, b, g))
        ref = np.percentile(x, q)
        assert ref.dtype == np.float32
        assert ours.tobytes() == ref.tobytes(), (n, q, ours, ref)


def test_percentile_ranks_collapse_at_large_n():
    """Above 2**24 values numpy forms the virtual index in float32: the two
    neighbours coincide and gamma is 0."""
    from mi355x_pfa.display import percentile_ranks
    k0, k1, g = percentile_ranks(8192 * 8192, 99.7)
    assert k0 == k1 and float(g) == 0.0
    k0, k1, g = percentile_ranks(1234, 99.7)
    assert (k0, k1) == (1229, 1230) and 0.2 < float(g) < 0.4


def test_normal_score_two_levels_closed_form():
    """Two values with equal counts: midpoint CDF 1/4 and 3/4, z = -/+ 0.6745,
    grey levels 128 -/+ 127 * 0.6745 / 3."""
    img = np.array([[5, 9], [9, 5]], np.uint16)
    out = rd.normal_score(img)
    z = 0.6744897501960817
    assert out[0, 0] == int(128 - 127 * z / 3) and out[0, 1] == int(128 + 127 * z / 3)
    assert out.dtype == np.uint8 and out.shape == img.shape


def test_normal_score_zero_fill_excluded():
    rs = np.random.RandomState(3)
    core = rs.randint(1, 4000, size=(64, 80)).astype(np.uint16)
    padded = np.zeros((128, 128), np.uint16)
    padded[:64, :80] = core
    a = rd.normal_score(padded)
    assert not a[64:, :].any() and not a[:, 80:].any()
    np.testing.assert_array_equal(a[:64, :80], rd.normal_score(core))
    assert not rd.normal_score(np.zeros((7, 9), np.uint16)).any()


def test_normal_score_is_monotone_and_gaussian():
    rs = np.random.RandomState(4)
    img = rs.permutation(65536).astype(np.uint16).reshape(256, 256) + np.uint16(0)
    img[img == 0] = 1                                   # no zero-fill: every value ranked
    out = rd.normal_score(img)
    order = np.argsort(img.ravel(), kind="stable")
    assert np.all(np.diff(out.ravel()[order].astype(int)) >= 0)
    v = out.astype(np.float64)
    assert abs(v.mean() - 128) < 0.6 and abs(v.std() - 127 / 3) < 1.0


def test_geotiff_u8_closed_form():
    amp = (np.arange(1001, dtype=np.float32) / np.float32(1000))
    data, hi = rd.geotiff_u8(amp)
    assert hi == np.float32(0.997)
    assert data[-1] == 255 and data[0] == 0
    expect = np.minimum(np.floor(amp.astype(np.float64) / 0.997 * 255), 255)
    assert np.max(np.abs(data.astype(int) - expect)) <= 1
    assert np.all(data[amp >= hi] == 255)


def test_normal_score_lut_without_scipy_follows_the_reference_fallback(monkeypatch):
    """sar_display.py:46-55 inverts the normal CDF with scipy's erfinv and,
    where scipy is missing, by np.interp on a 200001-point erf table over
    [-8, 8].  The product keeps that fallback for parity; with scipy.special
    made unimportable it runs, and its LUT stays within one level of the
    erfinv LUT (the table's spacing, 8e-5 in z, is far below one level of
    127 / 3 per unit z)."""
    import sys
    from mi355x_pfa import display
    rs = np.random.RandomState(4)
    img = np.minimum(rs.rayleigh(900.0, 200000), 65535).astype(np.uint16)
    img[:500] = 0
    hist = np.bincount(img, minlength=65536)
    lut_scipy, lo, blank = display.normal_score_lut(hist)
    monkeypatch.setitem(sys.modules, "scipy.special", None)
    lut_fb, lo2, blank2 = display.normal_score_lut(hist)
    assert (lo, blank) == (lo2, blank2) and lut_fb.shape == lut_scipy.shape
    d = np.abs(lut_fb.astype(int) - lut_scipy.astype(int))
    assert d.max() <= 1 and (d > 0).mean() < 0.01
    assert lut_fb[0 - lo] == 0 if lo == 0 else True
