"""Product passes after detection on the GPU, bit-exact against the oracle
(sar_display.normal_score, the uint8 GeoTIFF data of form_image_pfa.py:
286-291) and against numpy (np.bincount, np.partition, np.percentile):
host and device entry points, ragged and misaligned inputs, ties, infinities
and NaN, and the 8192 x 8192 frame size."""
import numpy as np
import pytest

from oracle import ref_display as rd

pytestmark = pytest.mark.gpu


def _images():
    rs = np.random.RandomState(11)
    yield "uniform", rs.randint(0, 65536, size=(300, 517)).astype(np.uint16)
    yield "rayleigh", np.minimum(rs.rayleigh(300.0, size=(1024, 1031)), 46341).astype(np.uint16)
    z = np.zeros((512, 512), np.uint16)
    z[:300, :280] = rs.randint(1, 2000, size=(300, 280))
    yield "zero_fill", z
    yield "one_value", np.full((70001,), 77, np.uint16)
    yield "blank", np.zeros((33, 65), np.uint16)
    yield "single", np.array([[40000]], np.uint16)


@pytest.mark.parametrize("name,img", list(_images()), ids=[n for n, _ in _images()])
def test_histogram_and_normal_score_equal_oracle(name, img):
    import torch
    from mi355x_pfa.display import histogram_u16, normal_score
    print(f"normal score {name} {img.shape}", flush=True)
    ref_h = np.bincount(img.ravel(), minlength=65536)
    assert np.array_equal(histogram_u16(img), ref_h)
    ref = rd.normal_score(img)
    out = normal_score(img)
    assert out.dtype == np.uint8 and out.shape == img.shape
    np.testing.assert_array_equal(out, ref)
    dev = torch.from_numpy(img.astype(np.int32)).to("cuda").to(torch.uint16)
    out_d = normal_score(dev)
    np.testing.assert_array_equal(out_d.cpu().numpy(), ref)
    assert np.array_equal(histogram_u16(dev), ref_h)


def test_device_inputs_misaligned_and_ragged():
    """A tensor view starting one element in (2-byte aligned) takes the scalar
    kernels; results equal the aligned ones."""
    import torch
    from mi355x_pfa.display import histogram_u16, normal_score, stretch_u8
    rs = np.random.RandomState(12)
    img = rs.randint(0, 5000, size=100_003).astype(np.uint16)
    dev = torch.from_numpy(img.astype(np.int32)).to("cuda").to(torch.uint16)
    view = dev[1:]
    assert view.data_ptr() % 16 != 0
    assert np.array_equal(histogram_u16(view), np.bincount(img[1:], minlength=65536))
    np.testing.assert_array_equal(normal_score(view).cpu().numpy(), rd.normal_score(img[1:]))
    amp = rs.rayleigh(size=70_001).astype(np.float32)
    a_dev = torch.from_numpy(amp).to("cuda")[3:]
    data, hi = stretch_u8(a_dev)
    ref, rhi = rd.geotiff_u8(amp[3:])
    assert hi.tobytes() == rhi.tobytes()
    np.testing.assert_array_equal(data.cpu().numpy(), ref)


@pytest.mark.parametrize("n", [1, 2, 7, 1000, 4_000_003])
def test_select_equals_partition(n):
    from mi355x_pfa import _lib
    from mi355x_pfa._lib import context, ptr
    import ctypes as C
    rs = np.random.RandomState(n % 97)
    x = (rs.standard_normal(n) * 10).astype(np.float32)
    if n > 10:
        x[rs.randint(0, n, size=n // 5)] = np.float32(0.25)         # ties
        x[rs.randint(0, n, size=3)] = np.inf
        x[rs.randint(0, n, size=3)] = -np.inf
    s = np.sort(x)
    ctx = context(0)
    ranks = sorted({0, n - 1, n // 2, (n * 3) // 7, max(n - 2, 0)})
    for k0 in ranks:
        k1 = min(k0 + 1, n - 1)
        v0, v1, nn = C.c_float(), C.c_float(), C.c_int64()
        ctx.check(_lib.lib().pfa_select_f32(ctx.handle, ptr(x), n, k0, k1, C.byref(v0),
                                             C.byref(v1), C.byref(nn)), "select")
        assert np.float32(v0.value).tobytes() == s[k0].tobytes(), (k0, v0.value, s[k0])
        assert np.float32(v1.value).tobytes() == s[k1].tobytes()
        assert nn.value == 0


def test_select_orders_nan_last_and_percentile_is_nan():
    from mi355x_pfa.display import percentile
    x = np.array([3.0, np.nan, -1.0, 2.0, np.nan, 0.5], np.float32)
    assert np.isnan(percentile(x, 99.7)) and np.isnan(np.percentile(x, 99.7))
    assert percentile(x[[0, 2, 3, 5]], 50.0) == np.percentile(x[[0, 2, 3, 5]], 50.0)


@pytest.mark.parametrize("n", [1, 5, 1234, 1_000_003])
def test_percentile_and_stretch_equal_numpy(n):
    import torch
    from mi355x_pfa.display import percentile, stretch_u8
    rs = np.random.RandomState(n % 89)
    amp = rs.rayleigh(size=n).astype(np.float32)
    for q in (0.0, 50.0, 99.7, 100.0):
        assert percentile(amp, q).tobytes() == np.percentile(amp, q).tobytes(), q
    data, hi = stretch_u8(amp)
    ref, rhi = rd.geotiff_u8(amp)
    assert hi.tobytes() == rhi.tobytes()
    np.testing.assert_array_equal(data, ref)
    data_d, hi_d = stretch_u8(torch.from_numpy(amp).to("cuda"))
    np.testing.assert_array_equal(data_d.cpu().numpy(), ref)


def test_full_frame_products():
    """8192 x 8192: the uint8 GeoTIFF data of a float frame and the normal
    score of a uint16 frame, from device tensors, equal the oracle's."""
    import torch
    from mi355x_pfa.display import normal_score, stretch_u8
    g = torch.Generator(device="cuda").manual_seed(8)
    re = torch.randn((8192, 8192), device="cuda", generator=g)
    im = torch.randn((8192, 8192), device="cuda", generator=g)
    amp = torch.sqrt(re * re + im * im)
    del re, im
    print("full-frame stretch", flush=True)
    data, hi = stretch_u8(amp)
    ref, rhi = rd.geotiff_u8(amp.cpu().numpy())
    assert hi.tobytes() == rhi.tobytes()
    assert torch.equal(data.cpu(), torch.from_numpy(ref))
    print("full-frame normal score", flush=True)
    i32 = torch.clamp(amp * 300, max=46341).to(torch.int32)
    i32[:, 6000:] = 0                                    # zero-fill columns
    u16 = i32.to(torch.uint16)
    del i32
    out = normal_score(u16)
    ref = rd.normal_score(u16.cpu().numpy())
    assert torch.equal(out.cpu(), torch.from_numpy(ref))
