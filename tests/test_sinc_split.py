"""CPU checks of the arithmetic behind the fixed-point gathers' one-dot-product
form (csrc/pfa_fixed.hip, sinc_q15s): the reference's 32-tap sum
gen_resample_vectors.gather_sinc (:257-296), acc = sum c_t x_t, then
(acc + 2^14) >> 15, computed from two int32 accumulators (taps 0..15 and
16..31) and a rounding formed from 15-bit halves."""
import numpy as np

from conftest import GOLDEN


def _wrap32(v):
    v = np.asarray(v, np.int64) & 0xFFFFFFFF
    return np.where(v >= 1 << 31, v - (1 << 32), v)


def _round_pair15(a, b):
    lo = ((a & 32767) + (b & 32767) + 16384) >> 15
    return np.clip((a >> 15) + (b >> 15) + lo, -32768, 32767)


def test_table_half_sums_allow_int32_accumulators():
    tab = np.load(GOLDEN / "sinc_table_q15.npy").astype(np.int64)
    a = np.abs(tab)
    assert a[:, :16].sum(1).max() <= 65535 and a[:, 16:].sum(1).max() <= 65535
    # the worst input for each phase and half: x_t = +-32768 with the taps' signs
    worst = np.maximum(a[:, :16].sum(1), a[:, 16:].sum(1)) * 32768
    assert worst.max() < 2 ** 31


def test_split_rounding_equals_the_exact_sum():
    rs = np.random.RandomState(1)
    tab = np.load(GOLDEN / "sinc_table_q15.npy").astype(np.int64)
    for trial in range(20):
        x = rs.randint(-32768, 32768, size=(4096, 32)).astype(np.int64)
        if trial % 4 == 0:                        # saturated inputs signed like the taps
            ph = rs.randint(0, 256, 4096)
            x = np.where(tab[ph] >= 0, 32767, -32768) * (1 if trial % 8 == 0 else -1)
        else:
            ph = rs.randint(0, 256, 4096)
        c = tab[ph]
        exact = np.clip(((c * x).sum(1) + (1 << 14)) >> 15, -32768, 32767)
        # each half accumulated in int32 with wrap-around (as the dot products do)
        a = _wrap32((c[:, :16] * x[:, :16]).sum(1))
        b = _wrap32((c[:, 16:] * x[:, 16:]).sum(1))
        assert np.array_equal(a, (c[:, :16] * x[:, :16]).sum(1))   # never wrapped
        assert np.array_equal(b, (c[:, 16:] * x[:, 16:]).sum(1))
        assert np.array_equal(_round_pair15(a, b), exact)


def test_split_rounding_at_the_int32_edges():
    v = np.array([-(1 << 31), -(1 << 31) + 1, -16385, -16384, -1, 0, 1, 16383, 16384,
                  (1 << 31) - 1], np.int64)
    a, b = np.meshgrid(v, v)
    a, b = a.ravel(), b.ravel()
    exact = np.clip((a + b + (1 << 14)) >> 15, -32768, 32767)
    assert np.array_equal(_round_pair15(a, b), exact)
