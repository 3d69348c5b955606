"""The bracket-fraction division of the resample kernels (div_bracket in
csrc/pfa_div.hpp, float and fixed passes): reciprocal, two Newton steps, the quotient and one
residual correction, every FMA and product rounded once to float64.

For 0 <= n <= d of the magnitudes positions take, the result must be the
correctly rounded n / d -- what np.interp's division gives -- whatever the last
bits of the first reciprocal are (the hardware reciprocal is an
approximation; here it is perturbed by up to 2^-14 relative).  Every
operation is formed exactly in rationals and rounded once."""
from fractions import Fraction

import numpy as np


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _mul(a, b):
    return float(Fraction(a) * Fraction(b))


def _div_bracket(n, d, r0):
    r1 = _fma(_fma(-d, r0, 1.0), r0, r0)
    r2 = _fma(_fma(-d, r1, 1.0), r1, r1)
    q = _mul(n, r2)
    return _fma(_fma(-d, q, n), r2, q)


def _cases(rng, count):
    out = []
    for _ in range(count):
        d = float(10.0 ** rng.uniform(-6, 3))
        out.append((float(d * rng.uniform(0.0, 1.0)), d))
    for d in (1.0, 0.5, 3.0, 1e-6, 1e3, 0.1):
        out += [(0.0, d), (float(np.nextafter(d, 0.0)), d), (d / 3.0, d), (d * 0.5, d), (d, d)]
    return [(n, d) for n, d in out if 0.0 <= n <= d]


def test_div_bracket_is_correctly_rounded():
    rng = np.random.default_rng(11)
    for n, d in _cases(rng, 3000):
        want = float(Fraction(n) / Fraction(d))
        for rel in (0.0, 2.0 ** -14, -(2.0 ** -14), float(rng.uniform(-2.0 ** -14, 2.0 ** -14))):
            r0 = (1.0 / d) * (1.0 + rel)
            assert _div_bracket(n, d, r0) == want, (n, d, rel)


def test_a_mistake_is_caught():
    """One Newton step instead of two leaves quotients off by an ulp for a
    perturbed start: the check above would see a wrong sequence."""
    rng = np.random.default_rng(12)
    wrong = 0
    for n, d in _cases(rng, 500):
        r0 = (1.0 / d) * (1.0 + 2.0 ** -14)
        r1 = _fma(_fma(-d, r0, 1.0), r0, r0)
        q = _mul(n, r1)
        got = _fma(_fma(-d, q, n), r1, q)
        wrong += got != float(Fraction(n) / Fraction(d))
    assert wrong > 0
