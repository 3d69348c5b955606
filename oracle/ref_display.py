"""numpy restatement of the reference's product passes after detection.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``normal_score``  -- mpfs/host/sar_display.py:24-60 (65536-bin histogram
                       CDF, midpoint ranks through the inverse normal CDF,
                       zero-fill excluded and kept black)
* ``geotiff_u8``    -- src/form_image_pfa.py:286-291 (the uint8 data of
                       save_detected_geotiff: 99.7-percentile clip level)

Both are the reference's numpy / scipy operations on the whole image; the
tests pin them with closed forms (two equal-count levels, a uniform ramp,
the Gaussian shape of the output histogram, percentile values of known
arrays) rather than by re-reading this code.
"""
from __future__ import annotations

from math import sqrt

import numpy as np


def normal_score(img, sigma=3.0, exclude_zero=True):
    """sar_display.py:24-60."""
    a = np.asarray(img)
    flat = a.ravel()
    lo, hi = int(flat.min()), int(flat.max())
    nb = hi - lo + 1
    counts = np.bincount((flat - lo).astype(np.int64), minlength=nb).astype(np.float64)
    if exclude_zero and lo <= 0 < lo + nb:
        counts[0 - lo] = 0.0
    tot = counts.sum()
    if tot <= 0:
        return np.zeros(a.shape, np.uint8)
    cdf = (np.cumsum(counts) - 0.5 * counts) / tot
    cdf = np.clip(cdf, 1e-9, 1 - 1e-9)
    from scipy.special import erfinv
    z = sqrt(2.0) * erfinv(2.0 * cdf - 1.0)
    lut = np.clip(128.0 + 127.0 * z / sigma, 0, 255).astype(np.uint8)
    if exclude_zero and lo <= 0 < lo + nb:
        lut[0 - lo] = 0
    return lut[(a - lo).astype(np.int64)]


def geotiff_u8(amp, pct=99.7):
    """form_image_pfa.py:286-291 for dtype="uint8": (data, hi)."""
    amp = np.asarray(amp, np.float32)
    hi = np.percentile(amp, pct)
    return np.clip(amp / (hi + 1e-12) * 255.0, 0, 255).astype(np.uint8), hi
