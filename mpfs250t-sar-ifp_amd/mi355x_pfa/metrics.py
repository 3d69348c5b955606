"""Image-quality figures used to state the error of a focus result against a
higher-precision reference of the same scene: peak-normalised RMS error,
point-target peak positions and shift, and PSLR / ISLR of a point target.

Pure numpy on host arrays (or anything np.asarray accepts)."""
from __future__ import annotations

import numpy as np


def pn_rms(img, ref):
    """rms|img - ref| / max|ref| over the whole (complex or real) image."""
    a = np.asarray(img).astype(np.complex128)
    b = np.asarray(ref).astype(np.complex128)
    return float(np.sqrt(np.mean(np.abs(a - b) ** 2)) / np.max(np.abs(b)))


def top_peaks(img, k, guard=3):
    """The k largest local maxima of |img|, each masking a (2 guard + 1)^2 box."""
    a = np.abs(np.asarray(img)).astype(np.float64).copy()
    out = []
    for _ in range(k):
        r, c = np.unravel_index(np.argmax(a), a.shape)
        out.append((int(r), int(c)))
        a[max(r - guard, 0):r + guard + 1, max(c - guard, 0):c + guard + 1] = 0
    return out


def peak_shift(img, ref, k=8, guard=3):
    """Largest pixel distance between the k strongest peaks of ref and the
    nearest of the k strongest peaks of img (0 = identical positions)."""
    pi, pr = top_peaks(img, k, guard), top_peaks(ref, k, guard)
    worst = 0.0
    for r, c in pr:
        d = min(np.hypot(r - a, c - b) for a, b in pi)
        worst = max(worst, float(d))
    return worst


def _first_nulls(cut, p):
    """Indices of the first local minima of a 1-D magnitude cut either side of p."""
    lo = p
    while lo > 0 and cut[lo - 1] < cut[lo]:
        lo -= 1
    hi = p
    while hi < cut.size - 1 and cut[hi + 1] < cut[hi]:
        hi += 1
    return lo, hi


def pslr_islr(img, peak, half=32):
    """PSLR and ISLR (dB) of the point target at `peak` from a (2 half + 1)^2
    window of |img|^2.  The main lobe is the box between the first nulls of the
    range and cross cuts through the peak; PSLR = largest side-lobe / peak,
    ISLR = side-lobe energy / main-lobe energy."""
    a = np.abs(np.asarray(img)).astype(np.float64)
    r, c = peak
    r0, r1 = max(r - half, 0), min(r + half + 1, a.shape[0])
    c0, c1 = max(c - half, 0), min(c + half + 1, a.shape[1])
    w = a[r0:r1, c0:c1] ** 2
    pr, pc = r - r0, c - c0
    lo_r, hi_r = _first_nulls(w[:, pc], pr)
    lo_c, hi_c = _first_nulls(w[pr, :], pc)
    main = np.zeros_like(w, dtype=bool)
    main[lo_r:hi_r + 1, lo_c:hi_c + 1] = True
    side = w[~main]
    pk = w[pr, pc]
    pslr = 10 * np.log10(side.max() / pk) if side.size and pk > 0 else float("nan")
    islr = 10 * np.log10(side.sum() / w[main].sum()) if w[main].sum() > 0 else float("nan")
    return float(pslr), float(islr)


def error_report(img, ref, k=8):
    """Dict of the figures above for img against ref (brightest target of ref)."""
    pk = top_peaks(ref, 1)[0]
    p_i, i_i = pslr_islr(img, pk)
    p_r, i_r = pslr_islr(ref, pk)
    return {"pn_rms": pn_rms(img, ref), "peak_shift_px": peak_shift(img, ref, k),
            "pslr_db": round(p_i, 3), "islr_db": round(i_i, 3),
            "pslr_ref_db": round(p_r, 3), "islr_ref_db": round(i_r, 3)}
