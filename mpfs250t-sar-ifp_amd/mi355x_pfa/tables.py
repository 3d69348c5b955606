"""Host-side table preparation for the focus library (O(m + grid) work).

* ``sinc_table_q15``  -- the 256x32 Q15 polyphase sinc of
  mpfs/fpga/tb/gen_resample_vectors.py:206-254 (scene independent).
* ``hamming_q15``     -- floor(np.hamming(length) * 32768) padded to the grid,
  as silicon_emulator.py:185-188 builds the window tapers.
* ``fixed_tables``    -- the per-scene tables of silicon_emulator.py:128-140
  (the reference's on-board DDR roles, serialize_inputs.py:203-230).
"""
from __future__ import annotations

import math
from fractions import Fraction
from functools import lru_cache

import numpy as np

C_LIGHT = 299792458.0


@lru_cache(maxsize=4)
def sinc_table_q15(phases: int = 256, taps: int = 32) -> np.ndarray:
    """DC-normalised truncated sinc per fractional phase, rounded half up to Q15,
    clamped to int16 and corrected to sum to exactly 2**15 per phase by moving
    the residual onto taps in order of decreasing magnitude (ties keep tap
    order), each only as far as its int16 headroom allows."""
    first = -(taps // 2 - 1)
    tab = np.zeros((phases, taps), np.int16)
    for ph in range(phases):
        frac = Fraction(ph, phases)
        vals = []
        for t in range(taps):
            d = frac - (first + t)
            vals.append(Fraction(1) if d == 0 else
                        Fraction(math.sin(math.pi * float(d))) / (Fraction(math.pi) * d))
        norm = sum(vals)
        q = [min(max(math.floor(v / norm * 32768 + Fraction(1, 2)), -32768), 32767) for v in vals]
        rest = 32768 - sum(q)
        for t in sorted(range(taps), key=lambda u: -abs(q[u])):
            if not rest:
                break
            head = (32767 - q[t]) if rest > 0 else (-32768 - q[t])
            mv = rest if abs(rest) <= abs(head) else head
            q[t] += mv
            rest -= mv
        if rest:
            raise ValueError("sinc table cannot reach unity gain within int16")
        tab[ph] = q
    tab.setflags(write=False)
    return tab


def hamming_q15(length: int, extent: int) -> np.ndarray:
    """int32[extent]: floor(np.hamming(length)[k] * 32768) for k < length, else 0."""
    h = np.zeros(extent)
    h[:length] = np.hamming(length)
    return np.floor(h * 32768).astype(np.int32)


def fixed_tables(tables: dict, m: int, n: int, grid: int) -> dict:
    """Scene tables the fixed datapath consumes, from a prepare_tables() dict
    (needs ax, ay, freq, KR, KC, tan_phi)."""
    ax = np.asarray(tables["ax"], np.float64)
    ay = np.asarray(tables["ay"], np.float64)
    freq = np.asarray(tables["freq"], np.float64)
    KR = np.asarray(tables["KR"], np.float64)
    KC = np.asarray(tables["KC"], np.float64)
    tan_phi = np.asarray(tables["tan_phi"], np.float64)
    if freq.shape != (m, n):
        raise ValueError(f"freq has shape {freq.shape}, expected {(m, n)}")
    mx, my = ax.mean(), ay.mean()
    nrm = math.hypot(mx, my)
    out = {
        "f0": np.ascontiguousarray(freq[:, 0].astype(np.float32)),
        "df": np.ascontiguousarray((freq[:, 1] - freq[:, 0]).astype(np.float32)),
        "pr": (ax * (mx / nrm) + ay * (my / nrm)).astype(np.float32),
    }
    order = np.argsort(tan_phi)
    out["tan_s"] = np.ascontiguousarray(tan_phi[order].astype(np.float32))
    out["inv_order"] = np.argsort(order).astype(np.int32)
    pad_r = float(KR.max() + (KR.max() - KR.min()) + 1.0)
    pad_c = float(KC.max() + (KC.max() - KC.min()) + 1.0)
    krp = np.full(grid, pad_r, np.float32)
    krp[:n] = KR
    kcp = np.full(grid, pad_c, np.float32)
    kcp[:m] = KC
    out["krp"], out["kcp"] = krp, kcp
    out["hamr"] = hamming_q15(n, grid)
    out["hamc"] = hamming_q15(m, grid)
    return out
