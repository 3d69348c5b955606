"""Deterministic fixed-point inputs at the production shapes, regenerated bit for
bit wherever the tests run (test helper; the digests in
tests/golden/fixed_digests.json were computed from these inputs by the oracle).

The phase history is int16 I/Q made by integer arithmetic only: a sum of
integer chirps, each a quadratic phase index into a 4096-entry cosine table
rounded to integers, plus hashed integer noise, saturated to int16.  The same
integer operations run in numpy (here) or torch (on the GPU), so both give the
same codes; the tables are the fixed datapath's scene tables of a synthetic
straight-line collect of that shape, formed by the product's table code.
"""
from __future__ import annotations

import hashlib
import math

import numpy as np

COS_BITS = 12
COS_AMP = 1024
SHAPES = {"C3": (8167, 8192), "C2": (5634, 4319)}


def cos_table():
    n = 1 << COS_BITS
    return np.array([int(round(math.cos(2.0 * math.pi * j / n) * COS_AMP)) for j in range(n)],
                    np.int64)


def chirp_params(seed, count=24):
    """(a, b, c, d, amp) of each chirp from a 64-bit LCG (plain Python ints)."""
    x = (seed * 6364136223846793005 + 1442695040888963407) % (1 << 64)
    out = []
    for _ in range(count):
        vals = []
        for _ in range(5):
            x = (x * 6364136223846793005 + 1442695040888963407) % (1 << 64)
            vals.append(x >> 33)
        a, b, c, d, amp = vals
        out.append((a % 40000, b % 9000, c % 64, d % 16, 400 + amp % 1000))
    return out


def phase_history_iq(m, n, seed, device=None):
    """int16 [m][n][2] codes.  device=None: numpy; else a torch device."""
    if device is None:
        xp, tab = np, cos_table()
        i = np.arange(m, dtype=np.int64)[:, None]
        k = np.arange(n, dtype=np.int64)[None, :]
        zeros = lambda: np.zeros((m, n), np.int64)  # noqa: E731
    else:
        import torch
        xp = torch
        tab = torch.as_tensor(cos_table(), device=device)
        i = torch.arange(m, dtype=torch.int64, device=device)[:, None]
        k = torch.arange(n, dtype=torch.int64, device=device)[None, :]
        zeros = lambda: torch.zeros((m, n), dtype=torch.int64, device=device)  # noqa: E731
    mask = (1 << (COS_BITS + 4)) - 1
    quarter = 1 << (COS_BITS - 2)
    acc_r, acc_i = zeros(), zeros()
    ik = i * k
    ii = i * i
    for a, b, c, d, amp in chirp_params(seed):
        p = ((a * k + b * i + c * ik + d * ii) & mask) >> 4
        acc_r += (amp * tab[p]) >> 10
        acc_i += (amp * tab[(p - quarter) & ((1 << COS_BITS) - 1)]) >> 10
    h = ((i * 73856093) ^ (k * 19349663) ^ (seed * 83492791)) & 0xFFFFFFFF
    h = ((h ^ (h >> 13)) * 1274126177) & 0xFFFFFFFF
    h = h ^ (h >> 16)
    acc_r += (h & 1023) - 512
    acc_i += ((h >> 10) & 1023) - 512
    if device is None:
        out = np.empty((m, n, 2), np.int16)
        out[..., 0] = np.clip(acc_r, -32768, 32767)
        out[..., 1] = np.clip(acc_i, -32768, 32767)
        return out
    return xp.stack([acc_r.clamp(-32768, 32767), acc_i.clamp(-32768, 32767)],
                    dim=-1).to(xp.int16)


def scene_tables(config, grid=8192):
    """The fixed datapath's tables of a synthetic collect of this shape
    (float64 geometry in numpy; no device)."""
    from mi355x_pfa import synth
    from mi355x_pfa.pfa import float_geometry
    from mi355x_pfa.tables import fixed_tables
    m, n = SHAPES[config]
    seed = {"C3": 3, "C2": 2}[config]
    sc = synth.make_scene(synth.SceneSpec(m, n, seed=seed), signal=False)
    ax, ay = sc.look()
    geom = float_geometry(ax, ay, f0=sc.sc0, df=sc.scss, n=n)
    freq2 = np.zeros((m, n))
    freq2[:, 0] = sc.sc0
    freq2[:, 1] = sc.sc0 + sc.scss
    tabs = fixed_tables({"ax": ax, "ay": ay, "freq": freq2, "KR": geom["KR"], "KC": geom["KC"],
                         "tan_phi": geom["tan_phi"]}, m, n, grid)
    return m, n, tabs


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()
