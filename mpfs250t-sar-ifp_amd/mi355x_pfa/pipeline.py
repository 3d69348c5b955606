"""Table preparation of the reference's host pipeline without the O(m n) host
arrays (mpfs/host/sar_pipeline.py:36-115, ``prepare_tables``).

The reference builds, per scene, the (m, n) float64 frequency array, the
(m, n) kr array whose row extrema set KR's extent, the (m, n) product behind
kc_max and the (m, n) float32 outer Hamming taper: at 8167 x 8192 that is
about 1.9 GB of host arrays and seconds of numpy before the focus starts.
Here every one of them is O(m + n):

* ``freq`` is a ``LinearFreq``: per pulse ``sc0`` and ``step = scss * nu`` and
  the first kept sample ``k0``; element [i, k] is ``sc0[i] + (k0 + k) *
  step[i]``, rounded as numpy rounds the reference's expression.  The focus
  kernels evaluate it per sample on the GPU; ``np.asarray(freq)`` builds the
  reference's array for a consumer that needs it.
* each kr row is monotone in k (an affine frequency times a per-pulse
  constant, every operation rounded monotonically), so its extrema are its end
  points and KR's extent is exact from 2 m values;
* ``max |KR[j] tan_phi[i]|`` is ``max |KR| * max |tan_phi|`` (rounding is
  monotone), with KR's largest magnitude at an end of the linspace;
* ``window`` is an ``OuterWindow`` of the two np.hamming vectors; the focus
  forms ``f32(hm[q] * hn[j])`` in the kernel, and ``np.asarray(window)``
  gives the reference's float32 outer product.

The O(m) look geometry (los, unit vectors, projections on uIAX / uIAY) uses
the reference's numpy operations, so ax, ay, tan_phi, KR, KC and geo are the
reference's values bit for bit.
"""
from __future__ import annotations

import numpy as np

from .pfa import float_geometry


class LinearFreq:
    """The (m, n) RF frequency table freq[i, k] = sc0[i] + (k0 + k) * step[i]
    (sar_pipeline.py:88-89) held as its per-pulse parameters."""

    def __init__(self, sc0, step, k0, n):
        self.sc0 = np.ascontiguousarray(sc0, np.float64)
        self.step = np.ascontiguousarray(step, np.float64)
        self.k0 = int(k0)
        self.shape = (self.sc0.size, int(n))
        self.dtype = np.dtype(np.float64)
        self.ndim = 2

    def __array__(self, dtype=None, copy=None):
        k = self.k0 + np.arange(self.shape[1])
        a = self.sc0[:, None] + k[None, :] * self.step[:, None]
        return a if dtype is None else a.astype(dtype)

    def __len__(self):
        return self.shape[0]


class OuterWindow:
    """The (m, n) taper np.outer(np.hamming(m), np.hamming(n)).astype(float32)
    (sar_pipeline.py:108) held as its two vectors."""

    def __init__(self, m, n):
        self.hm, self.hn = np.hamming(m), np.hamming(n)
        self.shape = (m, n)
        self.dtype = np.dtype(np.float32)
        self.ndim = 2

    def __array__(self, dtype=None, copy=None):
        a = np.outer(self.hm, self.hn).astype(np.float32)
        return a if dtype is None else a.astype(dtype)

    def __len__(self):
        return self.shape[0]


def prepare_tables(reader, meta, deci_pulse=1, deci_sample=1, crop_sample=0):
    """sar_pipeline.prepare_tables: the same dict (ax, ay, freq, sgn, KR, KC,
    tan_phi, window, geo, center_ecef, uiax, uiay, dims, deci, n_vec, n_samp,
    samp0) with ``freq`` a ``LinearFreq`` and ``window`` an ``OuterWindow``.
    ``crop_sample`` keeps that many centred (post-decimation) samples and
    ``samp0`` is the raw index of the first one, as in the reference."""
    mu, nu = max(1, deci_pulse), max(1, deci_sample)
    n_vec, n_samp = reader.get_data_size_as_tuple()[0]
    sgn = int(meta.Global.SGN)
    pvp = {k: np.asarray(reader.read_pvp_variable(k, 0))[::mu]
           for k in ("TxPos", "RcvPos", "SRPPos", "SC0", "SCSS")}
    apc = 0.5 * (pvp["TxPos"] + pvp["RcvPos"])
    planar = meta.SceneCoordinates.ReferenceSurface.Planar
    uiax = np.array([planar.uIAX.X, planar.uIAX.Y, planar.uIAX.Z])
    uiay = np.array([planar.uIAY.X, planar.uIAY.Y, planar.uIAY.Z])
    srp = pvp["SRPPos"]
    los = srp - apc
    u = los / np.linalg.norm(los, axis=1, keepdims=True)
    ax, ay = u @ uiax, u @ uiay
    m, n = len(range(0, n_vec, mu)), len(range(0, n_samp, nu))
    d0 = 0
    if crop_sample and crop_sample < n:
        d0 = (n - crop_sample) // 2
        n = crop_sample
    step = pvp["SCSS"] * nu
    geom = float_geometry(ax, ay, f0=pvp["SC0"], df=step, n=n, k0=d0, _hypot=np.hypot)
    KR, KC = geom["KR"], geom["KC"]
    geo = {"dc": 1.0 / (KC[-1] - KC[0]), "dr": 1.0 / (KR[-1] - KR[0]),
           "dhat": geom["geo"]["dhat"], "chat": geom["geo"]["chat"]}
    return {"ax": ax, "ay": ay, "freq": LinearFreq(pvp["SC0"], step, d0, n), "sgn": sgn,
            "KR": KR, "KC": KC, "tan_phi": geom["tan_phi"], "window": OuterWindow(m, n),
            "geo": geo, "center_ecef": srp[0], "uiax": uiax, "uiay": uiay, "dims": (m, n),
            "deci": (mu, nu), "n_vec": n_vec, "n_samp": n_samp, "samp0": d0 * nu}
