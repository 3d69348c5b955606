"""FP64 CPU reference of the autofocus sharpness and its gradient (test
infrastructure): the oracle's float focus (oracle/ref_float, complex128) with
the pass-2 gathers written as explicit per-range-bin matrices, evaluated in
torch float64 on the CPU so torch.autograd gives dS/dphi independently of the
library's hand-written adjoint."""
from __future__ import annotations

import numpy as np

from oracle import ref_fixed, ref_float as rfl


def pass1(sig, freq, ax, ay, interp, table):
    """Range pass in FP64 (oracle/ref_float.resample_kspace's first loop)."""
    g = rfl.geometry(freq, ax, ay)
    m, n = sig.shape
    g1 = np.empty((m, n), np.complex128)
    for i in range(m):
        xp, row = g["kr"][i], np.asarray(sig[i], np.complex128)
        if xp[0] > xp[-1]:
            xp, row = xp[::-1], row[::-1]
        g1[i] = rfl._interp_line(g["KR"], xp, row, interp, table)
    return g1, g


def pass2_matrix(query, xp, interp, table):
    """Z with Z @ f == the oracle's line gather of f at `query` (sinc32 or np.interp)."""
    S, Qn = xp.size, query.size
    Z = np.zeros((Qn, S))
    J = np.searchsorted(xp, query, side="right") - 1
    for q in range(Qn):
        j, x = J[q], query[q]
        if j < 0 or x > xp[-1]:
            continue
        if j == S - 1:
            Z[q, S - 1] = 1.0
            continue
        fr = (x - xp[j]) / (xp[j + 1] - xp[j])
        if interp == "sinc32" and 15 <= j <= S - 17:
            wq = min(int(np.rint(fr * 32768.0)), 32767)
            Z[q, j - 15:j + 17] = table[wq >> 7].astype(np.float64) / 32768.0
        else:
            Z[q, j] += 1.0 - fr
            Z[q, j + 1] += fr
    return Z


class Reference:
    """S(phi) in torch float64 on the CPU for one scene."""

    def __init__(self, sig, freq, ax, ay, sgn, interp="sinc32", window=True):
        import torch
        self.torch = torch
        table = ref_fixed.sinc_table_q15() if interp == "sinc32" else None
        g1, g = pass1(sig, freq, ax, ay, interp, table)
        m, n = g1.shape
        self.m, self.n, self.sgn = m, n, sgn
        self.order = g["order"]
        tan_s = g["tan_s"]
        Zs = np.stack([pass2_matrix(g["KC"], g["KR"][j] * tan_s, interp, table) for j in range(n)])
        self.check_line = (Zs[n // 2], g["KC"], g["KR"][n // 2] * tan_s, interp, table)
        self.Z = torch.from_numpy(Zs)                       # (n, q, p)
        self.g1 = torch.from_numpy(g1)
        w = np.outer(np.hamming(m), np.hamming(n)) if window else np.ones((m, n))
        self.w = torch.from_numpy(w)
        self.m2, self.n2 = rfl.to_pow2(m), rfl.to_pow2(n)

    def image(self, phi):
        t = self.torch
        g1p = self.g1 * t.exp(1j * phi.to(t.float64))[:, None]
        g1s = g1p[t.from_numpy(self.order.astype(np.int64))]          # (p, j)
        g2 = t.einsum("jqp,pj->qj", self.Z.to(t.complex128), g1s) * self.w
        pad = t.zeros((self.m2, self.n2), dtype=t.complex128)
        pad[:self.m, :self.n] = g2
        fwd = t.fft.fft2 if self.sgn < 0 else t.fft.ifft2
        return t.fft.fftshift(fwd(t.fft.ifftshift(pad)))

    def sharpness(self, phi):
        a = self.image(phi).abs() ** 2
        return (a * a).sum() / a.sum() ** 2

    def gradient(self, phi):
        t = self.torch
        p = t.tensor(np.asarray(phi, np.float64), requires_grad=True)
        s = self.sharpness(p)
        s.backward()
        return float(s.detach()), p.grad.numpy().copy()
