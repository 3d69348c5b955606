"""Gradient autofocus on the MI355X float focus path (no counterpart in the
reference).

Per-pulse phase corrections phi (radians, original pulse order) give the
corrected history sig'[i, k] = sig[i, k] e^{j phi_i}; the image is the PFA
focus of sig', and the sharpness

    S(phi) = sum |I|^4 / (sum |I|^2)^2

is maximised.  ``AutofocusPlan`` keeps the range pass of the history as its
checkpoint (the phase commutes with it) and runs the forward image + S and the
hand-written adjoint (libmi355x_pfa: pfa_af_forward / pfa_af_backward) on
torch CUDA tensors.  ``Sharpness`` is the torch.autograd.Function over them.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import PfaError, ptr
from .pfa import FloatPlan


class AutofocusPlan:
    """A float plan with the history's range pass checkpointed on the device.

    sig: torch complex64 CUDA tensor (m, n); f0/df compact frequencies (or a
    full freq CUDA float64 tensor)."""

    def __init__(self, sig, geom, f0=None, df=None, freq=None, sgn=-1, interp="sinc32",
                 window=True, device=0):
        import torch
        self.torch = torch
        m, n = sig.shape
        self.plan = FloatPlan(m, n, sgn, geom, f0=f0, df=df, window=window, interp=interp,
                              device=device)
        self.ctx = self.plan.ctx
        self.m, self.n, self.m2, self.n2 = m, n, self.plan.m2, self.plan.n2
        self.ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        self._freq = freq
        self.ctx.check(_lib.lib().pfa_af_checkpoint(self.plan._h, ptr(sig), ptr(freq)),
                       "af_checkpoint")

    def _phi(self, phi):
        t = self.torch
        if not (isinstance(phi, t.Tensor) and phi.is_cuda and phi.dtype == t.float32
                and phi.is_contiguous() and phi.shape == (self.m,)):
            raise PfaError(-1, f"phi must be a contiguous CUDA float32 tensor of shape ({self.m},)")
        return phi

    def forward(self, phi):
        """(S as a 0-d float64 tensor, complex image (m2, n2), stats [E2, E4, S])."""
        t = self.torch
        phi = self._phi(phi)
        img = t.empty((self.m2, self.n2), dtype=t.complex64, device=phi.device)
        stats = t.empty(3, dtype=t.float64, device=phi.device)
        self.ctx.check(_lib.lib().pfa_af_forward(self.plan._h, ptr(phi), ptr(img), ptr(stats)),
                       "af_forward")
        return stats[2], img, stats

    def backward(self, phi, img, stats):
        """dS/dphi (float32, (m,)) for the image and stats of forward(phi)."""
        t = self.torch
        phi = self._phi(phi)
        grad = t.empty(self.m, dtype=t.float32, device=phi.device)
        self.ctx.check(_lib.lib().pfa_af_backward(self.plan._h, ptr(phi), ptr(img), ptr(stats),
                                                  ptr(grad)), "af_backward")
        return grad

    def close(self):
        self.plan.close()


def _autograd_fn():
    import torch

    class Sharpness(torch.autograd.Function):
        """S(phi) with the library's adjoint as its backward."""

        @staticmethod
        def forward(ctx, phi, plan):
            S, img, stats = plan.forward(phi.detach().contiguous())
            ctx.plan = plan
            ctx.save_for_backward(phi.detach(), img, stats)
            return S.to(torch.float32)

        @staticmethod
        def backward(ctx, gS):
            phi, img, stats = ctx.saved_tensors
            g = ctx.plan.backward(phi.contiguous(), img, stats)
            return gS.to(torch.float32) * g, None

    return Sharpness


_FN = None


def sharpness(phi, plan):
    """Differentiable S(phi) (torch.autograd)."""
    global _FN
    if _FN is None:
        _FN = _autograd_fn()
    return _FN.apply(phi, plan)


def autofocus(plan, steps=50, lr=0.05, phi0=None, record=False):
    """Gradient ascent on S with Adam; returns (phi, history of S)."""
    import torch
    phi = (torch.zeros(plan.m, dtype=torch.float32, device="cuda") if phi0 is None
           else phi0.detach().clone().float().cuda())
    phi.requires_grad_(True)
    opt = torch.optim.Adam([phi], lr=lr)
    hist = []
    for _ in range(steps):
        opt.zero_grad()
        s = sharpness(phi, plan)
        (-s).backward()
        opt.step()
        hist.append(float(s.detach()))
    return phi.detach(), np.array(hist)
