"""Deterministic synthetic spotlight phase histories (FX domain, SGN = -1).

No Umbra CPHD scene is available offline, so the focus path is exercised on
synthetic collects with the same structure the reference reads from a CPHD
(form_image_pfa.py:326-360, sar_pipeline.py:61-89): per-pulse Tx/Rcv/SRP
positions, SC0/SCSS, the image-plane axes uIAX/uIAY, and a complex64 signal
``s[i, k] = sum_t a_t exp(-j 4 pi f_ik (|p_i - r_t| - |p_i - srp|) / c) + noise``.

Geometry: straight-line flight at a 10 km ground standoff and 3 km altitude,
aperture sized for ~0.5 m cross-range resolution, scene reference point at the
origin, image plane = ground (x, y).  Random draws use the legacy
``np.random.RandomState`` stream, which is stable across numpy versions.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

C_LIGHT = 299792458.0


@dataclass
class SceneSpec:
    pulses: int
    samples: int
    targets: int = 16
    clutter: int = 0               # weak scatterers with exponential power
    clutter_power: float = 1e-3    # mean |a|^2 of a clutter scatterer
    scene_half: float = 250.0      # targets uniform in [-half, half]^2 (m)
    clutter_half: float = 1000.0
    snr_db: float = 20.0
    f0: float = 9.6e9
    bandwidth: float = 300e6
    standoff: float = 10_000.0
    altitude: float = 3_000.0
    res_cross: float = 0.5
    seed: int = 1
    phase_error: dict | None = None    # {"quad_rad": 3.0, "walk_sigma": 0.05}


# the configurations the focus path is measured on
C1 = SceneSpec(1024, 1024, targets=16, seed=1)
C2 = SceneSpec(5634, 4319, targets=256, clutter=512, seed=2)
C3 = SceneSpec(8167, 8192, targets=256, clutter=512, seed=3)
C5 = SceneSpec(4096, 4096, targets=64, seed=5,
               phase_error={"quad_rad": 3.0, "walk_sigma": 0.05})


@dataclass
class Scene:
    spec: SceneSpec
    tx: np.ndarray              # (m, 3)
    rcv: np.ndarray
    srp: np.ndarray
    sc0: np.ndarray             # (m,)
    scss: np.ndarray            # (m,)
    uiax: np.ndarray
    uiay: np.ndarray
    sgn: int
    targets: np.ndarray         # (T, 3) positions
    amps: np.ndarray            # (T,) complex
    phase_error: np.ndarray | None
    sig: object = None          # complex64 (m, n) numpy array or torch tensor
    extra: dict = field(default_factory=dict)

    # ---- derived inputs of the focus stage (form_image_pfa.py:332-343, :359-360)
    def look(self):
        apc = 0.5 * (self.tx + self.rcv)
        los = self.srp - apc
        u = los / np.linalg.norm(los, axis=1, keepdims=True)
        return u @ self.uiax, u @ self.uiay

    def freq(self):
        k = np.arange(self.spec.samples)
        return self.sc0[:, None] + k[None, :] * (self.scss[:, None] * 1)


def _geometry(spec: SceneSpec):
    m = spec.pulses
    lam = C_LIGHT / spec.f0
    slant = math.hypot(spec.standoff, spec.altitude)
    aperture = slant * lam / (2.0 * spec.res_cross)
    y = np.linspace(-aperture / 2, aperture / 2, m)
    apc = np.stack([np.full(m, -spec.standoff), y, np.full(m, spec.altitude)], axis=1)
    srp = np.zeros((m, 3))
    sc0 = np.full(m, spec.f0 - spec.bandwidth / 2)
    scss = np.full(m, spec.bandwidth / (spec.samples - 1))
    return apc, srp, sc0, scss


def _scatterers(spec: SceneSpec, rs):
    T = spec.targets
    pos = np.zeros((T + spec.clutter, 3))
    pos[:T, :2] = rs.uniform(-spec.scene_half, spec.scene_half, size=(T, 2))
    amp = np.empty(T + spec.clutter, np.complex128)
    amp[:T] = rs.rayleigh(1.0, size=T) * np.exp(2j * np.pi * rs.uniform(size=T))
    if spec.clutter:
        C = spec.clutter
        pos[T:, :2] = rs.uniform(-spec.clutter_half, spec.clutter_half, size=(C, 2))
        power = rs.exponential(spec.clutter_power, size=C)
        amp[T:] = np.sqrt(power) * np.exp(2j * np.pi * rs.uniform(size=C))
    return pos, amp


def _phase_error(spec: SceneSpec, rs):
    if not spec.phase_error:
        return None
    m = spec.pulses
    t = np.linspace(-1.0, 1.0, m)
    quad = spec.phase_error.get("quad_rad", 0.0) * t * t      # edge-to-edge quad
    walk = np.cumsum(rs.normal(0.0, spec.phase_error.get("walk_sigma", 0.0), size=m))
    walk -= walk.mean()
    return quad + walk


def make_scene(spec: SceneSpec, device=None, chunk_targets=32, signal=True) -> Scene:
    """Build the scene.  ``device=None`` synthesises the signal with numpy
    (float64 phases); a torch device string synthesises it there in float64
    (used for the large configurations).  ``signal=False`` builds the
    geometry, scatterers and phase error only (``sig`` stays None)."""
    rs = np.random.RandomState(spec.seed)
    apc, srp, sc0, scss = _geometry(spec)
    pos, amp = _scatterers(spec, rs)
    noise_rs = np.random.RandomState(spec.seed + 1000)
    perr = _phase_error(spec, rs)
    scene = Scene(spec, apc.copy(), apc.copy(), srp, sc0, scss,
                  np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0]), -1,
                  pos, amp, perr)
    m, n = spec.pulses, spec.samples
    if not signal:
        return scene
    # range difference per (pulse, scatterer)
    r0 = np.linalg.norm(apc - srp, axis=1)                              # (m,)
    dR = np.linalg.norm(apc[:, None, :] - pos[None, :, :], axis=2) - r0[:, None]
    freq = scene.freq()
    if device is None:
        sig = np.zeros((m, n), np.complex128)
        for c0 in range(0, pos.shape[0], chunk_targets):
            c1 = min(c0 + chunk_targets, pos.shape[0])
            for t in range(c0, c1):
                ph = (-4.0 * np.pi / C_LIGHT) * freq * dR[:, t:t + 1]
                sig += amp[t] * np.exp(1j * ph)
        p_sig = float(np.mean(np.abs(sig) ** 2))
        sigma = math.sqrt(p_sig / 10 ** (spec.snr_db / 10) / 2)
        sig += sigma * (noise_rs.standard_normal((m, n)) + 1j * noise_rs.standard_normal((m, n)))
        if perr is not None:
            sig *= np.exp(1j * perr)[:, None]
        scene.sig = sig.astype(np.complex64)
    else:
        import torch
        dev = torch.device(device)
        fr = torch.as_tensor(freq, dtype=torch.float64, device=dev)
        acc_r = torch.zeros((m, n), dtype=torch.float64, device=dev)
        acc_i = torch.zeros((m, n), dtype=torch.float64, device=dev)
        dR_t = torch.as_tensor(dR, dtype=torch.float64, device=dev)
        k = -4.0 * math.pi / C_LIGHT
        for t in range(pos.shape[0]):
            ph = torch.remainder(fr * (k * dR_t[:, t:t + 1]), 2 * math.pi)
            a = amp[t]
            c, s = torch.cos(ph), torch.sin(ph)
            acc_r += a.real * c - a.imag * s
            acc_i += a.real * s + a.imag * c
            del ph, c, s
        p_sig = float((acc_r * acc_r + acc_i * acc_i).mean().item())
        sigma = math.sqrt(p_sig / 10 ** (spec.snr_db / 10) / 2)
        g = torch.Generator(device=dev)
        g.manual_seed(spec.seed + 1000)
        acc_r += sigma * torch.randn((m, n), generator=g, dtype=torch.float64, device=dev)
        acc_i += sigma * torch.randn((m, n), generator=g, dtype=torch.float64, device=dev)
        if perr is not None:
            pe = torch.as_tensor(perr, dtype=torch.float64, device=dev)[:, None]
            c, s = torch.cos(pe), torch.sin(pe)
            acc_r, acc_i = acc_r * c - acc_i * s, acc_r * s + acc_i * c
        scene.sig = torch.complex(acc_r.float(), acc_i.float())
        del acc_r, acc_i
    return scene


class _NS:
    def __init__(self, **kw):
        self.__dict__.update(kw)


class SceneReader:
    """A scene presented as the CPHD reader the reference's table preparation
    reads (sar_pipeline.py:61-75, form_image_pfa.py:326-343, 346-348):
    ``get_data_size_as_tuple()``, ``read_pvp_variable(name, 0)`` for TxPos,
    RcvPos, SRPPos, SC0 and SCSS, ``read_chip(dim1, dim2, index=0)`` with
    (start, stop, step) ranges, ``cphd_meta`` (Global.SGN, Global.DomainType,
    SceneCoordinates.ReferenceSurface.Planar.uIAX / uIAY) and ``close()``."""

    def __init__(self, scene: Scene):
        self.scene = scene
        vec = lambda v: _NS(X=float(v[0]), Y=float(v[1]), Z=float(v[2]))  # noqa: E731
        planar = _NS(uIAX=vec(scene.uiax), uIAY=vec(scene.uiay))
        self.cphd_meta = _NS(Global=_NS(SGN=scene.sgn, DomainType="FX"),
                             SceneCoordinates=_NS(ReferenceSurface=_NS(Planar=planar)))
        self._pvp = {"TxPos": scene.tx, "RcvPos": scene.rcv, "SRPPos": scene.srp,
                     "SC0": scene.sc0, "SCSS": scene.scss}

    def get_data_size_as_tuple(self):
        return ((self.scene.spec.pulses, self.scene.spec.samples),)

    def read_pvp_variable(self, name, index=0):
        if index != 0:
            raise IndexError("one channel")
        return np.array(self._pvp[name], copy=True)

    def read_chip(self, dim1_range, dim2_range, index=0):
        if index != 0:
            raise IndexError("one channel")
        sig = self.scene.sig
        if not isinstance(sig, np.ndarray):
            sig = sig.cpu().numpy()
        return np.ascontiguousarray(sig[slice(*dim1_range), slice(*dim2_range)])

    def close(self):
        pass
