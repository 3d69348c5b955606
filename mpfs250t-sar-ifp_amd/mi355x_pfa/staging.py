"""The board's on-disk input contract, written and read for the fixed datapath
(mpfs/host/serialize_inputs.py:150-277, the role files of ddr_layout.py).

A staging directory holds one little-endian binary per DDR role plus
``layout.json``:

    sig.bin       int16 I/Q interleaved, m x n x 2 (block-floating-point codes)
    f0.bin df.bin pr.bin tans.bin   float32[m]
    invorder.bin  int32[m]
    krgrid.bin    float32[Np]   (KR, then an out-of-range pad so the pad is zero)
    kcgrid.bin    float32[Mp]
    hamr.bin      int16[Np]     round(hamming(n) * 32768) clipped to 32767, 0 in the pad
    hamc.bin      int16[Mp]
    kr.bin kc.bin tanphi.bin    float32 (off-board geocoding / debug)

``write_staging`` forms them from a signal and a table-preparation dict the way
serialize_inputs.serialize does (the quantiser on the GPU); ``load_staging``
reads a directory back with size and CRC-32 checks (the board's own check) and
``form_image_staged`` runs the bit-exact fixed datapath on it.  The board's
tapers are the rounded int16 ones above, where silicon_emulator.form_image
uses floor values; the datapath takes either (int32 tapers).  The JTAG load
script, job descriptor and DDR addresses are board tooling and are not
written.
"""
from __future__ import annotations

import json
import math
import zlib
from pathlib import Path

import numpy as np

from ._lib import PfaError
from .emulator import FixedPlan, quantize_iq16

# role -> (dtype, length key)
ROLES = {"f0": (np.float32, "M"), "df": (np.float32, "M"), "pr": (np.float32, "M"),
         "tans": (np.float32, "M"), "invorder": (np.int32, "M"),
         "krgrid": (np.float32, "R"), "kcgrid": (np.float32, "A"),
         "hamr": (np.int16, "R"), "hamc": (np.int16, "A")}


def crc32(buf) -> int:
    """IEEE 802.3 CRC-32 (ddr_layout.py:341-343)."""
    return zlib.crc32(buf) & 0xFFFFFFFF


def _pow2(v: int) -> int:
    return 1 << max(0, math.ceil(math.log2(v)))


def _freq_column(freq, k):
    if hasattr(freq, "sc0"):                      # LinearFreq: column k without the array
        kk = np.array([freq.k0 + k])
        return (freq.sc0[:, None] + kk[None, :] * freq.step[:, None])[:, 0]
    return np.asarray(freq)[:, k]


def staging_tables(tables: dict, grid: int = 0) -> dict:
    """The role arrays of serialize_inputs.py:196-227 from a table-preparation
    dict (ax, ay, freq, KR, KC, tan_phi, dims)."""
    m, n = tables["dims"]
    Mp, Np = grid or _pow2(m), grid or _pow2(n)
    freq, ax, ay = tables["freq"], np.asarray(tables["ax"]), np.asarray(tables["ay"])
    c0, c1 = _freq_column(freq, 0), _freq_column(freq, 1)
    dx, dy = ax.mean(), ay.mean()
    dn = np.hypot(dx, dy)
    tan_phi = np.asarray(tables["tan_phi"])
    order = np.argsort(tan_phi)
    KR, KC = np.asarray(tables["KR"], float), np.asarray(tables["KC"], float)
    oob_r = float(KR.max() + (KR.max() - KR.min()) + 1.0)
    oob_c = float(KC.max() + (KC.max() - KC.min()) + 1.0)
    krp = np.full(Np, oob_r, np.float32)
    krp[:n] = KR
    kcp = np.full(Mp, oob_c, np.float32)
    kcp[:m] = KC
    hr, hc = np.zeros(Np), np.zeros(Mp)
    hr[:n], hc[:m] = np.hamming(n), np.hamming(m)
    return {"f0": c0.astype(np.float32), "df": (c1 - c0).astype(np.float32),
            "pr": (ax * (dx / dn) + ay * (dy / dn)).astype(np.float32),
            "tans": tan_phi[order].astype(np.float32),
            "invorder": np.argsort(order).astype(np.int32), "krgrid": krp, "kcgrid": kcp,
            "hamr": np.clip(np.round(hr * 32768.0), 0, 32767).astype(np.int16),
            "hamc": np.clip(np.round(hc * 32768.0), 0, 32767).astype(np.int16)}


def write_staging(out_dir, signal, tables: dict, grid: int = 0, scene: str = "scene",
                  device: int = 0, codes=None) -> dict:
    """serialize_inputs.serialize from the table-preparation step on: quantise
    the complex64 signal to int16 I/Q (on the GPU), write the role binaries and
    layout.json; returns the layout.  ``codes`` = (int16 [m, n, 2], exponent)
    supplies an already quantised signal instead (signal is then unused)."""
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    m, n = tables["dims"]
    Mp, Np = grid or _pow2(m), grid or _pow2(n)
    if codes is None:
        signal = np.ascontiguousarray(signal, np.complex64)
        if signal.shape != (m, n):
            raise PfaError(-1, f"signal shape {signal.shape} != tables dims {(m, n)}")
        codes, exp = quantize_iq16(signal, device)
    else:
        codes, exp = np.ascontiguousarray(codes[0], np.int16), int(codes[1])
        if codes.shape != (m, n, 2):
            raise PfaError(-1, f"codes shape {codes.shape} != {(m, n, 2)}")
    sig_bytes = codes.tobytes()
    (out / "sig.bin").write_bytes(sig_bytes)
    roles = staging_tables(tables, grid)
    sizes, crcs = {"sig": len(sig_bytes), "out_uint16": Mp * Np * 2}, {"sig": crc32(sig_bytes)}
    for name, arr in roles.items():
        b = arr.tobytes()
        (out / f"{name}.bin").write_bytes(b)
        sizes[name], crcs[name] = len(b), crc32(b)
    np.asarray(tables["KR"], np.float32).tofile(out / "kr.bin")
    np.asarray(tables["KC"], np.float32).tofile(out / "kc.bin")
    np.asarray(tables["tan_phi"], np.float32).tofile(out / "tanphi.bin")
    geo = tables["geo"]
    mu, nu = tables.get("deci", (1, 1))
    layout = {
        "scene": scene, "dims": {"M": m, "N": n}, "fft_len": {"R": Np, "A": Mp},
        "deci": {"pulse": mu, "sample": nu},
        "bfp_input": {"lsb": 2.0 ** exp, "exp": exp},
        "sizes": sizes, "crc32": crcs,
        "geocode": {"dc": geo["dc"], "dr": geo["dr"], "dhat": list(geo["dhat"]),
                    "chat": list(geo["chat"]),
                    "center_ecef": [float(v) for v in tables.get("center_ecef", (0, 0, 0))],
                    "uiax": [float(v) for v in tables.get("uiax", (1, 0, 0))],
                    "uiay": [float(v) for v in tables.get("uiay", (0, 1, 0))]},
    }
    (out / "layout.json").write_text(json.dumps(layout, indent=2))
    return layout


def load_staging(stage_dir, verify_crc: bool = True):
    """Read a staging directory: (sig_iq int16 [m, n, 2] memory-mapped, role
    arrays, layout).  Sizes are checked against layout.json, and the CRC-32 of
    every role when verify_crc is set; a mismatch raises PfaError."""
    d = Path(stage_dir)
    layout = json.loads((d / "layout.json").read_text())
    m, n = layout["dims"]["M"], layout["dims"]["N"]
    lens = {"M": m, "R": layout["fft_len"]["R"], "A": layout["fft_len"]["A"]}

    def read(name, dtype, count):
        p = d / f"{name}.bin"
        want = count * np.dtype(dtype).itemsize
        got = p.stat().st_size
        if got != want or layout["sizes"].get(name, want) != want:
            raise PfaError(-1, f"{p.name}: {got} bytes, layout and dims want {want}")
        a = np.memmap(p, dtype=dtype, mode="r", shape=(count,))
        if verify_crc and name in layout.get("crc32", {}) and crc32(a) != layout["crc32"][name]:
            raise PfaError(-1, f"{p.name}: CRC-32 mismatch")
        return a

    sig = read("sig", np.int16, m * n * 2).reshape(m, n, 2)
    roles = {name: np.asarray(read(name, dt, lens[key])) for name, (dt, key) in ROLES.items()}
    return sig, roles, layout


def form_image_staged(stage_dir, range_sinc=False, az_sinc=False, verify_crc=True,
                      device: int = 0):
    """The fixed datapath (silicon_emulator.form_image after quantisation,
    bit-exact) on a staging directory: uint16 image [range][cross] of the
    fabric grid, and the two passes' row exponents.  The grid must be square
    (Np == Mp), as the fabric's is."""
    sig, r, layout = load_staging(stage_dir, verify_crc)
    m, n = layout["dims"]["M"], layout["dims"]["N"]
    G = layout["fft_len"]["R"]
    if layout["fft_len"]["A"] != G:
        raise PfaError(-2, f"staged grid {layout['fft_len']} is not square")
    tabs = {"f0": r["f0"], "df": r["df"], "pr": r["pr"], "tan_s": r["tans"],
            "inv_order": r["invorder"], "krp": r["krgrid"], "kcp": r["kcgrid"],
            "hamr": r["hamr"].astype(np.int32), "hamc": r["hamc"].astype(np.int32)}
    plan = FixedPlan(tabs, m, n, G, range_sinc, az_sinc, device=device)
    try:
        return plan.form_image(np.ascontiguousarray(sig))
    finally:
        plan.close()
