"""The board's on-disk input contract (serialize_inputs.py:150-277) on the CPU:
the role arrays against the emulator's scene tables (the same quantities,
silicon_emulator.py:128-140) and the closed-form rounded tapers, a write/read
round trip, and the size and CRC-32 checks that refuse a damaged file."""
import dataclasses
import json

import numpy as np
import pytest

from oracle import ref_fixed as rf


@pytest.fixture(scope="module")
def scene():
    from mi355x_pfa import synth
    from mi355x_pfa.pipeline import prepare_tables
    spec = dataclasses.replace(synth.C1, pulses=301, samples=256, targets=4, seed=21)
    sc = synth.make_scene(spec)
    rd = synth.SceneReader(sc)
    return sc, prepare_tables(rd, rd.cphd_meta)


def test_role_arrays_equal_emulator_tables(scene):
    from mi355x_pfa.staging import staging_tables
    sc, t = scene
    r = staging_tables(t, grid=512)
    e = rf.fixed_tables(t["ax"], t["ay"], np.asarray(t["freq"]), t["KR"], t["KC"], t["tan_phi"],
                        301, 256, 512)
    for a, b in (("f0", "f0"), ("df", "df"), ("tans", "tan_s"), ("krgrid", "krp"),
                 ("kcgrid", "kcp")):
        np.testing.assert_array_equal(r[a], e[b])
    np.testing.assert_array_equal(r["invorder"], e["inv_order"])
    # serialize_inputs uses np.hypot where the emulator uses math.hypot
    assert np.max(np.abs(r["pr"].view(np.int32) - e["pr"].view(np.int32))) <= 1
    assert r["krgrid"].size == 512 and r["kcgrid"].size == 512


def test_rounded_tapers_closed_form(scene):
    from mi355x_pfa.staging import staging_tables
    _, t = scene
    r = staging_tables(t, grid=512)
    k = np.arange(256)
    h = 0.54 - 0.46 * np.cos(2 * np.pi * k / 255)
    assert np.max(np.abs(r["hamr"][:256] - np.minimum(np.round(h * 32768), 32767))) <= 1
    assert not r["hamr"][256:].any() and not r["hamc"][301:].any()
    assert r["hamc"][150] == 32767                     # odd length: centre 1.0, clipped
    assert r["hamr"].dtype == np.int16 and r["hamc"].dtype == np.int16


def test_write_load_round_trip_and_checks(scene, tmp_path):
    from mi355x_pfa._lib import PfaError
    from mi355x_pfa.staging import load_staging, staging_tables, write_staging
    sc, t = scene
    re, im, e = rf.quantize_iq16(sc.sig)
    codes = np.stack([re, im], axis=-1).astype(np.int16)
    lay = write_staging(tmp_path, None, t, grid=512, codes=(codes, e))
    assert lay["dims"] == {"M": 301, "N": 256} and lay["fft_len"] == {"R": 512, "A": 512}
    assert lay["bfp_input"]["exp"] == e
    sig, roles, layout = load_staging(tmp_path)
    np.testing.assert_array_equal(sig, codes)
    for k, v in staging_tables(t, grid=512).items():
        np.testing.assert_array_equal(roles[k], v)
    assert json.loads((tmp_path / "layout.json").read_text()) == layout
    # a flipped byte fails the CRC, a short file the size check
    b = bytearray((tmp_path / "pr.bin").read_bytes())
    b[5] ^= 0x40
    (tmp_path / "pr.bin").write_bytes(bytes(b))
    with pytest.raises(PfaError, match="CRC"):
        load_staging(tmp_path)
    load_staging(tmp_path, verify_crc=False)
    (tmp_path / "kcgrid.bin").write_bytes(b"\0" * 100)
    with pytest.raises(PfaError, match="bytes"):
        load_staging(tmp_path, verify_crc=False)
