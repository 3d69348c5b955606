"""The fixed datapath on a staging directory (the board's DDR role files,
serialize_inputs.py:150-277) is bit-exact against the oracle's form_image
(silicon_emulator.py:216-253) run on the same staged tables, with the board's
rounded tapers; the GPU quantiser writes the same codes as the oracle's."""
import dataclasses

import numpy as np
import pytest

from oracle import ref_fixed as rf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def staged(tmp_path_factory):
    from mi355x_pfa import synth
    from mi355x_pfa.pipeline import prepare_tables
    from mi355x_pfa.staging import write_staging
    spec = dataclasses.replace(synth.C1, pulses=301, samples=256, targets=6, seed=22)
    sc = synth.make_scene(spec)
    rd = synth.SceneReader(sc)
    t = prepare_tables(rd, rd.cphd_meta)
    d = tmp_path_factory.mktemp("stage")
    write_staging(d, sc.sig, t, grid=512)
    return sc, d


def test_staged_codes_equal_oracle_quantiser(staged):
    from mi355x_pfa.staging import load_staging
    sc, d = staged
    sig, _, layout = load_staging(d)
    re, im, e = rf.quantize_iq16(sc.sig)
    assert layout["bfp_input"]["exp"] == e
    np.testing.assert_array_equal(sig[..., 0], re)
    np.testing.assert_array_equal(sig[..., 1], im)


@pytest.mark.parametrize("range_sinc,az_sinc", [(False, False), (True, True)])
def test_form_image_staged_bit_exact(staged, range_sinc, az_sinc):
    from mi355x_pfa.staging import form_image_staged, load_staging
    sc, d = staged
    print(f"staged form_image sinc={range_sinc},{az_sinc}", flush=True)
    img, ea, eb = form_image_staged(d, range_sinc, az_sinc)
    sig, r, _ = load_staging(d)
    tabs = {"f0": r["f0"], "df": r["df"], "pr": r["pr"], "tan_s": r["tans"],
            "inv_order": r["invorder"].astype(np.int64), "krp": r["krgrid"], "kcp": r["kcgrid"]}
    ref, rea, reb = rf.form_image_fixed(sig[..., 0].astype(np.int64), sig[..., 1].astype(np.int64),
                                        tabs, 512, range_sinc, az_sinc,
                                        hr=r["hamr"].astype(np.int64), hc=r["hamc"].astype(np.int64))
    assert img.shape == (512, 512) and img.dtype == np.uint16
    np.testing.assert_array_equal(img, ref)
    np.testing.assert_array_equal(ea, rea)
    np.testing.assert_array_equal(eb, reb)
    assert img.max() > 0
