"""CPU-side checks of the C ABI library and the host-side table code (no GPU calls)."""
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import REPO, GOLDEN, load_golden
from oracle import ref_fixed as rf


def _header_symbols():
    text = "".join((REPO / "include").glob("*.h").__next__().read_text())
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pfa_[a-z0-9_]+)\s*\(", text)))


def test_library_builds_and_exports_every_declared_symbol():
    from mi355x_pfa import _lib
    L = _lib.lib()
    syms = _header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(syms) == set(_lib.PROTOS), set(syms) ^ set(_lib.PROTOS)
    assert L.pfa_abi_version() == 1


def test_no_device_is_a_loud_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from mi355x_pfa import _lib
    with pytest.raises(_lib.PfaError):
        _lib.Context(0)


@pytest.mark.parametrize("bits", [16, 18])
def test_c_twiddles_equal_numpy_twiddles(bits):
    from mi355x_pfa import _lib
    for n in (64, 128, 1024, 4096, 8192):
        out = np.zeros(n, np.int32)
        assert _lib.lib().pfa_twiddle_codes(n, bits, out.ctypes.data) == 0
        wr, wi = rf.twiddle_codes(n, bits)
        assert np.array_equal(out[0::2], wr) and np.array_equal(out[1::2], wi), n


def test_twiddle_bad_args_rejected():
    from mi355x_pfa import _lib
    out = np.zeros(16, np.int32)
    assert _lib.lib().pfa_twiddle_codes(12, 16, out.ctypes.data) == -1
    assert _lib.lib().pfa_twiddle_codes(16, 16, None) == -1


def test_product_sinc_table_equals_reference_fixture():
    from mi355x_pfa.tables import sinc_table_q15
    ref = np.load(GOLDEN / "sinc_table_q15.npy")
    assert np.array_equal(sinc_table_q15(), ref)


def test_product_fixed_tables_equal_oracle():
    from mi355x_pfa.tables import fixed_tables
    g = load_golden("fixed_small.npz")
    grid = int(g["grid"])
    m, n = g["sig"].shape
    tables = {k: g["tab_" + k] for k in ("ax", "ay", "freq", "KR", "KC", "tan_phi")}
    mine = fixed_tables(tables, m, n, grid)
    ref = rf.fixed_tables(g["tab_ax"], g["tab_ay"], g["tab_freq"], g["tab_KR"], g["tab_KC"],
                          g["tab_tan_phi"], m, n, grid)
    for k in ("f0", "df", "pr", "tan_s", "inv_order", "krp", "kcp"):
        assert np.array_equal(mine[k], ref[k]), k
    assert np.array_equal(mine["hamr"], rf.hamming_q15(n, grid))
    assert np.array_equal(mine["hamc"], rf.hamming_q15(m, grid))
