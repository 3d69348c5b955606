"""Refusals at the C ABI: every entry point validates its tables and sizes on
the host and returns a status with a message instead of launching a kernel on
inputs it cannot handle.  Each case names the status and the message."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ARG, DIMS, GEOM = -1, -2, -3


def _float_geom(m, n):
    return {"pr": np.ones(m), "order": np.arange(m, dtype=np.int32),
            "tan_s": np.linspace(-0.01, 0.01, m), "KR": np.linspace(1.0, 2.0, n),
            "KC": np.linspace(-1.0, 1.0, m), "geo": {}}


def _plan_float(geom, m, n, sgn=-1, f0=True, flags=0, sinc=None):
    from mi355x_pfa import _lib
    from mi355x_pfa._lib import context, ptr
    ctx = context(0)
    keep = {k: np.ascontiguousarray(v, np.int32 if k == "order" else np.float64)
            for k, v in geom.items() if k != "geo"}
    f = np.ones(m) if f0 else None
    h = C.c_void_p()
    st = _lib.lib().pfa_plan_float(ctx.handle, C.byref(h), m, n, sgn, ptr(keep["pr"]),
                                   ptr(keep["order"]), ptr(keep["tan_s"]), ptr(keep["KR"]),
                                   ptr(keep["KC"]), ptr(f), ptr(f), None, None, ptr(sinc), flags)
    msg = _lib.lib().pfa_last_error(ctx.handle).decode()
    if st == 0:
        _lib.lib().pfa_plan_destroy(h)
    return st, msg


@pytest.mark.parametrize("case,status,text", [
    ("order_repeat", GEOM, "permutation"),
    ("order_range", GEOM, "permutation"),
    ("tan_descending", GEOM, "ascending"),
    ("kr_decreasing", GEOM, "non-decreasing"),
    ("kr_nonpositive", GEOM, "positive"),
    ("kc_decreasing", GEOM, "non-decreasing"),
    ("pr_nan", GEOM, "non-finite"),
    ("pr_zero", GEOM, "zero"),
    ("sgn_zero", ARG, "sgn"),
    ("tiny", DIMS, "at least 33"),
    ("one_pulse", DIMS, "m, n >= 2"),
    ("unknown_flag", ARG, "unknown flag"),
    ("sinc_no_table", ARG, "tap table"),
    ("bf16_and_fp16", ARG, "choose one"),
    ("split_alone", ARG, "FFT_SPLIT"),
    ("bf16_2048", DIMS, "matrix-core"),
])
def test_plan_float_refuses(case, status, text):
    m, n = 100, 90
    g = _float_geom(m, n)
    kw = {}
    if case == "order_repeat":
        g["order"] = np.arange(m, dtype=np.int32)
        g["order"][5] = 4
    elif case == "order_range":
        g["order"] = np.arange(m, dtype=np.int32)
        g["order"][0] = m
    elif case == "tan_descending":
        g["tan_s"] = g["tan_s"][::-1].copy()
    elif case == "kr_decreasing":
        g["KR"] = g["KR"][::-1].copy()
    elif case == "kr_nonpositive":
        g["KR"] = np.linspace(-1.0, 2.0, n)
    elif case == "kc_decreasing":
        g["KC"] = g["KC"][::-1].copy()
    elif case == "pr_nan":
        g["pr"] = np.ones(m)
        g["pr"][3] = np.nan
    elif case == "pr_zero":
        g["pr"] = np.ones(m)
        g["pr"][3] = 0.0
    elif case == "sgn_zero":
        kw["sgn"] = 0
    elif case == "tiny":
        m, n = 20, 90
        g = _float_geom(m, n)
    elif case == "one_pulse":
        m = 1
        g = _float_geom(m, n)
    elif case == "unknown_flag":
        kw["flags"] = 0x100
    elif case == "sinc_no_table":
        kw["flags"] = 0x1
    elif case == "bf16_and_fp16":
        kw["flags"] = 0x2 | 0x4
    elif case == "split_alone":
        kw["flags"] = 0x8
    elif case == "bf16_2048":
        m, n = 2000, 300
        g = _float_geom(m, n)
        kw["flags"] = 0x2
    st, msg = _plan_float(g, m, n, **kw)
    print(f"{case}: {st} {msg}", flush=True)
    assert st == status and text in msg


def test_focus_without_frequencies_and_shape_checks():
    from mi355x_pfa._lib import PfaError
    from mi355x_pfa.pfa import FloatPlan
    m, n = 100, 90
    plan = FloatPlan(m, n, -1, _float_geom(m, n))              # no compact f0 / df
    with pytest.raises(PfaError, match="compact frequencies"):
        plan.focus(np.zeros((m, n), np.complex64))
    with pytest.raises(PfaError, match="signal shape"):
        plan.focus(np.zeros((m, n + 1), np.complex64), np.ones((m, n + 1)))
    plan.close()


def _fixed_tabs(m, n, G):
    return {"f0": np.full(m, 9.45e9, np.float32), "df": np.full(m, 1e5, np.float32),
            "pr": np.ones(m, np.float32), "tan_s": np.linspace(-0.01, 0.01, m).astype(np.float32),
            "inv_order": np.arange(m, dtype=np.int32),
            "krp": np.linspace(60.0, 70.0, G).astype(np.float32),
            "kcp": np.linspace(-1.0, 1.0, G).astype(np.float32),
            "hamr": np.full(G, 100, np.int32), "hamc": np.full(G, 100, np.int32)}


@pytest.mark.parametrize("case,status,text", [
    ("grid_not_pow2", DIMS, "power of two"),
    ("m_above_grid", DIMS, "<= grid"),
    ("inv_order", GEOM, "permutation"),
    ("taper", GEOM, "taper"),
    ("nan_table", GEOM, "non-finite"),
    ("bad_sinc", GEOM, "2^15"),
])
def test_plan_fixed_refuses(case, status, text):
    from mi355x_pfa import _lib
    from mi355x_pfa._lib import context, ptr
    from mi355x_pfa.tables import sinc_table_q15
    m, n, G = 100, 90, 128
    t = _fixed_tabs(m, n, G)
    sinc, flags = None, 0
    if case == "grid_not_pow2":
        G = 100
        t = _fixed_tabs(m, n, G)
    elif case == "m_above_grid":
        m = 200
        t = _fixed_tabs(m, n, G)
    elif case == "inv_order":
        t["inv_order"][7] = 8
    elif case == "taper":
        t["hamr"][3] = 40000
    elif case == "nan_table":
        t["f0"][2] = np.nan
    elif case == "bad_sinc":
        sinc = np.array(sinc_table_q15(), copy=True)       # the library's table is read-only
        sinc[10, 3] += 1
        flags = 0x1
    ctx = context(0)
    h = C.c_void_p()
    st = _lib.lib().pfa_plan_fixed(ctx.handle, C.byref(h), m, n, G, ptr(t["f0"]), ptr(t["df"]),
                                   ptr(t["pr"]), ptr(t["tan_s"]), ptr(t["inv_order"]),
                                   ptr(t["krp"]), ptr(t["kcp"]), ptr(t["hamr"]), ptr(t["hamc"]),
                                   ptr(sinc), flags)
    msg = _lib.lib().pfa_last_error(ctx.handle).decode()
    if st == 0:
        _lib.lib().pfa_plan_destroy(h)
    print(f"{case}: {st} {msg}", flush=True)
    assert st == status and text in msg


def test_stage_and_product_entry_points_refuse():
    from mi355x_pfa import _lib
    from mi355x_pfa._lib import context, ptr
    L, ctx = _lib.lib(), context(0)
    x = np.zeros((4, 96, 2), np.int16)
    y = np.zeros_like(x)
    e = np.zeros(4, np.int32)
    assert L.pfa_fft_pass_bfp(ctx.handle, ptr(x), 4, 96, ptr(y), ptr(e)) == DIMS     # not 2^k
    img = np.zeros(64, np.uint16)
    out = np.zeros(64, np.uint8)
    lut = np.zeros(16, np.uint8)
    assert L.pfa_lut_u16_u8(ctx.handle, ptr(img), 64, ptr(lut), 0, 0, ptr(out)) == ARG
    assert L.pfa_lut_u16_u8(ctx.handle, ptr(img), 64, ptr(lut), 0, 70000, ptr(out)) == ARG
    f = np.zeros(10, np.float32)
    v0, v1 = C.c_float(), C.c_float()
    assert L.pfa_select_f32(ctx.handle, ptr(f), 10, 10, 0, C.byref(v0), C.byref(v1), None) == ARG
    assert L.pfa_select_f32(ctx.handle, ptr(f), 0, 0, 0, C.byref(v0), C.byref(v1), None) == ARG
    assert "ranks" in L.pfa_last_error(ctx.handle).decode()
