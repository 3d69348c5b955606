"""Index and twiddle identities the LDS FFTs rely on, checked exhaustively on
the CPU for every size the kernels are built for.

The kernels address a butterfly's elements, and the matrix-core stages their
outputs, as one base address plus compile-time offsets; that is only right if
the padded layouts split additively for every lane.  Each test below forms the
position (or twiddle exponent) directly, the way the layout defines it, and
through the decomposition the kernel uses, and requires them to agree.

  FP32 FFT (csrc/pfa_fft_f32.hpp):   padf(i) = i + i / 32, Stockham passes of
      radix min(E, remaining), E = N / NT elements per thread.
  Matrix-core FFT (csrc/pfa_fft_mfma.hpp): pp(p) = p + 4 (p / 32) + 4 (p / 512),
      radix plans {256: 16.16, 512: 32.16, 1024: 32.32, 4096: 16.16.16,
      8192: 2 (FP32 registers, put_r2) then 16.16.16 from LS = 2}, element
      e = k + r N/R of a radix-R stage at word k R + r.
  The FP32 azimuth kernel runs a radix-G first pass in registers (G = the
      thread's output groups) and then Stockham passes from LS = G (fft_rest).
"""
import numpy as np
import pytest


def padf(i):
    return i + (i >> 5)


def cfg(L, emax):
    n = 1 << L
    nt = n // emax if n // emax > 64 else (n // 2 if n // 2 < 64 else 64)
    e = n // nt
    radices, rem = [], n
    while rem > 1:
        r = min(rem, e)
        radices.append(r)
        rem //= r
    return n, nt, radices


F32_CASES = [(L, emax) for emax in (8, 16, 32) for L in range(6, 15)]


@pytest.mark.parametrize("L,emax", F32_CASES)
def test_f32_butterfly_offsets(L, emax):
    """padf(k + r N/R) = padf(k) + r N/R + (r N/R)/32 for k < N/R (loads) and
    padf(o + r LS) = padf(o) + r LS + (r LS)/32 for o = (k/LS) LS R + k % LS
    (stores), every pass."""
    n, _, radices = cfg(L, emax)
    ls = 1
    for R in radices:
        nr = n // R
        k = np.arange(nr)
        for r in range(R):
            assert np.array_equal(padf(k + r * nr), padf(k) + r * nr + (r * nr) // 32)
        o = (k // ls) * ls * R + k % ls
        for r in range(R):
            assert np.array_equal(padf(o + r * ls), padf(o) + r * ls + (r * ls) // 32)
        ls *= R


@pytest.mark.parametrize("L,emax", F32_CASES)
def test_f32_twiddle_lookups_in_table(L, emax):
    """Radix >= 8 passes look up W^(j step) and W^(4 j step): both exponents
    stay inside the N-entry two-level table, and (4a + c) j step is the
    exponent the pass needs for r = 4a + c."""
    n, _, radices = cfg(L, emax)
    ls = 1
    for R in radices:
        if ls > 1 and R >= 8:
            step = n // (ls * R)
            j = np.arange(ls)
            assert (4 * j * step).max() < n // 2
            for r in range(1, R):
                a, c = divmod(r, 4)
                assert np.array_equal(a * (4 * j * step) + c * (j * step), j * r * step)
        ls *= R


def _twiddle_tables(n):
    t = np.arange(n // 64)
    hi = np.exp(-2j * np.pi * 64 * t / n).astype(np.complex64)
    lo = np.exp(-2j * np.pi * np.arange(64) / n).astype(np.complex64)
    return hi, lo


def _f32(x):
    return np.float32(x)


def _cmul(a, b):
    """cmul of csrc/pfa_fft_f32.hpp: (fma(a.x, b.x, -(a.y b.y)), fma(a.x, b.y, a.y b.x)),
    each FMA formed exactly in float64 (float32 products are exact there) and
    rounded once to float32."""
    ax, ay = np.float64(a.real), np.float64(a.imag)
    bx, by = np.float64(b.real), np.float64(b.imag)
    p1 = np.float64(_f32(ay * by))
    p2 = np.float64(_f32(ay * bx))
    return np.complex64(complex(_f32(ax * bx - p1), _f32(ax * by + p2)))


@pytest.mark.parametrize("L", [12, 13])
def test_f32_product_twiddles_error(L):
    """The products w4^a w_c (at most four deep) stay within 8 float32 ulp of
    the exact twiddle for every (pass, j, r) of the radix-16 passes."""
    n, _, radices = cfg(L, 16)
    hi, lo = _twiddle_tables(n)

    def tw(t):
        return _cmul(hi[t >> 6], lo[t & 63])

    worst, ls = 0.0, 1
    for R in radices:
        if ls > 1 and R >= 8:
            step = n // (ls * R)
            for j in range(0, ls, max(1, ls // 64)):
                w1, w4 = tw(j * step), tw(4 * j * step)
                w2 = _cmul(w1, w1)
                wc = [np.complex64(1), w1, w2, _cmul(w2, w1)]
                wa = w4
                for a in range(R // 4):
                    if a >= 2:
                        wa = _cmul(wa, w4)
                    for c in range(4):
                        r = 4 * a + c
                        if r == 0:
                            continue
                        w = wc[c] if a == 0 else (wa if c == 0 else _cmul(wa, wc[c]))
                        exact = np.exp(-2j * np.pi * j * r * step / n)
                        worst = max(worst, abs(complex(w) - exact))
        ls *= R
    assert worst < 8 * 2.0 ** -24, worst


def pp(p):
    return p + ((p >> 5) << 2) + ((p >> 9) << 2)


MPLAN = {8: [16, 16], 9: [32, 16], 10: [32, 32], 12: [16, 16, 16], 13: [16, 16, 16]}
LS0 = {13: 2}                 # radices the caller runs in registers before the put (put_r2)


def _pos0(L, e):
    R = MPLAN[L][0]
    nv = (1 << L) // R
    return (e % nv) * R + e // nv


@pytest.mark.parametrize("L", sorted(MPLAN))
def test_mfma_stage_epilogue(L):
    """Output m = 16 mt + 4 q + i of lane (k, q) in stage S: its next-stage word
    pp(k2 R2 + r2) = pp(p0) + pp((16 mt + i) LS R2) and its twiddle exponent
    (k2 % LS2) r2 N / (LS2 R2) = T0 + (16 mt + i) D, with p0, T0, D the lane's
    own; the exponent stays below N."""
    Rs, n = MPLAN[L], 1 << L
    for S in range(len(Rs) - 1):
        R, R2 = Rs[S], Rs[S + 1]
        ls = LS0.get(L, 1) * (int(np.prod(Rs[:S], dtype=np.int64)) if S else 1)
        nv, nv2, ls2 = n // R, n // R2, ls * R
        assert nv2 % ls2 == 0
        kb, C = nv2 // ls2, n // (ls2 * R2)
        k = np.arange(nv)[:, None]
        q = np.arange(4)[None, :]
        j, h = k % ls, k // ls
        r2 = h // kb
        kq = j + 4 * q * ls
        p0 = ((h % kb) * ls2 + kq) * R2 + r2
        T0, D = kq * r2 * C, ls * r2 * C
        # the lanes look up W^T0, W^D and (R = 32) W^(16 D): all inside the table
        assert T0.max() < n and D.max() < n and (R // 16 == 1 or (16 * D).max() < n)
        obase = (k // ls) * ls * R + j
        for mt in range(R // 16):
            for i in range(4):
                m = 16 * mt + 4 * q + i
                o = obase + m * ls
                k2, rr2 = o % nv2, o // nv2
                T = (k2 % ls2) * rr2 * C
                assert np.array_equal(rr2, np.broadcast_to(r2, rr2.shape))
                assert np.array_equal(pp(k2 * R2 + rr2), pp(p0) + pp((16 * mt + i) * ls * R2))
                assert np.array_equal(T, T0 + (16 * mt + i) * D)
                assert T.max() < n


def _nw_for(L, ncol, bw):
    return min(bw, min(ncol * ((1 << L) // R) // 16 for R in MPLAN[L]))


@pytest.mark.parametrize("L", [L for L in sorted(MPLAN) if LS0.get(L, 1) == 1])
def test_mfma_first_stage_stores(L):
    """pp(pos(e0 + d)) = pp(pos(e0)) + pp(pos(d)) for the azimuth kernel's
    e0 = 4 t, d = 4 g NB + u and FFT-B's e0 = t, d = i NB (plans without a
    register pass; put_r2 forms each position whole)."""
    n = 1 << L
    _, nt, _ = cfg(L, 8)
    nb = max(64, nt)
    G = (n + 4 * nb - 1) // (4 * nb)
    assert 4 * G * nb == n
    t = np.arange(nb)
    for g in range(G):
        for u in range(4):
            d = 4 * g * nb + u
            assert np.array_equal(pp(_pos0(L, 4 * t + d)), pp(_pos0(L, 4 * t)) + pp(_pos0(L, d)))
    nbf = 64 * _nw_for(L, 2, 8)
    assert n % nbf == 0
    t = np.arange(nbf)
    for i in range(n // nbf):
        d = i * nbf
        assert np.array_equal(pp(_pos0(L, t + d)), pp(_pos0(L, t)) + pp(_pos0(L, d)))


# ------------------------------------------------------------------ register first passes
def stockham(x, radices, ls=1, pre_twiddled=False):
    """Stockham passes as the kernels index them, in complex128: pass (R, LS)
    reads elements k + r N/R (k < N/R), multiplies by W_N^((k % LS) r N/(LS R))
    (unless the caller applied it: pre_twiddled, first pass only), runs the
    R-point DFT and writes outputs (k / LS) LS R + k % LS + r LS."""
    n = len(x)
    buf = np.asarray(x, np.complex128)
    for i, R in enumerate(radices):
        nr = n // R
        k = np.arange(nr)
        j = k % ls
        X = np.stack([buf[k + r * nr] for r in range(R)])
        if ls > 1 and not (pre_twiddled and i == 0):
            X = X * np.exp(-2j * np.pi * (j[None, :] * np.arange(R)[:, None] * (n // (ls * R))) / n)
        Y = np.fft.fft(X, axis=0)
        out = np.empty_like(buf)
        o = (k // ls) * ls * R + j
        for r in range(R):
            out[o + r * ls] = Y[r]
        buf, ls = out, ls * R
    return buf


def rest_radices(L, emax, ls):
    """fft_rest: radix min(E, N / LS) from LS on, E = N / NT of Cfg<L, EMAX>."""
    n, nt, _ = cfg(L, emax)
    e = n // nt
    out = []
    while ls < n:
        out.append(min(e, n // ls))
        ls *= out[-1]
    return out


@pytest.mark.parametrize("L", [9, 10, 11, 12, 13, 14])
def test_f32_register_first_pass(L):
    """The azimuth kernel's FP32 FFT: a thread's G output groups hold points
    k + g N/G (k = 4 t + u), so a radix-G pass with LS = 1 runs in registers and
    stores outputs G k + r; fft_rest then continues from LS = G.  The whole is
    the DFT."""
    n = 1 << L
    emax = 16 if L >= 14 else 8                      # az_emax
    _, nt, _ = cfg(L, emax)
    nb = max(64, nt)
    G = (n + 4 * nb - 1) // (4 * nb)
    assert G >= 2 and 4 * G * nb == n
    x = np.random.default_rng(L).standard_normal(n) + 1j * np.random.default_rng(L + 1).standard_normal(n)
    y = stockham(x, [G] + rest_radices(L, 16, G))
    assert np.allclose(y, np.fft.fft(x), rtol=0, atol=1e-9 * np.abs(x).sum())


def test_mfma_put_r2_twiddle():
    """put_r2 at 8192 points: the odd output o of the radix-2 pass gets
    W_32^r, r = o / (N/16), which is the first matrix stage's input twiddle
    W_N^((k % 2) r step) of element o (k = o % (N/16), step = N/32), and the
    even outputs get none; with (a +- b)/2 and 1/16 per matrix stage the
    transform is DFT/N."""
    L, n = 13, 1 << 13
    nv = n // 16
    e = np.arange(n)
    jj, r = (e % nv) % 2, e // nv
    assert np.array_equal(jj, e % 2)
    put = np.where(e % 2 == 1, r, 0) * (n // 32) % n               # exponent put_r2 applies
    stage = (jj * r * (n // 32)) % n                                # the stage's own exponent
    assert np.array_equal(put, stage)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    a, b = x[: n // 2], x[n // 2:]
    line = np.empty(n, np.complex128)
    line[0::2] = (a + b) / 2
    line[1::2] = (a - b) / 2 * np.exp(-2j * np.pi * (np.arange(1, n, 2) // nv) / 32)
    y = stockham(line, MPLAN[L], ls=LS0[L], pre_twiddled=True) / 16 ** len(MPLAN[L])
    assert np.allclose(y, np.fft.fft(x) / n, rtol=0, atol=1e-12 * np.abs(x).sum())
