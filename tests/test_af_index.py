"""The autofocus adjoint's first-output table, checked on the CPU.

k_af_az_adj (csrc/pfa_af.hip) finds, for its sources p0 .. p0 + 3, the outputs
q whose 32-tap windows cover them as the run [first[p0 - 15], first[p0 + 20]),
where first[p] is the first output whose bracket count c(q) reaches p.  The
kernel fills the table without searching: output q >= 1 writes the counts in
(c(q - 1), c(q)], and all threads write the runs up to c(0) (output 0) and
past the last output's count (M2, no output).  Counts are clamped to
M2 + FO - 1 and the table holds p in [-FO, M2 + FO).  These tests build the
table by those rules and compare it with a lower-bound search (numpy
searchsorted on the unclamped counts) at every position a thread looks up.
"""
import numpy as np
import pytest

FO = 32


def kernel_table(c, M2):
    cc = np.minimum(c, M2 + FO - 1)
    first = np.full(M2 + 2 * FO, -1, np.int64)
    pp = np.arange(M2 + 2 * FO) - FO
    first[pp <= cc[0]] = 0
    first[pp > cc[-1]] = M2
    for q in range(1, M2):
        first[cc[q - 1] + 1 + FO: cc[q] + 1 + FO] = q
    return first


def counts(M2, m, c0, rng, step_max=3, plateau_at=None):
    """Non-decreasing bracket counts of M2 outputs over m sources: real outputs
    q < m in [0, m], padded outputs q >= m at m + 64 (beyond every window)."""
    inc = rng.integers(0, step_max + 1, size=m)
    c = np.minimum(c0 + np.cumsum(inc) - inc[0], m)
    if plateau_at is not None:
        c[plateau_at:] = c[plateau_at]
    return np.concatenate([c, np.full(M2 - m, m + 64)]).astype(np.int64)


CASES = [
    # M2, m, first count, largest step, plateau from
    (64, 64, 0, 2, None),
    (256, 200, 0, 3, None),
    (4096, 4096, 0, 2, None),
    (4096, 3900, 5, 2, None),
    (4096, 4096, 700, 1, None),       # the first output's window far into the sources
    (4096, 4000, 0, 1, 1500),         # outputs past the last source: counts stay at m
    (4096, 4096, 0, 40, None),        # sparse outputs: long runs between neighbours
    (1024, 1000, 990, 1, None),       # c(0) near m: nearly every count maps to output 0
]


@pytest.mark.parametrize("M2,m,c0,step_max,plateau", CASES)
def test_first_table_equals_lower_bound(M2, m, c0, step_max, plateau):
    rng = np.random.default_rng(M2 + m + c0)
    c = counts(M2, m, c0, rng, step_max, plateau)
    assert np.all(np.diff(c) >= 0)
    first = kernel_table(c, M2)
    p0 = 4 * np.arange(M2 // 4)
    for p in (p0 - 15, p0 + 20):
        got = first[p + FO]
        assert np.all(got >= 0), "a looked-up entry was never written"
        assert np.array_equal(got, np.searchsorted(c, p, side="left"))


def test_planted_mistake_is_caught():
    """Filling (c(q - 1), c(q)) without its upper end -- the off-by-one a
    half-open loop would make -- disagrees with the lower bound."""
    M2, m = 256, 256
    c = counts(M2, m, 3, np.random.default_rng(1))
    first = kernel_table(c, M2)
    bad = first.copy()
    cc = np.minimum(c, M2 + FO - 1)
    for q in range(1, M2):
        if cc[q] > cc[q - 1]:
            bad[cc[q] + FO] = q + 1 if q + 1 < M2 else M2
    p = 4 * np.arange(M2 // 4) + 20
    assert not np.array_equal(bad[p + FO], np.searchsorted(c, p, side="left"))
