"""CPU checks of the error bound behind the fixed-point BBP path.

The device computes each series exactly to 64 fractional bits,
F_j = sum floor(2^64 r_k / m_k) + sum_i floor(2^(64-4i) / m_i) (mod 2^64),
and accepts T = 4F1 - 2F4 - F5 - F6 as the reference's digit only when T is
farther than 2 Mu(d) from a digit boundary, Mu(d) = 24584 (d+1) + 262144.
These tests measure, against the oracle's restatement of the reference's
FP64 fold (include/bbp_kernel.h:99-122), how far the reference's per-series
value really is from the exact one, and check that decided digits equal the
reference's digits."""
from __future__ import annotations

from fractions import Fraction

import pytest

import oracle

M64 = 1 << 64


def fx_series(j: int, d: int) -> int:
    f = 0
    for k in range(d + 1):
        m = 8 * k + j
        f += (pow(16, d - k, m) << 64) // m
    for i in range(1, 15):
        f += (1 << (64 - 4 * i)) // (8 * (d + i) + j)
    return f % M64


def per_series_bound(d: int) -> int:
    # head 3*2^10 per term, tail below 15*2^11, floors below d+15 units
    return 3072 * (d + 1) + 30720 + (d + 15)


def circ(a: Fraction, b: int) -> Fraction:
    x = (a - b) % M64
    return min(x, M64 - x)


@pytest.mark.parametrize("d", list(range(0, 40)) + [63, 64, 255, 1000, 4095, 20000])
def test_reference_fold_within_bound(d):
    for j in (1, 4, 5, 6):
        s = oracle.olib().orc_bbp_series(j, d)
        dev = circ(Fraction(s) * M64, fx_series(j, d))
        assert dev <= per_series_bound(d), (j, d, float(dev))


def test_bound_is_not_vacuous():
    """The measured deviation grows with d but stays well inside the bound."""
    d = 20000
    worst = max(circ(Fraction(oracle.olib().orc_bbp_series(j, d)) * M64, fx_series(j, d)) for j in (1, 4, 5, 6))
    assert 0 < worst < per_series_bound(d) / 4


def fx_digit(pos: int):
    d = pos - 1
    F = [fx_series(j, d) for j in (1, 4, 5, 6)]
    T = (4 * F[0] - 2 * F[1] - F[2] - F[3]) % M64
    cell = 1 << 60
    x = T & (cell - 1)
    dist = min(x, cell - x)
    mu = 24584 * (d + 1) + 262144
    return (T >> 60) if dist > 2 * mu else None


def test_decided_digits_equal_reference(golden):
    positions = list(range(1, 401)) + [1000, 4096, 10000, 65536]
    want = oracle.bbp_values_positions(positions)
    undecided = 0
    for p, w in zip(positions, want):
        got = fx_digit(p)
        if got is None:
            undecided += 1
        else:
            assert got == w, p
    assert undecided == 0
    # the reference's 64 hex digits of pi
    assert "".join("0123456789ABCDEF"[fx_digit(p)] for p in range(1, 65)) == golden["pi_hex_64"]


def test_workload_mulmod_count_matches_oracle():
    """bench.py's mulmod count (package code) equals the oracle's closed form
    and its brute-force count of the reference's bbp_mulmod calls."""
    import numpy as np
    import workloads as W
    for pos in (1, 2, 3, 17, 64, 1000, 4097):
        assert W.bbp_mulmod_count([pos]) == oracle.bbp_mulmod_count_brute(pos) == oracle.bbp_mulmod_count(pos)
    ps = np.array([5, 99, 65536, 10**6, 10**8, 1 << 28], dtype=np.uint64)
    assert W.bbp_mulmod_count(ps) == sum(oracle.bbp_mulmod_count(int(p)) for p in ps)


# ---- digit blocks (k_bbp_fx_block): one evaluation decides G positions ----

def block_digits(dmax: int) -> int:
    """The host's G: the largest G <= 8 with 16^(G-1) e1(dmax) <= 2^46."""
    e1 = 4 * (dmax + 15) + 2
    g = 1
    while g < 8 and (e1 << (4 * g)) <= (1 << 46):
        g += 1
    return g


def fx_block(anchor: int, G: int):
    """Digits of positions anchor .. anchor + G - 1 from the anchor's T, None
    where the margin 16^i e1 + Mu(d + i) (2 Mu(d) for i = 0) is not cleared."""
    d = anchor - 1
    F = [fx_series(j, d) for j in (1, 4, 5, 6)]
    T = (4 * F[0] - 2 * F[1] - F[2] - F[3]) % M64
    e1 = 4 * (d + 15) + 2
    cell = 1 << 60
    out = []
    for i in range(G):
        Ti = (T << (4 * i)) % M64
        x = Ti & (cell - 1)
        dist = min(x, cell - x)
        margin = 2 * (24584 * (d + 1) + 262144) if i == 0 else (e1 << (4 * i)) + 24584 * (d + i + 1) + 262144
        assert margin < cell // 2
        out.append(Ti >> 60 if dist > margin else None)
    return out


def test_block_digits_choice():
    assert block_digits(65535) == 7 and block_digits(10**8 - 1) == 5 and block_digits(0) == 8
    assert block_digits((1 << 28) - 1) == 4


@pytest.mark.parametrize("start,count", [(1, 700), (4090, 140), (20000, 40)])
def test_block_digits_equal_reference(start, count):
    """Every digit a block decides equals the reference's digit at that
    position; nearly all are decided."""
    G = block_digits(start + count - 2)
    got = []
    for a in range(start, start + count, G):
        got += fx_block(a, G)
    got = got[:count]
    want = oracle.bbp_values_positions(list(range(start, start + count)))
    decided = [i for i, g in enumerate(got) if g is not None]
    assert len(decided) >= count - 2
    for i in decided:
        assert got[i] == want[i], start + i

