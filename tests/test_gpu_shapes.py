"""Every launchable transform shape (engine, column/row split, column tile,
radix) gives the same, bit-exact product; the autotuner only changes speed.

Products with N = 2^n > 2^13 run on one of two engines: 0, the Goldilocks
four-step over 2^n points (the reference's field and digit size), and 1, the
three-prime four-step over 2^(n-1) points of 32-bit digits (the default),
whose rows run either as radix-8 levels through LDS (lr_row 3) or as
register-resident radix-16 sets (lr_row 4)."""
from __future__ import annotations

import hashlib

import numpy as np
import pytest

import oracle
import workloads as W

pytestmark = pytest.mark.gpu


def lds_words(e):
    return e + (e >> 3)


def shapes(n, eng):
    out = []
    if n <= 13:
        # one workgroup: Goldilocks at radix 8 or 16, or the three primes
        return [(0, 1, 3, lr) for lr in (3, 4)] if eng == 0 else [(0, 1, 3, 3)]
    pts = n - eng
    for L1 in range(1, pts - 2):
        L2 = pts - L1
        if L2 < 3 or L2 > (13 if eng == 0 else 14):
            continue
        for T in ((4, 8, 16, 32, 64) if eng == 0 else (4, 8, 16, 32)):
            if T > (1 << L2):
                continue
            if eng == 0 and lds_words((1 << L1) * T) * 8 > 147456:
                continue
            if eng == 1 and 3 * lds_words((1 << L1) * T) * 4 > 147456:
                continue
            if eng == 0:
                radices = [(lc, lr) for lc in (3, 4) for lr in (3, 4)]
            else:
                radices = [(3, 3)] + ([(3, 4)] if L2 >= 8 else [])
            for lc, lr in radices:
                out.append((L1, T, lc, lr))
    return out


@pytest.fixture
def restore_shapes(pdh, tmp_path):
    """Every shape and whether it was set by hand (a hand-set shape is also
    used for batches) as they were before the test."""
    path = str(tmp_path / "shapes.txt")
    pdh.tune_save(path)
    yield
    pdh.tune_load(path)


CASES = [(5, 0), (5, 1), (12, 0), (12, 1), (13, 0), (13, 1), (14, 0), (14, 1), (15, 1), (17, 0), (17, 1), (19, 0),
         (19, 1)]


@pytest.mark.parametrize("n,eng", CASES)
def test_every_shape_is_exact(pdh, restore_shapes, n, eng):
    N = 1 << n
    a, b = W.operand(500 + n, N // 4), W.operand(600 + n, N // 4 - 1)
    want = oracle.mul_ntt(a, b)
    want_sq = oracle.mul_ntt(a, a)
    assert pdh.tune_engine(n, eng) == eng
    for sh in shapes(n, eng):
        pdh.tune_set(n, *sh)
        assert np.array_equal(pdh.mul_ntt_hip(a, b), want), (n, eng, sh)
        assert np.array_equal(pdh.sqr_ntt_hip(a), want_sq), (n, eng, sh)


@pytest.mark.parametrize("eng", [0, 1])
def test_every_shape_2pow22(pdh, restore_shapes, golden, eng):
    g = golden["c2_mul"]
    a, b = W.operand(g["seed_a"], g["limbs"]), W.operand(g["seed_b"], g["limbs"])
    pdh.tune_engine(24, eng)
    for sh in shapes(24, eng):
        pdh.tune_set(24, *sh)
        got = pdh.mul_ntt_hip(a, b)
        assert hashlib.sha256(got.astype("<u4").tobytes()).hexdigest() == g["sha256"], (eng, sh)


def test_invalid_shape_rejected(pdh, restore_shapes):
    pdh.tune_engine(24, 0)
    with pytest.raises(pdh.PdhError):
        pdh.tune_set(24, 12, 64, 3, 3)        # tile beyond the LDS
    pdh.tune_engine(20, 0)
    with pytest.raises(pdh.PdhError):
        pdh.tune_set(20, 5, 8, 3, 3)          # row pass longer than 2^13
    pdh.tune_engine(24, 1)
    with pytest.raises(pdh.PdhError):
        pdh.tune_set(24, 10, 64, 3, 3)        # three residue tiles beyond the LDS
    with pytest.raises(pdh.PdhError):
        pdh.tune_set(24, 10, 8, 4, 3)         # engine 1 columns are radix 8 only
    with pytest.raises(pdh.PdhError):
        pdh.tune_set(24, 8, 8, 3, 3)          # engine 1 rows longer than 2^14
    with pytest.raises(pdh.PdhError):
        pdh.tune_set(14, 6, 4, 3, 4)          # radix-16 row sets need rows of 2^8 or more
    pdh.tune_engine(13, 1)                    # one workgroup on the three primes
    with pytest.raises(pdh.PdhError):
        pdh.tune_set(13, 0, 1, 3, 4)          # which runs radix-8 levels only
    with pytest.raises(pdh.PdhError):
        pdh.tune_engine(1, 1)                 # no product has N = 2


def test_engines_agree(pdh, restore_shapes):
    rng = np.random.default_rng(31)
    for n in (8, 12, 14, 16, 18, 20, 22):
        N = 1 << n
        na = int(rng.integers(N // 8, N // 4))
        a, b = W.operand(900 + n, na), W.operand(950 + n, N // 2 - na)
        outs = []
        for eng in (0, 1):
            pdh.tune_engine(n, eng)
            outs.append((pdh.mul_ntt_hip(a, b), pdh.sqr_ntt_hip(a)))
        assert np.array_equal(outs[0][0], outs[1][0]), n
        assert np.array_equal(outs[0][1], outs[1][1]), n
        if n <= 18:
            assert np.array_equal(outs[1][0], oracle.mul_ntt(a, b)), n


def test_autotune_keeps_parity(pdh, restore_shapes, tmp_path):
    pdh.autotune(13, 16, 1)
    tuned = {n: pdh.tune_get(n) for n in range(2, 25)}
    path = str(tmp_path / "tuned.txt")
    pdh.tune_save(path)                                 # a later process loads the tuned shapes
    pdh.tune_set(14, *pdh.tune_get(14))
    pdh.tune_set(16, 8, 4, 3, 3)
    pdh.tune_load(path)
    assert {n: pdh.tune_get(n) for n in range(2, 25)} == tuned
    for n in (13, 14, 16):
        N = 1 << n
        a, b = W.operand(700 + n, N // 4), W.operand(800 + n, N // 4)
        assert np.array_equal(pdh.mul_ntt_hip(a, b), oracle.mul_ntt(a, b)), n
