"""Every launchable transform shape (column/row split, column tile, radix)
gives the same, bit-exact product; the autotuner only changes speed."""
from __future__ import annotations

import hashlib

import numpy as np
import pytest

import oracle
import workloads as W

pytestmark = pytest.mark.gpu


def shapes(n):
    out = []
    if n <= 13:
        return [(0, 1, 3, lr) for lr in (3, 4)]
    for L1 in range(max(1, n - 13), min(12, n - 1) + 1):
        for T in (4, 8, 16, 32, 64):
            if T > (1 << (n - L1)) or (((1 << L1) * T) * 9 // 8) * 8 > 147456:
                continue
            for lc in (3, 4):
                for lr in (3, 4):
                    out.append((L1, T, lc, lr))
    return out


@pytest.fixture
def restore_shapes(pdh):
    saved = {n: pdh.tune_get(n) for n in range(2, 25)}
    yield
    for n, (L1, T, lc, lr) in saved.items():
        pdh.tune_set(n, L1, T, lc, lr)


@pytest.mark.parametrize("n", [5, 12, 13, 14, 17, 19])
def test_every_shape_is_exact(pdh, restore_shapes, n):
    N = 1 << n
    a, b = W.operand(500 + n, N // 4), W.operand(600 + n, N // 4 - 1)
    want = oracle.mul_ntt(a, b)
    for sh in shapes(n):
        pdh.tune_set(n, *sh)
        assert np.array_equal(pdh.mul_ntt_hip(a, b), want), (n, sh)
        assert np.array_equal(pdh.sqr_ntt_hip(a), oracle.mul_ntt(a, a)), (n, sh)


def test_every_shape_2pow22(pdh, restore_shapes, golden):
    g = golden["c2_mul"]
    a, b = W.operand(g["seed_a"], g["limbs"]), W.operand(g["seed_b"], g["limbs"])
    for sh in shapes(24):
        pdh.tune_set(24, *sh)
        got = pdh.mul_ntt_hip(a, b)
        assert hashlib.sha256(got.astype("<u4").tobytes()).hexdigest() == g["sha256"], sh


def test_invalid_shape_rejected(pdh, restore_shapes):
    with pytest.raises(pdh.PdhError):
        pdh.tune_set(24, 12, 64, 3, 3)        # tile beyond the LDS
    with pytest.raises(pdh.PdhError):
        pdh.tune_set(20, 5, 8, 3, 3)          # row pass longer than 2^13


def test_autotune_keeps_parity(pdh, restore_shapes):
    pdh.autotune(13, 16, 1)
    for n in (13, 14, 16):
        N = 1 << n
        a, b = W.operand(700 + n, N // 4), W.operand(800 + n, N // 4)
        assert np.array_equal(pdh.mul_ntt_hip(a, b), oracle.mul_ntt(a, b)), n
