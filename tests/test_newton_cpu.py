"""CPU tests of the exact floor division and square root behind the
reference's div_fast / isqrt slots (src/bignum_div.cpp:170-178, :189-204),
run through pdh_divmod_cb / pdh_isqrt_cb with every multiply done by the
oracle's ntt-cpu, against Python's integer divmod and math.isqrt.

The Newton estimates are within a unit or two, so on ordinary inputs the
exact correction loops rarely run. The test switch pdh_newton_set_perturb(k)
moves every estimate by k units first; the results must not change, and the
correction counter shows the loops ran. The same switch also moves the
decimal conversion's quotient estimates, whose fix-up is checked through
whole Chudnovsky runs against the reference's 3,500-digit table."""
from __future__ import annotations

import ctypes
import math
import random

import numpy as np
import pytest

import oracle
import pdh


@pytest.fixture(scope="module")
def mul_fn():
    return ctypes.cast(oracle.olib().orc_mul_ntt, ctypes.c_void_p)


@pytest.fixture(autouse=True)
def no_perturb():
    pdh.newton_set_perturb(0)
    yield
    pdh.newton_set_perturb(0)


def L(x: int) -> np.ndarray:
    return pdh.int_to_limbs(x)


def I(a) -> int:
    return pdh.limbs_to_int(a)


def div_cases():
    rng = random.Random(7)
    cases = []
    for bits_a, bits_d in [(40, 20), (64, 64), (100, 37), (200, 64), (600, 250), (2000, 1000), (5000, 63),
                           (8000, 4000), (12000, 11990), (3000, 3100)]:
        a = rng.getrandbits(bits_a) | (1 << (bits_a - 1))
        d = rng.getrandbits(bits_d) | (1 << (bits_d - 1))
        cases.append((a, d))
    # divisor divides the dividend, and one below / above such a multiple
    d = rng.getrandbits(1500) | (1 << 1499)
    q = rng.getrandbits(2100) | (1 << 2099)
    cases += [(q * d, d), (q * d - 1, d), (q * d + d - 1, d)]
    # divisors 2^k - 1 and all-ones dividends (maximal borrows)
    for k in (64, 65, 977, 2048):
        cases.append(((1 << 4096) - 1, (1 << k) - 1))
        cases.append(((1 << 3001) + 12345, (1 << k) - 1))
    # a == d, a < d, d == 1
    cases += [(d, d), (d - 1, d), (d, 1)]
    return cases


@pytest.mark.parametrize("k", [0, 1, 3, -1, -3])
def test_divmod_matches_python(mul_fn, k):
    pdh.newton_set_perturb(k)
    before = pdh.newton_corrections()
    for a, d in div_cases():
        q, r = pdh.divmod_hip(L(a), L(d), mul_fn)
        assert (I(q), I(r)) == divmod(a, d), (a.bit_length(), d.bit_length(), k)
    if k:
        assert pdh.newton_corrections() > before, "the correction loop never ran"


def sqrt_cases():
    rng = random.Random(11)
    cases = []
    for bits in (101, 128, 200, 333, 1000, 4001, 9000):
        x = rng.getrandbits(bits // 2) | (1 << (bits // 2 - 1))
        cases += [x * x, x * x - 1, x * x + 2 * x, x * x + x]       # exact squares and both edges
        cases.append(rng.getrandbits(bits) | (1 << (bits - 1)))
    cases += [(1 << 4000) - 1, (1 << 4001) - 1, 1 << 4000, 3, 10005 * 10 ** 700]
    return cases


@pytest.mark.parametrize("k", [0, 1, 3, -1, -3])
def test_isqrt_matches_python(mul_fn, k):
    pdh.newton_set_perturb(k)
    before = pdh.newton_corrections()
    for s in sqrt_cases():
        x = pdh.isqrt_hip(L(s), mul_fn)
        assert I(x) == math.isqrt(s), (s.bit_length(), k)
    if k:
        assert pdh.newton_corrections() > before, "the correction loop never ran"


def test_zero_and_errors(mul_fn):
    q, r = pdh.divmod_hip(L(0), L(12345), mul_fn)
    assert I(q) == 0 and I(r) == 0
    assert I(pdh.isqrt_hip(L(0), mul_fn)) == 0
    with pytest.raises(pdh.PdhError):
        pdh.divmod_hip(L(5), L(0), mul_fn)
    # leading zero limbs are ignored
    a = np.concatenate([L(10 ** 50), np.zeros(3, np.uint32)])
    d = np.concatenate([L(10 ** 20 + 7), np.zeros(2, np.uint32)])
    q, r = pdh.divmod_hip(a, d, mul_fn)
    assert (I(q), I(r)) == divmod(10 ** 50, 10 ** 20 + 7)


def test_capacity_is_reported(mul_fn):
    import ctypes as C
    a, d = L(1 << 300), L(3)
    q = np.zeros(2, np.uint32)
    ql = C.c_size_t()
    rc = pdh.pilib().pdh_divmod_cb(mul_fn, pdh._ptr(a, C.c_uint32), len(a), pdh._ptr(d, C.c_uint32), len(d),
                                   pdh._ptr(q, C.c_uint32), len(q), C.byref(ql), None, 0, None)
    assert rc == pdh.PDH_E_CAPACITY and ql.value == 10


@pytest.mark.parametrize("k", [2, -1])
def test_pi_digits_with_perturbed_estimates(golden, mul_fn, k):
    """Every quotient, root and decimal-level estimate off by k: the digits
    (the final division, the square root, the decimal fix-up) still equal
    the reference's table."""
    from test_pi_cpu import pi_cpu
    pdh.newton_set_perturb(k)
    before = pdh.newton_corrections()
    s = pi_cpu(3500)
    assert s[2:] == golden["pi_dec_3500"]
    assert pdh.newton_corrections() > before
