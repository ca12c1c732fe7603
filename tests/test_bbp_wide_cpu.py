"""CPU restatement of the BBP residue and fraction steps for moduli from 2^30
up to 2^48 (positions up to 2^45, the range where the reference's
bbp_mulmod is exact, include/bbp_kernel.h:63-83), as bbp_kernels.hip runs
them: modpow16_exact, the FP64 square-and-multiply ladder with signed
residues |r| <= 0.54 m, and frac64 = floor(r 2^64 / m) by two 32-bit
quotient steps with the remainder formed mod 2^64.

The FP64 operations are IEEE double: Python floats for products and sums,
fused multiply-adds emulated exactly (the exact rational a b + c, rounded
once). 1/m is taken correctly rounded and, to cover the hardware reciprocal
refined by two Newton steps, also two units in the last place off. Checked
against pow(16, e, m) and floor((16^e mod m) 2^64 / m) on Python integers,
and the ladder's residue bound is asserted at every step."""
from __future__ import annotations

import math
import random
from fractions import Fraction

import numpy as np
import pytest


def fma(a: float, b: float, c: float) -> float:
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def modpow16_exact(e: int, m: float, minv: float, bound=None) -> float:
    """bbp_kernels.hip modpow16_exact."""
    r = 1.0
    if e:
        bit = e.bit_length() - 1
        if bit >= 2:
            t = e >> (bit - 2)
            r = math.ldexp(1.0, 4 * t)
            r = fma(-round(r * minv), m, r)
            bit -= 3
        while bit >= 0:
            p = r * r
            err = fma(r, r, -p)
            q = float(round(p * minv))
            r = fma(-q, m, p) + err
            if bound is not None:
                bound(r)
            if (e >> bit) & 1:
                v = r * 16.0
                r = fma(-float(round(v * minv)), m, v)
                if bound is not None:
                    bound(r)
            bit -= 1
    q = float(math.floor(r * minv))
    r = fma(-q, m, r)
    if r < 0.0:
        r += m
    if r >= m:
        r -= m
    return r


def div_step(r: int, m: int, c32: float):
    """bbp_kernels.hip div_step: floor(r 2^32 / m) and the remainder."""
    q = int(float(r) * c32) & 0xFFFFFFFF
    t = ((r << 32) - q * m) & ((1 << 64) - 1)
    if t >= 1 << 63:
        t -= 1 << 64                     # the wrapped difference read as signed
    if t < 0:
        q -= 1
        t += m
    elif t >= m:
        q += 1
        t -= m
    return q, t


def frac64(r: int, m: int, minv: float) -> int:
    c32 = minv * 4294967296.0
    hi, rem = div_step(r, m, c32)
    lo, rem = div_step(rem, m, c32)
    return (hi << 32) | lo


def bbp_moduli(rng, n, lo_bits=30, hi_bits=48):
    out = []
    for _ in range(n):
        bits = rng.randint(lo_bits, hi_bits)
        k = rng.getrandbits(bits - 3) | (1 << (bits - 4))
        out.append(8 * k + rng.choice((1, 4, 5, 6)))
    return out


@pytest.mark.parametrize("ulps", [0, 2, -2])
def test_ladder_is_exact_with_bounded_residues(ulps):
    rng = random.Random(45 + ulps)
    mods = bbp_moduli(rng, 300) + [(1 << 48) - 2, (1 << 48) - 3, 8 * ((1 << 45) - 1) + 6, (1 << 30) + 1]
    for m in mods:
        assert m < 1 << 48
        mf = float(m)
        minv = 1.0 / mf
        if ulps:
            minv = float(np.nextafter(minv, np.inf if ulps > 0 else 0.0))
            minv = float(np.nextafter(minv, np.inf if ulps > 0 else 0.0))
        for e in (1, 2, 3, 7, rng.getrandbits(20), rng.getrandbits(45), (m - 6) // 8):
            worst = [0.0]

            def bound(r, m=m, worst=worst):
                worst[0] = max(worst[0], abs(r) / m)
                assert r == int(r)
            r = modpow16_exact(e, mf, minv, bound)
            assert int(r) == pow(16, e, m), (e, m, ulps)
            assert worst[0] <= 0.54, (e, m, worst[0])


def test_fraction_is_floor():
    rng = random.Random(9)
    for m in bbp_moduli(rng, 2000) + [(1 << 48) - 2, (1 << 31) + 3]:
        minv = 1.0 / float(m)
        for r in (0, 1, m - 1, rng.randrange(m), rng.randrange(m)):
            assert frac64(r, m, minv) == (r << 64) // m, (r, m)


def test_digit_step_above_2p28():
    """One whole term as the wide kernel forms it, r = 16^(d-k) mod m and its
    64-bit fraction, at positions up to 2^45."""
    rng = random.Random(3)
    for d in ((1 << 28) + 5, 10 ** 9, (1 << 45) - 1):
        for j in (1, 4, 5, 6):
            for k in (0, 1, d // 3, d - 1):
                m = 8 * k + j
                if m == 1:
                    continue
                e = d - k
                r = int(modpow16_exact(e, float(m), 1.0 / float(m)))
                assert r == pow(16, e, m)
                assert frac64(r, m, 1.0 / float(m)) == (r << 64) // m
