"""The BBP kernels' integer residue 16^e mod m (bbp_kernels.hip,
modpow16_mont), restated step for step on Python integers with the same
32-bit wrap-arounds, against pow(16, e, m); the asserts pin the lazy bounds
the kernel relies on (no final correction after a Montgomery square, one
unsigned min after a double). Moduli are the reference's 8k + j,
j in {1, 4, 5, 6} (include/bbp_kernel.h:99-108), below 2^30."""
from __future__ import annotations

import random

M32 = (1 << 32) - 1


def modpow16_mont(e: int, m: int) -> int:
    if e == 0:
        return 1
    s = (m & -m).bit_length() - 1
    assert s <= 2
    mo = m >> s
    E = 4 * e - s
    inv = mo
    for _ in range(4):
        inv = (inv * ((2 - mo * inv) & M32)) & M32
    assert (inv * mo) & M32 == 1
    ninv = (-inv) & M32
    q = int(4294967296.0 * ((1.0 / m) * (1 << s)))
    rr = (1 << 32) - q * mo
    if rr < 0:
        rr += mo
    if rr >= mo:
        rr -= mo
    assert 0 <= rr < mo
    r, two_mo = rr, 2 * mo
    for b in range(E.bit_length() - 1, -1, -1):
        t = r * r
        assert t < (1 << 32) * mo
        mq = ((t & M32) * ninv) & M32
        r = (mq * mo + t) >> 32
        assert r < two_mo
        r = (r << ((E >> b) & 1)) & M32
        d = (r - two_mo) & M32
        r = min(r, d)
        assert r < two_mo
    mq = (r * ninv) & M32
    y = (mq * mo + r) >> 32
    assert y <= mo
    if y >= mo:
        y -= mo
    return y << s


def test_small_moduli_every_exponent():
    for m in range(2, 600):
        if m % 8 not in (1, 4, 5, 6):
            continue
        for e in range(0, 70):
            assert modpow16_mont(e, m) == pow(16, e, m), (e, m)


def test_random_large():
    rng = random.Random(7)
    for _ in range(20000):
        m = 8 * rng.randrange(0, 1 << 27) + rng.choice((1, 4, 5, 6))
        if m < 2 or m >= 1 << 30:
            continue
        e = rng.randrange(0, 1 << 28)
        assert modpow16_mont(e, m) == pow(16, e, m), (e, m)


def test_edges():
    top = (1 << 27) - 1
    for j in (1, 4, 5, 6):
        m = 8 * top + j
        if m < 1 << 30:
            for e in (1, 2, (1 << 28) - 1, 1 << 27):
                assert modpow16_mont(e, m) == pow(16, e, m), (e, m)
