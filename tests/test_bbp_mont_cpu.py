"""The BBP kernels' integer residue 16^e mod m (bbp_kernels.hip,
modpow16_mont), restated step for step on Python integers with the same
32-bit wrap-arounds, against pow(16, e, m); the asserts pin the lazy bounds
the kernel relies on (no final correction after a Montgomery square, one
unsigned min after a double). Moduli are the reference's 8k + j,
j in {1, 4, 5, 6} (include/bbp_kernel.h:99-108), below 2^30. The same for
the fixed-point path's fraction (frac64_mont_wave), against
floor((16^e mod m) 2^64 / m) on Python integers."""
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


# ---- frac64_mont_wave: the fixed-point fraction floor((16^e mod m) 2^64 / m) ----

def frac64_mont_wave(e: int, m: int, q_off: int = 0, emax: int | None = None, fold: bool = False) -> int:
    """Step for step as the kernel: inverse from (3 mo) xor 2, the Montgomery
    one from a quotient q within one of floor(2^32 / mo) (q_off models the
    single-precision reciprocal's error), the ladder on E = 4e - s + 32 run
    for the wavefront's longest bit count emax, and the two halves
    F_hi = rem1 (-mo^-1), F_lo = rem2 (-mo^-1) mod 2^32."""
    assert e >= 1 and 1 < m < (1 << 30)
    s = (m & -m).bit_length() - 1
    mo = m >> s
    E = 4 * e - s + 32
    inv = ((3 * mo) ^ 2) & M32
    for _ in range(3):
        inv = (inv * ((2 - mo * inv) & M32)) & M32
    assert (inv * mo) & M32 == 1
    ninv = (-inv) & M32
    q = ((1 << 32) // mo + q_off) & M32 if mo >= 2048 else ((1 << 32) // mo) & M32
    r = (-(q * mo)) & M32
    if r >= 1 << 31:                       # (int32_t) r < 0
        r = (r + mo) & M32
    assert r < 2 * mo and r % mo == (1 << 32) % mo
    two_mo = 2 * mo
    bits = (emax if emax is not None else E).bit_length()
    assert E.bit_length() <= bits
    if fold:
        assert mo < 1 << 29
    for b in range(bits - 1, -1, -1):
        assert r < two_mo
        if fold:
            # the doubling folded into the square: r (r << bit) < 8 mo^2 < 2^32 mo
            a = r << ((E >> b) & 1)
            assert a < 1 << 32
            t = r * a
            assert t < mo << 32
            mq = ((t & M32) * ninv) & M32
            r = (mq * mo + t) >> 32
            assert r < two_mo
        else:
            t = r * r
            mq = ((t & M32) * ninv) & M32
            r = (mq * mo + t) >> 32
            assert r < two_mo
            r = (r << ((E >> b) & 1)) & M32
            d = (r - two_mo) & M32
            r = d if d < r else r
            assert r < two_mo
    rem2 = r - mo if r >= mo else r
    mq = (r * ninv) & M32
    rem1 = (mq * mo + r) >> 32
    assert rem1 <= mo
    if rem1 >= mo:
        rem1 -= mo
    return (((rem1 * ninv) & M32) << 32) | ((rem2 * ninv) & M32)


def frac_exact(e: int, m: int) -> int:
    return (pow(16, e, m) << 64) // m


def test_frac64_small_moduli_every_exponent():
    for m in range(2, 700):
        if m % 8 not in (1, 4, 5, 6):
            continue
        for e in range(1, 60):
            assert frac64_mont_wave(e, m) == frac_exact(e, m), (e, m)


def test_frac64_random_with_quotient_error_and_wave_padding():
    """Random terms up to m < 2^30, e < 2^28, with the Montgomery one's
    quotient off by -1, 0 or +1 and the ladder padded to a longer exponent of
    the same wavefront."""
    rng = random.Random(11)
    for _ in range(6000):
        m = 8 * rng.randrange(256, 1 << 27) + rng.choice((1, 4, 5, 6))
        if m >= 1 << 30:
            continue
        e = rng.randrange(1, 1 << 28)
        q_off = rng.choice((-1, 0, 1))
        emax = 4 * (e + rng.randrange(0, 256)) + 32
        assert frac64_mont_wave(e, m, q_off, emax) == frac_exact(e, m), (e, m, q_off)


def test_frac64_fold_form():
    """Steps whose terms all have k <= 2^26 - 2 (every m below 2^29): each
    doubling folded into the next square, up to the largest modulus the
    kernel gives this form, 8 (2^26 - 2) + 6."""
    rng = random.Random(13)
    for m in [8 * ((1 << 26) - 2) + j for j in (1, 4, 5, 6)] + list(range(2, 200)):
        if m % 8 in (1, 4, 5, 6):
            for e in (1, 2, 7, 64, (1 << 25) - 2, (1 << 28) - 1):
                assert frac64_mont_wave(e, m, fold=True) == frac_exact(e, m), (e, m)
    for _ in range(4000):
        m = 8 * rng.randrange(256, (1 << 26) - 1) + rng.choice((1, 4, 5, 6))
        e = rng.randrange(1, 1 << 28)
        emax = 4 * (e + rng.randrange(0, 256)) + 32
        assert frac64_mont_wave(e, m, rng.choice((-1, 0, 1)), emax, fold=True) == frac_exact(e, m), (e, m)


def test_frac64_edges():
    for m in (4, 12, 20, 2045 * 8 + 4, 2048 * 8 + 1, 8 * ((1 << 27) - 1) + 6, (1 << 30) - 3):
        if m % 8 not in (1, 4, 5, 6) or m >= 1 << 30:
            continue
        for e in (1, 2, 3, 7, 8, 9, (1 << 28) - 1):
            assert frac64_mont_wave(e, m) == frac_exact(e, m), (e, m)


def frac64_mont_uniform(e: int, m: int, q_off: int = 0, fold: bool = False) -> int:
    """The positions-in-lanes form: the same set-up, the leading exponent bit
    taken as one doubling of the Montgomery one, then the remaining bits
    left-aligned in a 32-bit word, four per trip and the rest one by one."""
    s = (m & -m).bit_length() - 1
    mo = m >> s
    E = 4 * e - s + 32
    inv = ((3 * mo) ^ 2) & M32
    for _ in range(3):
        inv = (inv * ((2 - mo * inv) & M32)) & M32
    ninv = (-inv) & M32
    q = ((1 << 32) // mo + q_off) & M32 if mo >= 2048 else ((1 << 32) // mo) & M32
    r = (-(q * mo)) & M32
    if r >= 1 << 31:
        r = (r + mo) & M32
    two_mo = 2 * mo

    def sq(r):
        t = r * r
        mq = ((t & M32) * ninv) & M32
        r = (mq * mo + t) >> 32
        assert r < two_mo
        return r

    def dbl(r):
        r = (r << 1) & M32
        d = (r - two_mo) & M32
        r = d if d < r else r
        assert r < two_mo
        return r

    n = E.bit_length() - 1
    assert 5 <= n <= 31
    x = (E << (32 - n)) & M32
    if fold:
        # mo < 2^29: the leading one doubled and reduced, then each bit's
        # doubling folded into its square, r (r << bit) < 8 mo^2 < 2^32 mo
        assert mo < 1 << 29
        r = dbl(r)
        while n > 0:
            assert r < two_mo
            a = r << (x >> 31)
            assert a < 1 << 32
            t = r * a
            assert t < mo << 32
            mq = ((t & M32) * ninv) & M32
            r = (mq * mo + t) >> 32
            assert r < two_mo
            n -= 1
            x = (x << 1) & M32
    else:
        r = dbl(r)
    while n >= 4:
        for i in range(4):
            r = sq(r)
            if (x >> (31 - i)) & 1:
                r = dbl(r)
        n -= 4
        x = (x << 4) & M32
    while n > 0:
        r = sq(r)
        if x >> 31:
            r = dbl(r)
        n -= 1
        x = (x << 1) & M32
    rem2 = r - mo if r >= mo else r
    mq = (r * ninv) & M32
    rem1 = (mq * mo + r) >> 32
    if rem1 >= mo:
        rem1 -= mo
    return (((rem1 * ninv) & M32) << 32) | ((rem2 * ninv) & M32)


def test_frac64_uniform_form():
    rng = random.Random(12)
    for m in range(2, 400):
        if m % 8 in (1, 4, 5, 6):
            for e in range(1, 40):
                assert frac64_mont_uniform(e, m) == frac_exact(e, m), (e, m)
    for _ in range(4000):
        m = 8 * rng.randrange(256, 1 << 27) + rng.choice((1, 4, 5, 6))
        if m < 1 << 30:
            e = rng.randrange(1, 1 << 28)
            assert frac64_mont_uniform(e, m, rng.choice((-1, 0, 1))) == frac_exact(e, m), (e, m)
    # the folded form, up to the largest modulus it takes (8 (2^26 - 2) + 6),
    # with exponents up to the largest position in this kernel (2^28)
    for m in list(range(2, 300)) + [8 * ((1 << 26) - 2) + j for j in (1, 4, 5, 6)]:
        if m % 8 in (1, 4, 5, 6):
            for e in (1, 2, 3, 5, 8, 13, 1000, (1 << 25) - 2, (1 << 28) - 1):
                assert frac64_mont_uniform(e, m, fold=True) == frac_exact(e, m), (e, m)
    for _ in range(4000):
        m = 8 * rng.randrange(256, (1 << 26) - 1) + rng.choice((1, 4, 5, 6))
        e = rng.randrange(1, 1 << 28)
        assert frac64_mont_uniform(e, m, rng.choice((-1, 0, 1)), fold=True) == frac_exact(e, m), (e, m)


# ---- moduli in lanes: 64 terms k per wavefront stepping through sorted positions ----

def moduli_block_sums(keys, j: int, k0: int, dmax: int, r1_high: bool = False) -> list:
    """k_bbp_fx_moduli's k-block k0 .. k0 + 63 of series j over the sorted
    positions - 1 `keys`, lane by lane as the kernel: the Montgomery set-up,
    the tables U0[n] = form of 16^n and U1[n] = form of 16^(16 n) and V = form
    of 16^256, the form h of 2^-s, each lane started at its first position by
    its own ladder (folded below 2^29), a gap n0 + 16 n1 + 256 r by U0[n0],
    U1[n1] and V^r (square-and-multiply), the fraction from rem1 = w and
    rem2 = w U0[8] (frac64 for e = 0), and the sum over the 64
    lanes of each position mod 2^64. Every lazy value's bound is asserted.
    r1_high takes the Montgomery one as 2^32 mod mo + mo, what a quotient
    estimate one too low gives."""
    s = 2 if j == 4 else (1 if j == 6 else 0)
    fold = min(k0 + 63, dmax) <= (1 << 26) - 2
    nv = len(keys)
    first = next(i for i in range(nv) if keys[i] >= k0)
    sums = [0] * nv
    for lane in range(64):
        k = k0 + lane
        mi = 8 * k + j
        if not (k <= dmax and mi != 1):
            continue
        mo = mi >> s
        assert mo < (1 << 29 if fold else 1 << 30)
        inv = ((3 * mo) ^ 2) & M32
        for _ in range(3):
            inv = (inv * ((2 - mo * inv) & M32)) & M32
        ninv = (-inv) & M32
        two_mo = 2 * mo
        r1 = (-(((1 << 32) // mo) * mo)) & M32
        r1 = r1 if r1 < 1 << 31 else (r1 + mo) & M32
        if r1_high:
            r1 += mo
        assert r1 < two_mo and r1 % mo == (1 << 32) % mo

        def mont(a, b):
            assert a < two_mo and b < 1 << 32 and a * b < mo << 32
            t = a * b
            r = (((t & M32) * ninv & M32) * mo + t) >> 32
            assert r < two_mo
            return r

        def dbl(r):
            r <<= 1
            assert r < 1 << 32
            return r - two_mo if r >= two_mo else r

        U0 = [r1, dbl(dbl(dbl(dbl(r1))))]
        for _ in range(14):
            U0.append(mont(U0[-1], U0[1]))
        U1 = [r1, mont(U0[15], U0[1])]
        for _ in range(14):
            U1.append(mont(U1[-1], U1[1]))
        V = mont(U1[15], U1[1])

        def gap_mul(w, gp):
            n0, n1, r = gp & 15, (gp >> 4) & 15, gp >> 8
            w = mont(mont(w, U0[n0]), U1[n1])       # entry 0 is the form of one
            if r:
                p = V
                for b in range(r.bit_length() - 2, -1, -1):
                    p = mont(p, p)
                    if (r >> b) & 1:
                        p = mont(p, V)
                w = mont(w, p)
            return w
        h = r1
        for _ in range(s):
            h = (h + mo) >> 1 if h & 1 else h >> 1

        def ladder(E):                       # form of 2^E, E >= 1
            r = r1
            for b in range(E.bit_length() - 1, -1, -1):
                bit = (E >> b) & 1
                if fold:
                    r = mont(r, r << bit)
                else:
                    r = dbl(mont(r, r)) if bit else mont(r, r)
            return r

        started, w, dprev = False, 0, 0
        for i in range(first, nv):
            di = keys[i]
            if i > first and started:
                w = gap_mul(w, di - dprev)
            dprev = di
            if not started and k <= di:
                e = di - k
                w = h if e == 0 else mont(ladder(4 * e), h)
                started = True
            if started:
                if k == di:
                    F = (1 << 64) // mi
                else:
                    rem1 = w - mo if w >= mo else w
                    rem2 = mont(w, U0[8])
                    rem2 = rem2 - mo if rem2 >= mo else rem2
                    F = (((rem1 * ninv) & M32) << 32) | ((rem2 * ninv) & M32)
                sums[i] = (sums[i] + F) & ((1 << 64) - 1)
    return sums


def moduli_block_direct(keys, j: int, k0: int, dmax: int) -> list:
    out = []
    for d in keys:
        t = 0
        for k in range(k0, min(k0 + 63, dmax, d) + 1):
            m = 8 * k + j
            if m != 1:
                t += (pow(16, d - k, m) << 64) // m
        out.append(t & ((1 << 64) - 1))
    return out


def test_moduli_layout_restated():
    """Small positions (k = 0 and e = 0 terms, m = 4), blocks that start
    inside the group (lanes started late), gaps of one and two nibbles and
    longer ones (the power of 16^256), repeats, and k-blocks on both sides of the folded ladder's bound (2^26 - 2)
    and at the layout's largest position - 1 (2^27 - 2)."""
    cases = [
        (list(range(0, 70)), 0),
        (list(range(0, 70)), 64),
        ([5, 5, 9, 300, 301, 900, 4000, 4000, 4255, 70000, 70256, 70272, 70288 + 3 * 2 ** 20], 0),
        ([5, 5, 9, 300, 301, 900, 4000, 4000, 4255, 70000], 4032),
        ([(1 << 26) - 40 + 37 * i for i in range(12)], (1 << 26) - 64),
        ([(1 << 26) - 40 + 37 * i for i in range(12)], (1 << 26) - 40),
        ([(1 << 26) + 900, (1 << 26) + 1300, (1 << 26) + 1301], 1 << 25),
        ([(1 << 27) - 1000, (1 << 27) - 301, (1 << 27) - 300, (1 << 27) - 2], (1 << 27) - 64),
        ([(1 << 27) - 1000, (1 << 27) - 301, (1 << 27) - 300, (1 << 27) - 2], (1 << 27) - 1024),
    ]
    for keys, k0 in cases:
        keys = sorted(keys)
        dmax = keys[-1]
        for j in (1, 4, 5, 6):
            want = moduli_block_direct(keys, j, k0, dmax)
            for high in (False, True):
                assert moduli_block_sums(keys, j, k0, dmax, high) == want, (keys[:3], j, k0, high)
