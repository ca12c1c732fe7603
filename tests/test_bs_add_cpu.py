"""CPU model of the device signed add that merges binary-splitting T values
(bs_kernels.hip: k_bs_add_status, k_bs_add_apply), checked against Python
integers: T = sa*A + sb*B with A = T_L Q_R, B = P_L T_R, the merge rule of
src/chudnovsky.cpp:100.

The model keeps the kernels' structure -- tiles of NT threads x K limbs,
per-segment carry/borrow statuses (GEN, PROP, KILL), the carry into a
segment taken from the nearest lower non-PROP segment, the order of |A| and
|B| from the highest non-PROP tile, GEN/KILL flipped for B - A -- at small
NT and K so that many tile and segment boundaries are crossed. Cases that
would catch a plausible slip: carries and borrows rippling across every
tile, exact cancellation, |A| < |B| by one unit, zero operands."""
from __future__ import annotations

import random

import pytest

KILL, PROP, GEN = 0, 1, 2
MASK = (1 << 32) - 1


def seg_status(x, y, sub):
    if not sub:
        c, allones = 0, True
        for a, b in zip(x, y):
            c += a + b
            allones &= (c & MASK) == MASK
            c >>= 32
        return GEN if c else (PROP if allones else KILL)
    for a, b in zip(reversed(x), reversed(y)):
        if a != b:
            return GEN if a < b else KILL
    return PROP


def nearest_lower(statuses, j):
    for s in reversed(statuses[:j]):
        if s != PROP:
            return s
    return PROP


def model(A, la, sa, B, lb, sb, cap, NT=4, K=3):
    tile = NT * K
    ntiles = (cap + tile - 1) // tile
    la, lb = (la if sa else 0), (lb if sb else 0)
    sub = bool(sa and sb and sa != sb)

    def seg(arr, n, base):
        return [arr[base + k] if base + k < n else 0 for k in range(K)]

    # pass 1: one status per tile
    tstat = []
    for j in range(ntiles):
        segs = [seg_status(seg(A, la, j * tile + t * K), seg(B, lb, j * tile + t * K), sub) for t in range(NT)]
        tstat.append(nearest_lower(segs, NT))
    # pass 2
    T = [0] * cap
    top = 0
    swap = sub and nearest_lower(tstat, ntiles) == GEN
    flip = {GEN: KILL, KILL: GEN, PROP: PROP}
    for j in range(ntiles):
        tin = nearest_lower(tstat, j)
        if swap:
            tin = flip[tin]
        X, lx, Y, ly = (B, lb, A, la) if swap else (A, la, B, lb)
        xs = [seg(X, lx, j * tile + t * K) for t in range(NT)]
        ys = [seg(Y, ly, j * tile + t * K) for t in range(NT)]
        segs = [seg_status(xs[t], ys[t], sub) for t in range(NT)]
        for t in range(NT):
            st = nearest_lower(segs, t)
            if st == PROP:
                st = tin
            c = 1 if st == GEN else 0
            for k in range(K):
                if sub:
                    d = xs[t][k] - ys[t][k] - c
                    r, c = d & MASK, 1 if d < 0 else 0
                else:
                    s = xs[t][k] + ys[t][k] + c
                    r, c = s & MASK, s >> 32
                i = j * tile + t * K + k
                if i < cap:
                    T[i] = r
                    if r:
                        top = max(top, i + 1)
    sign = (sa if sa else sb) if not sub else (sb if swap else sa)
    return T, top, (sign if top else 0)


def limbs(v, n=None):
    out = []
    while v:
        out.append(v & MASK)
        v >>= 32
    if n is not None:
        out += [0] * (n - len(out))
    return out


def value(arr, n):
    return sum(arr[i] << (32 * i) for i in range(n))


def check(a, sa, b, sb, NT=4, K=3):
    """a, b >= 0 magnitudes with signs sa, sb (0 when the magnitude is 0)."""
    A, B = limbs(a), limbs(b)
    cap = max(len(A), len(B)) + 1
    # garbage beyond the lengths must be ignored
    Ag = A + [0xDEADBEEF] * 5
    Bg = B + [0x12345678] * 5
    T, n, s = model(Ag, len(A), sa, Bg, len(B), sb, cap, NT, K)
    want = sa * a + sb * b
    assert n == len(limbs(abs(want)))
    assert s * value(T, n) == want
    assert s == (0 if want == 0 else (1 if want > 0 else -1))


@pytest.mark.parametrize("seed", range(40))
def test_random(seed):
    rng = random.Random(seed)
    for _ in range(20):
        a = rng.getrandbits(rng.randint(0, 32 * 60))
        b = rng.getrandbits(rng.randint(0, 32 * 60))
        sa = rng.choice((-1, 1)) if a else 0
        sb = rng.choice((-1, 1)) if b else 0
        check(a, sa, b, sb, NT=rng.choice((1, 2, 4)), K=rng.choice((1, 2, 3, 5)))


@pytest.mark.parametrize("n", [1, 7, 12, 13, 36, 100])
def test_ripples_across_every_tile(n):
    big = 1 << (32 * n)
    check(big - 1, 1, 1, 1)            # carry through every limb, length + 1
    check(big, 1, 1, -1)               # borrow through every limb
    check(big, -1, big - 1, 1)         # |A| > |B| by one: T = -1
    check(big - 1, 1, big, -1)         # |A| < |B| by one: swapped, T = -1
    check(big - 1, -1, big, 1)         # T = +1
    check(big + 12345, 1, big + 12345, -1)   # exact cancellation: T = 0, sign 0
    check(0, 0, big - 1, -1)           # A = 0
    check(big - 1, 1, 0, 0)            # B = 0
    check(0, 0, 0, 0)


def test_equal_high_tiles_decide_low():
    """|A| and |B| agree on every tile but the lowest, in both orders."""
    rng = random.Random(5)
    hi = rng.getrandbits(32 * 40) << (32 * 8)
    for lo_a, lo_b in [(5, 9), (9, 5), (7, 7)]:
        check(hi + lo_a, 1, hi + lo_b, -1)
        check(hi + lo_a, -1, hi + lo_b, 1)


def test_merge_rule_with_products():
    """T = T_L Q_R + P_L T_R on consecutive Chudnovsky ranges."""
    A_, B_, C = 13591409, 545140134, 10939058860032000

    def leaf(k):
        if k == 0:
            return 1, 1, A_
        p = (6 * k - 5) * (2 * k - 1) * (6 * k - 1)
        return p, k ** 3 * C, (A_ + B_ * k) * p * (-1 if k & 1 else 1)

    def rng_pqt(lo, hi):
        p, q, t = leaf(lo)
        for k in range(lo + 1, hi):
            pk, qk, tk = leaf(k)
            p, q, t = p * pk, q * qk, t * qk + p * tk
        return p, q, t

    for lo, mid, hi in [(0, 1, 2), (0, 64, 128), (64, 128, 192), (128, 192, 200), (3, 40, 77)]:
        pl, ql, tl = rng_pqt(lo, mid)
        pr, qr, tr = rng_pqt(mid, hi)
        sgn = lambda v: (v > 0) - (v < 0)
        check(abs(tl) * qr, sgn(tl), pl * abs(tr), sgn(tr))
