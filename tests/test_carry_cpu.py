"""CPU model of the carry scan the multiply kernels run (ntt_common.h,
ntt_kernels.hip): a chunk of KL limbs with carry-in 0 gives digits D and a
carry-out C; with carry-in c its carry-out is f(c) = C + [c >= thr], thr =
2^(32 KL) - D when that fits in 64 bits (limbs 2..KL-1 of D all ones), else
never. Compositions keep that form, each workgroup's chunks compose into one
aggregate, and a workgroup's carry-in is the composition of the aggregates
below it down to the nearest constant one, taken in windows (the look-back
of k_carry_apply). The model runs the same steps on Python integers and is
checked against the carry of the whole number done sequentially, on random
coefficients and on long runs of all-ones limbs that make the look-back
cross several windows."""
from __future__ import annotations

import random

import pytest

NEVER = (1 << 64) - 1
M64 = (1 << 64) - 1


def chunk_fn(D, C, KL):
    ones = all(d == 0xFFFFFFFF for d in D[2:KL])
    lo = D[0] | (D[1] << 32)
    thr = ((1 << 64) - lo) & M64 if (ones and lo != 0) else NEVER
    return (thr, C, C + 1)


def apply(f, c):
    return f[2] if c >= f[0] else f[1]


def compose(later, earlier):
    return (earlier[0], apply(later, earlier[1]), apply(later, earlier[2]))


def constant(f):
    return f[0] == NEVER or f[1] == f[2]


PAD = (NEVER, 0, 0)


def scan(fs):
    out, acc = [], None
    for f in fs:
        acc = f if acc is None else compose(f, acc)
        out.append(acc)
    return out


def lookback(agg, blk, window):
    """carry_in_lookback of ntt_kernels.hip, windows of `window` aggregates."""
    prev = agg[blk - 1]
    if constant(prev):
        return prev[1]
    acc, hi = None, blk
    while True:
        lo = max(0, hi - window)
        Fw = scan(agg[lo:hi])[-1]
        acc = Fw if acc is None else compose(acc, Fw)
        if lo == 0 or constant(Fw):
            return apply(acc, 0)
        hi = lo


def model_carry(coeffs, KL, per_block, window):
    """coeffs: base-2^32 coefficients (each < 2^86), chunks of KL limbs. Returns the limbs."""
    n = len(coeffs)
    nch = (n + KL - 1) // KL
    Ds, Cs = [], []
    for ch in range(nch):
        carry, D = 0, []
        for q in range(KL):
            i = ch * KL + q
            v = coeffs[i] if i < n else 0
            s = (carry & 0xFFFFFFFF) + (v & 0xFFFFFFFF)
            D.append(s & 0xFFFFFFFF)
            carry = (s >> 32) + (carry >> 32) + (v >> 32)
        Ds.append(D)
        Cs.append(carry)
    fns = [chunk_fn(D, C, KL) for D, C in zip(Ds, Cs)]
    nblk = (nch + per_block - 1) // per_block
    agg = []
    for b in range(nblk):
        part = fns[b * per_block:(b + 1) * per_block]
        agg.append(scan(part)[-1])
    out = []
    for b in range(nblk):
        cin = 0 if b == 0 else lookback(agg, b, window)
        for ch in range(b * per_block, min(nch, (b + 1) * per_block)):
            D = Ds[ch]
            v = sum(d << (32 * q) for q, d in enumerate(D)) + cin
            out += [(v >> (32 * q)) & 0xFFFFFFFF for q in range(KL)]
            cin = apply(fns[ch], cin)
    return out[:n]


def reference(coeffs):
    total = sum(c << (32 * i) for i, c in enumerate(coeffs))
    return [(total >> (32 * i)) & 0xFFFFFFFF for i in range(len(coeffs))]


@pytest.mark.parametrize("KL,per_block,window", [(16, 4, 3), (2, 8, 4), (4, 2, 2), (16, 1, 5)])
def test_random_coefficients(KL, per_block, window):
    rng = random.Random(KL * 100 + per_block)
    coeffs = [rng.getrandbits(86) for _ in range(700)] + [0] * 5
    assert model_carry(coeffs, KL, per_block, window) == reference(coeffs)


@pytest.mark.parametrize("KL,per_block,window", [(16, 4, 3), (2, 8, 4), (4, 2, 2)])
def test_all_ones_runs_cross_windows(KL, per_block, window):
    """(2^(32n) - 1)^2 style digits: long all-ones runs make every aggregate
    non-constant, so the look-back walks several windows down to block 0;
    a carry generated at the bottom ripples to the top."""
    n = 640
    coeffs = [0xFFFFFFFF] * n
    coeffs[0] = (1 << 40) + 5                  # a carry out of the lowest chunk
    assert model_carry(coeffs, KL, per_block, window) == reference(coeffs)
    # and a run that ends in the middle: the look-back stops at its first constant window
    coeffs = [rng_v for rng_v in random.Random(3).choices([0, 0xFFFFFFFF], k=n)]
    coeffs[:300] = [random.Random(5).getrandbits(40) for _ in range(300)]
    coeffs[300:] = [0xFFFFFFFF] * (n - 300)
    coeffs[310] = 0xFFFFFFFF + (1 << 33)
    assert model_carry(coeffs, KL, per_block, window) == reference(coeffs)


def test_lookback_stops_at_a_constant_aggregate():
    agg = [(NEVER, 7, 8), (5, 1, 2), (NEVER, 3, 3), (9, 4, 5), (2, 0, 1)]
    # block 5: agg[4] not constant -> compose agg[4], agg[3] onto constant agg[2] (value 3)
    want = apply(agg[4], apply(agg[3], 3))
    for w in (1, 2, 3, 5):
        assert lookback(agg, 5, w) == want


def test_lookback_equals_sequential_application():
    """Random aggregates (mostly non-constant, small thresholds so both
    branches occur): the windowed look-back equals applying agg[0..blk-1] in
    order to carry-in 0, for every block and window size."""
    rng = random.Random(17)
    for trial in range(200):
        n = rng.randint(1, 40)
        agg = []
        for _ in range(n):
            v0 = rng.randint(0, 5)
            if rng.random() < 0.1:
                agg.append((NEVER, v0, v0 + 1))
            else:
                agg.append((rng.randint(0, 6), v0, v0 + rng.randint(0, 1)))
        for blk in range(1, n + 1):
            c = 0
            for f in agg[:blk]:
                c = apply(f, c)
            for w in range(1, 8):
                assert lookback(agg, blk, w) == c, (trial, blk, w)
