"""Seeded synthetic inputs for the multiply and BBP workloads.

All generators are deterministic functions of their seed so the GPU runs, the
CPU baseline and the committed golden fixtures see identical inputs.

* ``operand(seed, n)``: n uniform random uint32 limbs, top limb forced
  nonzero (a normalized magnitude of exactly n limbs).
* ``batch_pairs(seed, count, lo, hi)``: operand pairs with limb counts drawn
  independently log-uniform over [lo, hi].
* ``sampled_positions(seed, count, lo, hi)``: BBP positions uniform in
  [lo, hi].
"""
from __future__ import annotations

import numpy as np

# Workload shapes (limb counts / positions).
C1_LIMBS = 60_000                 # plain multiply, the reference's CPU-runnable case
C2_LIMBS = 1 << 22                # large single multiply (N = 2^24) and square
C3_PAIRS = 4096                   # batched uneven multiplies
C3_LO, C3_HI = 64, 65_536
C4_RANGE = (1, 65_536)            # BBP contiguous positions
C4_SAMPLED = 1 << 20              # BBP sampled positions in [1e6, 1e8]
C4_LO, C4_HI = 10**6, 10**8


def operand(seed: int, n: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 1 << 32, size=n, dtype=np.uint32)
    if n and a[-1] == 0:
        a[-1] = 1
    return a


def batch_sizes(seed: int, count: int, lo: int, hi: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    u = rng.random((count, 2))
    return np.floor(lo * (hi / lo) ** u).astype(np.int64).clip(lo, hi)


def batch_pairs(seed: int, count: int, lo: int = C3_LO, hi: int = C3_HI):
    sizes = batch_sizes(seed, count, lo, hi)
    return [(operand(seed * 1_000_003 + 2 * i, int(na)), operand(seed * 1_000_003 + 2 * i + 1, int(nb)))
            for i, (na, nb) in enumerate(sizes)]


def sampled_positions(seed: int, count: int, lo: int = C4_LO, hi: int = C4_HI) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.integers(lo, hi + 1, size=count, dtype=np.uint64)


def bbp_mulmod_count(positions) -> int:
    """Exact number of bbp_mulmod calls the reference makes for these 1-based
    positions: per series j in {1,4,5,6} and k = 0..d (d = pos - 1),
    bbp_modpow16(d - k, 8k + j) squares once per bit of the exponent and
    multiplies once per set bit (include/bbp_kernel.h:86-96); the j = 1,
    k = 0 term has modulus 1 and makes none. Closed form over the exponent
    range [0, d] per bit position."""
    d = np.atleast_1d(np.asarray(positions, dtype=np.uint64)) - np.uint64(1)
    x = d + np.uint64(1)                          # exponents e in [0, x)
    total = 0
    for b in range(63):
        period, half = np.uint64(1 << (b + 1)), np.uint64(1 << b)
        full, rem = x // period, x % period
        ones = full * half + np.where(rem > half, rem - half, np.uint64(0))   # popcount prefix
        lo = np.uint64(1 << b)                    # e in [2^b, 2^(b+1)) has bit length b + 1
        hi = np.uint64(1 << (b + 1))
        span = np.clip(x, lo, hi) - lo
        total += int(ones.sum(dtype=np.uint64)) + (b + 1) * int(span.sum(dtype=np.uint64))
    bits = np.zeros_like(d)
    ones_d = np.zeros_like(d)
    for b in range(64):
        sh = d >> np.uint64(b)
        bits += (sh > 0).astype(np.uint64)
        ones_d += sh & np.uint64(1)
    return 4 * total - int((bits + ones_d).sum(dtype=np.uint64))


def bs_level_sizes(terms: int, depths=range(8, 15)):
    """Operand limb counts of the binary-splitting products that build the
    tree nodes at the given depths for a Chudnovsky run of `terms` terms
    (src/chudnovsky.cpp:64-102; depth 0 is the root, a node [a, b) splits at
    (a + b) / 2). Each node multiplies its children: Q_L Q_R, T_L Q_R,
    P_L T_R and P_L P_R. Child sizes are the sums of their leaves' bit
    lengths, P(k) = (6k-5)(2k-1)(6k-1), Q(k) = k^3 C^3/24, T(k) = (A + Bk) P(k)."""
    k = np.arange(terms, dtype=np.float64)
    kk = np.maximum(k, 1.0)
    p_bits = np.where(k == 0, 1.0, np.log2(6 * kk - 5) + np.log2(2 * kk - 1) + np.log2(6 * kk - 1))
    q_bits = np.where(k == 0, 1.0, 3 * np.log2(kk) + np.log2(10939058860032000.0))
    t_bits = p_bits + np.log2(13591409.0 + 545140134.0 * k)
    cum = {n: np.concatenate([[0.0], np.cumsum(b)]) for n, b in (("P", p_bits), ("Q", q_bits), ("T", t_bits))}

    def limbs(kind, a, b):
        return max(1, int(np.ceil((cum[kind][b] - cum[kind][a]) / 32.0)))

    out = []
    for d in depths:
        edges = [0, terms]
        for _ in range(d):
            edges = sorted(set(edges + [(x + y) // 2 for x, y in zip(edges[:-1], edges[1:])]))
        for a, b in zip(edges[:-1], edges[1:]):
            m = (a + b) // 2
            out += [(limbs("Q", a, m), limbs("Q", m, b)), (limbs("T", a, m), limbs("Q", m, b)),
                    (limbs("P", a, m), limbs("T", m, b)), (limbs("P", a, m), limbs("P", m, b))]
    return out


def bs_level_pairs(seed: int, terms: int, depths=range(8, 15)):
    """Seeded random operands (top limbs nonzero) with the limb counts of
    bs_level_sizes, drawn from one generator."""
    sizes = bs_level_sizes(terms, depths)
    rng = np.random.default_rng(seed)
    flat = rng.integers(1, 1 << 32, size=sum(a + b for a, b in sizes), dtype=np.uint32)
    out, o = [], 0
    for na, nb in sizes:
        out.append((flat[o:o + na], flat[o + na:o + na + nb]))
        o += na + nb
    return out
