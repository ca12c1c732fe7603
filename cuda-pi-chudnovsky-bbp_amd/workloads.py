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
