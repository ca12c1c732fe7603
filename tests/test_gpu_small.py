"""Products that fit one workgroup (N = 2^n <= 2^13) on both of their engines:
Goldilocks over 2^n points of 16-bit digits and the three primes over
2^(n-1) points of 32-bit limbs (k3_small_mul), against the oracle's
schoolbook and NTT products (the reference's self-test draws random sizes
the same way, tests/self_test.cu:64-78).

The three-prime path's carry hands each thread at least two limbs, because
a chunk's carry-out is a function of the (threshold, C, C + 1) form only
when the carry-in (below 2^57 with 32-bit digits) is smaller than the
chunk; the closed forms with every limb 0xFFFFFFFF make every carry ripple
through every chunk, and the products of a few limbs put most threads on
padding."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture
def restore_shapes(pdh, tmp_path):
    path = str(tmp_path / "shapes.txt")
    pdh.tune_save(path)
    yield
    pdh.tune_load(path)


def sizes_for(n, rng, k):
    """k (na, nb) pairs whose transform length is 2^n."""
    out = []
    while len(out) < k:
        total = int(rng.integers(2 ** (n - 2) + 1, 2 ** (n - 1) + 1))   # 2(na + nb) in (2^(n-1), 2^n]
        na = int(rng.integers(1, total))
        nb = total - na
        if nb >= 1:
            out.append((na, nb))
    return out


@pytest.mark.parametrize("eng", [0, 1])
def test_random_sizes_every_n(pdh, restore_shapes, eng):
    rng = np.random.default_rng(1200 + eng)
    for n in range(3, 14):
        assert pdh.tune_engine(n, eng) == eng
        for k, (na, nb) in enumerate(sizes_for(n, rng, 3)):
            assert pdh.ntt_length(na, nb) == 1 << n
            a, b = W.operand(9000 + 7 * n + k, na), W.operand(9500 + 7 * n + k, nb)
            assert np.array_equal(pdh.mul_ntt_hip(a, b), oracle.mul_ntt(a, b)), (eng, n, na, nb)
        a = W.operand(9900 + n, 2 ** (n - 2))
        assert np.array_equal(pdh.sqr_ntt_hip(a), oracle.mul_ntt(a, a)), (eng, n, len(a))


@pytest.mark.parametrize("eng", [0, 1])
def test_tiny_and_schoolbook(pdh, restore_shapes, eng):
    for n in range(2, 14):
        pdh.tune_engine(n, eng)
    rng = np.random.default_rng(77 + eng)
    for na in (1, 2, 3, 5, 17):
        for nb in (1, 2, 7, 64):
            a = rng.integers(0, 1 << 32, size=na, dtype=np.uint64).astype(np.uint32)
            b = rng.integers(0, 1 << 32, size=nb, dtype=np.uint64).astype(np.uint32)
            a[-1] |= 1
            b[-1] |= 1
            assert np.array_equal(pdh.mul_ntt_hip(a, b), oracle.mul_schoolbook(a, b)), (eng, na, nb)


@pytest.mark.parametrize("eng", [0, 1])
def test_all_ones_closed_form(pdh, restore_shapes, eng):
    """(2^(32 na) - 1)(2^(32 nb) - 1) = 2^(32(na+nb)) - 2^(32 na) - 2^(32 nb) + 1."""
    for n in range(3, 14):
        pdh.tune_engine(n, eng)
    for na, nb in ((1, 1), (2, 1), (3, 3), (100, 37), (511, 512), (2047, 2048)):
        a = np.full(na, 0xFFFFFFFF, dtype=np.uint32)
        b = np.full(nb, 0xFFFFFFFF, dtype=np.uint32)
        want = (1 << (32 * (na + nb))) - (1 << (32 * na)) - (1 << (32 * nb)) + 1
        got = pdh.mul_ntt_hip(a, b)
        assert int.from_bytes(got.astype("<u4").tobytes(), "little") == want, (eng, na, nb)
        assert got[-1] != 0


def test_batch_of_small_products(pdh, restore_shapes):
    """A batch of uneven small products (every class N <= 2^13) on the
    three-prime one-workgroup path against the oracle, product by product."""
    import torch
    for n in range(2, 14):
        pdh.tune_engine(n, 1)
    rng = np.random.default_rng(31)
    pairs = []
    for i in range(300):
        na, nb = int(rng.integers(1, 1500)), int(rng.integers(1, 1500))
        pairs.append((W.operand(20000 + 2 * i, na), W.operand(20001 + 2 * i, nb)))
    na = np.array([len(x) for x, _ in pairs], dtype=np.int64)
    nb = np.array([len(y) for _, y in pairs], dtype=np.int64)
    a_off = np.concatenate([[0], np.cumsum(na)[:-1]])
    b_off = np.concatenate([[0], np.cumsum(nb)[:-1]])
    o_off = np.concatenate([[0], np.cumsum(na + nb)[:-1]])
    A = torch.from_numpy(np.concatenate([x for x, _ in pairs]).view(np.int32)).cuda()
    B = torch.from_numpy(np.concatenate([y for _, y in pairs]).view(np.int32)).cuda()
    out = torch.zeros(int((na + nb).sum()), dtype=torch.int32, device="cuda")
    lens = torch.zeros(len(pairs), dtype=torch.int64, device="cuda")
    descs = [(A.data_ptr() + 4 * int(a_off[i]), int(na[i]), B.data_ptr() + 4 * int(b_off[i]), int(nb[i]),
              out.data_ptr() + 4 * int(o_off[i]), lens.data_ptr() + 8 * i) for i in range(len(pairs))]
    pdh.mul_batch_dev(descs)
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint32)
    L = lens.cpu().numpy()
    for i, (x, y) in enumerate(pairs):
        want = oracle.mul_ntt(x, y)
        assert L[i] == len(want), i
        assert np.array_equal(got[o_off[i]:o_off[i] + L[i]], want), i
