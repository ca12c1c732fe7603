"""GPU parity of the ntt-hip multiply against the oracle and golden fixtures.

Mirrors the reference's tests/self_test.cu:64-78 (random sizes vs schoolbook)
and extends it to every transform length the library supports, closed-form
operands, the 60,000-limb and 2^22-limb workloads, squaring, the
device-resident entry points and the error paths.
"""
from __future__ import annotations

import hashlib

import ctypes

import numpy as np
import pytest

import oracle
import workloads as W

pytestmark = pytest.mark.gpu


def sha(limbs) -> str:
    return hashlib.sha256(np.ascontiguousarray(limbs, dtype="<u4").tobytes()).hexdigest()


class XorShift:
    """The reference self-test's generator (tests/self_test.cu:38-43)."""

    def __init__(self, s=0x9E3779B97F4A7C15):
        self.s = s

    def rnd64(self):
        s = self.s
        s ^= (s << 13) & 0xFFFFFFFFFFFFFFFF
        s ^= s >> 7
        s ^= (s << 17) & 0xFFFFFFFFFFFFFFFF
        self.s = s
        return s

    def rnd32(self):
        return self.rnd64() & 0xFFFFFFFF

    def randbig(self, n):
        a = np.array([self.rnd32() for _ in range(n)], dtype=np.uint32)
        nz = np.nonzero(a)[0]
        return a[: nz[-1] + 1] if len(nz) else a[:0]


def test_selftest_random_pairs(pdh):
    """120 pairs, na, nb in [1, 1500]: ntt-hip == schoolbook == ntt-cpu."""
    rng = XorShift()
    before = pdh.launch_count()
    for _ in range(120):
        na, nb = 1 + rng.rnd32() % 1500, 1 + rng.rnd32() % 1500
        a, b = rng.randbig(na), rng.randbig(nb)
        sb = oracle.mul_schoolbook(a, b)
        assert np.array_equal(oracle.mul_ntt(a, b), sb)
        got = pdh.mul_ntt_hip(a, b)
        assert np.array_equal(got, sb), (len(a), len(b))
    assert pdh.launch_count() > before, "the device path did not run"


@pytest.mark.parametrize("n", list(range(2, 23)))
def test_every_transform_length(pdh, n):
    """Balanced and lopsided operands whose transform length is exactly 2^n."""
    N = 1 << n
    total = N // 2                       # 2na + 2nb == N  ->  na + nb == N/2
    cases = [(max(1, total // 2), max(1, total - total // 2))]
    if total >= 3:
        cases.append((1, total - 1))
        cases.append((total - total // 3, total // 3))
    for k, (na, nb) in enumerate(cases):
        a, b = W.operand(1000 + 7 * n + k, na), W.operand(2000 + 7 * n + k, nb)
        assert pdh.ntt_length(na, nb) == N
        assert np.array_equal(pdh.mul_ntt_hip(a, b), oracle.mul_ntt(a, b)), (n, na, nb)


def test_all_ones_closed_form(pdh):
    """(2^(32k) - 1)^2 = 2^(64k) - 2^(32k+1) + 1: maximal carry runs."""
    for k in (1, 2, 3, 31, 1000, 65536, 1 << 20):
        a = np.full(k, 0xFFFFFFFF, dtype=np.uint32)
        want = np.full(2 * k, 0xFFFFFFFF, dtype=np.uint32)
        want[0] = 1
        want[1:k] = 0
        want[k] = 0xFFFFFFFE
        assert np.array_equal(pdh.mul_ntt_hip(a, a), want), k
        assert np.array_equal(pdh.sqr_ntt_hip(a), want), k


def test_powers_of_two_and_small(pdh):
    for ka, kb in [(0, 0), (31, 31), (32, 5), (1000, 2047), (40000, 1)]:
        A, B = 1 << ka, 1 << kb
        got = pdh.mul_ntt_hip(oracle.int_to_limbs(A), oracle.int_to_limbs(B))
        assert oracle.limbs_to_int(got) == A * B
    rs = np.random.default_rng(5)
    for _ in range(50):
        x, y = int(rs.integers(1, 1 << 62)), int(rs.integers(1, 1 << 62))
        got = pdh.mul_ntt_hip(oracle.int_to_limbs(x), oracle.int_to_limbs(y))
        assert oracle.limbs_to_int(got) == x * y


def test_edge_cases(pdh):
    a = W.operand(7, 100)
    assert len(pdh.mul_ntt_hip(a, np.zeros(0, np.uint32))) == 0
    assert len(pdh.mul_ntt_hip(np.zeros(0, np.uint32), a)) == 0
    assert len(pdh.mul_ntt_hip(np.zeros(5, np.uint32), a)) == 0       # zero value, not normalized
    padded = np.concatenate([a, np.zeros(37, np.uint32)])              # leading-zero limbs
    assert np.array_equal(pdh.mul_ntt_hip(padded, a), oracle.mul_ntt(a, a))
    # batch plans hold one transform per product: larger products are refused there
    dev = [pdh.MulDesc(8, 1 << 22, 8, (1 << 22) + 1, 8, 8)]
    plan = ctypes.c_void_p()
    assert pdh.lib().pdh_batch_create((pdh.MulDesc * 1)(*dev), 1, ctypes.byref(plan)) == pdh.PDH_E_TOO_LARGE


def _to_int(limbs) -> int:
    return int.from_bytes(np.ascontiguousarray(limbs, dtype=np.uint32).tobytes(), "little")


def _split_product(pdh, a, b, blk):
    """a*b summed from products of blk-limb pieces, each within one transform."""
    acc = 0
    for i in range(0, len(a), blk):
        for j in range(0, len(b), blk):
            acc += _to_int(pdh.mul_ntt_hip(a[i:i + blk], b[j:j + blk])) << (32 * (i + j))
    return acc


def test_products_beyond_one_transform(pdh):
    """na + nb > 2^23 limbs: the library multiplies 2^22-limb blocks and sums
    them on the device (the reference's seam takes any size,
    src/ntt_cuda.cu:184-185). Checked against a sum over an independent
    2^21-limb split and, for all-ones operands (maximal carries across the
    block seams), the closed form 2^(A+B) - 2^A - 2^B + 1."""
    a = W.operand(901, (1 << 22) + 5000)
    b = W.operand(902, (1 << 22) + 3000)
    got = pdh.mul_ntt_hip(a, b)
    assert len(got) in (len(a) + len(b), len(a) + len(b) - 1) and got[-1] != 0
    assert _to_int(got) == _split_product(pdh, a, b, 1 << 21)
    sq = pdh.sqr_ntt_hip(b)
    assert _to_int(sq) == _split_product(pdh, b, b, 1 << 21)
    n, m = (1 << 22) + 7, (1 << 22) + 1
    ones = pdh.mul_ntt_hip(np.full(n, 0xFFFFFFFF, np.uint32), np.full(m, 0xFFFFFFFF, np.uint32))
    A, B = 32 * n, 32 * m                       # (2^A - 1)(2^B - 1) without a big multiply
    assert _to_int(ones) == (1 << (A + B)) - (1 << A) - (1 << B) + 1


def test_square_matches_multiply(pdh):
    for n in (1, 2, 17, 1000, 4096, 60000, 300001):
        a = W.operand(300 + n, n)
        assert np.array_equal(pdh.sqr_ntt_hip(a), oracle.mul_ntt(a, a)), n


def test_c1_60000_limbs(pdh, golden):
    g = golden["c1_mul"]
    a, b = W.operand(g["seed_a"], g["limbs"]), W.operand(g["seed_b"], g["limbs"])
    got = pdh.mul_ntt_hip(a, b)
    assert len(got) == g["len"] and sha(got) == g["sha256"]
    assert np.array_equal(got, oracle.mul_ntt(a, b))


def _mod(limbs, q):
    """value of a limb array mod q (q < 2^31) without big integers."""
    x = np.asarray(limbs, dtype=np.uint64) % np.uint64(q)
    base = np.uint64((1 << 32) % q)
    pw = np.ones(len(x), dtype=np.uint64)
    # powers of 2^32 mod q by repeated doubling of the prefix
    filled = 1
    while filled < len(x):
        step = min(filled, len(x) - filled)
        pw[filled: filled + step] = pw[:step] * pw[filled - 1] % np.uint64(q) * base % np.uint64(q)
        filled += step
    return int((x * pw % np.uint64(q)).sum() % np.uint64(q))


@pytest.mark.parametrize("key", ["c2_mul", "c2_sqr"])
def test_c2_2pow22_limbs(pdh, golden, key):
    g = golden[key]
    a = W.operand(g["seed_a"], g["limbs"])
    b = a if key == "c2_sqr" else W.operand(g["seed_b"], g["limbs"])
    got = pdh.sqr_ntt_hip(a) if key == "c2_sqr" else pdh.mul_ntt_hip(a, b)
    assert len(got) == g["len"] and sha(got) == g["sha256"]
    for q in (2147483647, 2147483629, 1000000007):
        assert _mod(got, q) == _mod(a, q) * _mod(b, q) % q


def test_device_resident_entry_points(pdh):
    import torch
    a, b = W.operand(11, 50000), W.operand(12, 70001)
    da = torch.from_numpy(a.view(np.int32)).cuda()
    db = torch.from_numpy(b.view(np.int32)).cuda()
    out = torch.zeros(len(a) + len(b), dtype=torch.int32, device="cuda")
    ln = torch.zeros(1, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    pdh.mul_dev(da.data_ptr(), len(a), db.data_ptr(), len(b), out.data_ptr(), ln.data_ptr(), stream)
    torch.cuda.synchronize()
    want = oracle.mul_ntt(a, b)
    n = int(ln.item())
    assert n == len(want)
    assert np.array_equal(out.cpu().numpy().view(np.uint32)[:n], want)
    sq = torch.zeros(2 * len(a), dtype=torch.int32, device="cuda")
    pdh.sqr_dev(da.data_ptr(), len(a), sq.data_ptr(), ln.data_ptr(), stream)
    torch.cuda.synchronize()
    want = oracle.mul_ntt(a, a)
    assert int(ln.item()) == len(want)
    assert np.array_equal(sq.cpu().numpy().view(np.uint32)[: len(want)], want)
