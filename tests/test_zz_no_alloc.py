"""The allocation-free multiply (pdh_mul_dev_noalloc) against pdh_mul_dev,
through the C ABI: bit-exact products on two calls in a row (the second
smaller, so stale scratch from the first would show), the empty product,
balanced and lopsided operands up to the 2^20-limb limit, and the limit
itself rejected before any device work."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import workloads as W

pytestmark = pytest.mark.gpu


def run(pdh, fn, a, b):
    import torch
    A = torch.from_numpy(a.view(np.int32)).cuda() if len(a) else torch.zeros(1, dtype=torch.int32, device="cuda")
    B = torch.from_numpy(b.view(np.int32)).cuda() if len(b) else torch.zeros(1, dtype=torch.int32, device="cuda")
    O = torch.full((max(len(a) + len(b), 1),), 0x77777777, dtype=torch.int32, device="cuda")
    L = torch.full((1,), 12345, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    fn(A.data_ptr(), len(a), B.data_ptr(), len(b), O.data_ptr(), L.data_ptr(), 0)
    torch.cuda.synchronize()
    n = int(L.item())
    return O.cpu().numpy().view(np.uint32)[:n].copy()


def test_noalloc_equals_mul_dev(pdh):
    assert pdh.device_available()
    cases = [(1 << 19, 1 << 19), (300_000, 17), (70_000, 50_000), (3000, 5000), (1, 1)]
    for seed, (na, nb) in enumerate(cases):
        a, b = W.operand(900 + seed, na), W.operand(950 + seed, nb)
        print("noalloc", na, nb, flush=True)
        got = run(pdh, pdh.mul_dev_noalloc, a, b)
        want = run(pdh, pdh.mul_dev, a, b)
        assert np.array_equal(got, want), (na, nb)
        if na + nb <= 20_000:
            assert np.array_equal(got, oracle.mul_ntt(a, b))


def test_second_smaller_call(pdh):
    big_a, big_b = W.operand(31, 1 << 19), W.operand(32, (1 << 19) - 11)
    small_a, small_b = W.operand(33, 40_000), W.operand(34, 23_456)
    first = run(pdh, pdh.mul_dev_noalloc, big_a, big_b)
    second = run(pdh, pdh.mul_dev_noalloc, small_a, small_b)
    assert np.array_equal(first, run(pdh, pdh.mul_dev, big_a, big_b))
    assert np.array_equal(second, run(pdh, pdh.mul_dev, small_a, small_b))
    assert np.array_equal(second, oracle.mul_ntt(small_a, small_b))


def test_empty_and_limit(pdh):
    a = W.operand(41, 100)
    assert len(run(pdh, pdh.mul_dev_noalloc, a, np.zeros(0, np.uint32))) == 0
    with pytest.raises(pdh.PdhError) as e:
        pdh.mul_dev_noalloc(1, pdh.NOALLOC_MAX_LIMBS, 1, 1, 1, 1, 0)
    assert e.value.code == pdh.PDH_E_TOO_LARGE
