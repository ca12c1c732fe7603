"""The exact floor division and square root on the device (pdh_divmod,
pdh_isqrt and their device-buffer forms), the reference's div_fast / isqrt
slots (src/bignum_div.cpp:170-178, :189-204), against Python's divmod and
math.isqrt. The test switch pdh_newton_set_perturb moves every estimate off
by k units so the exact corrections (quotient steps, square-root steps, the
decimal conversion's fix-up) run; results must not change."""
from __future__ import annotations

import hashlib
import math
import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def no_perturb(pdh):
    pdh.newton_set_perturb(0)
    yield
    pdh.newton_set_perturb(0)


def cases():
    rng = random.Random(5)
    out = []
    for bits_a, bits_d in [(90, 40), (300, 130), (5000, 2000), (70000, 30000), (400000, 150000),
                           (2_000_000, 900_000), (1_000_000, 999_000), (300, 900)]:
        a = rng.getrandbits(bits_a) | (1 << (bits_a - 1))
        d = rng.getrandbits(bits_d) | (1 << (bits_d - 1))
        out.append((a, d))
    d = rng.getrandbits(200_000) | (1 << 199_999)
    q = rng.getrandbits(300_000) | (1 << 299_999)
    out += [(q * d, d), (q * d - 1, d), (q * d + d - 1, d)]
    for k in (64, 4097, 100_000):
        out.append(((1 << 400_000) - 1, (1 << k) - 1))
    return out


@pytest.mark.parametrize("k", [0, 2, -2])
def test_divmod_host_buffers(pdh, k):
    pdh.newton_set_perturb(k)
    before = pdh.newton_corrections()
    for a, d in cases():
        print("divmod", a.bit_length(), d.bit_length(), k, flush=True)
        q, r = pdh.divmod_hip(pdh.int_to_limbs(a), pdh.int_to_limbs(d))
        assert (pdh.limbs_to_int(q), pdh.limbs_to_int(r)) == divmod(a, d)
    if k:
        assert pdh.newton_corrections() > before


@pytest.mark.parametrize("k", [0, 2, -2])
def test_isqrt_host_buffers(pdh, k):
    pdh.newton_set_perturb(k)
    rng = random.Random(9)
    before = pdh.newton_corrections()
    vals = []
    for bits in (150, 3000, 64001, 1_000_001, 3_000_000):
        x = rng.getrandbits(bits // 2) | (1 << (bits // 2 - 1))
        vals += [x * x, x * x - 1, x * x + 2 * x, rng.getrandbits(bits) | (1 << (bits - 1))]
    vals += [(1 << 200_000) - 1, 10005 * 10 ** 20000]
    for s in vals:
        print("isqrt", s.bit_length(), k, flush=True)
        assert pdh.limbs_to_int(pdh.isqrt_hip(pdh.int_to_limbs(s))) == math.isqrt(s)
    if k:
        assert pdh.newton_corrections() > before


def test_device_buffers(pdh):
    """pdh_divmod_dev / pdh_isqrt_dev on torch buffers with zero top limbs
    and garbage beyond the capacity the call may write."""
    import torch
    rng = random.Random(13)
    a = rng.getrandbits(500_000) | 1 << 499_999
    d = rng.getrandbits(123_456) | 1 << 123_455
    la, ld = pdh.int_to_limbs(a), pdh.int_to_limbs(d)
    A = torch.from_numpy(np.concatenate([la, np.zeros(5, np.uint32)]).view(np.int32)).cuda()
    D = torch.from_numpy(np.concatenate([ld, np.zeros(2, np.uint32)]).view(np.int32)).cuda()
    qcap, rcap = len(la) - len(ld) + 1, len(ld)
    Q = torch.full((qcap + 4,), 0x5A5A5A5A, dtype=torch.int32, device="cuda")
    R = torch.full((rcap + 4,), 0x5A5A5A5A, dtype=torch.int32, device="cuda")
    QL = torch.zeros(1, dtype=torch.int64, device="cuda")
    RL = torch.zeros(1, dtype=torch.int64, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    rc = pdh.pilib().pdh_divmod_dev(A.data_ptr(), A.numel(), D.data_ptr(), D.numel(), Q.data_ptr(), qcap,
                                    QL.data_ptr(), R.data_ptr(), rcap, RL.data_ptr(), s)
    assert rc == 0, pdh.lib().pdh_last_error()
    torch.cuda.synchronize()
    qv = Q.cpu().numpy().view(np.uint32)
    rv = R.cpu().numpy().view(np.uint32)
    q, r = divmod(a, d)
    assert pdh.limbs_to_int(qv[: int(QL.item())]) == q and pdh.limbs_to_int(rv[: int(RL.item())]) == r
    assert not qv[int(QL.item()):qcap].any() and not rv[int(RL.item()):rcap].any()   # zero-padded to capacity
    assert (qv[qcap:] == 0x5A5A5A5A).all() and (rv[rcap:] == 0x5A5A5A5A).all()      # nothing beyond it
    X = torch.full((len(la) // 2 + 3,), 0x5A5A5A5A, dtype=torch.int32, device="cuda")
    XL = torch.zeros(1, dtype=torch.int64, device="cuda")
    xcap = len(la) // 2 + 1
    rc = pdh.pilib().pdh_isqrt_dev(A.data_ptr(), A.numel(), X.data_ptr(), xcap, XL.data_ptr(), s)
    assert rc == 0
    torch.cuda.synchronize()
    xv = X.cpu().numpy().view(np.uint32)
    assert pdh.limbs_to_int(xv[: int(XL.item())]) == math.isqrt(a)
    assert (xv[xcap:] == 0x5A5A5A5A).all()


@pytest.mark.parametrize("k", [3, -1])
def test_pi_with_perturbed_estimates(pdh, golden, k):
    """Every estimate of the device engine off by k: the final division, the
    square root and the decimal levels' host fix-up correct them, and the
    digits still equal the independent 10^5 fixture."""
    pdh.newton_set_perturb(k)
    before = pdh.newton_corrections()
    s, _ = pdh.pi_chudnovsky(100_000)
    assert hashlib.sha256(s[2:].encode()).hexdigest() == golden["pi_dec_sha256"]["100000"]["sha256"]
    assert pdh.newton_corrections() > before
