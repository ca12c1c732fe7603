"""Device-resident magnitude operations (pdh_sadd_batch_dev, pdh_shr_batch_dev,
pdh_decimal_leaves_dev), checked against Python integers: the building
blocks of divide-and-conquer decimal conversion (the reference's to_decimal,
src/bignum_div.cpp:239-267, computes the same digits by repeated division
by 10^9)."""
from __future__ import annotations

import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def limbs(v: int) -> np.ndarray:
    n = (v.bit_length() + 31) // 32
    return np.frombuffer(v.to_bytes(4 * n, "little"), dtype=np.uint32).copy() if n else np.zeros(0, np.uint32)


class Dev:
    def __init__(self):
        import torch
        self.torch = torch
        self.keep = []

    def mag(self, v: int, cap: int = 0, garbage: bool = True):
        """Device magnitude of v in a buffer of max(cap, len) limbs (garbage
        above the length when asked) and its device length."""
        a = limbs(v)
        n = max(cap, len(a), 1)
        host = np.full(n, 0xA5A5A5A5, np.uint32) if garbage else np.zeros(n, np.uint32)
        host[: len(a)] = a
        t = self.torch.from_numpy(host.view(np.int32)).cuda()
        ln = self.torch.tensor([len(a)], dtype=self.torch.int64, device="cuda")
        self.keep += [t, ln]
        return t, ln

    def out(self, cap: int):
        t = self.torch.full((max(cap, 1),), 0x3C3C3C3C, dtype=self.torch.int32, device="cuda")
        ln = self.torch.full((1,), 999, dtype=self.torch.int64, device="cuda")
        sg = self.torch.full((1,), 7, dtype=self.torch.int32, device="cuda")
        self.keep += [t, ln, sg]
        return t, ln, sg

    @staticmethod
    def value(t, ln) -> int:
        n = int(ln.item())
        return int.from_bytes(t[:n].cpu().numpy().view(np.uint32).tobytes(), "little")


def sgn(v: int) -> int:
    return (v > 0) - (v < 0)


def _sadd_case(pdh, dev, a, b, sa, sb, cap=None):
    A, AL = dev.mag(abs(a))
    B, BL = dev.mag(abs(b))
    want = sa * abs(a) + sb * abs(b)
    cap = cap or max(len(limbs(abs(a))), len(limbs(abs(b)))) + 1
    O, OL, OS = dev.out(cap)
    d = pdh.SaddDesc(A.data_ptr(), AL.data_ptr(), sa if a else 0, B.data_ptr(), BL.data_ptr(), sb if b else 0,
                     O.data_ptr(), cap, OL.data_ptr(), OS.data_ptr())
    return d, (O, OL, OS), want


def test_sadd_batch_random_and_edges(pdh):
    rng = random.Random(3)
    dev = Dev()
    cases = []
    for _ in range(40):
        a = rng.getrandbits(32 * rng.randint(0, 9000))
        b = rng.getrandbits(32 * rng.randint(0, 9000))
        cases.append((a, b, rng.choice((-1, 1)), rng.choice((-1, 1))))
    big = 1 << (32 * 8193)
    cases += [(big - 1, 1, 1, 1), (big, 1, 1, -1), (big, big - 1, -1, 1), (big - 1, big, 1, -1),
              (12345 << 9000, 12345 << 9000, 1, -1), (0, big - 1, 1, -1), (0, 0, 1, 1)]
    descs, outs, wants = [], [], []
    for a, b, sa, sb in cases:
        d, o, w = _sadd_case(pdh, dev, a, b, sa, sb)
        descs.append(d)
        outs.append(o)
        wants.append(w)
    pdh.sadd_batch_dev(descs)
    for (O, OL, OS), w in zip(outs, wants):
        mag = Dev.value(O, OL)
        assert mag == abs(w)
        if w:
            assert int(OS.item()) == sgn(w)
        n = int(OL.item())
        if n:
            assert int(O[n - 1].item()) != 0


def test_sadd_in_place_and_narrow_cap(pdh):
    """out may be a itself, and a cap below the operands' lengths is enough
    when the result fits (a remainder x - q d of a long x)."""
    dev = Dev()
    rng = random.Random(8)
    x = rng.getrandbits(32 * 6000)
    d = rng.getrandbits(32 * 3000) | (1 << (32 * 3000 - 1))
    q = x // d
    qd = q * d
    X, XL = dev.mag(x, cap=len(limbs(x)) + 1)
    QD, QDL = dev.mag(qd)
    R, RL, RS = dev.out(3001)
    pdh.sadd_batch_dev([pdh.SaddDesc(X.data_ptr(), XL.data_ptr(), 1, QD.data_ptr(), QDL.data_ptr(), -1,
                                     R.data_ptr(), 3001, RL.data_ptr(), RS.data_ptr())])
    assert Dev.value(R, RL) == x - qd
    # in place: X = X - QD over X's own buffer, lengths in separate slots
    L2 = dev.torch.zeros(1, dtype=dev.torch.int64, device="cuda")
    S2 = dev.torch.zeros(1, dtype=dev.torch.int32, device="cuda")
    pdh.sadd_batch_dev([pdh.SaddDesc(X.data_ptr(), XL.data_ptr(), 1, QD.data_ptr(), QDL.data_ptr(), -1,
                                     X.data_ptr(), len(limbs(x)) + 1, L2.data_ptr(), S2.data_ptr())])
    assert Dev.value(X, L2) == x - qd and int(S2.item()) == 1


@pytest.mark.parametrize("seed", [0, 1])
def test_shr_batch(pdh, seed):
    rng = random.Random(seed)
    dev = Dev()
    descs, outs, wants = [], [], []
    cases = [(rng.getrandbits(32 * rng.randint(1, 7000)), rng.randint(0, 32 * 7000)) for _ in range(30)]
    cases += [((1 << 200) + 5, 200), ((1 << 200) + 5, 201), (3 << (32 * 50), 32 * 50 + 1), (1, 0), (7, 64),
              ((1 << (32 * 4096)) - 1, 31), ((1 << (32 * 4096)) - 1, 32)]
    for a, bits in cases:
        A, AL = dev.mag(a)
        w = a >> bits
        cap = max(len(limbs(a)) - bits // 32, 1)
        O, OL, _ = dev.out(cap)
        descs.append(pdh.ShrDesc(A.data_ptr(), AL.data_ptr(), O.data_ptr(), cap, OL.data_ptr(), bits))
        outs.append((O, OL, cap))
        wants.append(w)
    pdh.shr_batch_dev(descs)
    for (O, OL, cap), w in zip(outs, wants):
        assert int(OL.item()) == len(limbs(w))
        full = int.from_bytes(O[:cap].cpu().numpy().view(np.uint32).tobytes(), "little")
        assert full == w                      # zeros above the length, up to cap


@pytest.mark.parametrize("width,pad", [(9, 0), (18, 0), (1224, 0), (2304, 0), (1224, 400)])
def test_decimal_leaves(pdh, width, pad):
    """Both kernels: values staged in LDS (stride up to 480 limbs) and, with
    a padded stride, converted in place in global memory."""
    import torch
    rng = random.Random(width)
    vals = [rng.randrange(10 ** width) for _ in range(70)] + [0, 1, 10 ** width - 1, 10 ** (width - 1)]
    stride = (width * 3322 // 1000) // 32 + 3 + pad
    X = np.zeros(len(vals) * stride, np.uint32)
    L = np.zeros(len(vals), np.int64)
    for i, v in enumerate(vals):
        a = limbs(v)
        X[i * stride: i * stride + len(a)] = a
        L[i] = len(a)
    dX = torch.from_numpy(X.view(np.int32)).cuda()
    dL = torch.from_numpy(L).cuda()
    dO = torch.zeros(len(vals) * width, dtype=torch.uint8, device="cuda")
    pdh.decimal_leaves_dev(dX.data_ptr(), stride, dL.data_ptr(), len(vals), width, dO.data_ptr())
    torch.cuda.synchronize()
    s = dO.cpu().numpy().tobytes().decode()
    for i, v in enumerate(vals):
        assert s[i * width:(i + 1) * width] == str(v).zfill(width), i
    with pytest.raises(pdh.PdhError):
        pdh.decimal_leaves_dev(dX.data_ptr(), stride, dL.data_ptr(), 1, 10, dO.data_ptr())


def test_single_shift_both_directions(pdh):
    """pdh_shift_dev: a * 2^k for k >= 0 and floor(a / 2^-k) for k < 0, on one
    stream with no host synchronization between calls."""
    import torch
    rng = random.Random(21)
    dev = Dev()
    stream = torch.cuda.current_stream().cuda_stream
    cases = [(rng.getrandbits(32 * rng.randint(1, 5000)), rng.randint(-32 * 5000, 32 * 300)) for _ in range(40)]
    cases += [(1, 0), (1, 31), (1, 32), (0xFFFFFFFF, 1), (0xFFFFFFFF, -31), (7, -3), (7, -64), (0, 5), (0, -5),
              ((1 << (32 * 2048)) - 1, 33), (1 << 4000, -4000)]
    outs = []
    for a, k in cases:
        A, AL = dev.mag(a)
        want = a << k if k >= 0 else a >> -k
        cap = max(len(limbs(a)) + (k // 32 + 1 if k >= 0 else -((-k) // 32)), 1)
        O, OL, _ = dev.out(cap)
        assert pdh.lib().pdh_shift_dev(A.data_ptr(), AL.data_ptr(), k, O.data_ptr(), cap, OL.data_ptr(),
                                       stream) == 0
        outs.append((O, OL, cap, want))
    torch.cuda.synchronize()
    for O, OL, cap, want in outs:
        assert int(OL.item()) == len(limbs(want))
        assert int.from_bytes(O[:cap].cpu().numpy().view(np.uint32).tobytes(), "little") == want


def test_single_sadd_chain(pdh):
    """pdh_sadd_dev calls chained on one stream: each sum feeds the next,
    with lengths and signs only on the device until the end."""
    import torch
    rng = random.Random(22)
    dev = Dev()
    stream = torch.cuda.current_stream().cuda_stream
    vals = [rng.getrandbits(32 * rng.randint(1, 6000)) for _ in range(12)]
    signs = [rng.choice((-1, 1)) for _ in vals]
    acc = vals[0]
    A, AL = dev.mag(vals[0], cap=6100)
    S = dev.torch.zeros(1, dtype=dev.torch.int32, device="cuda")
    for v, sg in zip(vals[1:], signs[1:]):
        B, BL = dev.mag(v)
        O, OL, OS = dev.out(6100)
        sa = 1 if acc >= 0 else -1
        # |acc| is on the device with its sign known only to the host's model; pass it
        assert pdh.lib().pdh_sadd_dev(A.data_ptr(), AL.data_ptr(), sa if acc else 0, B.data_ptr(), BL.data_ptr(),
                                      sg, O.data_ptr(), 6100, OL.data_ptr(), OS.data_ptr(), stream) == 0
        acc = acc + sg * v
        A, AL, S = O, OL, OS
    torch.cuda.synchronize()
    assert Dev.value(A, AL) == abs(acc)
    if acc:
        assert int(S.item()) == sgn(acc)
