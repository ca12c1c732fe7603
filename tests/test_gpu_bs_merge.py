"""pdh_bs_merge_dev: one binary-splitting merge level in device memory,
checked against Python integers (the merge rule of src/chudnovsky.cpp:93-100:
P = P_L P_R, Q = Q_L Q_R, T = T_L Q_R + P_L T_R with T signed), and a whole
tree from pdh_bs_groups to the root against the definition of P, Q, T."""
from __future__ import annotations

import random

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MASK = (1 << 32) - 1


def limbs(v: int) -> np.ndarray:
    n = (v.bit_length() + 31) // 32
    return np.frombuffer(v.to_bytes(4 * n, "little"), dtype=np.uint32).copy() if n else np.zeros(0, np.uint32)


def value(arr: np.ndarray) -> int:
    return int.from_bytes(arr.astype(np.uint32).tobytes(), "little")


class Arena:
    """Device buffers kept alive for the test's duration."""

    def __init__(self):
        import torch
        self.torch = torch
        self.keep = []

    def put(self, v: int, extra: int = 0):
        t = self.torch.zeros(max(1, (v.bit_length() + 31) // 32 + extra), dtype=self.torch.int32, device="cuda")
        a = limbs(v)
        if len(a):
            t[: len(a)] = self.torch.from_numpy(a.view(np.int32)).cuda()
        self.keep.append(t)
        return t.data_ptr(), len(a)

    def buf(self, n: int, fill: int = 0x5A5A5A5A):
        t = self.torch.full((max(1, n),), fill - (1 << 32) if fill >= (1 << 31) else fill,
                            dtype=self.torch.int32, device="cuda")
        self.keep.append(t)
        return t

    def get(self, ptr_tensor, n: int) -> int:
        return value(ptr_tensor[:n].cpu().numpy().view(np.uint32)) if n else 0


def sgn(v: int) -> int:
    return (v > 0) - (v < 0)


def run_level(pdh, ar: Arena, nodes, skip_p=False):
    """nodes: list of (P, Q, T) Python ints. Returns the merged list of
    (P, Q, T) (P = None when skipped) read back from the device."""
    ins, outs, bufs = [], [], []
    for P, Q, T in nodes:
        pp, npl = ar.put(P)
        qp, nq = ar.put(Q)
        tp, nt = ar.put(abs(T))
        ins.append(pdh.BsNode(pp, npl, qp, nq, tp, nt, sgn(T)))
    for i in range(0, len(nodes) - 1, 2):
        (PL, QL, TL), (PR, QR, TR) = nodes[i], nodes[i + 1]
        nPL, nPR = len(limbs(PL)), len(limbs(PR))
        nQL, nQR = len(limbs(QL)), len(limbs(QR))
        nTL, nTR = len(limbs(abs(TL))), len(limbs(abs(TR)))
        b = (ar.buf(nPL + nPR), ar.buf(nQL + nQR), ar.buf(max(nTL + nQR, nPL + nTR) + 1))
        bufs.append(b)
        outs.append(pdh.BsNode(0 if skip_p else b[0].data_ptr(), 0, b[1].data_ptr(), 0, b[2].data_ptr(), 0, 0))
    if len(nodes) & 1:
        P, Q, T = nodes[-1]
        b = (ar.buf(len(limbs(P))), ar.buf(len(limbs(Q))), ar.buf(max(1, len(limbs(abs(T))))))
        bufs.append(b)
        outs.append(pdh.BsNode(0 if skip_p else b[0].data_ptr(), 0, b[1].data_ptr(), 0, b[2].data_ptr(), 0, 0))
    stream = ar.torch.cuda.current_stream().cuda_stream
    res = pdh.bs_merge_dev(ins, outs, skip_p, stream)
    out = []
    for r, b in zip(res, bufs):
        P = None if skip_p else ar.get(b[0], r.nP)
        out.append((P, ar.get(b[1], r.nQ), r.sT * ar.get(b[2], r.nT)))
        if r.nT:
            assert ar.get(b[2], r.nT) >> (32 * (r.nT - 1)) != 0       # normalized
        assert r.sT in (-1, 0, 1) and (r.sT == 0) == (r.nT == 0)
    return out


def merged(nodes, skip_p=False):
    out = []
    for i in range(0, len(nodes) - 1, 2):
        (PL, QL, TL), (PR, QR, TR) = nodes[i], nodes[i + 1]
        out.append((None if skip_p else PL * PR, QL * QR, TL * QR + PL * TR))
    if len(nodes) & 1:
        P, Q, T = nodes[-1]
        out.append((None if skip_p else P, Q, T))
    return out


def rand_node(rng, max_limbs):
    P = rng.getrandbits(32 * rng.randint(1, max_limbs)) | 1
    Q = rng.getrandbits(32 * rng.randint(1, max_limbs)) | 1
    t = rng.getrandbits(32 * rng.randint(0, max_limbs)) if rng.random() > 0.1 else 0
    return P, Q, rng.choice((-1, 1)) * t


@pytest.mark.parametrize("seed,count,max_limbs", [(0, 2, 3), (1, 37, 200), (2, 9, 3000), (3, 4, 20000)])
def test_random_levels(pdh, seed, count, max_limbs):
    rng = random.Random(seed)
    nodes = [rand_node(rng, max_limbs) for _ in range(count)]
    ar = Arena()
    assert run_level(pdh, ar, nodes) == merged(nodes)


@pytest.mark.parametrize("n", [1, 4095, 4096, 4097, 13000])
def test_ripples_and_cancellation(pdh, n):
    """Carries and borrows through every 4,096-limb tile, |A| and |B| equal
    but for the lowest limb in both orders, exact cancellation (T = 0) and a
    zero operand. With P_L = Q_R = 1 the sum is T_L + T_R itself."""
    big = 1 << (32 * n)
    rng = random.Random(n)
    hi = rng.getrandbits(32 * n) << 64
    cases = [
        (big - 1, 1),            # carry through every limb
        (big, -1),               # borrow through every limb
        (-big, big - 1),         # T = -1
        (big - 1, -big),         # |A| < |B| by one: T = -1
        (-(big - 1), big),       # T = +1
        (hi + 5, -(hi + 9)),
        (-(hi + 9), hi + 5),
        (hi + 7, -(hi + 7)),     # T = 0
        (0, -(big - 1)),
        (big - 1, 0),
    ]
    nodes = []
    for tl, tr in cases:
        nodes.append((1, 3, tl))
        nodes.append((5, 1, tr))
    want = [(5, 3, tl + tr) for tl, tr in cases]
    ar = Arena()
    assert run_level(pdh, ar, nodes) == want


def test_root_skips_p_and_odd_copy(pdh):
    rng = random.Random(11)
    nodes = [rand_node(rng, 500) for _ in range(3)]
    ar = Arena()
    assert run_level(pdh, ar, nodes, skip_p=True) == merged(nodes, skip_p=True)
    assert run_level(pdh, ar, nodes[2:]) == nodes[2:]


def test_invalid_input_raises(pdh):
    ar = Arena()
    p, n = ar.put(7)
    bad = [pdh.BsNode(p, n, p, 0, p, n, 1), pdh.BsNode(p, n, p, n, p, n, 1)]
    out = [pdh.BsNode(p, 0, p, 0, p, 0, 0)]
    with pytest.raises(pdh.PdhError) as e:
        pdh.bs_merge_dev(bad, out)
    assert e.value.code == pdh.PDH_E_ARG


def _leaf(k):
    A, B, C = 13591409, 545140134, 10939058860032000
    if k == 0:
        return 1, 1, A
    p = (6 * k - 5) * (2 * k - 1) * (6 * k - 1)
    return p, k ** 3 * C, (A + B * k) * p * (-1 if k & 1 else 1)


def _pqt(lo, hi):
    if hi - lo == 1:
        return _leaf(lo)
    m = (lo + hi) // 2
    pl, ql, tl = _pqt(lo, m)
    pr, qr, tr = _pqt(m, hi)
    return pl * pr, ql * qr, tl * qr + pl * tr


@pytest.mark.parametrize("n_terms", [130, 5000])
def test_tree_from_groups_to_root(pdh, n_terms):
    """pdh_bs_groups then pdh_bs_merge_dev level by level to the root (P
    skipped there) gives Q(0, N) and T(0, N) of the definition."""
    import torch
    G = 6
    groups = (n_terms + 63) >> 6
    dP = torch.zeros(groups * (3 << G), dtype=torch.int32, device="cuda")
    dQ = torch.zeros(groups * (5 << G), dtype=torch.int32, device="cuda")
    dT = torch.zeros(groups * (6 << G), dtype=torch.int32, device="cuda")
    dL = torch.zeros(groups * 3, dtype=torch.int64, device="cuda")
    dS = torch.zeros(groups, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    pdh.bs_groups(n_terms, G, dP.data_ptr(), dQ.data_ptr(), dT.data_ptr(), dL.data_ptr(), dS.data_ptr(), stream)
    torch.cuda.synchronize()
    L, S = dL.cpu().numpy(), dS.cpu().numpy()
    lvl = [pdh.BsNode(dP.data_ptr() + 4 * g * (3 << G), int(L[3 * g]), dQ.data_ptr() + 4 * g * (5 << G),
                      int(L[3 * g + 1]), dT.data_ptr() + 4 * g * (6 << G), int(L[3 * g + 2]), int(S[g]))
           for g in range(groups)]
    keep = []
    while len(lvl) > 1:
        root = len(lvl) == 2
        outs = []
        for i in range(0, len(lvl), 2):
            if i + 1 < len(lvl):
                a, b = lvl[i], lvl[i + 1]
                caps = (a.nP + b.nP, a.nQ + b.nQ, max(a.nT + b.nQ, a.nP + b.nT) + 1)
            else:
                a = lvl[i]
                caps = (a.nP, a.nQ, max(a.nT, 1))
            bufs = [torch.zeros(max(1, c), dtype=torch.int32, device="cuda") for c in caps]
            keep.append(bufs)
            outs.append(pdh.BsNode(0 if root else bufs[0].data_ptr(), 0, bufs[1].data_ptr(), 0,
                                   bufs[2].data_ptr(), 0, 0))
        lvl = pdh.bs_merge_dev(lvl, outs, root, stream)
        keep = keep[-len(outs):]
    q_want, t_want = _pqt(0, n_terms)[1:]
    r = lvl[0]
    qb, tb = keep[-1][1], keep[-1][2]
    assert value(qb[: r.nQ].cpu().numpy().view(np.uint32)) == q_want
    assert r.sT * value(tb[: r.nT].cpu().numpy().view(np.uint32)) == t_want
