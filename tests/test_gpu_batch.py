"""GPU parity of the batched multiply (pdh_batch_*) against the oracle.

Every product of an uneven batch must equal the oracle's ntt-cpu product of
that pair computed alone; a captured hipGraph must give the same products on
every replay, with other device work between the replays.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
import workloads as W

pytestmark = pytest.mark.gpu


class DevBatch:
    """Operands packed in one device buffer, outputs in another."""

    def __init__(self, pairs):
        import torch
        self.pairs = pairs
        na = np.array([len(a) for a, _ in pairs], dtype=np.int64)
        nb = np.array([len(b) for _, b in pairs], dtype=np.int64)
        self.a_off = np.concatenate([[0], np.cumsum(na)[:-1]])
        self.b_off = np.concatenate([[0], np.cumsum(nb)[:-1]])
        self.o_off = np.concatenate([[0], np.cumsum(na + nb)[:-1]])
        A = np.concatenate([a for a, _ in pairs] + [np.zeros(1, np.uint32)])
        B = np.concatenate([b for _, b in pairs] + [np.zeros(1, np.uint32)])
        self.A = torch.from_numpy(A.view(np.int32)).cuda()
        self.B = torch.from_numpy(B.view(np.int32)).cuda()
        self.out = torch.zeros(int((na + nb).sum()) + 1, dtype=torch.int32, device="cuda")
        self.lens = torch.zeros(len(pairs), dtype=torch.int64, device="cuda")
        pa, pb, po, pl = self.A.data_ptr(), self.B.data_ptr(), self.out.data_ptr(), self.lens.data_ptr()
        self.descs = [(pa + 4 * int(self.a_off[i]), int(na[i]), pb + 4 * int(self.b_off[i]), int(nb[i]),
                       po + 4 * int(self.o_off[i]), pl + 8 * i) for i in range(len(pairs))]

    def results(self):
        out = self.out.cpu().numpy().view(np.uint32)
        lens = self.lens.cpu().numpy()
        return [out[int(o): int(o) + int(n)] for o, n in zip(self.o_off, lens)]

    def clear(self):
        self.out.zero_()
        self.lens.fill_(-1)


def _check(pairs, got):
    want = oracle.mul_ntt_batch(pairs, threads=8)
    for i, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(g, w), (i, len(pairs[i][0]), len(pairs[i][1]))


def test_uneven_batch_one_shot(pdh):
    import torch
    pairs = W.batch_pairs(77, 300, 1, 70000)
    pairs += [(W.operand(5, 1), W.operand(6, 1)), (np.full(4096, 0xFFFFFFFF, np.uint32),) * 2]
    db = DevBatch(pairs)
    pdh.mul_batch_dev(db.descs, torch.cuda.current_stream().cuda_stream)
    _check(pairs, db.results())


def test_batch_with_empty_operand(pdh):
    import torch
    pairs = [(W.operand(1, 100), np.zeros(0, np.uint32)), (W.operand(2, 3000), W.operand(3, 20000))]
    db = DevBatch(pairs)
    pdh.mul_batch_dev(db.descs, torch.cuda.current_stream().cuda_stream)
    got = db.results()
    assert len(got[0]) == 0
    assert np.array_equal(got[1], oracle.mul_ntt(*pairs[1]))


def test_graph_replays_match_eager(pdh):
    import torch
    pairs = W.batch_pairs(91, 120, 64, 65536)
    db = DevBatch(pairs)
    s = torch.cuda.Stream()
    plan = pdh.Batch(db.descs)
    with torch.cuda.stream(s):
        plan.run(s.cuda_stream)                      # eager
    s.synchronize()
    eager = [r.copy() for r in db.results()]
    _check(pairs, eager)
    plan.capture(s.cuda_stream)
    other_a, other_b = W.operand(8, 30000), W.operand(9, 40000)
    for rep in range(3):
        db.clear()
        torch.cuda.synchronize()
        plan.run(s.cuda_stream)                      # graph replay
        s.synchronize()
        got = db.results()
        for g, e in zip(got, eager):
            assert np.array_equal(g, e), rep
        # unrelated eager device work between replays
        assert np.array_equal(pdh.mul_ntt_hip(other_a, other_b), oracle.mul_ntt(other_a, other_b))
    plan.close()


@pytest.mark.parametrize("which", ["c3_4096_pairs", "bs_levels_1e7"])
def test_full_size_batches_vs_oracle(pdh, which):
    """The two batched workloads at their full sizes, each run as one captured
    plan (a hipGraph) and replayed, against the oracle's ntt-cpu product of
    every pair computed one at a time (the reference's self-test compares
    its device products the same way, tests/self_test.cu:64-78): the 4,096
    pairs with limb counts log-uniform over [64, 65,536] (seed 3), and the
    P, Q, T products of the 10^7-digit tree's depths 8-14 (src/chudnovsky.cpp:
    64-102 level shapes; 130,048 products)."""
    import torch
    if which == "c3_4096_pairs":
        pairs = W.batch_pairs(3, W.C3_PAIRS, W.C3_LO, W.C3_HI)
        assert len(pairs) == 4096
    else:
        pairs = W.bs_level_pairs(7, int(pdh.pilib().pdh_chudnovsky_terms(10 ** 7)))
        assert len(pairs) == 130_048
    print(which, len(pairs), "pairs", flush=True)
    db = DevBatch(pairs)
    s = torch.cuda.Stream()
    plan = pdh.Batch(db.descs)
    plan.capture(s.cuda_stream)
    for rep in range(2):
        db.clear()
        torch.cuda.synchronize()
        plan.run(s.cuda_stream)
        s.synchronize()
        got = db.results()
        if rep == 0:
            want = oracle.mul_ntt_batch(pairs, threads=min(16, os.cpu_count() or 1))
        bad = [i for i, (g, w) in enumerate(zip(got, want)) if not np.array_equal(g, w)]
        assert not bad, (rep, bad[:5])
    plan.close()


def test_chained_one_shot_batches_are_stream_ordered(pdh):
    """pdh_mul_batch_dev returns without waiting: three batches chained on
    one side stream (each multiplies the previous batch's products by new
    operands, read from the device before the host has waited for
    anything) must give the oracle's products once the stream is synchronized."""
    import torch
    rng = np.random.default_rng(404)
    k = 48
    sizes = [(int(x), int(y)) for x, y in zip(rng.integers(1, 3000, k), rng.integers(1, 3000, k))]
    ops = [[W.operand(40000 + 3 * i + j, sizes[i][j % 2] if j < 2 else int(rng.integers(1, 5000)))
            for j in range(4)] for i in range(k)]
    # stage 0: a*b, stage 1: (a*b)*c, stage 2: ((a*b)*c)*d
    caps = [[len(o[0]) + len(o[1])] for o in ops]
    for i, o in enumerate(ops):
        caps[i].append(caps[i][0] + len(o[2]))
        caps[i].append(caps[i][1] + len(o[3]))
    dev = [[torch.from_numpy(x.view(np.int32)).cuda() for x in o] for o in ops]
    outs = [[torch.zeros(c, dtype=torch.int32, device="cuda") for c in cs] for cs in caps]
    lens = torch.zeros((3, k), dtype=torch.int64, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    for st in range(3):
        descs = []
        for i in range(k):
            a = dev[i][0] if st == 0 else outs[i][st - 1]
            na = len(ops[i][0]) if st == 0 else caps[i][st - 1]
            b = dev[i][st + 1]
            descs.append((a.data_ptr(), na, b.data_ptr(), len(ops[i][st + 1]), outs[i][st].data_ptr(),
                          lens[st].data_ptr() + 8 * i))
        pdh.mul_batch_dev(descs, s.cuda_stream)
    s.synchronize()
    L = lens.cpu().numpy()
    for i, o in enumerate(ops):
        want = o[0]
        for st in range(3):
            want = oracle.mul_ntt(want, o[st + 1])
            got = outs[i][st].cpu().numpy().view(np.uint32)[: L[st, i]]
            assert L[st, i] == len(want) and np.array_equal(got, want), (i, st)
