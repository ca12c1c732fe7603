"""Device-resident calls from several host threads on several streams share
the library's workspace in turn (an event per piece of scratch, waited on by
the next stream): products enqueued concurrently on two streams equal the
oracle's, and BBP calls overlap with multiplies without changing digits."""
from __future__ import annotations

import threading

import numpy as np
import pytest

import oracle
import workloads as W

pytestmark = pytest.mark.gpu


def residues(x_limbs, primes):
    v = int.from_bytes(np.asarray(x_limbs, np.uint32).tobytes(), "little")
    return [v % p for p in primes]


def test_two_streams_two_threads(pdh):
    import torch
    lib = pdh.lib()
    big_a, big_b = W.operand(71, 1 << 21), W.operand(72, (1 << 21) - 3)
    small = [(W.operand(100 + i, n), W.operand(200 + i, m))
             for i, (n, m) in enumerate(W.batch_sizes(77, 48, 1 << 10, 1 << 15))]
    results = {}
    errors = []

    def worker(name, pairs, rounds):
        try:
            st = pdh.lib()
            import ctypes
            s = ctypes.c_void_p()
            assert st.pdh_stream_create(ctypes.byref(s)) == 0
            keep = []
            for r in range(rounds):
                for i, (a, b) in enumerate(pairs):
                    A = torch.from_numpy(a.view(np.int32)).cuda()
                    B = torch.from_numpy(b.view(np.int32)).cuda()
                    O = torch.zeros(len(a) + len(b), dtype=torch.int32, device="cuda")
                    L = torch.zeros(1, dtype=torch.int64, device="cuda")
                    torch.cuda.synchronize()                 # uploads done before the library's stream reads them
                    rc = lib.pdh_mul_dev(A.data_ptr(), len(a), B.data_ptr(), len(b), O.data_ptr(), L.data_ptr(), s)
                    assert rc == 0, lib.pdh_last_error()
                    keep.append((name, r, i, A, B, O, L))
            assert st.pdh_stream_sync(s) == 0
            for name_, r, i, A, B, O, L in keep:
                results[(name_, r, i)] = O.cpu().numpy().view(np.uint32)[: int(L.item())].copy()
            st.pdh_stream_destroy(s)
        except Exception as e:       # surfaced in the main thread
            errors.append(e)

    t1 = threading.Thread(target=worker, args=("big", [(big_a, big_b)], 6))
    t2 = threading.Thread(target=worker, args=("small", small, 3))
    t1.start(); t2.start(); t1.join(); t2.join()
    assert not errors, errors
    want_small = [oracle.mul_ntt(a, b) for a, b in small]
    for r in range(3):
        for i in range(len(small)):
            assert np.array_equal(results[("small", r, i)], want_small[i]), (r, i)
    primes = [(1 << 61) - 1, (1 << 59) - 55, (1 << 57) - 13]
    ra, rb = residues(big_a, primes), residues(big_b, primes)
    want = [x * y % p for x, y, p in zip(ra, rb, primes)]
    first = results[("big", 0, 0)]
    assert residues(first, primes) == want
    for r in range(1, 6):
        assert np.array_equal(results[("big", r, 0)], first), r


def test_bbp_overlaps_multiplies(pdh, golden):
    """A BBP range on one thread while another thread runs multiplies: both
    results equal the ones computed alone."""
    a, b = W.operand(5, 1 << 20), W.operand(6, 1 << 20)
    alone = pdh.mul_ntt_hip(a, b)
    out = {}

    def bbp():
        out["bbp"] = pdh.bbp_hex_digits_hip(1, 4096)

    def muls():
        out["mul"] = [pdh.mul_ntt_hip(a, b) for _ in range(8)]

    t1, t2 = threading.Thread(target=bbp), threading.Thread(target=muls)
    t1.start(); t2.start(); t1.join(); t2.join()
    assert out["bbp"] == golden_hex(golden)[:4096]
    for m in out["mul"]:
        assert np.array_equal(m, alone)


def golden_hex(golden):
    import os
    from conftest import GOLDEN_DIR
    with open(os.path.join(GOLDEN_DIR, "bbp_1_65536.txt")) as f:
        return f.read().strip()
