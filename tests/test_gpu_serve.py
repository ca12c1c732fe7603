"""The serving loop returns, for every request in a mix submitted together,
exactly what the same request returns alone (multiplies, Chudnovsky pi,
BBP digits), with requests admitted while others are on the device."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import workloads as W

pytestmark = pytest.mark.gpu


def test_mix_equals_alone(pdh, golden):
    rng = np.random.default_rng(21)
    muls = [(W.operand(1000 + i, int(na)), W.operand(2000 + i, int(nb)))
            for i, (na, nb) in enumerate(W.batch_sizes(5, 30, 1 << 6, 1 << 17))]
    muls.append((W.operand(3000, 1 << 21), W.operand(3001, (1 << 21) + 5)))
    muls.append((W.operand(3002, 17), np.zeros(0, np.uint32)))
    pis = [3500, 20000]
    bbps = [(1, 64), (500, 64), (1_000_000, 10), (int(rng.integers(10**5, 10**6)), 16)]
    alone_mul = [pdh.mul_ntt_hip(a, b) for a, b in muls]
    alone_pi = [pdh.pi_chudnovsky(d)[0] for d in pis]
    alone_bbp = [pdh.bbp_values_range(s, c) for s, c in bbps]

    srv = pdh.Server(200 << 30, 16)
    tickets = []
    for i, (a, b) in enumerate(muls):                  # interleave kinds as they would arrive
        tickets.append(("mul", i, srv.submit_mul(a, b)))
        if i % 10 == 3 and i // 10 < len(pis):
            tickets.append(("pi", i // 10, srv.submit_pi(pis[i // 10])))
        if i % 8 == 1 and i // 8 < len(bbps):
            tickets.append(("bbp", i // 8, srv.submit_bbp(*bbps[i // 8])))
    lat = [srv.wait(t) for _, _, t in tickets]
    for kind, i, t in tickets:
        got = srv.result(t)
        if kind == "mul":
            assert np.array_equal(got, alone_mul[i]), i
        elif kind == "pi":
            assert got == alone_pi[i]
        else:
            assert np.array_equal(got, alone_bbp[i]), bbps[i]
    srv.close()
    assert all(x > 0 for x in lat)
    assert alone_pi[0][2:] == golden["pi_dec_3500"]
    assert pdh.to_hex(alone_bbp[0]) == golden["pi_hex_64"]
    assert np.array_equal(alone_mul[0], oracle.mul_ntt(*muls[0]))


def _mul_bytes(pdh, na, nb):
    """The server's admission estimate of a multiply (serve_hip.cpp mul_bytes)."""
    return 8 * (na + nb) + 16 + 16 * pdh.ntt_length(na, nb) + (1 << 20)


def test_budget_admission_and_backfill(pdh):
    """A budget that holds two large multiplies at a time: the others wait
    for memory and are admitted as it frees up, small requests pass the
    waiting large ones, the budget is never exceeded, every product equals
    the oracle's."""
    big = [(W.operand(600 + i, 1 << 19), W.operand(700 + i, (1 << 19) - i)) for i in range(8)]
    small = [(W.operand(800 + i, 1000 + 37 * i), W.operand(900 + i, 777)) for i in range(12)]
    budget = 2 * _mul_bytes(pdh, 1 << 19, 1 << 19) + 3 * (1 << 20)
    srv = pdh.Server(budget, 16)
    # the small requests right behind the third large one, which waits while
    # the first two hold the budget (submitted after all eight, they could
    # find the budget free again or the waiting one past its starvation limit)
    order = big[:3] + small + big[3:]
    tickets = [srv.submit_mul(a, b) for a, b in order]
    got = []
    for t in tickets:
        srv.wait(t)
        got.append(srv.result(t))
    st = srv.stats()
    srv.close()
    want = oracle.mul_ntt_batch(order, threads=16)
    for i, (g, w) in enumerate(zip(got, want)):
        assert np.array_equal(g, w), i
    assert st["completed"] == len(tickets)
    # how many of the large ones wait depends on how fast the box moves
    # their operands; the third always does, and the small ones pass it
    assert st["waited_for_memory"] >= 1, st
    assert st["backfilled"] >= 1, st
    assert st["peak_bytes"] <= budget, st


def _residues(limbs, primes):
    v = int.from_bytes(np.asarray(limbs, np.uint32).tobytes(), "little")
    return [v % p for p in primes]


def test_c5_trace_results_vs_independent_checks(pdh, golden):
    """A seeded trace of the serving mix (70% multiplies of 2^10..2^22 limbs,
    20% Chudnovsky of 10^4..10^7 digits, 10% BBP blocks of 16-256 digits at
    10^5..10^8) submitted as one burst. Each result is checked against a
    source other than the server: multiplies up to 2^18 limbs against the
    oracle, larger ones by their residues mod three primes and their length;
    pi strings as prefixes of a 10^7-digit string pinned by the independent
    fixture; BBP blocks against the per-position path, and against the
    oracle where the block starts below 2*10^6."""
    import hashlib
    import sys
    from conftest import ROOT
    sys.path.insert(0, ROOT)
    import bench
    reqs = bench.serve_trace(W, 64, 1000.0, seed=91) + [(0.0, "bbp", (1_000_000, 16))]
    kinds = [k for _, k, _ in reqs]
    assert kinds.count("mul") > 30 and kinds.count("pi") > 5 and kinds.count("bbp") >= 3
    pi7, _ = pdh.pi_chudnovsky(10 ** 7)
    assert hashlib.sha256(pi7[2:].encode()).hexdigest() == golden["pi_dec_sha256"]["10000000"]["sha256"]
    srv = pdh.Server(200 << 30, 64)
    tickets = [srv.submit_mul(*arg) if k == "mul" else srv.submit_pi(arg) if k == "pi" else srv.submit_bbp(*arg)
               for _, k, arg in reqs]
    results = []
    for t in tickets:
        srv.wait(t)
        results.append(srv.result(t))
    srv.close()
    primes = [(1 << 61) - 1, (1 << 59) - 55, (1 << 57) - 13]
    small = [(i, arg) for i, (_, k, arg) in enumerate(reqs) if k == "mul" and len(arg[0]) + len(arg[1]) <= 1 << 18]
    want_small = oracle.mul_ntt_batch([arg for _, arg in small], threads=16)
    for (i, _), w in zip(small, want_small):
        assert np.array_equal(results[i], w), i
    for i, (_, k, arg) in enumerate(reqs):
        r = results[i]
        if k == "mul" and len(arg[0]) + len(arg[1]) > 1 << 18:
            a, b = arg
            want = [x * y % p for x, y, p in zip(_residues(a, primes), _residues(b, primes), primes)]
            assert _residues(r, primes) == want, i
            assert len(r) in (len(a) + len(b) - 1, len(a) + len(b)) and r[-1] != 0
        elif k == "pi":
            assert r == pi7[: arg + 2], (i, arg)
        elif k == "bbp":
            start, count = arg
            per_pos = pdh.bbp_values_positions(np.arange(start, start + count, dtype=np.uint64))
            assert np.array_equal(r, per_pos), arg
            if start <= 2 * 10 ** 6:
                assert pdh.to_hex(r) == oracle.bbp_hex_range(start, count, threads=16), arg
