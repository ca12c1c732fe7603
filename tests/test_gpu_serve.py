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
