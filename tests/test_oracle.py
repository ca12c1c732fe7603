"""CPU tests that pin the oracle (oracle/cpu_ref.cpp) before it is trusted.

Each check is chosen so a plausible slip in the restatement (a dropped
carry fold, a wrong twiddle, a transposed operand, an off-by-one in the
BBP loops) fails:

* field: closed forms (2^64 == EPS, 2^96 == -1, order of 2 is 192), brute
  force with Python integers, and the reference self-test's identities
  (tests/self_test.cu:46-62);
* NTT multiply: schoolbook on random sizes (self_test.cu:64-78), Python big
  integers, closed-form all-ones operands;
* repack / carry: against Python integer arithmetic;
* BBP: the reference's 64 hex digits and Bailey's checkpoint at 10^6;
* the closed-form mulmod count against brute force;
* everything against oracle/_ref (the reference compiled from source) where
  that library has been built.
"""
from __future__ import annotations

import random

import numpy as np
import pytest

import oracle
import workloads as W

P = (1 << 64) - (1 << 32) + 1
O = oracle.olib


def test_field_closed_forms():
    L = O()
    assert L.orc_gl_pow(2, 64) == (1 << 32) - 1
    assert L.orc_gl_pow(2, 96) == P - 1
    assert L.orc_gl_pow(2, 192) == 1
    assert all(L.orc_gl_pow(2, k) != 1 for k in (64, 96, 128))
    assert L.orc_gl_pow(7, P - 1) == 1                                   # self_test.cu:55
    for lg in range(1, 23):                                              # self_test.cu:56-61
        w = L.orc_gl_root_of_unity(1 << lg)
        assert L.orc_gl_pow(w, 1 << lg) == 1
        assert L.orc_gl_pow(w, 1 << (lg - 1)) == P - 1


def test_field_brute_force():
    L = O()
    rng = random.Random(7)
    special = [0, 1, 2, P - 1, P - 2, (1 << 32) - 1, 1 << 32, 1 << 63, P - (1 << 32)]
    for t in range(20000):
        a = special[t % len(special)] if t < 200 else rng.randrange(P)
        b = special[(t // 9) % len(special)] if t < 200 else rng.randrange(P)
        assert L.orc_gl_add(a, b) == (a + b) % P
        assert L.orc_gl_sub(a, b) == (a - b) % P
        assert L.orc_gl_mul(a, b) == (a * b) % P
        lo, hi = rng.randrange(1 << 64), rng.randrange(1 << 64)
        assert L.orc_gl_reduce128(lo, hi) == ((hi << 64) | lo) % P
        if a:
            assert L.orc_gl_mul(a, L.orc_gl_inv(a)) == 1


def test_ntt_length_and_repack():
    assert O().orc_ntt_length(1, 1) == 4
    assert O().orc_ntt_length(60000, 60000) == 1 << 18
    assert O().orc_ntt_length(1 << 22, 1 << 22) == 1 << 24
    assert O().orc_ntt_length(3, 5) == 16
    a = W.operand(3, 77)
    out = np.zeros(256, dtype=np.uint64)
    O().orc_repack_to16(oracle._ptr(a, oracle.ctypes.c_uint32), len(a),
                        oracle._ptr(out, oracle.ctypes.c_uint64), 256)
    v = sum(int(d) << (16 * i) for i, d in enumerate(out))
    assert v == oracle.limbs_to_int(a)
    assert out.max() < (1 << 16) and not out[154:].any()


def test_carry_propagate_against_integers():
    rng = np.random.default_rng(11)
    for n in (1, 2, 7, 64, 1000):
        conv = rng.integers(0, 1 << 55, size=n, dtype=np.uint64)
        if n > 3:
            conv[-2:] = 0
        out = np.zeros(n // 2 + 4, dtype=np.uint32)
        ln = oracle._sz(0)
        assert O().orc_carry_propagate(oracle._ptr(conv, oracle.ctypes.c_uint64), n,
                                       oracle._ptr(out, oracle.ctypes.c_uint32), len(out),
                                       oracle.ctypes.byref(ln)) == 0
        want = sum(int(c) << (16 * i) for i, c in enumerate(conv))
        assert oracle.limbs_to_int(out[: ln.value]) == want


def test_transform_is_the_dft():
    """The radix-2 NTT equals sum x_n w^(nk) with w = 7^((p-1)/N)."""
    for n in (2, 8, 32):
        x = np.array([random.Random(n).randrange(P) for _ in range(n)], dtype=np.uint64)
        y = x.copy()
        O().orc_ntt_transform(oracle._ptr(y, oracle.ctypes.c_uint64), n, 0)
        w = pow(7, (P - 1) // n, P)
        want = [sum(int(x[j]) * pow(w, j * k, P) for j in range(n)) % P for k in range(n)]
        assert [int(v) for v in y] == want
        O().orc_ntt_transform(oracle._ptr(y, oracle.ctypes.c_uint64), n, 1)
        assert np.array_equal(y, x)


class _XS:
    def __init__(self):
        self.s = 0x9E3779B97F4A7C15

    def r32(self):
        s = self.s
        s ^= (s << 13) & 0xFFFFFFFFFFFFFFFF
        s ^= s >> 7
        s ^= (s << 17) & 0xFFFFFFFFFFFFFFFF
        self.s = s
        return s & 0xFFFFFFFF


def test_ntt_multiply_vs_schoolbook_and_integers():
    x = _XS()
    for _ in range(40):                                                  # self_test.cu:67-77
        na, nb = 1 + x.r32() % 1500, 1 + x.r32() % 1500
        a = np.array([x.r32() for _ in range(na)], dtype=np.uint32)
        b = np.array([x.r32() for _ in range(nb)], dtype=np.uint32)
        sb = oracle.mul_schoolbook(a, b)
        assert np.array_equal(oracle.mul_ntt(a, b), sb)
        assert oracle.limbs_to_int(sb) == oracle.limbs_to_int(a) * oracle.limbs_to_int(b)
    assert len(oracle.mul_ntt(np.zeros(0, np.uint32), np.ones(3, np.uint32))) == 0


def test_all_ones_closed_form():
    for k in (1, 5, 4096):
        a = np.full(k, 0xFFFFFFFF, dtype=np.uint32)
        assert oracle.limbs_to_int(oracle.mul_ntt(a, a)) == ((1 << (32 * k)) - 1) ** 2


def test_bbp_reference_digits(golden):
    assert oracle.bbp_hex_range(1, 64, threads=8) == golden["pi_hex_64"]
    assert oracle.bbp_hex_range(1_000_000, 10, threads=8) == golden["bailey_1e6"]


def test_bbp_modpow_brute():
    for e, m in [(0, 7), (1, 7), (5, 1), (123, 1000), (99999, 2**31 - 1), (10**8, 8 * 10**7 + 5)]:
        assert O().orc_bbp_modpow16(e, m) == (pow(16, e, m) if m > 1 else 0)


def test_bbp_series_head_small_d():
    """For d = 0 the head term is 16^0 mod j / j exactly; the tail sums 16^-i/(8i+j)."""
    for j in (1, 4, 5, 6):
        s = 0.0 if j == 1 else 1.0 / j
        s -= np.floor(s)
        p16, k = 1.0 / 16.0, 1
        while p16 > 1e-18:
            s += p16 / (8 * k + j)
            p16 /= 16.0
            k += 1
        s -= np.floor(s)
        assert O().orc_bbp_series(j, 0) == s


@pytest.mark.parametrize("pos", [1, 2, 3, 17, 1000, 65536, 123457])
def test_mulmod_count_closed_form(pos):
    assert oracle.bbp_mulmod_count(pos) == oracle.bbp_mulmod_count_brute(pos)


needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference checkout)")


@needs_ref
def test_oracle_equals_reference_multiply():
    rng = np.random.default_rng(99)
    for na, nb in [(1, 1), (2, 3), (1499, 17), (3000, 3000), (60000, 60000), (65537, 31)]:
        a, b = W.operand(int(rng.integers(1 << 30)), na), W.operand(int(rng.integers(1 << 30)), nb)
        assert np.array_equal(oracle.mul_ntt(a, b), oracle.ref_mul_ntt(a, b))


@needs_ref
def test_oracle_equals_reference_bbp():
    assert oracle.bbp_hex_range(1, 2048, threads=8) == oracle.ref_bbp_hex_range(1, 2048, threads=8)
    pos = W.sampled_positions(5, 8, 10**5, 3 * 10**5)
    assert oracle.bbp_values_positions(pos, threads=8).tolist() == \
        [int(c, 16) for c in oracle.ref_bbp_hex_positions(pos, threads=8)]
