"""CPU tests of the Chudnovsky engine's algorithms (Newton division and
square root, batched binary splitting, divide-and-conquer decimal output)
with every multiply done by the oracle's ntt-cpu through pdh_pi_chudnovsky_cb.
The digits are checked against the reference's 3,500-digit table and
independent SHA-256 fixtures of pi (tests/golden/make_pi_golden.py)."""
from __future__ import annotations

import ctypes
import hashlib
import os

import pytest

import oracle
from conftest import PKG

_L = None


def lib():
    global _L
    if _L is None:
        _L = ctypes.CDLL(os.path.join(PKG, "libpidigits_hip.so"))
        _L.pdh_pi_chudnovsky_cb.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t,
                                            ctypes.POINTER(ctypes.c_double)]
        _L.pdh_pi_chudnovsky_cb.restype = ctypes.c_int
        _L.pdh_chudnovsky_terms.argtypes = [ctypes.c_uint64]
        _L.pdh_chudnovsky_terms.restype = ctypes.c_uint64
    return _L


def pi_cpu(digits: int, mul: str = "orc_mul_ntt") -> str:
    fn = ctypes.cast(getattr(oracle.olib(), mul), ctypes.c_void_p)
    buf = ctypes.create_string_buffer(digits + 8)
    assert lib().pdh_pi_chudnovsky_cb(digits, fn, buf, len(buf), None) == 0
    return buf.value.decode()


def test_terms_rule():
    # reference src/chudnovsky.cpp:54-59 with prec = digits + 16
    for d in (0, 1, 100, 3500, 10**6, 10**7):
        assert lib().pdh_chudnovsky_terms(d) == max(2, int((d + 16) / 14.181647462725477) + 2)


@pytest.mark.parametrize("digits", [0, 1, 2, 50, 100, 1000, 3500])
def test_reference_table(golden, digits):
    s = pi_cpu(digits)
    if digits == 0:
        assert s == "3"
        return
    assert s[:2] == "3." and len(s) == digits + 2
    assert s[2:] == golden["pi_dec_3500"][:digits]


def test_schoolbook_multiplier_agrees():
    assert pi_cpu(1200, "orc_mul_schoolbook") == pi_cpu(1200)


@pytest.mark.parametrize("digits", [10000, 100000])
def test_independent_fixture(golden, digits):
    g = golden["pi_dec_sha256"][str(digits)]
    s = pi_cpu(digits)
    assert hashlib.sha256(s[2:].encode()).hexdigest() == g["sha256"]
