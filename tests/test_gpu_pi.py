"""Chudnovsky pi with every multiply on the device (pdh_pi_chudnovsky):
digits equal the reference's table and independent fixtures of pi. Mirrors
the reference self-test's forced-GPU run (tests/self_test.cu:103-122)."""
from __future__ import annotations

import ctypes
import hashlib
import os

import pytest

from conftest import PKG

pytestmark = pytest.mark.gpu


def pi_dev(pdh, digits):
    L = ctypes.CDLL(os.path.join(PKG, "libpidigits_hip.so"))
    L.pdh_pi_chudnovsky.argtypes = [ctypes.c_uint64, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_double)]
    L.pdh_pi_chudnovsky.restype = ctypes.c_int
    buf = ctypes.create_string_buffer(digits + 8)
    ph = (ctypes.c_double * 6)()
    rc = L.pdh_pi_chudnovsky(digits, buf, len(buf), ph)
    assert rc == 0, pdh.lib().pdh_last_error().decode()
    return buf.value.decode(), list(ph)


@pytest.mark.parametrize("digits", [100, 1000, 3500, 8000])
def test_reference_table(pdh, golden, digits):
    before = pdh.launch_count()
    s, _ = pi_dev(pdh, digits)
    n = min(digits, 3500)
    assert len(s) == digits + 2 and s[2:2 + n] == golden["pi_dec_3500"][:n]
    assert pdh.launch_count() > before


@pytest.mark.parametrize("digits", [10000, 100000, 1000000])
def test_independent_fixture(pdh, golden, digits):
    key = str(digits)
    if key not in golden.get("pi_dec_sha256", {}):
        pytest.skip(f"no fixture for {digits} digits")
    s, ph = pi_dev(pdh, digits)
    print(digits, "phases ms", [round(x, 1) for x in ph])
    assert hashlib.sha256(s[2:].encode()).hexdigest() == golden["pi_dec_sha256"][key]["sha256"]
