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


def test_1e7_digits_engines_agree(pdh, golden):
    """10^7 digits: the first 10^6 equal the independent fixture, and the
    whole string is the same whether every multiply runs on the three-prime
    engine or on the reference's Goldilocks NTT (a different exact algorithm
    for every product); with a 10^7 fixture present, its hash too."""
    digits = 10 ** 7
    s1, _ = pi_dev(pdh, digits)
    assert hashlib.sha256(s1[2:2 + 10 ** 6].encode()).hexdigest() == golden["pi_dec_sha256"]["1000000"]["sha256"]
    saved = {n: pdh.tune_engine(n) for n in range(14, 25)}
    try:
        for n in range(14, 25):
            pdh.tune_engine(n, 0)
        s0, _ = pi_dev(pdh, digits)
    finally:
        for n, e in saved.items():
            pdh.tune_engine(n, e)
    assert s0 == s1
    ref = golden["pi_dec_sha256"].get(str(digits))
    if ref:
        assert hashlib.sha256(s1[2:].encode()).hexdigest() == ref["sha256"]


@pytest.mark.parametrize("digits", [3 * 10 ** 7, 10 ** 8])
def test_beyond_1e7_prefix(pdh, golden, digits):
    """3*10^7 and 10^8 digits: products beyond one transform (the library's
    blocked path), decimal levels and merge levels too large for a batch.
    Pi's digits are a prefix property, so the first 10^6 and 10^7 must equal
    the independent fixtures."""
    s, _ = pi_dev(pdh, digits)
    assert len(s) == digits + 2 and s[:2] == "3."
    assert hashlib.sha256(s[2:2 + 10 ** 6].encode()).hexdigest() == golden["pi_dec_sha256"]["1000000"]["sha256"]
    ref = golden["pi_dec_sha256"].get("10000000")
    if ref:
        assert hashlib.sha256(s[2:2 + 10 ** 7].encode()).hexdigest() == ref["sha256"]


def _leaf(k):
    A, B, C = 13591409, 545140134, 10939058860032000
    if k == 0:
        return 1, 1, A
    p = (6 * k - 5) * (2 * k - 1) * (6 * k - 1)
    return p, k ** 3 * C, (A + B * k) * p * (-1 if k & 1 else 1)


@pytest.mark.parametrize("n_terms,G", [(1, 1), (7, 2), (64, 6), (1000, 6), (777, 5), (5000, 6)])
def test_bs_groups_match_the_definition(pdh, n_terms, G):
    """pdh_bs_groups: P, Q, T of every group of 2^G terms equal the products
    and the sum that define them (src/chudnovsky.cpp:64-102), computed here
    with Python integers, including the partial last group and the sign."""
    import numpy as np
    import torch
    groups = (n_terms + (1 << G) - 1) >> G
    dP = torch.zeros(groups * (3 << G), dtype=torch.int32, device="cuda")
    dQ = torch.zeros(groups * (5 << G), dtype=torch.int32, device="cuda")
    dT = torch.zeros(groups * (6 << G), dtype=torch.int32, device="cuda")
    dL = torch.zeros(groups * 3, dtype=torch.int64, device="cuda")
    dS = torch.zeros(groups, dtype=torch.int32, device="cuda")
    pdh.bs_groups(n_terms, G, dP.data_ptr(), dQ.data_ptr(), dT.data_ptr(), dL.data_ptr(), dS.data_ptr(),
                  torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    P, Q, T = (x.cpu().numpy().view(np.uint32) for x in (dP, dQ, dT))
    L, S = dL.cpu().numpy(), dS.cpu().numpy()

    def val(arr, off, n):
        return sum(int(arr[off + i]) << (32 * i) for i in range(n))
    for g in range(groups):
        lo, hi = g << G, min(n_terms, (g + 1) << G)
        p, q, t = _leaf(lo)
        for k in range(lo + 1, hi):
            pk, qk, tk = _leaf(k)
            p, q, t = p * pk, q * qk, t * qk + p * tk
        assert val(P, g * (3 << G), L[3 * g]) == p, g
        assert val(Q, g * (5 << G), L[3 * g + 1]) == q, g
        assert int(S[g]) * val(T, g * (6 << G), L[3 * g + 2]) == t, g
    with pytest.raises(pdh.PdhError):
        pdh.bs_groups(10, 7, dP.data_ptr(), dQ.data_ptr(), dT.data_ptr(), dL.data_ptr(), dS.data_ptr())
