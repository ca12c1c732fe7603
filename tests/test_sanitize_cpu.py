"""The host side of the Chudnovsky engine (host/chudnovsky_hip.cpp, through
its callback entry points) and the oracle's multiply, built with
AddressSanitizer and UndefinedBehaviorSanitizer (`make -C oracle sanitize`,
host code only) and run without a device: pi strings against the fixtures,
with and without perturbed Newton estimates; exact division and square root
checked by schoolbook products; and a callback multiply that fails part way,
which must make the entry point return an error, leak nothing and leave the
engine working. Any sanitizer report fails the run (halt_on_error)."""
from __future__ import annotations

import hashlib
import os
import subprocess

import pytest

from conftest import ROOT

EXE = os.path.join(ROOT, "oracle", "_san", "pdh_sanitize")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
           UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")


@pytest.fixture(scope="module")
def exe():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("sanitizer build failed:\n" + r.stdout[-2000:] + r.stderr[-2000:])
    return EXE


def run(exe, *args, timeout=600):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, env=ENV, timeout=timeout)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    return r.stdout


@pytest.mark.parametrize("perturb", [0, 3, -3])
def test_pi_3500(exe, golden, perturb):
    s = run(exe, "pi", 3500, perturb).strip()
    assert s[:2] == "3." and s[2:] == golden["pi_dec_3500"]


def test_pi_1e5(exe, golden):
    s = run(exe, "pi", 100000).strip()
    assert hashlib.sha256(s[2:].encode()).hexdigest() == golden["pi_dec_sha256"]["100000"]["sha256"]


def test_division_and_square_root(exe):
    assert run(exe, "arith").strip() == "arith ok 81"


@pytest.mark.parametrize("k", [1, 2, 40, 300])
def test_failing_callback_multiply(exe, golden, k):
    out = run(exe, "fail", k).splitlines()
    rc, calls = (int(x.split("=")[1]) for x in out[0].split()[1:])
    assert out[0].startswith("fail") and rc == -2 and calls == k    # PDH_E_HIP at the failing call, no later call
    assert out[1].startswith("after rc=0")
    assert out[2][2:] == golden["pi_dec_3500"][:1000]
