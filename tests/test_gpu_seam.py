"""The reference's own programs on the ntt-hip seam, and this build's CLI.

oracle/Makefile's `seam` target compiles the reference's unchanged
tests/self_test.cu and src/main.cpp with its bignum.cpp (mul_dispatch),
bignum_div.cpp, ntt_cpu.cpp, chudnovsky.cpp and bbp_cpu.cpp against
libpdh_seam.so, which defines the reference's mul_ntt_cuda, cuda_available
and bbp_hex_digits_cuda on the device. The reference's self-test then
checks every device product against its schoolbook, Karatsuba and ntt-cpu,
its Chudnovsky digits (including a run with every multiply forced onto the
device backend) against its digit table, and its GPU BBP digits against its
CPU BBP and Bailey's checkpoint. A probe linked in prints how many kernels
libpdh.so launched, so a host fallback cannot pass.

The programs are built in the container that holds the reference checkout
and travel with the tree; where they were not built the tests skip."""
from __future__ import annotations

import os
import re
import subprocess

import pytest

from conftest import PKG, ROOT

pytestmark = pytest.mark.gpu

REF = os.path.join(ROOT, "oracle", "_ref")


def run(args, timeout=300):
    print("run:", " ".join(os.path.basename(a) if i == 0 else a for i, a in enumerate(args)), flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    m = re.search(r"pdh_launch_count=(\d+)", p.stderr)
    return p, int(m.group(1)) if m else None


def need(path):
    if not os.path.exists(path):
        pytest.skip(f"{os.path.relpath(path, ROOT)} not built (needs the reference checkout at build time)")
    return path


def test_reference_selftest_on_seam():
    p, launches = run([need(os.path.join(REF, "ref_selftest_seam"))], timeout=600)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-2000:]
    assert "cuda_available = yes" in p.stdout
    m = re.search(r"(\d+) checks passed, (\d+) failed", p.stdout)
    assert m and int(m.group(2)) == 0 and int(m.group(1)) >= 83000, p.stdout[-2000:]
    assert "ALL TESTS PASSED" in p.stdout
    assert launches and launches > 1000, launches


def test_reference_cli_on_seam(golden):
    exe = need(os.path.join(REF, "ref_pi_seam"))
    p, launches = run([exe, "--backend", "ntt-cuda", "--digits", "3500", "--verify"])
    assert p.returncode == 0, p.stderr
    assert "verify : PASS (checked 3500 digits against reference)" in p.stdout
    assert "3." + golden["pi_dec_3500"] in p.stdout
    assert launches and launches > 100
    p, launches = run([exe, "--algo", "bbp", "--bbp-start", "1", "--bbp-count", "64", "--verify"])
    assert p.returncode == 0 and "verify : PASS" in p.stdout and golden["pi_hex_64"] in p.stdout
    assert "device : gpu" in p.stdout and launches and launches > 0
    p, _ = run([exe, "--algo", "bbp", "--bbp-start", "1000000", "--bbp-count", "10", "--quiet"])
    assert p.returncode == 0 and p.stdout.strip() == "26C65E52CB"


def test_pi_hip_cli(golden, tmp_path):
    exe = os.path.join(PKG, "pi_hip")
    p, _ = run([exe, "--digits", "20000", "--verify"])
    assert p.returncode == 0 and "verify : PASS (checked 1000 digits)" in p.stdout, p.stdout + p.stderr
    out = tmp_path / "pi.txt"
    p, _ = run([exe, "--digits", "3500", "--quiet", "--out", str(out)])
    assert p.returncode == 0 and out.read_text().strip() == "3." + golden["pi_dec_3500"]
    p, _ = run([exe, "--algo", "bbp", "--bbp-start", "1", "--bbp-count", "64", "--verify", "--bbp-device", "gpu"])
    assert p.returncode == 0 and "verify : PASS" in p.stdout and golden["pi_hex_64"] in p.stdout
    p, _ = run([exe, "--backend", "ntt-cpu"])
    assert p.returncode == 2 and "CPU backend" in p.stderr
    p, _ = run([exe, "--algo", "bbp", "--bbp-device", "cpu"])
    assert p.returncode == 2


def test_pdh_selftest():
    p, _ = run([os.path.join(PKG, "pdh_selftest")])
    assert p.returncode == 0, p.stdout[-2000:]
    assert re.search(r"\d+ checks passed, 0 failed", p.stdout)
