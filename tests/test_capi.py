"""CPU tests of the C-ABI library: it loads, exports every symbol the public
header declares, and fails loudly (no CPU fallback) when no device is usable.
No compute call is made without a GPU."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "pdh_capi.h")
PI_HEADER = os.path.join(ROOT, "include", "pdh_pi.h")
DIAG_HEADER = os.path.join(ROOT, "include", "pdh_diag.h")


def declared(header=HEADER, text=None):
    text = open(header).read() if text is None else text
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pdh_[a-z_0-9]+)\s*\(", text)))


def declared_diag():
    """The test and diagnostic entry points of pdh_diag.h, split by the
    library that exports them."""
    text = open(DIAG_HEADER).read()
    core, pi = text.split("exported by libpidigits_hip.so")
    return declared(text=core), declared(text=pi)


def test_header_declarations_are_exported():
    import pdh
    names = declared()
    assert "pdh_mul" in names and "pdh_bbp_positions" in names
    diag = declared_diag()[0]
    assert "pdh_bbp_set_exact" in diag and "pdh_bbp_set_exact" not in names
    L = pdh.lib()
    for n in names + diag:
        assert hasattr(L, n), n
    assert set(names) | set(diag) == set(pdh.EXPORTED)


def test_pi_header_declarations_are_exported():
    import pdh
    names = [n for n in declared(PI_HEADER) if n != "pdh_host_mul_fn"]
    diag = declared_diag()[1]
    assert diag == ["pdh_newton_corrections", "pdh_newton_set_perturb"]
    L = pdh.pilib()
    for n in names + diag:
        assert hasattr(L, n), n
    assert set(names) | set(diag) == set(pdh.PI_EXPORTED)


def test_ntt_length_is_the_reference_rule():
    import pdh
    assert pdh.ntt_length(1, 1) == 4
    assert pdh.ntt_length(60000, 60000) == 1 << 18
    assert pdh.ntt_length(1 << 22, 1 << 22) == 1 << 24


def test_no_device_fails_loudly():
    """Without a gfx950 device every compute entry point reports
    PDH_E_NODEVICE; nothing silently runs on the host."""
    import numpy as np
    import pdh
    if pdh.device_available():
        pytest.skip("a device is present; covered by the gpu tests")
    with pytest.raises(pdh.PdhError) as e:
        pdh.mul_ntt_hip(np.ones(3, np.uint32), np.ones(3, np.uint32))
    assert e.value.code == pdh.PDH_E_NODEVICE
    with pytest.raises(pdh.PdhError) as e:
        pdh.bbp_values_range(1, 4)
    assert e.value.code == pdh.PDH_E_NODEVICE
    assert pdh.launch_count() == 0


def test_argument_errors_before_device():
    import numpy as np
    import pdh
    # empty operands give an empty product without touching the device
    assert len(pdh.mul_ntt_hip(np.zeros(0, np.uint32), np.ones(4, np.uint32))) == 0
    out = (ctypes.c_uint32 * 2)()
    n = ctypes.c_size_t(0)
    a = (ctypes.c_uint32 * 2)(1, 2)
    rc = pdh.lib().pdh_mul(a, 2, a, 2, out, 2, ctypes.byref(n))     # capacity 2 < 4
    assert rc == pdh.PDH_E_CAPACITY


def test_bs_merge_argument_errors_before_device():
    """pdh_bs_merge_dev rejects null arrays, an input without P or Q and a
    sign outside {-1, 0, 1} before any device work; an empty level is a no-op."""
    import pdh
    L = pdh.lib()
    assert L.pdh_bs_merge_dev(None, 0, None, 0, None) == pdh.PDH_OK
    assert L.pdh_bs_merge_dev(None, 2, None, 0, None) == pdh.PDH_E_ARG
    ins = (pdh.BsNode * 2)(pdh.BsNode(8, 1, 8, 0, 8, 1, 1), pdh.BsNode(8, 1, 8, 1, 8, 1, 1))
    outs = (pdh.BsNode * 1)(pdh.BsNode(8, 0, 8, 0, 8, 0, 0))
    assert L.pdh_bs_merge_dev(ins, 2, outs, 0, None) == pdh.PDH_E_ARG          # nQ = 0
    ins[0].nQ = 1
    ins[1].sT = 2
    assert L.pdh_bs_merge_dev(ins, 2, outs, 0, None) == pdh.PDH_E_ARG          # sign 2
    ins[1].sT = -1
    outs[0].T = None
    assert L.pdh_bs_merge_dev(ins, 2, outs, 0, None) == pdh.PDH_E_ARG          # no T buffer
    assert pdh.launch_count() == 0


REFERENCE_DIR = "/root/reference"   # read only where the checkout exists (this container)


def test_seam_library_exports_the_reference_symbols():
    """libpdh_seam.so defines the reference's mul_ntt_cuda, cuda_available and
    bbp_hex_digits_cuda (include/bignum.h:200-205, include/bbp.h:34) with the
    reference's C++ signatures."""
    import subprocess
    from conftest import PKG
    so = os.path.join(PKG, "libpdh_seam.so")
    out = subprocess.run(["nm", "-DC", "--defined-only", so], capture_output=True, text=True, check=True).stdout
    for sig in ("pidigits::mul_ntt_cuda(std::vector<unsigned int, std::allocator<unsigned int> > const&, "
                "std::vector<unsigned int, std::allocator<unsigned int> > const&)",
                "pidigits::cuda_available()",
                "pidigits::bbp_hex_digits_cuda[abi:cxx11](unsigned long, unsigned int)"):
        assert sig in out, sig


@pytest.mark.skipif(not os.path.isdir(REFERENCE_DIR), reason="reference checkout not present")
def test_mirror_header_compiles_beside_reference_headers(tmp_path):
    """A reference translation unit can include the C++ mirror next to its own
    bignum.h, bbp.h and chudnovsky.h (no redefinition) and catch HipError."""
    import subprocess
    from conftest import PKG
    src = tmp_path / "tu.cpp"
    src.write_text('#include "bignum.h"\n#include "bbp.h"\n#include "chudnovsky.h"\n#include "pidigits_hip.h"\n'
                   "int f() { try { return (int)pidigits::mul_ntt_hip({1u}, {2u}).size(); }\n"
                   "          catch (const pidigits::HipError& e) { return e.code; } }\n"
                   "bool g() { return pidigits::hip_available() || pidigits::cuda_available(); }\n")
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(REFERENCE_DIR, "include"),
                    "-I", os.path.join(PKG, "host"), str(src)], check=True)


@pytest.mark.skipif(not os.path.isdir(REFERENCE_DIR), reason="reference checkout not present")
def test_reference_programs_link_against_the_seam():
    """oracle/Makefile's seam target built the reference's unchanged
    self-test and CLI against libpdh_seam.so (run on the GPU by
    tests/test_gpu_seam.py)."""
    import subprocess
    for exe in ("ref_selftest_seam", "ref_pi_seam"):
        path = os.path.join(ROOT, "oracle", "_ref", exe)
        assert os.path.exists(path), exe
        ldd = subprocess.run(["ldd", path], capture_output=True, text=True, check=True).stdout
        assert "libpdh_seam.so" in ldd and "libpdh.so" in ldd and "not found" not in ldd


def test_shape_file_round_trip_and_rejects_bad_files(tmp_path):
    """pdh_tune_save / pdh_tune_load need no device: shapes written are read
    back, and a malformed file, an unlaunchable shape or a repeated size is
    refused with nothing changed."""
    import pdh
    before = {n: pdh.tune_get(n) for n in range(1, 25)}
    path = tmp_path / "shapes.txt"
    pdh.tune_save(str(path))
    lines = path.read_text().splitlines()
    assert lines[0] == "pdh-shapes 1 gfx950" and len(lines) == 25
    orig = path.read_text()
    try:
        pdh.tune_set(16, 5, 16, 3, 4)
        assert pdh.tune_get(16) == (5, 16, 3, 4)
        pdh.tune_load(str(path))
        assert {n: pdh.tune_get(n) for n in range(1, 25)} == before
        pdh.tune_set(16, 5, 16, 3, 4)
        bad = {
            "magic": orig.replace("pdh-shapes 1 gfx950", "pdh-shapes 2 gfx950"),
            "unlaunchable": orig.replace(lines[16], "16 1 5 10 7 3 4 1"),
            "repeated": orig + lines[16] + "\n",
            "trailing": orig.replace(lines[16], lines[16] + " 9"),
            "custom flag": orig.replace(lines[16], lines[16][:-1] + "2"),
        }
        for why, text in bad.items():
            p = tmp_path / "bad.txt"
            p.write_text(text)
            with pytest.raises(pdh.PdhError) as e:
                pdh.tune_load(str(p))
            assert e.value.code == pdh.PDH_E_ARG, why
            assert pdh.tune_get(16) == (5, 16, 3, 4), why
        with pytest.raises(pdh.PdhError):
            pdh.tune_load(str(tmp_path / "missing.txt"))
    finally:
        pdh.tune_load(str(path))
    assert {n: pdh.tune_get(n) for n in range(1, 25)} == before
