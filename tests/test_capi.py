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


def declared(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pdh_[a-z_0-9]+)\s*\(", text)))


def test_header_declarations_are_exported():
    import pdh
    names = declared()
    assert "pdh_mul" in names and "pdh_bbp_positions" in names
    L = pdh.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(pdh.EXPORTED)


def test_pi_header_declarations_are_exported():
    import pdh
    names = [n for n in declared(PI_HEADER) if n != "pdh_host_mul_fn"]
    L = pdh.pilib()
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(pdh.PI_EXPORTED)


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
