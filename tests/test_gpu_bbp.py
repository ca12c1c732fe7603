"""GPU parity of the BBP kernel against the reference's digits.

Mirrors tests/self_test.cu:124-136 (positions 1..64 vs PI_REF_HEX, CPU ==
GPU at 500, Bailey's checkpoint at 10^6) and adds the full contiguous range
1..65536 and sampled positions in [10^6, 10^8], both produced by the
reference's own bbp_cpu.cpp (tests/golden/make_golden.py).
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import oracle
import workloads as W
from conftest import GOLDEN_DIR

pytestmark = pytest.mark.gpu


def test_first_64_hex_digits(pdh, golden):
    assert pdh.bbp_hex_digits_hip(1, 64) == golden["pi_hex_64"]


def test_block_at_500_matches_cpu(pdh):
    assert pdh.bbp_hex_digits_hip(500, 64) == oracle.bbp_hex_range(500, 64, threads=8)


def test_bailey_checkpoint_1e6(pdh, golden):
    assert pdh.bbp_hex_digits_hip(1_000_000, 10) == golden["bailey_1e6"]


def test_contiguous_1_to_65536(pdh, golden):
    g = golden["bbp_range"]
    want = open(os.path.join(GOLDEN_DIR, g["file"])).read().strip()
    got = pdh.bbp_hex_digits_hip(g["start"], g["count"])
    bad = [i for i in range(len(want)) if got[i] != want[i]]
    assert not bad, f"{len(bad)} mismatches, first at position {bad[0] + 1}"


def test_positions_api_any_order(pdh):
    rng = np.random.default_rng(3)
    pos = rng.integers(1, 3000, size=500).astype(np.uint64)
    got = pdh.bbp_values_positions(pos)
    rng_vals = pdh.bbp_values_range(1, 3000)
    assert np.array_equal(got, rng_vals[pos.astype(np.int64) - 1])


def test_sampled_positions(pdh, golden):
    g = golden["bbp_sampled"]
    pos = np.array(g["positions"], dtype=np.uint64)
    assert np.array_equal(pos, W.sampled_positions(g["seed"], g["count"], g["lo"], g["hi"]))
    got = pdh.to_hex(pdh.bbp_values_positions(pos))
    assert got == g["hex"]


def test_bad_positions_raise(pdh):
    with pytest.raises(pdh.PdhError):
        pdh.bbp_values_positions([0])
    with pytest.raises(pdh.PdhError):
        pdh.bbp_values_range(0, 4)
    with pytest.raises(pdh.PdhError):
        pdh.bbp_values_positions([(1 << 28) + 1])
