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


def test_range_prefix(pdh, golden):
    """Positions 1..40,000 in one call equal the reference's digits."""
    g = golden["bbp_range"]
    want = open(os.path.join(GOLDEN_DIR, g["file"])).read().strip()
    assert pdh.bbp_hex_digits_hip(1, 40000) == want[:40000]


def test_any_order_60000(pdh, golden):
    """60,000 positions in random order equal the reference's digits at
    those positions."""
    g = golden["bbp_range"]
    want = open(os.path.join(GOLDEN_DIR, g["file"])).read().strip()
    rng = np.random.default_rng(11)
    pos = rng.integers(1, g["count"] + 1, size=60000).astype(np.uint64)
    got = pdh.to_hex(pdh.bbp_values_positions(pos))
    assert got == "".join(want[int(p) - 1] for p in pos)


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


def test_sampled_4096_positions(pdh):
    """Config C4's parity subset: 4,096 positions uniform in [10^6, 10^8]
    (seed 401), position for position against digits computed on the host
    in the reference's FP64 order (tests/golden/bbp_sampled_4096.json; its
    generator agrees with the reference's own bbp_cpu on 1..65,536 and on the
    64-position prefix). Also counts how many positions needed the ordered
    fallback."""
    import json
    with open(os.path.join(GOLDEN_DIR, "bbp_sampled_4096.json")) as f:
        g = json.load(f)
    pos = np.array(g["positions"], dtype=np.uint64)
    assert len(pos) == 4096 and len(g["hex"]) == 4096
    before = pdh.bbp_exact_count()
    got = pdh.to_hex(pdh.bbp_values_positions(pos))
    bad = [i for i in range(len(pos)) if got[i] != g["hex"][i]]
    assert not bad, [(int(pos[i]), got[i], g["hex"][i]) for i in bad[:8]]
    print("ordered fallback positions:", pdh.bbp_exact_count() - before)


@pytest.mark.parametrize("start,count", [(1_000_000, 256), (50_000_000, 200), (100_000_000, 64),
                                         ((1 << 28) - 99, 100)])
def test_range_blocks_equal_per_position(pdh, golden, start, count):
    """Digit blocks (pdh_bbp_range_dev, pdh_bbp_range: one fixed-point
    evaluation per G positions) against the per-position path at the same
    positions, up to the largest position the entry points take."""
    import torch
    want = pdh.bbp_values_positions(np.arange(start, start + count, dtype=np.uint64))
    if start == 1_000_000:
        assert pdh.to_hex(want[:10]) == golden["bailey_1e6"]
    dv = torch.full((count,), 0x77, dtype=torch.uint8, device="cuda")
    pdh.bbp_range_dev(start, count, dv.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert np.array_equal(dv.cpu().numpy(), want)
    assert np.array_equal(pdh.bbp_values_range(start, count), want)


def test_bad_positions_raise(pdh):
    """Positions outside [1, 2^45] (moduli from 2^48 on, where the
    reference's bbp_mulmod stops being exact): the host entry points reject
    them before any device work; the device-positions entry point computes
    the valid ones, writes 0xFF for the others and reports PDH_E_ARG."""
    import torch
    dv = torch.zeros(16, dtype=torch.uint8, device="cuda")
    for start in ((1 << 45) - 5, 0):
        with pytest.raises(pdh.PdhError, match="outside"):
            pdh.bbp_range_dev(start, 10, dv.data_ptr())
    with pytest.raises(pdh.PdhError, match="null"):
        pdh.bbp_range_dev(1, 10, 0)
    with pytest.raises(pdh.PdhError):
        pdh.bbp_values_positions([0])
    with pytest.raises(pdh.PdhError):
        pdh.bbp_values_range(0, 4)
    with pytest.raises(pdh.PdhError):
        pdh.bbp_values_positions([(1 << 45) + 1])
    pos = torch.tensor([1, 0, 2, (1 << 45) + 1], dtype=torch.int64, device="cuda")
    out = torch.zeros(4, dtype=torch.uint8, device="cuda")
    with pytest.raises(pdh.PdhError, match="outside"):
        pdh.bbp_positions_dev(pos.data_ptr(), 4, out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert out.cpu().tolist() == [2, 0xFF, 4, 0xFF]


def test_bad_positions_in_a_sorted_batch(pdh, golden):
    """More than 64 positions take the device sort and the scatter back:
    invalid positions among them still come back as 0xFF at their own index,
    with PDH_E_ARG, and the valid ones equal the reference's digits."""
    import torch
    rng = np.random.default_rng(5)
    vals = np.concatenate([np.arange(1, 99, dtype=np.int64), np.array([0, (1 << 45) + 1], dtype=np.int64)])
    perm = rng.permutation(len(vals))
    pos = torch.from_numpy(vals[perm]).cuda()
    out = torch.zeros(len(vals), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    pdh.timing_enable(True)
    with pytest.raises(pdh.PdhError, match="outside"):
        pdh.bbp_positions_dev(pos.data_ptr(), len(vals), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    names = pdh.timing_read()
    pdh.timing_enable(False)
    assert "k_bbp_scatter" in names, names
    want = open(os.path.join(GOLDEN_DIR, golden["bbp_range"]["file"])).read().strip()
    got = out.cpu().numpy()
    for i, p in enumerate(vals[perm]):
        if 1 <= p <= 98:
            assert got[i] == int(want[p - 1], 16), (i, p)
        else:
            assert got[i] == 0xFF, (i, p)


@pytest.fixture
def exact_mode(pdh):
    yield pdh.bbp_set_exact
    pdh.bbp_set_exact(0)


def test_sampled_positions_fixed_point_decides_all(pdh, golden):
    """The 64 sampled positions are all decided by the fixed-point path (no
    FP64-order fallback), through the host and the device entry points."""
    import torch
    g = golden["bbp_sampled"]
    before = pdh.bbp_exact_count()
    pos = np.array(g["positions"], dtype=np.uint64)
    assert pdh.to_hex(pdh.bbp_values_positions(pos)) == g["hex"]
    dpos = torch.from_numpy(pos.view(np.int64)).cuda()
    dval = torch.zeros(len(pos), dtype=torch.uint8, device="cuda")
    pdh.bbp_positions_dev(dpos.data_ptr(), len(pos), dval.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert pdh.to_hex(dval.cpu().numpy()) == g["hex"]
    assert pdh.bbp_exact_count() == before


def test_ordered_path_forced(pdh, golden, exact_mode):
    """Mode 1: the FP64-order path alone gives the reference's digits."""
    exact_mode(1)
    before = pdh.bbp_exact_count()
    assert pdh.bbp_hex_digits_hip(1, 64) == golden["pi_hex_64"]
    assert pdh.bbp_hex_digits_hip(1_000_000, 10) == golden["bailey_1e6"]
    assert pdh.bbp_hex_digits_hip(500, 64) == oracle.bbp_hex_range(500, 64, threads=8)
    assert pdh.bbp_exact_count() == before + 138


def test_lane_per_series_path(pdh, golden, exact_mode):
    """Mode 3: one lane per series in the reference's FP64 order, the
    wavefronts fed by the work queue, for the contiguous range 1..65,536
    (host entry point) and for random positions (device entry point)."""
    import torch
    exact_mode(3)
    g = golden["bbp_range"]
    want = open(os.path.join(GOLDEN_DIR, g["file"])).read().strip()
    assert pdh.bbp_hex_digits_hip(g["start"], g["count"]) == want
    rng = np.random.default_rng(12)
    pos = rng.integers(1, g["count"] + 1, size=5000).astype(np.uint64)
    dpos = torch.from_numpy(pos.view(np.int64)).cuda()
    dval = torch.zeros(len(pos), dtype=torch.uint8, device="cuda")
    pdh.bbp_positions_dev(dpos.data_ptr(), len(pos), dval.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert pdh.to_hex(dval.cpu().numpy()) == "".join(want[int(p) - 1] for p in pos)


def test_fallback_plumbing(pdh, golden, exact_mode):
    """Mode 2: every odd index (sorted order) is recomputed by the fallback
    and scattered back; host and device entry points give the same digits."""
    import torch
    exact_mode(2)
    before = pdh.bbp_exact_count()
    assert pdh.bbp_hex_digits_hip(1, 64) == golden["pi_hex_64"]
    assert pdh.bbp_exact_count() == before + 32
    dv = torch.zeros(64, dtype=torch.uint8, device="cuda")
    pdh.bbp_range_dev(1, 64, dv.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert pdh.to_hex(dv.cpu().numpy()) == golden["pi_hex_64"]
    assert pdh.bbp_exact_count() == before + 64
    rng = np.random.default_rng(17)
    pos = rng.integers(1, 200_000, size=301).astype(np.uint64)
    want = oracle.bbp_values_positions(pos, threads=8)
    assert np.array_equal(pdh.bbp_values_positions(pos), want)
    dpos = torch.from_numpy(pos.view(np.int64)).cuda()
    dval = torch.zeros(len(pos), dtype=torch.uint8, device="cuda")
    pdh.bbp_positions_dev(dpos.data_ptr(), len(pos), dval.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert np.array_equal(dval.cpu().numpy(), want)
    with pytest.raises(pdh.PdhError):
        pdh.bbp_set_exact(4)


@pytest.fixture
def layout(pdh):
    yield pdh.bbp_set_layout
    pdh.bbp_set_layout(0)


@pytest.mark.parametrize("mode", [1, 2, 3])
def test_both_term_layouts_give_the_reference_digits(pdh, golden, layout, mode):
    """Terms in lanes (1), positions in lanes (2, one exponent per
    wavefront) and moduli in lanes (3, 64 terms k per wavefront stepping
    through the sorted positions): the contiguous range 1..65,536 by
    positions, the 64 sampled golden positions (spread over [10^6, 10^8], so
    lanes idle in layout 2 and the gaps pass 2^8 in layout 3) and digit
    blocks at 10^6."""
    import torch
    layout(mode)
    g = golden["bbp_range"]
    want = open(os.path.join(GOLDEN_DIR, g["file"])).read().strip()
    pos = torch.arange(1, g["count"] + 1, dtype=torch.int64, device="cuda")
    dv = torch.zeros(g["count"], dtype=torch.uint8, device="cuda")
    pdh.bbp_positions_dev(pos.data_ptr(), g["count"], dv.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = pdh.to_hex(dv.cpu().numpy())
    bad = [i for i in range(len(want)) if got[i] != want[i]]
    assert not bad, f"{len(bad)} mismatches, first at position {bad[0] + 1}"
    s = golden["bbp_sampled"]
    assert pdh.to_hex(pdh.bbp_values_positions(np.array(s["positions"], dtype=np.uint64))) == s["hex"]
    assert pdh.bbp_hex_digits_hip(1_000_000, 10) == golden["bailey_1e6"]
    assert pdh.bbp_hex_digits_hip(1_000_000, 40)[:10] == golden["bailey_1e6"]


def test_layout_choice(pdh, golden, layout):
    """The default takes moduli in lanes for more than 64 positions up to
    position 2^27 - 1 (a contiguous range); otherwise positions in lanes for
    close positions and terms in lanes for spread ones (the 64 sampled
    positions, and positions above 2^27)."""
    import torch
    layout(0)
    s = golden["bbp_sampled"]
    above = np.uint64((1 << 27) + 1)
    for positions, kernel in ((np.arange(1, 4097, dtype=np.uint64), "k_bbp_fx_moduli"),
                              (np.array(s["positions"], dtype=np.uint64), "k_bbp_fx_terms"),
                              (above + np.arange(0, 256, dtype=np.uint64), "k_bbp_fx_lanes"),
                              (np.uint64(1) + np.arange(0, 64, dtype=np.uint64) * np.uint64(1 << 22),
                               "k_bbp_fx_terms")):
        dpos = torch.from_numpy(positions.view(np.int64)).cuda()
        dv = torch.zeros(len(positions), dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        pdh.timing_enable(True)
        pdh.bbp_positions_dev(dpos.data_ptr(), len(positions), dv.data_ptr(),
                              torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        names = pdh.timing_read()
        pdh.timing_enable(False)
        assert kernel in names, names
    with pytest.raises(pdh.PdhError):
        pdh.bbp_set_layout(4)


def test_moduli_layout_edges(pdh, golden, layout):
    """Moduli in lanes against terms in lanes, position by position, where
    its cases differ: positions in no order with repeats (more than 64 are
    sorted on the device and the digits scattered back; the kernel sorts each
    group of 64), positions 1 and 2 (the terms k = 0 and e = 0, m = 4 among
    them), ragged group and batch sizes, gaps below and above 2^8, moduli on
    both sides of 2^29 (the folded ladder's bound) and the largest position
    the layout takes, 2^27 - 1, beside one it hands to the default."""
    import torch
    rng = np.random.default_rng(77)
    s = np.array(golden["bbp_sampled"]["positions"], dtype=np.uint64)
    sets = {
        "sampled, shuffled, repeats": np.concatenate([rng.permutation(s), s[:5], s[3:4]]),
        "1..70 reversed": np.arange(70, 0, -1, dtype=np.uint64),
        "around 2^26 by 37": (np.uint64(1 << 26) - np.uint64(600) + np.arange(0, 37 * 45, 37, dtype=np.uint64)),
        "near 2^27 - 1": np.array([(1 << 27) - 1, (1 << 27) - 1000, (1 << 27) - 300, 5, (1 << 27) - 301],
                                  dtype=np.uint64),
        "past the layout": np.array([(1 << 27) + 5, 1000], dtype=np.uint64),
    }
    for name, positions in sets.items():
        want = None
        for mode in (1, 3):
            layout(mode)
            dpos = torch.from_numpy(positions.astype(np.uint64).view(np.int64)).cuda()
            dv = torch.zeros(len(positions), dtype=torch.uint8, device="cuda")
            torch.cuda.synchronize()
            pdh.timing_enable(True)
            pdh.bbp_positions_dev(dpos.data_ptr(), len(positions), dv.data_ptr(),
                                  torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            names = pdh.timing_read()
            pdh.timing_enable(False)
            got = dv.cpu().numpy()
            if mode == 1:
                want = got
            else:
                assert ("k_bbp_fx_moduli" in names) == (name != "past the layout"), (name, names)
                # more than one group of 64 in moduli in lanes: sorted on the device, scattered back
                sorted_path = len(positions) > 64 and name != "past the layout"
                assert ("k_bbp_scatter" in names) == sorted_path, (name, names)
                bad = np.nonzero(got != want)[0]
                assert not len(bad), f"{name}: {len(bad)} differ, first position {positions[bad[0]]}"
        print(f"{name}: {len(positions)} positions equal", flush=True)
    s64 = golden["bbp_sampled"]
    layout(3)
    assert pdh.to_hex(pdh.bbp_values_positions(np.array(s64["positions"], dtype=np.uint64)[::-1].copy())) == \
        s64["hex"][::-1]
