"""BBP digits above position 2^27, where moduli 8k + j pass 2^30 (the FP64
ladder instead of the Montgomery one), 2^31 (past the former limit of the
entry points), and 2^32 (the reference's 128-bit bbp_mulmod branch,
include/bbp_kernel.h:70-83), against tests/golden/golden.json "bbp_high":
eight positions in (2^27, 2^28] and two in (2^28, 2^29] from the
reference's own bbp_cpu, two more (6*10^8 and 10^9) from bbp_fast.c, whose
residues were checked against the reference's bbp_modpow16 up to 2^48.

Every path that takes these positions is compared: the fixed-point path
with terms in lanes (32-bit indices up to 2^28, the wide kernel above),
positions in lanes (up to 2^28), the FP64-order chunked path (forced), the
fallback plumbing, digit blocks of a range, and device positions. The
lane-per-series kernel (exact mode 3) walks d + 1 terms in one lane, about
two minutes per position here, and is tested at lower positions only."""
from __future__ import annotations

import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def high(golden):
    g = golden["bbp_high"]
    pos = np.array(g["positions"], dtype=np.uint64)
    return pos, g["hex"], g["from_reference"]


@pytest.fixture
def modes(pdh):
    yield pdh
    pdh.bbp_set_exact(0)
    pdh.bbp_set_layout(0)


def test_fixture_shape(high):
    pos, hexd, nref = high
    assert len(pos) == len(hexd) == 12 and nref == 10
    assert all((1 << 27) < int(p) <= (1 << 28) for p in pos[:8])
    assert all((1 << 28) < int(p) <= (1 << 29) for p in pos[8:10])
    assert all(8 * (int(p) - 1) + 6 >= (1 << 32) for p in pos[10:])


@pytest.mark.parametrize("layout", [0, 1, 2, 3])
def test_fixed_point_layouts(modes, high, layout):
    pdh = modes
    pos, hexd, _ = high
    sel = pos if layout != 2 else pos[:8]               # positions in lanes take 32-bit indices
    pdh.bbp_set_layout(layout)
    t = time.time()
    got = pdh.to_hex(pdh.bbp_values_positions(sel))
    print(f"layout {layout}: {len(sel)} positions in {time.time() - t:.2f} s", flush=True)
    assert got == hexd[: len(sel)]


def test_ordered_path_forced(modes, high):
    pdh = modes
    pos, hexd, _ = high
    pdh.bbp_set_exact(1)
    before = pdh.bbp_exact_count()
    t = time.time()
    got = pdh.to_hex(pdh.bbp_values_positions(pos[:10]))
    print(f"FP64-order path: 10 positions in {time.time() - t:.2f} s", flush=True)
    assert got == hexd[:10]
    assert pdh.bbp_exact_count() - before == 10


def test_fallback_plumbing(modes, high):
    pdh = modes
    pos, hexd, _ = high
    pdh.bbp_set_exact(2)                                # every odd index through the ordered fallback
    assert pdh.to_hex(pdh.bbp_values_positions(pos[:10])) == hexd[:10]


def test_ranges_and_device_positions(pdh, high):
    import torch
    pos, hexd, _ = high
    for i in (0, 7, 9, 11):
        p = int(pos[i])
        got = pdh.bbp_values_range(p, 24)              # digit blocks from this position on
        want = pdh.bbp_values_positions(np.arange(p, p + 24, dtype=np.uint64))
        assert np.array_equal(got, want), p
        assert pdh.to_hex(got[:1]) == hexd[i]
    d = torch.from_numpy(pos.view(np.int64)).cuda()
    out = torch.zeros(len(pos), dtype=torch.uint8, device="cuda")
    pdh.bbp_positions_dev(d.data_ptr(), len(pos), out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert pdh.to_hex(out.cpu().numpy()) == hexd
