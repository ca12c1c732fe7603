"""CPU checks of the 4,096-position BBP fixture (tests/golden/bbp_sampled_4096.json):
its positions are the seeded draw over [10^6, 10^8], its 64-position prefix
is the fixture written by the reference's own bbp_cpu, and its digits at the
smallest positions equal the oracle's restatement of the reference
(include/bbp_kernel.h:86-137), recomputed here."""
from __future__ import annotations

import json
import os

import numpy as np

import oracle
import workloads as W
from conftest import GOLDEN_DIR


def load():
    with open(os.path.join(GOLDEN_DIR, "bbp_sampled_4096.json")) as f:
        return json.load(f)


def test_fixture_positions_and_prefix(golden):
    g = load()
    pos = np.array(g["positions"], dtype=np.uint64)
    assert len(pos) == g["count"] == 4096 and len(g["hex"]) == 4096
    assert np.array_equal(pos, W.sampled_positions(g["seed"], g["count"], g["lo"], g["hi"]))
    assert set(g["hex"]) <= set("0123456789ABCDEF")
    s64 = golden["bbp_sampled"]
    assert [int(x) for x in pos[:64]] == s64["positions"]
    assert g["hex"][:64] == s64["hex"]


def test_fixture_smallest_positions_match_the_oracle():
    g = load()
    pos = np.array(g["positions"], dtype=np.uint64)
    idx = np.argsort(pos)[:4]
    vals = oracle.bbp_values_positions(pos[idx], threads=4)
    got = "".join("0123456789ABCDEF"[int(v)] for v in vals)
    assert got == "".join(g["hex"][i] for i in idx)
