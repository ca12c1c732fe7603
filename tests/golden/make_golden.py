"""Writes the golden fixtures in this directory from the reference's own code.

The products and BBP digits come from oracle/_ref/libref.so -- the
reference's src/ntt_cpu.cpp and src/bbp_cpu.cpp compiled from source by
``make -C oracle ref`` -- and the decimal/hex digits of pi from the data in
the reference's include/pi_reference.h. Run it in a container that has the
reference checkout; the GPU tests read only the files it writes.

    python tests/golden/make_golden.py --part mul          # C1/C2 products (~1 min)
    python tests/golden/make_golden.py --part bbp-range    # positions 1..65536 (~2 min, 8 threads)
    python tests/golden/make_golden.py --part bbp-sampled --count 64
    python tests/golden/make_golden.py --part pi-data
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import re
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "cuda-pi-chudnovsky-bbp_amd"))
import oracle  # noqa: E402
import workloads as W  # noqa: E402

GOLDEN = os.path.join(HERE, "golden.json")
REFERENCE_DIR = os.environ.get("REFERENCE_DIR", "/root/reference")


def sha(limbs: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(limbs, dtype="<u4").tobytes()).hexdigest()


def load() -> dict:
    if os.path.exists(GOLDEN):
        with open(GOLDEN) as f:
            return json.load(f)
    return {}


def save(g: dict) -> None:
    with open(GOLDEN, "w") as f:
        json.dump(g, f, indent=1, sort_keys=True)
        f.write("\n")


def part_mul(g: dict) -> None:
    for key, sa, sb, n, square in [("c1_mul", 101, 102, W.C1_LIMBS, False),
                                   ("c2_mul", 201, 202, W.C2_LIMBS, False),
                                   ("c2_sqr", 201, 201, W.C2_LIMBS, True)]:
        a = W.operand(sa, n)
        b = a if square else W.operand(sb, n)
        t = time.time()
        p = oracle.ref_mul_ntt(a, b)
        g[key] = {"seed_a": sa, "seed_b": sb, "limbs": n, "len": int(len(p)), "sha256": sha(p),
                  "top_limb": int(p[-1]), "low_limb": int(p[0])}
        print(key, len(p), f"{time.time() - t:.1f}s", flush=True)


def part_bbp_range(g: dict, threads: int) -> None:
    lo, hi = W.C4_RANGE
    t = time.time()
    s = oracle.ref_bbp_hex_range(lo, hi - lo + 1, threads)
    with open(os.path.join(HERE, "bbp_1_65536.txt"), "w") as f:
        f.write(s + "\n")
    g["bbp_range"] = {"start": lo, "count": hi - lo + 1, "file": "bbp_1_65536.txt",
                      "sha256": hashlib.sha256(s.encode()).hexdigest()}
    print("bbp range", f"{time.time() - t:.1f}s", flush=True)


def part_bbp_sampled(g: dict, count: int, threads: int) -> None:
    pos = W.sampled_positions(401, count)
    t = time.time()
    s = oracle.ref_bbp_hex_positions(pos, threads)
    g["bbp_sampled"] = {"seed": 401, "count": count, "lo": W.C4_LO, "hi": W.C4_HI,
                        "positions": [int(x) for x in pos], "hex": s}
    print("bbp sampled", count, f"{time.time() - t:.1f}s", flush=True)


def part_pi_data(g: dict) -> None:
    text = open(os.path.join(REFERENCE_DIR, "include", "pi_reference.h")).read()

    def literal(name):
        body = text.split(name + " =", 1)[1].split(";", 1)[0]
        return "".join(re.findall(r'"([0-9A-F]*)"', body))
    dec = literal("PI_REF_DEC")
    hexd = literal("PI_REF_HEX")
    assert len(dec) == 3500 and len(hexd) == 64
    g["pi_dec_3500"] = dec
    g["pi_hex_64"] = hexd
    g["bailey_1e6"] = "26C65E52CB"     # reference tests/self_test.cu:135


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--part", required=True, choices=["mul", "bbp-range", "bbp-sampled", "pi-data"])
    ap.add_argument("--count", type=int, default=64)
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    assert oracle.ref_available(), "build oracle/_ref first: make -C oracle ref"
    g = load()
    if args.part == "mul":
        part_mul(g)
    elif args.part == "bbp-range":
        part_bbp_range(g, args.threads)
    elif args.part == "bbp-sampled":
        part_bbp_sampled(g, args.count, args.threads)
    else:
        part_pi_data(g)
    # re-read before writing so parts run in parallel do not drop each other's keys
    cur = load()
    cur.update({k: v for k, v in g.items()})
    save(cur)


if __name__ == "__main__":
    main()
