"""Writes SHA-256 fixtures of pi's decimal digits computed independently of
both the reference and this build (mpmath's own pi, pure-Python backend).

The reference prints the first `digits` fractional digits of
floor(pi * 10^(digits+16)) (src/chudnovsky.cpp:116-152); with 16 guard
digits these are pi's own digits, which the reference's self-test checks
against its 3,500-digit table (tests/self_test.cu:103-122). Beyond 3,500
digits pi itself is the fixture.

    python tests/golden/make_pi_golden.py 100000 1000000
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import mpmath

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden.json")


def pi_digits(n: int) -> str:
    mpmath.mp.dps = n + 30
    s = mpmath.nstr(+mpmath.pi, n + 20, strip_zeros=False)     # "3.14159..."
    return s[2: 2 + n]


def main() -> None:
    for arg in sys.argv[1:]:
        n = int(arg)
        t = time.time()
        d = pi_digits(n)
        g = json.load(open(GOLDEN))
        g.setdefault("pi_dec_sha256", {})[str(n)] = {
            "sha256": hashlib.sha256(d.encode()).hexdigest(), "head": d[:20], "tail": d[-20:]}
        with open(GOLDEN, "w") as f:
            json.dump(g, f, indent=1, sort_keys=True)
            f.write("\n")
        print(n, f"{time.time() - t:.1f}s", d[-20:], flush=True)


if __name__ == "__main__":
    main()
