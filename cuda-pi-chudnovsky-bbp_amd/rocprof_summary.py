"""Short text summaries of rocprofv3 result databases (rocpd SQLite) for the
profiles/ directory.

    python cuda-pi-chudnovsky-bbp_amd/rocprof_summary.py trace DB TITLE COMMAND
        per-kernel calls / total / average duration (top_kernels view)
    python cuda-pi-chudnovsky-bbp_amd/rocprof_summary.py pmc [--split KERNEL] DB [DB ...]
        per-kernel mean per dispatch of every counter collected (with
        --split, KERNEL's dispatches one by one, as "KERNEL [n]" in dispatch
        order, for a kernel that serves several workloads), and, where
        SQ_INSTS_VALU and GRBM_GUI_ACTIVE are among them, derived lines:
        the dispatch's shader cycles per XCD (GRBM_GUI_ACTIVE is summed over
        MI355X's 8 XCDs), the VALU issue fraction (SQ_INSTS_VALU counts wave
        instructions; each takes 4 cycles of one of the 1,024 SIMDs), LDS
        bank-conflict cycles per LDS instruction and the waves' wait fraction
    python cuda-pi-chudnovsky-bbp_amd/rocprof_summary.py traffic FETCH_DB WRITE_DB CALIB_BYTES COMMAND
        HBM bytes per dispatch from FETCH_SIZE and WRITE_SIZE (KiB), each
        scaled so that the calibration kernel (torch's elementwise negate,
        which reads and writes CALIB_BYTES) reads as CALIB_BYTES; one line
        per kernel: name fetch_bytes write_bytes (the format bench.py reads)
"""
from __future__ import annotations

import collections
import sqlite3
import sys

CALIB = "neg_kernel"
XCDS, SIMDS = 8, 1024          # MI355X: 8 XCDs, 256 CUs of 4 SIMDs


def short(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "")
    return name.split("(")[0].replace("void ", "").replace("pdh::", "")


def counters(db: str, split: str | None = None) -> dict:
    c = sqlite3.connect(db)
    per = collections.defaultdict(float)
    names = {}
    for disp, kname, cname, val in c.execute(
            "select dispatch_id, kernel_name, counter_name, value from counters_collection"):
        per[(disp, cname)] += val
        names[disp] = kname
    order = {d: n + 1 for n, d in enumerate(sorted(d for d in names if short(names[d]) == split))}
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for (disp, cname), v in per.items():
        k = short(names[disp])
        agg[f"{k} [{order[disp]}]" if k == split else k][cname].append(v)
    return {k: {cn: sum(v) / len(v) for cn, v in d.items()} for k, d in agg.items()}


def trace(db: str, title: str, cmd: str) -> None:
    c = sqlite3.connect(db)
    rows = list(c.execute("select name, total_calls, total_duration, average, percentage from top_kernels"))
    print(f"# {title}")
    print(f"# command: {cmd}")
    print("# rocprofv3 kernel trace (top_kernels view); durations in microseconds")
    print(f"{'kernel':<44} {'calls':>6} {'total_us':>11} {'avg_us':>9} {'pct':>6}")
    for name, calls, tot, avg, pct in rows:
        print(f"{short(name)[:44]:<44} {calls:>6} {tot:>11.1f} {avg:>9.2f} {pct:>6.2f}")


def pmc(dbs: list[str], split: str | None = None) -> None:
    out = collections.defaultdict(dict)
    for db in dbs:
        for k, d in counters(db, split).items():
            out[k].update(d)
    for k in sorted(out):
        print(k)
        c = out[k]
        for cn, v in sorted(c.items()):
            print(f"    {cn:<28} {v:>16.1f}")
        if c.get("GRBM_GUI_ACTIVE") and "SQ_INSTS_VALU" in c:
            cyc = c["GRBM_GUI_ACTIVE"] / XCDS
            print(f"    {'= cycles per XCD':<28} {cyc:>16.1f}")
            print(f"    {'= VALU issue fraction':<28} {c['SQ_INSTS_VALU'] * 4 / SIMDS / cyc:>16.3f}")
            if c.get("SQ_INSTS_LDS"):
                print(f"    {'= LDS conflict cyc/instr':<28} {c.get('SQ_LDS_BANK_CONFLICT', 0) / c['SQ_INSTS_LDS']:>16.2f}")
            if c.get("SQ_WAVE_CYCLES"):
                print(f"    {'= wait fraction':<28} {c.get('SQ_WAIT_ANY', 0) / c['SQ_WAVE_CYCLES']:>16.3f}")


def traffic(fetch_db: str, write_db: str, calib_bytes: float, cmd: str) -> None:
    f, w = counters(fetch_db), counters(write_db)
    cf = next(v["FETCH_SIZE"] for k, v in f.items() if CALIB in k)
    cw = next(v["WRITE_SIZE"] for k, v in w.items() if CALIB in k)
    sf, sw = calib_bytes / cf, calib_bytes / cw
    print("# HBM traffic per dispatch from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (two runs)")
    print(f"# command: {cmd}")
    print(f"# calibration: torch neg over {calib_bytes:.0f} B in / {calib_bytes:.0f} B out read "
          f"FETCH_SIZE {cf:.1f}, WRITE_SIZE {cw:.1f}; scale factors {sf:.1f} and {sw:.1f} B per unit")
    print(f"# {'kernel':<30} {'fetch_bytes':>14} {'write_bytes':>14}")
    for k in sorted(set(f) & set(w)):
        if k.startswith(("k_", "k3_")):
            base = k.split("<")[0]
            print(f"{base:<32} {f[k]['FETCH_SIZE'] * sf:>14.0f} {w[k]['WRITE_SIZE'] * sw:>14.0f}")


if __name__ == "__main__":
    mode = sys.argv[1]
    if mode == "trace":
        trace(sys.argv[2], sys.argv[3], sys.argv[4])
    elif mode == "pmc":
        if sys.argv[2] == "--split":
            pmc(sys.argv[4:], sys.argv[3])
        else:
            pmc(sys.argv[2:])
    elif mode == "traffic":
        traffic(sys.argv[2], sys.argv[3], float(sys.argv[4]), sys.argv[5])
    else:
        raise SystemExit(__doc__)
