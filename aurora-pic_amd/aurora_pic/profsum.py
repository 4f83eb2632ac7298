"""Summarise rocprofv3 result databases (kernel trace and PMC counters).

    python aurora-pic_amd/aurora_pic/profsum.py trace <run_results.db>
    python aurora-pic_amd/aurora_pic/profsum.py pmc <run_results.db> [kernel-substring]

trace: per-kernel calls, average and total duration.
pmc:   per-kernel mean of each counter (summed over the counter's hardware
       instances per dispatch) and the mean dispatch duration.
"""
from __future__ import annotations

import collections
import re
import sqlite3
import sys


def short(name: str) -> str:
    n = name.replace("(anonymous namespace)", "").split("(")[0]
    n = re.sub(r"^void ", "", n).replace("aurora::::", "")
    return n[:80]


def trace(db: str) -> str:
    cur = sqlite3.connect(db).cursor()
    rows = cur.execute("select name, count(*), avg(end-start), sum(end-start) from kernels "
                       "group by name order by sum(end-start) desc").fetchall()
    total = sum(r[3] for r in rows) or 1
    out = [f"{'kernel':80s} {'calls':>6s} {'avg_us':>10s} {'total_ms':>10s} {'pct':>6s}"]
    for name, n, avg, s in rows:
        out.append(f"{short(name):80s} {n:6d} {avg / 1e3:10.2f} {s / 1e6:10.3f} "
                   f"{100 * s / total:6.1f}")
    return "\n".join(out)


def pmc(db: str, match: str | None = None) -> str:
    cur = sqlite3.connect(db).cursor()
    rows = cur.execute("select name, dispatch_id, counter_name, counter_value, duration "
                       "from pmc_events").fetchall()
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    dur = {}
    for name, did, cn, val, d in rows:
        k = short(name)
        if match and match not in k:
            continue
        per[(k, did)][cn] += val
        dur[(k, did)] = d
    agg = collections.defaultdict(lambda: collections.defaultdict(list))
    for (k, did), cnts in per.items():
        for cn, v in cnts.items():
            agg[k][cn].append(v)
        agg[k]["duration_ns"].append(dur[(k, did)])
    out = []
    for k in sorted(agg):
        out.append(k)
        for cn, vals in sorted(agg[k].items()):
            out.append(f"    {cn:24s} dispatches={len(vals):4d} mean={sum(vals) / len(vals):.6g}")
    return "\n".join(out)


if __name__ == "__main__":
    mode, db = sys.argv[1], sys.argv[2]
    print(trace(db) if mode == "trace" else pmc(db, sys.argv[3] if len(sys.argv) > 3 else None))
