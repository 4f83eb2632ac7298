"""Summarise rocprofv3 result databases (kernel trace and PMC counters).

    python aurora-pic_amd/aurora_pic/profsum.py trace <run_results.db>
    python aurora-pic_amd/aurora_pic/profsum.py pmc <run_results.db> [kernel-substring]
    python aurora-pic_amd/aurora_pic/profsum.py traffic <fetch.db> <write.db> [algorithmic-bytes]

trace: per-kernel calls, average and total duration.
pmc:   per-kernel mean of each counter (summed over the counter's hardware
       instances per dispatch) and the mean dispatch duration.
traffic: HBM bytes per launch of each k_push_tile3d variant from a FETCH_SIZE
       run and a WRITE_SIZE run of `bench.py --counter-pass`, each counter
       scaled by what it reads for the flat copy kernel of the same run
       (1 GiB read and 1 GiB written per launch).
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


def _per_dispatch(db: str, counter: str):
    cur = sqlite3.connect(db).cursor()
    rows = cur.execute("select name, dispatch_id, counter_value, duration from pmc_events "
                       "where counter_name = ?", (counter,)).fetchall()
    per = collections.OrderedDict()
    for name, did, val, d in rows:
        e = per.setdefault(did, [name, 0.0, d])
        e[1] += val
    return list(per.values())


def traffic(fetch_db: str, write_db: str, algorithmic: float = 56 * 33554432) -> str:
    out = []
    scale = {}
    data = {}
    for counter, db in (("FETCH_SIZE", fetch_db), ("WRITE_SIZE", write_db)):
        disp = _per_dispatch(db, counter)
        copies = [v for n, v, _ in disp if "k_stream_copy_flat" in n]
        scale[counter] = (1 << 30) / (copies[-1] * 1024.0)
        out.append(f"calibration {counter}: the copy counts {copies[-1]:.0f} KiB for 1 GiB -> "
                   f"factor {scale[counter]:.4f}")
        for n, v, d in disp:
            m = re.search(r"k_push_tile3d<(\d+), (\d+), (\d+), (\d+), \d+, (\d+), (\d+)>", n)
            if not m:
                continue
            key = ",".join(m.groups())
            data.setdefault(key, {}).setdefault(counter, []).append(v * 1024.0 * scale[counter])
            data[key].setdefault("dur_" + counter, []).append(d)
    for key, v in data.items():
        f = sum(v["FETCH_SIZE"]) / len(v["FETCH_SIZE"])
        w = sum(v["WRITE_SIZE"]) / len(v["WRITE_SIZE"])
        out.append(f"variant {key} bytes_per_launch {int(round(f + w))} fetch_bytes {int(round(f))} "
                   f"write_bytes {int(round(w))} over_algorithmic {(f + w) / algorithmic:.4f} "
                   f"dispatches {len(v['FETCH_SIZE'])}")
    return "\n".join(out)


if __name__ == "__main__":
    mode, db = sys.argv[1], sys.argv[2]
    if mode == "traffic":
        print(traffic(db, sys.argv[3], *(float(a) for a in sys.argv[4:5])))
    else:
        print(trace(db) if mode == "trace" else pmc(db, sys.argv[3] if len(sys.argv) > 3 else None))
