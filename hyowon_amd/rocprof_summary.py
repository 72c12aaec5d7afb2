"""Summaries of rocprofv3 result databases (rocpd SQLite) for the PM step.

    python -m hyowon_amd.rocprof_summary stats   RUN_DB
    python -m hyowon_amd.rocprof_summary traffic FETCH_DB WRITE_DB COPY_BYTES

stats:   per-kernel dispatch count, mean / min / max / total duration (the
         --kernel-trace --stats table, recomputed from the dispatch records).
traffic: per-kernel HBM bytes per dispatch from a FETCH_SIZE run and a
         WRITE_SIZE run (KiB per dispatch), each scaled by the factor that
         maps the same counter onto the known bytes of a plain elementwise
         copy recorded in the same run (bench.py --calibrate-copy BYTES).
         Prints JSON.
"""
from __future__ import annotations

import glob
import json
import sqlite3
import sys


def _dbs(pattern: str):
    out = sorted(glob.glob(pattern))
    if not out:
        raise FileNotFoundError(pattern)
    return out


def kernel_stats(pattern: str):
    rows = {}
    for db in _dbs(pattern):
        c = sqlite3.connect(db)
        cols = [r[1] for r in c.execute("pragma table_info(kernels)")]
        name = "kernel_name" if "kernel_name" in cols else "name"
        dur = "duration" if "duration" in cols else "(end - start)"
        for k, n, s, mn, mx in c.execute(
                f"select {name}, count(*), sum({dur}), min({dur}), max({dur}) from kernels group by {name}"):
            r = rows.setdefault(k, [0, 0.0, float("inf"), 0.0])
            r[0] += n
            r[1] += s
            r[2] = min(r[2], mn)
            r[3] = max(r[3], mx)
        c.close()
    return rows


def counter_means(pattern: str, counter: str):
    out = {}
    for db in _dbs(pattern):
        c = sqlite3.connect(db)
        for k, n, v, d in c.execute(
                "select kernel_name, count(*), avg(value), avg(duration) from counters_collection "
                "where counter_name = ? group by kernel_name", (counter,)):
            out[k] = (n, v, d)
        c.close()
    return out


def traffic(fetch_pattern: str, write_pattern: str, copy_bytes: float) -> dict:
    fetch = counter_means(fetch_pattern, "FETCH_SIZE")
    write = counter_means(write_pattern, "WRITE_SIZE")
    cal = None
    for k, (n, v, d) in write.items():
        if "elementwise" in k and k in fetch and abs(v * 1024.0 - copy_bytes) < 0.01 * copy_bytes:
            cal = k
            break
    if cal is None:
        raise RuntimeError("calibration copy kernel not found in the WRITE_SIZE run")
    ff = copy_bytes / (fetch[cal][1] * 1024.0)
    wf = copy_bytes / (write[cal][1] * 1024.0)
    res = {"calibration": {"kernel": cal.split("(")[0], "bytes": copy_bytes,
                           "fetch_factor": ff, "write_factor": wf}, "kernels": {}}
    for k in sorted(set(fetch) & set(write)):
        if "hy::" not in k:
            continue
        fb, wb = fetch[k][1] * 1024.0 * ff, write[k][1] * 1024.0 * wf
        res["kernels"][k.split("(")[0]] = {"dispatches": fetch[k][0], "fetch_bytes": fb,
                                            "write_bytes": wb, "traffic_bytes": fb + wb,
                                            "avg_us": fetch[k][2] / 1e3}
    return res


def main(argv):
    if len(argv) >= 2 and argv[0] == "stats":
        rows = kernel_stats(argv[1])
        tot = sum(r[1] for r in rows.values())
        print(f"| kernel | calls | avg us | min us | max us | total ms | % |")
        print(f"|---|---:|---:|---:|---:|---:|---:|")
        for k, (n, s, mn, mx) in sorted(rows.items(), key=lambda kv: -kv[1][1])[:40]:
            print(f"| `{k.split('(')[0][:90]}` | {n} | {s / n / 1e3:.1f} | {mn / 1e3:.1f} | {mx / 1e3:.1f} | "
                  f"{s / 1e6:.2f} | {100 * s / tot:.2f} |")
    elif len(argv) >= 4 and argv[0] == "traffic":
        print(json.dumps(traffic(argv[1], argv[2], float(argv[3])), indent=1))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
