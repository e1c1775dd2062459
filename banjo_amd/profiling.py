"""Summaries of rocprofv3 result databases (``*_results.db``) for the lattice kernels.

    python -m banjo_amd.profiling trace  <results.db>
    python -m banjo_amd.profiling pmc    <results.db> [--copy <copy_results.db> --copy-bytes N]
    python -m banjo_amd.profiling traffic <fetch.db> --write-db <write.db> --copy <copy_fetch.db>
                                          --copy-write <copy_write.db> --copy-bytes N [--json out.json]

`trace` prints calls, total, mean, min and max duration per kernel (the
``--kernel-trace --stats`` view). `pmc` prints the mean value per dispatch of
every collected counter per kernel; with a copy-probe database of a copy of
known size it also prints the scale that turns the counter into bytes (what
the counter reads for the copy over the bytes the copy moved). `traffic`
combines a FETCH_SIZE and a WRITE_SIZE run, each scaled by its copy probe,
into HBM bytes per dispatch for every kernel, and can write them as JSON.
"""
from __future__ import annotations

import argparse
import json
import re
import sqlite3
import sys
from collections import defaultdict


def short(name: str) -> str:
    """Kernel name without template arguments or parameter lists."""
    base = re.sub(r"\(.*$", "", name)
    m = re.match(r"(?:void\s+)?(?:[\w:]+::)?(\w+)(<.*>)?", base)
    if not m:
        return base
    targs = m.group(2) or ""
    return m.group(1) + (targs if len(targs) <= 40 else "<...>")


def trace(db_path):
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, duration from kernels").fetchall()
    agg = defaultdict(list)
    for name, dur in rows:
        agg[short(name)].append(dur)
    total_all = sum(sum(v) for v in agg.values())
    out = []
    for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        out.append((k, len(v), sum(v) / 1e3, sum(v) / len(v) / 1e3, min(v) / 1e3, max(v) / 1e3, sum(v) / total_all))
    return out


def pmc(db_path):
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, counter_name, counter_value from pmc_events").fetchall()
    agg = defaultdict(list)
    for name, counter, value in rows:
        agg[(short(name), counter)].append(value)
    return {k: (len(v), sum(v) / len(v)) for k, v in agg.items()}


def scale_of(copy_db, counter, copy_bytes):
    """Counter units per byte moved, from the copy probe's dispatches."""
    for (k, c), (n, mean) in pmc(copy_db).items():
        if "copy_probe" in k and c == counter:
            return mean / copy_bytes
    raise SystemExit(f"no copy-probe dispatch with {counter} in {copy_db}")


def traffic(fetch_db, write_db, copy_fetch_db, copy_write_db, copy_bytes):
    fs = scale_of(copy_fetch_db, "FETCH_SIZE", copy_bytes)
    ws = scale_of(copy_write_db, "WRITE_SIZE", copy_bytes)
    out = {}
    for (k, c), (n, mean) in pmc(fetch_db).items():
        if c == "FETCH_SIZE":
            out.setdefault(k, {})["fetch_bytes"] = mean / fs
            out[k]["dispatches"] = n
    for (k, c), (n, mean) in pmc(write_db).items():
        if c == "WRITE_SIZE":
            out.setdefault(k, {})["write_bytes"] = mean / ws
    for v in out.values():
        v["bytes"] = v.get("fetch_bytes", 0.0) + v.get("write_bytes", 0.0)
    return {"fetch_scale": fs, "write_scale": ws, "kernels": out}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "pmc", "traffic"])
    ap.add_argument("db")
    ap.add_argument("--copy")
    ap.add_argument("--copy-write")
    ap.add_argument("--write-db")
    ap.add_argument("--copy-bytes", type=float, default=0)
    ap.add_argument("--json")
    a = ap.parse_args(argv)
    if a.mode == "traffic":
        t = traffic(a.db, a.write_db, a.copy, a.copy_write, a.copy_bytes)
        print(f"# FETCH_SIZE units per byte {t['fetch_scale']:.6g}, WRITE_SIZE units per byte {t['write_scale']:.6g}")
        print(f"{'kernel':58s} {'dispatches':>10s} {'fetch_MB':>10s} {'write_MB':>10s} {'total_MB':>10s}")
        for k, v in sorted(t["kernels"].items(), key=lambda kv: -kv[1]["bytes"]):
            print(f"{k:58s} {v.get('dispatches', 0):10d} {v.get('fetch_bytes', 0) / 1e6:10.2f} "
                  f"{v.get('write_bytes', 0) / 1e6:10.2f} {v['bytes'] / 1e6:10.2f}")
        if a.json:
            with open(a.json, "w") as f:
                json.dump(t, f, indent=1, sort_keys=True)
        return
    if a.mode == "trace":
        print(f"{'kernel':58s} {'calls':>7s} {'total_us':>12s} {'mean_us':>10s} {'min_us':>10s} {'max_us':>10s} {'share':>6s}")
        for k, n, tot, mean, mn, mx, share in trace(a.db):
            print(f"{k:58s} {n:7d} {tot:12.1f} {mean:10.2f} {mn:10.2f} {mx:10.2f} {share:6.1%}")
        return
    res = pmc(a.db)
    scale = {}
    if a.copy:
        for (k, counter), (n, mean) in pmc(a.copy).items():
            if "copy_probe" in k and a.copy_bytes:
                scale[counter] = mean / a.copy_bytes
                print(f"# copy probe: {counter} mean {mean:.1f} per dispatch for {a.copy_bytes:.0f} bytes "
                      f"-> {mean / a.copy_bytes:.6g} counter units per byte")
    print(f"{'kernel':58s} {'counter':>12s} {'dispatches':>10s} {'mean':>16s} {'bytes/dispatch':>16s}")
    for (k, counter), (n, mean) in sorted(res.items()):
        b = mean / scale[counter] if counter in scale and scale[counter] else float("nan")
        print(f"{k:58s} {counter:>12s} {n:10d} {mean:16.1f} {b:16.4g}")


if __name__ == "__main__":
    sys.exit(main())
