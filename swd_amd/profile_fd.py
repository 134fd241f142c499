"""Workload and summaries of the FD kernel profiles.

Run under rocprofv3 (one counter per pass):

    rocprofv3 --pmc FETCH_SIZE -d <dir>/fetch --output-format csv -- python -m swd_amd.profile_fd run
    rocprofv3 --pmc WRITE_SIZE -d <dir>/write --output-format csv -- python -m swd_amd.profile_fd run
    rocprofv3 --kernel-trace --stats -d <dir>/trace --output-format csv -- python -m swd_amd.profile_fd run
    rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU -d <dir>/sq --output-format csv -- python -m swd_amd.profile_fd run

``run`` copies a known byte count with the calibration kernel, then computes
one gradient of the production-shots workload (60 shots, 132 x 660 padded,
nt 6000) with fixed forward / adjoint launch shapes.  ``summarize <dir>``
reads the counter CSVs, scales the FD kernels' FETCH_SIZE / WRITE_SIZE by the
copy kernel's (bytes known), and writes the per-launch HBM traffic of each
time-blocked kernel, and its VALU instruction count per launch when the
``sq`` pass is there, to ``fd_traffic.json`` beside this file (read by the
benchmark; records of earlier runs at other launch shapes are kept, keyed by
kernel and shape, e.g. ``fd_fwd_tb:6x26``) and a markdown table to stdout.
``--shapes`` names the shapes of the run (default: both phases at 5 x 26).
``trace <dir>`` reads a kernel trace of a benchmark run and prints the mean,
min and max duration of the last dispatches of each time-blocked kernel, the
ones of the timed steps.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TRAFFIC_JSON = os.path.join(HERE, "fd_traffic.json")
COPY_BYTES = 1 << 30


def _shape(s):
    a, b = (int(x) for x in s.split(","))
    return a, b


def run(args):
    import ctypes
    from ._lib import check, load
    from . import problems as P
    from .engine import Engine
    ms = ctypes.c_double(0)
    check(load().swd_calib_copy(COPY_BYTES, args.copies, ctypes.byref(ms)))
    print(f"copy of {COPY_BYTES} B: {ms.value:.3f} ms ({2 * COPY_BYTES / ms.value / 1e6:.0f} GB/s)", flush=True)
    prob, mod = P.production_shots(nshots=args.shots, nt=args.nt)
    e = Engine(prob)
    e.set_wavelet(mod["s"])
    e.set_observed_picks(np.ones((e.npair, e.nshots)), np.ones((e.npair, e.nshots)))
    e.set_model(mod["vp"], mod["vs"])
    (nf, hf), (na, ha) = _shape(args.fwd), _shape(args.adj)
    e.set_time_block(nf, hf, phase=0)
    e.set_time_block(na, ha, phase=1)
    for _ in range(args.gradients):
        e.gradient()
    t = e.timings()
    info = e.info()
    e.close()
    print(json.dumps({"fwd": [nf, hf], "adj": [na, ha], "fwd_kernel_us": t["fwd_kernel_us"],
                      "adj_kernel_us": t["adj_kernel_us"], "nzbc": info["nzbc"], "nxbc": info["nxbc"],
                      "shots": args.shots, "nt": args.nt}), flush=True)


def _rows(d):
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no counter_collection.csv under {d}")
    for f in files:
        with open(f, newline="") as fh:
            yield from csv.DictReader(fh)


def _per_kernel(d, counter):
    """Mean counter value (KiB) per dispatch of each kernel, summed over the
    dimensions rocprofv3 reports it in."""
    per = {}
    for r in _rows(d):
        if r.get("Counter_Name") != counter:
            continue
        key = (r["Kernel_Name"], r.get("Dispatch_Id") or r.get("Correlation_Id"))
        per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
    out = {}
    for (k, _), v in per.items():
        out.setdefault(k, []).append(v)
    return {k: (float(np.mean(v)), len(v)) for k, v in out.items()}


def _short(name):
    for k in ("fd_fwd_tb", "fd_adj_tb", "fd_fwd_march", "fd_adj_march", "calib_copy"):
        if k in name:
            return k
    return None


def summarize(args):
    fetch = _per_kernel(os.path.join(args.dir, "fetch"), "FETCH_SIZE")
    write = _per_kernel(os.path.join(args.dir, "write"), "WRITE_SIZE")
    cf = [v for k, v in fetch.items() if _short(k) == "calib_copy"]
    cw = [v for k, v in write.items() if _short(k) == "calib_copy"]
    if not cf or not cw:
        raise RuntimeError("calibration copy kernel missing from the counter files")
    kib = COPY_BYTES / 1024.0
    f_fetch, f_write = kib / cf[0][0], kib / cw[0][0]
    res = {"copy_bytes": COPY_BYTES, "fetch_scale": f_fetch, "write_scale": f_write,
           "copy_fetch_kib": cf[0][0], "copy_write_kib": cw[0][0], "kernels": {}}
    lines = ["| kernel | dispatches | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch | HBM bytes/dispatch |",
             "|---|---|---|---|---|"]
    for k, (fv, n) in sorted(fetch.items()):
        s = _short(k)
        if s is None or s == "calib_copy":
            continue
        wv = write.get(k, (0.0, 0))[0]
        b = 1024.0 * (fv * f_fetch + wv * f_write)
        res["kernels"][s] = {"fetch_kib": fv, "write_kib": wv, "bytes_per_launch": b, "dispatches": n,
                             "name": k}
        lines.append(f"| {s} | {n} | {fv:.0f} | {wv:.0f} | {b / 1e6:.1f} MB |")
    sq_dir = os.path.join(args.dir, "sq")
    if os.path.isdir(sq_dir):
        valu = _per_kernel(sq_dir, "SQ_INSTS_VALU")
        for k, (v, n) in valu.items():
            s = _short(k)
            if s in res["kernels"] and res["kernels"][s]["name"] == k:
                res["kernels"][s]["valu_insts"] = v
                lines.append(f"| {s} | {n} | VALU instructions per dispatch {v / 1e6:.2f} M | | |")
    if args.shapes:
        res["shapes"] = json.loads(args.shapes)
    # records by kernel and launch shape, merged with those of earlier runs at
    # other shapes (the benchmark looks up the shape its autotuner picked)
    by_shape = {}
    if os.path.exists(TRAFFIC_JSON):
        with open(TRAFFIC_JSON) as fh:
            by_shape = json.load(fh).get("by_shape", {})
    for k, rec in res["kernels"].items():
        shp = res.get("shapes", {}).get(k)
        if shp:
            by_shape[f"{k}:{shp[0]}x{shp[1]}"] = dict(rec, copy_fetch_kib=cf[0][0], copy_write_kib=cw[0][0])
    res["by_shape"] = by_shape
    with open(TRAFFIC_JSON, "w") as fh:
        json.dump(res, fh, indent=1)
    print(f"calibration: copy of {COPY_BYTES} B reads {cf[0][0]:.0f} KiB (scale {f_fetch:.4f}), "
          f"writes {cw[0][0]:.0f} KiB (scale {f_write:.4f})")
    print("\n".join(lines))


def trace(args):
    """Durations of the last `--last` dispatches of each time-blocked FD
    kernel in a rocprofv3 kernel trace, in start order: the dispatches of the
    timed steps of a benchmark run, whose autotuning launches come earlier."""
    files = glob.glob(os.path.join(args.dir, "**", "*kernel_trace.csv"), recursive=True)
    if not files:
        raise FileNotFoundError(f"no kernel_trace.csv under {args.dir}")
    rows = []
    for f in files:
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                s = _short(r["Kernel_Name"])
                if s in ("fd_fwd_tb", "fd_adj_tb"):
                    t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
                    rows.append((t0, s, (t1 - t0) / 1e3))
    rows.sort()
    out = {}
    for k, n in (("fd_fwd_tb", args.last_fwd), ("fd_adj_tb", args.last_adj)):
        us = np.array([d for _, s, d in rows if s == k][-n:])
        out[k] = {"dispatches": int(us.size), "of": sum(1 for _, s, _ in rows if s == k),
                  "mean_us": round(float(us.mean()), 2) if us.size else None,
                  "min_us": round(float(us.min()), 2) if us.size else None,
                  "max_us": round(float(us.max()), 2) if us.size else None}
    print(json.dumps(out))


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--fwd", default="5,26", help="forward steps per launch, band rows")
    r.add_argument("--adj", default="5,26", help="adjoint steps per launch, band rows")
    r.add_argument("--shots", type=int, default=60)
    r.add_argument("--nt", type=int, default=6000)
    r.add_argument("--gradients", type=int, default=1)
    r.add_argument("--copies", type=int, default=5)
    s = sub.add_parser("summarize")
    s.add_argument("dir")
    s.add_argument("--shapes", default='{"fd_fwd_tb": [5, 26], "fd_adj_tb": [5, 26]}')
    t = sub.add_parser("trace")
    t.add_argument("dir")
    t.add_argument("--last-fwd", type=int, default=6000, help="forward dispatches to keep (5 steps x 1200)")
    t.add_argument("--last-adj", type=int, default=5995, help="adjoint dispatches to keep (5 steps x 1199)")
    a = ap.parse_args(argv)
    {"run": run, "summarize": summarize, "trace": trace}[a.cmd](a)


if __name__ == "__main__":
    main()
