#!/usr/bin/env python3
"""Summarise rocprofv3 results (rocpd SQLite databases) for libkgstencil.so.

    python3 kgprof.py --trace T.db [--fetch F.db --write W.db] [--cmd "..."]
                      [--out summary.txt] [--traffic-json traffic.json]

Per kernel: launches, mean duration (kernel trace), and -- when the two PMC
databases are given -- HBM bytes per launch from FETCH_SIZE and WRITE_SIZE
(kilobytes per dispatch), each scaled by the factor that makes the
calibration copy kernel (stream_copy: reads and writes exactly its buffer
size) read its known byte count.  Algorithmic bytes per launch use the
suite's bytes per grid point (each input array read once, each output written
once) at the benchmark size.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re
import sqlite3
import sys

N = 512 * 256 * 256
BYTES_PER_POINT = {
    "laplacian": 16, "divergence": 32, "gradient": 32, "wave13pt": 24, "uxx1": 48,
    "tricubic": 40, "tricubic2": 40, "lapgsrb": 16, "jacobi": 16, "gaussblur": 16,
    "whispering": 96, "gameoflife": 16,
}


def short_name(full: str) -> str:
    """Map a demangled kernel name to the suite's stencil name."""
    for key, pat in [("tricubic2", r"tricubic_k<2, 2, true>|Tricubic2Op|TricubicOpT<[^>]*true"),
                     ("tricubic", r"tricubic_k<1, 2, false>|TricubicOp"),
                     ("laplacian", r"laplacian_k|LaplacianOp"), ("gradient", r"gradient_k|GradientOp"),
                     ("wave13pt", r"wave13pt_k|Wave13ptOp"), ("lapgsrb", r"lapgsrb_k|LapgsrbOp"),
                     ("divergence", r"divergence2?_k"), ("uxx1", r"uxx1(_2)?_k"),
                     ("jacobi", r"jacobi_k|JacobiOp"), ("gaussblur", r"gaussblur_k|GaussblurOp"),
                     ("whispering", r"whispering_k|WhisperingOp"),
                     ("gameoflife", r"gameoflife_k|GameoflifeOp"),
                     ("stream_copy", r"stream_copy_k"), ("matmul", r"matmul_k")]:
        if re.search(pat, full):
            return key
    return full.split("(")[0][:40]


def durations(path: str) -> dict:
    """Kernel durations (ns) per kernel from a rocpd database or a
    *_kernel_trace.csv file."""
    out = collections.defaultdict(list)
    if path.endswith(".csv"):
        with open(path) as fh:
            for r in csv.DictReader(fh):
                out[short_name(r["Kernel_Name"])].append(
                    int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
        return out
    con = sqlite3.connect(path)
    for name, dur in con.execute("select name, duration from kernels"):
        out[short_name(name)].append(dur)
    return out


def counters(path: str, counter: str) -> dict:
    """Per-dispatch values of a counter (summed over its dimensions) per
    kernel, from a rocpd database or a *_counter_collection.csv file."""
    out = collections.defaultdict(list)
    if path.endswith(".csv"):
        per = collections.defaultdict(float)
        names = {}
        with open(path) as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] == counter:
                    per[r["Dispatch_Id"]] += float(r["Counter_Value"])
                    names[r["Dispatch_Id"]] = r["Kernel_Name"]
        for d, v in per.items():
            out[short_name(names[d])].append(v)
        return out
    con = sqlite3.connect(path)
    for name, val in con.execute(
            "select kernel_name, value from counters_collection where counter_name = ?", (counter,)):
        out[short_name(name)].append(val)
    return out


def main(argv=None) -> int:
    global N
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--cmd", default="")
    ap.add_argument("--copy-bytes", type=float, default=2 * N * 8,
                    help="bytes the calibration copy reads (= writes) per launch")
    ap.add_argument("--scale", type=float, nargs=2, default=None, metavar=("FETCH", "WRITE"),
                    help="counter scales when the trace has no calibration copy "
                         "(measured on MI355X with stream_copy: 2.0 1.0)")
    ap.add_argument("--nx", type=int, default=None, help="grid points per launch (default 512x256x256)")
    ap.add_argument("--match", help="report only kernels whose full name contains this text, "
                    "as one row named by it (e.g. a batch's stepping kernel)")
    ap.add_argument("--alg-bytes", type=float, help="algorithmic bytes per launch of --match")
    ap.add_argument("--out")
    ap.add_argument("--traffic-json")
    a = ap.parse_args(argv)
    if a.match:
        global short_name
        plain = short_name
        short_name = lambda full: a.match if a.match in full else "other:" + plain(full)
        BYTES_PER_POINT.clear()
    if a.nx:
        N = a.nx

    dur = durations(a.trace)
    fetch = counters(a.fetch, "FETCH_SIZE") if a.fetch else {}
    write = counters(a.write, "WRITE_SIZE") if a.write else {}
    fscale = wscale = None
    if fetch.get("stream_copy") and write.get("stream_copy"):
        mean = lambda v: sum(v) / len(v)
        fscale = a.copy_bytes / (mean(fetch["stream_copy"]) * 1024)
        wscale = a.copy_bytes / (mean(write["stream_copy"]) * 1024)
    elif a.scale and fetch and write:
        fscale, wscale = a.scale

    lines = []
    if a.cmd:
        lines.append(f"command: {a.cmd}")
    if fscale and a.scale and not fetch.get("stream_copy"):
        lines.append(f"counter scales (measured with the calibration copy on MI355X): "
                     f"FETCH_SIZE x {fscale:.4f}, WRITE_SIZE x {wscale:.4f}")
    elif fscale:
        lines.append(f"counter calibration (stream_copy, {a.copy_bytes / 2**20:.0f} MiB read + written): "
                     f"FETCH_SIZE x {fscale:.4f}, WRITE_SIZE x {wscale:.4f}")
    lines.append(f"{'kernel':12s} {'launches':>8s} {'mean us':>9s} {'alg MB':>8s} {'alg GB/s':>9s} "
                 f"{'%8TB/s':>7s} {'read MB':>8s} {'write MB':>9s} {'traffic/alg':>11s}")
    traffic = {}
    for k in sorted(dur, key=lambda k: -sum(dur[k])):
        d = dur[k]
        us = sum(d) / len(d) / 1e3
        if a.match and k.startswith("other:"):
            continue
        bpp = BYTES_PER_POINT.get(k)
        alg = bpp * N if bpp else (2 * a.copy_bytes if k == "stream_copy" else None)
        if a.match and k == a.match:
            alg = a.alg_bytes
        rd = wr = ratio = None
        if fscale and k in fetch and k in write:
            rd = sum(fetch[k]) / len(fetch[k]) * 1024 * fscale
            wr = sum(write[k]) / len(write[k]) * 1024 * wscale
            if alg:
                ratio = (rd + wr) / alg
                traffic[k] = {"read_bytes": round(rd), "write_bytes": round(wr),
                              "traffic_bytes": round(rd + wr), "algorithmic_bytes": alg,
                              "mean_launch_us": round(us, 2)}
        f = lambda v, w, p: (f"{v:{w}.{p}f}" if v is not None else " " * (w - 1) + "-")
        gbs = alg / (us * 1e3) if alg else None
        lines.append(f"{k:12s} {len(d):8d} {us:9.1f} {f(alg / 1e6 if alg else None, 8, 1)} "
                     f"{f(gbs, 9, 1)} {f(gbs / 80 if gbs else None, 7, 1)} "
                     f"{f(rd / 1e6 if rd else None, 8, 1)} {f(wr / 1e6 if wr else None, 9, 1)} "
                     f"{f(ratio, 11, 3)}")
    text = "\n".join(lines) + "\n"
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(text)
    if a.traffic_json and traffic:
        with open(a.traffic_json, "w") as fh:
            json.dump(traffic, fh, indent=1)
    sys.stdout.write(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
