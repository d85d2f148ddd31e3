#!/usr/bin/env python3
"""Throughput of the FP64 stencil suite on one MI355X.

Workload "suite-fp64": one step = one launch of each of the twelve FP64
stencils of the suite at the benchmark configuration (3D tests at
512x256x256, 2D tests at 512x65536 -- the reference's `./benchmark 512 256
256 ...` sizes), inputs already resident in HBM, ping-pong buffers swapped
between steps as the reference's main() does.  `value` is grid points
processed per second over the timed region (12 x 33,554,432 points per step).

Per-kernel durations come from HIP events recorded around every launch on
the launching stream (kg_timing_drain).  The roofline entry is for the kernel
with the largest share of the step: algorithmic bytes per point (each array
read once, each output written once) x points / mean launch duration, against
8 TB/s.  The CPU baseline times the reference's own kernel functions
(oracle/_ref/libref_fast.so: compiled from the reference sources with the
reference's gcc flags) on one host core over a bounded sample.

The line also carries `layers`: the batched, served, precision and Poisson
layers beside the FP64 drop-in, each at its own configuration (4096 grids x
100 steps FP64 and FP32; Poisson-arrival serving at several loads; 512^3 x
1000 steps FP64/FP32/BF16 with errors against FP64; Poisson 8192^2 to 1e-10),
and the uint8 Life throughput at 8192^2.  `--no-layers` skips them.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-layers]
"""
from __future__ import annotations

import argparse
import ctypes as C
import importlib.util
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
PKG = ROOT / "kernelgen-perf-tests_amd"
HBM_PEAK_GBS = 8000.0
METRIC = "grid-pts/s + HBM GB/s (% of 8 TB/s) per stencil; served p50/p99 latency"

NX, NY, NS = 512, 256, 256
N = NX * NY * NS
DIMS3 = (NX, NY, NS)
DIMS2 = (NX, NY * NS)

# name -> (dims, algorithmic bytes per point, coefficient values, array roles,
#          pairs rotated between steps).  Coefficients: the reference draws
# them from rand(); fixed representative values are used here (timing does
# not depend on them).
STENCILS = {
    "laplacian": (DIMS3, 16, (0.1, 0.2), ("w0", "w1"), [("w0", "w1")]),
    "divergence": (DIMS3, 32, (0.5, -0.25, 0.75), ("u", "ux", "uy", "uz"), []),
    "gradient": (DIMS3, 32, (0.5, -0.25, 0.75), ("u", "ux", "uy", "uz"), []),
    "wave13pt": (DIMS3, 24, (0.6, 0.1, -0.05), ("w0", "w1", "w2"), [("w0", "w1", "w2")]),
    "uxx1": (DIMS3, 48, (0.5, -0.3), ("u0", "u1", "d1", "xx", "xy", "xz"), [("u0", "u1")]),
    "tricubic": (DIMS3, 40, (), ("u0", "u1", "a", "b", "c"), [("u0", "u1")]),
    "tricubic2": (DIMS3, 40, (), ("u0", "u1", "a", "b", "c"), [("u0", "u1")]),
    "lapgsrb": (DIMS3, 16, (0.5, 0.05, -0.02, 0.03), ("w0", "w1"), [("w0", "w1")]),
    "jacobi": (DIMS2, 16, (0.5, 0.1, 0.02), ("w0", "w1"), [("w0", "w1")]),
    "gaussblur": (DIMS2, 16, (0.4, 0.3, 0.2, 0.1, 0.05, 0.02), ("w0", "w1"), [("w0", "w1")]),
    "whispering": (DIMS2, 96, (0.7, 0.3),
                   ("e", "e_next", "h0", "h1", "u_em0", "u_em1", "ca", "cb", "da", "db"),
                   [("e", "e_next"), ("h0", "h1"), ("u_em0", "u_em1")]),
    "gameoflife": (DIMS2, 16, (), ("u0", "u1"), [("u0", "u1")]),
}
# Kernel symbol the C ABI reports for each entry point.
KNAME = {"jacobi": "jacobi_"}
SIZES = {"e": 2, "e_next": 2}


def load_pkg():
    spec = importlib.util.spec_from_file_location("kgstencil", PKG / "kgstencil.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def rotate(bufs: dict, pairs) -> None:
    for grp in pairs:
        vals = [bufs[r] for r in grp]
        for r, v in zip(grp, vals[1:] + vals[:1]):
            bufs[r] = v


def make_inputs(rng, roles, sizes):
    out = {}
    for r in roles:
        n = N * sizes.get(r, 1)
        a = rng.uniform(-1.0, 1.0, n)
        if r == "d1":
            a = np.abs(a) + 1.0          # density in [1,2): dth/d stays finite
        out[r] = a
    return out


def calibrate(ctx, kg, reps: int = 5):
    """Plain copy of 2 x 256 MiB (exactly 512 MiB of HBM traffic per launch):
    achievable bandwidth, and the yardstick for scaling FETCH_SIZE/WRITE_SIZE."""
    n = 2 * N
    src = ctx.empty(n)
    dst = ctx.empty(n)
    ctx.call("stream_copy", dst, src, n)     # first touch
    ctx.sync()
    ctx.drain()
    for _ in range(reps):
        ctx.call("stream_copy", dst, src, n)
    t = float(np.mean(ctx.drain()))
    src.free()
    dst.free()
    return {"bytes_per_launch": 16 * n, "mean_launch_us": round(t * 1e6, 2),
            "gbs": round(16 * n / t / 1e9, 1)}


def gpu_leg(kg, steps: int, warmup: int, rng, do_calibrate=False):
    ctx = kg.Context(0)
    cal = calibrate(ctx, kg) if do_calibrate else None
    bufs_by_stencil = {}
    host_pool = {}
    for name, (dims, bpp, coefs, roles, pairs) in STENCILS.items():
        for r in roles:
            key = (r, SIZES.get(r, 1))
            if key not in host_pool:
                host_pool[key] = make_inputs(rng, [r], SIZES)[r]
        bufs_by_stencil[name] = {r: ctx.array(host_pool[(r, SIZES.get(r, 1))]) for r in roles}
    if "whispering" in bufs_by_stencil:   # identical e borders in both e buffers
        b = bufs_by_stencil["whispering"]
        b["e_next"].copy_from(b["e"])
    host_pool.clear()

    def one_step():
        for name, (dims, bpp, coefs, roles, pairs) in STENCILS.items():
            b = bufs_by_stencil[name]
            ctx.call(name, *dims, *coefs, *[b[r] for r in roles])
            rotate(b, pairs)

    for _ in range(warmup):
        one_step()
    ctx.sync()
    ctx.drain()
    t0 = time.perf_counter()
    for _ in range(steps):
        one_step()
    ctx.sync()
    wall = time.perf_counter() - t0
    times = ctx.drain(cap=len(STENCILS) * steps + 16)
    assert len(times) == len(STENCILS) * steps, (len(times), steps)
    per = {}
    names = list(STENCILS)
    arr = np.array(times).reshape(steps, len(names))
    for idx, name in enumerate(names):
        dims, bpp, *_ = STENCILS[name]
        t = float(arr[:, idx].mean())
        per[name] = {
            "us": round(t * 1e6, 2),
            "gpts_s": round(N / t / 1e9, 2),
            "gbs": round(bpp * N / t / 1e9, 1),
            "frac_hbm": round(bpp * N / t / 1e9 / HBM_PEAK_GBS, 4),
            "bytes_per_pt": bpp,
            "regs": ctx.regcount(KNAME.get(name, name)),
        }
    ctx.close()
    return wall, per, cal


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(reps: int, rng, omp: bool = False):
    """Reference kernel functions (reference gcc flags) on one core, each
    stencil `reps` times at the full configuration.  omp: the same kernels
    built with -fopenmp (oracle/_ref/libref_omp.so, their `omp parallel for`
    active) on OMP_NUM_THREADS host threads; jacobi (Fortran) and whispering
    (whose reference OpenMP loop is a race) are left out of that build."""
    lib_path = ROOT / "oracle" / "_ref" / ("libref_omp.so" if omp else "libref_fast.so")
    kind = "reference"
    if not lib_path.exists():
        if omp:
            return None
        lib_path = ROOT / "oracle" / "liboracle.so"
        kind = "port"
    if not lib_path.exists():
        return None
    lib = C.CDLL(str(lib_path))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)) if omp else 1
    pool = {}
    total_t = 0.0
    per = {}
    for name, (dims, bpp, coefs, roles, pairs) in STENCILS.items():
        if omp and name in ("jacobi", "whispering"):
            continue
        roles = [r for r in roles if r != "e_next"]
        arrs = {}
        for r in roles:
            key = (r, SIZES.get(r, 1))
            if key not in pool:
                pool[key] = make_inputs(rng, [r], SIZES)[r]
            arrs[r] = pool[key]
        fn_name = name + ("_" if name == "jacobi" else "") if kind == "reference" else "oracle_" + name
        fn = getattr(lib, fn_name)
        fn.restype = None
        t0 = time.perf_counter()
        for _ in range(reps):
            if kind == "reference" and name == "jacobi":
                ints = [C.c_int(d) for d in dims]
                cs = [C.c_double(c) for c in coefs]
                fn(*[C.byref(x) for x in ints], *[C.byref(x) for x in cs],
                   *[C.c_void_p(arrs[r].ctypes.data) for r in roles])
            else:
                fn(*[C.c_int(d) for d in dims], *[C.c_double(c) for c in coefs],
                   *[C.c_void_p(arrs[r].ctypes.data) for r in roles])
            rotate(arrs, [p for p in pairs if "e_next" not in p])
        dt = (time.perf_counter() - t0) / reps
        per[name] = round(N / dt / 1e9, 4)
        total_t += dt
    pool.clear()
    if not per:
        return None
    return {
        "value": round(len(per) * N / total_t, 1),
        "unit": "grid-pts/s",
        "cores": threads,
        "kind": kind,
        "cpu": cpu_model(),
        "sample": f"each of the {len(per)} stencils {sorted(per)} run {reps}x at 512x256x256 / "
                  f"512x65536 on {threads} host thread(s) ({lib_path.name}); "
                  f"per-stencil Gpts/s: {json.dumps(per)}",
    }


FP64_MFMA_PEAK_TFLOPS = 78.6


def matmul_leg(kg, reps: int = 20) -> dict:
    """matmul at the benchmark configuration (C(512,256) += A(512,256) B(256,256),
    FP64 on the matrix cores): mean launch time, achieved TFLOP/s and its
    fraction of the FP64 matrix peak."""
    nx, ny, ns = NX, NY, NS
    rng = np.random.default_rng(12)
    ctx = kg.Context(0)
    A = ctx.array(rng.uniform(-1, 1, nx * ny))
    B = ctx.array(rng.uniform(-1, 1, ny * ns))
    Cm = ctx.array(np.zeros(nx * ns))
    for _ in range(3):
        ctx.call("matmul", nx, ny, ns, A, B, Cm)
    ctx.sync()
    ctx.drain()
    for _ in range(reps):
        ctx.call("matmul", nx, ny, ns, A, B, Cm)
    t = float(np.mean(ctx.drain()))
    regs = ctx.regcount("matmul_")
    for a in (A, B, Cm):
        a.free()
    ctx.close()
    flop = 2.0 * nx * ny * ns
    return {"dims": [nx, ny, ns], "flop_per_launch": flop, "us": round(t * 1e6, 2),
            "tflops": round(flop / t / 1e12, 3),
            "frac_fp64_mfma_peak": round(flop / t / 1e12 / FP64_MFMA_PEAK_TFLOPS, 4), "regs": regs}


def load_layers():
    spec = importlib.util.spec_from_file_location("kglayers", PKG / "kglayers.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def life_leg(kg, ctx, n: int = 8192, reps: int = 20) -> dict:
    """uint8 Life (gameoflife X mode), n x n cells, Bernoulli p = 0.3."""
    rng = np.random.default_rng(5)
    board = (rng.random((n, n)) < 0.3).astype(np.uint8)
    a, b = ctx.array(board), ctx.array(board)
    for _ in range(3):
        ctx.call("life_u8", n, n, a, b)
        a, b = b, a
    ctx.sync()
    ctx.drain()
    for _ in range(reps):
        ctx.call("life_u8", n, n, a, b)
        a, b = b, a
    t = float(np.mean(ctx.drain()))
    a.free()
    b.free()
    return {"cells": n * n, "us": round(t * 1e6, 2), "gcells_s": round(n * n / t / 1e9, 1),
            "gbs": round(2 * n * n / t / 1e9, 1), "frac_hbm": round(2 * n * n / t / 1e9 / HBM_PEAK_GBS, 4)}


def plain_rate(kg, ctx, fn: str, coefs, nlev: int, reps: int = 10) -> float:
    """Grid points per second of one single-grid FP64 entry point at
    512x256x256 (levels rotated between launches)."""
    rng = np.random.default_rng(3)
    bufs = [ctx.array(rng.uniform(-1, 1, N)) for _ in range(nlev)]
    for _ in range(2):
        ctx.call(fn, *DIMS3, *coefs, *bufs)
        bufs = bufs[1:] + bufs[:1]
    ctx.sync()
    ctx.drain()
    for _ in range(reps):
        ctx.call(fn, *DIMS3, *coefs, *bufs)
        bufs = bufs[1:] + bufs[:1]
    t = float(np.mean(ctx.drain()))
    for b in bufs:
        b.free()
    return N / t


def layers_leg(kg, suite_per: dict) -> dict:
    """The layers beside the drop-in kernels, summarised for the JSON line."""
    L = load_layers()
    ctx = kg.Context(0)
    out = {}
    try:
        out["life_u8_8192"] = life_leg(kg, ctx)
        plain = suite_per.get("laplacian", {}).get("gpts_s")
        for name, dt in (("batch_f64", kg.F64), ("batch_f32", kg.F32)):
            r = L.run_batch(kg, ctx, dt)
            r.pop("shapes", None)
            out[name] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
        if plain:
            out["batch_f32_vs_plain_fp64_laplacian_512x256x256"] = round(
                out["batch_f32"]["gpts_s"] / plain, 3)
            out["batch_f32_vs_sequential_fp64_launches"] = round(
                out["batch_f32"]["gpts_s"] / out["batch_f64"]["sequential_gpts_s"], 2)
            # The same mix of kinds done by the plain FP64 single-grid kernels
            # at 512x256x256: the batch's points of each kind at that kind's
            # plain rate.
            rates = {"laplacian": plain * 1e9,
                     "wave13pt": suite_per["wave13pt"]["gpts_s"] * 1e9,
                     "gaussblur3": plain_rate(kg, ctx, "gaussblur3_f64", L.GAUSS, 2)}
            pts = {k: 0 for k in rates}
            for kind, (nx, ny, ns) in L.batch_grids(1, 4096):
                pts[kind] += nx * ny * ns
            t_plain = sum(pts[k] / rates[k] for k in pts)
            mix = sum(pts.values()) / t_plain
            out["plain_fp64_same_kinds_gpts_s"] = {
                "mix": round(mix / 1e9, 2), "points_share": {k: round(v / sum(pts.values()), 4)
                                                              for k, v in pts.items()},
                **{k: round(v / 1e9, 2) for k, v in rates.items()}}
            out["batch_f32_vs_plain_fp64_same_kinds"] = round(out["batch_f32"]["gpts_s"] * 1e9 / mix, 3)
        s = L.run_serve(kg, ctx)
        out["serve"] = {
            "jobs": s["jobs"], "budget_bytes": s["budget_bytes"],
            "capacity_jobs_s": round(s["capacity_jobs_s"], 1),
            "loads": [{"offered_load": r["offered_load"],
                       **{m: {k: round(r[m][k], 4) for k in ("p50_s", "p99_s", "gpts_s")}
                          for m in ("continuous", "sequential")},
                       "peak_active": r["continuous"]["peak_active"],
                       "checked": r["checked"]} for r in s["loads"]]}
        p = L.run_precision(kg, ctx)
        out["precision_512cubed_1000_steps"] = {
            kind: {"final": {d: {k: float("%.3g" % v) for k, v in p[kind]["checks"][-1][d].items()}
                             for d in ("f32", "bf16")},
                   "gpts_s": {d: round(p[kind]["timing"][d]["gpts_s"], 1) for d in p[kind]["timing"]},
                   "coef": p[kind]["coef"]}
            for kind in ("laplacian", "wave13pt")}
        q = L.run_poisson(kg, ctx)
        out["poisson_8192"] = {k: q[k] for k in ("outer", "vcycles", "levels", "rel_residual",
                                                 "seconds")}
    except Exception as e:   # a layer failure is reported, not hidden
        out["error"] = f"{type(e).__name__}: {e}"
    finally:
        ctx.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="also time a plain 512 MiB device copy (reported as copy_calibration)")
    ap.add_argument("--only", default="",
                    help="comma-separated subset of stencils (profiling runs)")
    ap.add_argument("--no-layers", action="store_true",
                    help="skip the batch / serving / precision / Poisson / uint8 Life layers")
    args = ap.parse_args()
    if args.only:
        keep = args.only.split(",")
        for k in list(STENCILS):
            if k not in keep:
                del STENCILS[k]
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the stencil path runs on a single device "
              "(multi-GPU z-slab sharding is not built)",
              file=sys.stderr)
        return 2
    rng = np.random.default_rng(2024)
    kg = load_pkg()
    wall, per, cal = gpu_leg(kg, args.steps, args.warmup, rng, args.calibrate)
    value = len(STENCILS) * N * args.steps / wall
    dom = max(per, key=lambda k: per[k]["us"])
    d = per[dom]
    # HBM bytes per launch of the dominant kernel, from the committed
    # rocprofv3 FETCH_SIZE/WRITE_SIZE passes (scaled by the copy kernel).
    traffic = None
    tfile = ROOT / "profiles" / "suite_fp64_512x256x256_traffic.json"
    if tfile.exists():
        entry = json.loads(tfile.read_text()).get(dom)
        if entry:
            traffic = entry["traffic_bytes"]
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "grid-pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic uniform[-1,1) grids (uxx1 density in [1,2)), resident in HBM",
        "config": {"workload": "suite-fp64: 12 FP64 stencils x 1 step, 3D 512x256x256, "
                               "2D 512x65536",
                   "grid_3d": list(DIMS3), "grid_2d": list(DIMS2), "stencils": list(STENCILS)},
        "roofline": {"bound": "hbm", "kernel": dom,
                     "achieved": d["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(d["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_launch": d["bytes_per_pt"] * N, "mean_launch_us": d["us"]},
        "per_stencil": per,
    }
    if cal:
        line["copy_calibration"] = cal
    if not args.only:
        line["matmul"] = matmul_leg(kg)
    if "tricubic" in per:
        # tricubic is bound by FP64 issue: 252 multiplies and adds per point
        # (no contraction), against 39.3 T non-FMA FP64 ops/s (78.6 TFLOP/s
        # counts an FMA as two)
        t = per["tricubic"]["us"] * 1e-6
        per["tricubic"]["dp_ops_per_pt"] = 252
        per["tricubic"]["frac_fp64_issue"] = round(252 * N / t / 39.3e12, 4)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_leg(args.cpu_reps, rng)
        omp = cpu_leg(args.cpu_reps, rng, omp=True)
        if omp:
            line["cpu_baseline_omp"] = omp
    if not args.no_layers and not args.only:
        line["layers"] = layers_leg(kg, per)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
