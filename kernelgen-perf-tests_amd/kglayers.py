"""Measurement drivers for the layers beside the FP64 drop-in kernels:
batching, serving and precision, each at the configuration it is specified
for.  Every function takes the kgstencil module (``kg``) and an open
``kg.Context``; all computation runs in libkgstencil.so on the device.

* ``run_batch``: 4096 independent 3D grids, every edge log-uniform in
  [16, 256] from a seed, kinds drawn uniformly from laplacian / gaussblur
  (per xy-plane) / wave13pt, 100 steps under one captured graph, tile shapes
  autotuned per size class; the same grids stepped by one plain launch per
  grid and step for comparison.
* ``run_serve``: Poisson arrivals of jobs (kind uniform over laplacian,
  wave13pt, gaussblur, jacobi-to-1e-8 and gameoflife; volume log-uniform from
  32^3 to 512x256x256; steps log-uniform in [10, 1000]; FP64 or FP32 with
  p = 0.5) admitted against a 250 GB HBM budget, at several offered loads,
  continuous batching against one-job-at-a-time plain launches.
* ``run_precision``: laplacian and wave13pt at 512^3 for 1000 steps in FP64,
  FP32 and BF16 storage, with errors of the FP32/BF16 fields against the FP64
  run (bit-identical to the oracle) and FP64 norms/energy.
* ``run_poisson``: 2D Poisson on 8192^2 cells with a random right-hand side,
  FP32 multigrid sweeps inside FP64 defect correction to a relative residual
  of 1e-10.
"""
from __future__ import annotations

import math
import time

import numpy as np

# Coefficients chosen stable for long runs (see run_precision):
#  laplacian alpha = 1/4, beta = 1/8: amplification alpha + beta*sum(2 cos) in [-1/2, 1];
#  wave13pt C = (c dt/h)^2 = 3/32 of the 4th-order leapfrog (stable for C <= 1/4):
#   m0 = 2 - 7.5 C, m1 = 4C/3, m2 = -C/12, all exact in FP32 and BF16;
#  gaussblur sigma = 1: s(r^2) = exp(-r^2/2);
#  jacobi: the 9-point average c0 = 0, c1 = 1/5, c2 = 1/20.
LAP = (0.25, 0.125)
WAVE = (1.296875, 0.125, -0.0078125)
GAUSS = tuple(math.exp(-r2 / 2.0) for r2 in (0, 1, 2, 4, 5, 8))
JAC = (0.0, 0.2, 0.05)
COEFS = {"laplacian": LAP, "wave13pt": WAVE, "gaussblur3": GAUSS, "jacobi2": JAC, "life2": ()}
HBM_PEAK_GBS = 8000.0


def _log_uniform(rng, lo, hi, size=None):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size))


# ----------------------------------------------------------------- batching
def batch_grids(seed: int = 1, ngrids: int = 4096, lo: int = 16, hi: int = 256):
    rng = np.random.default_rng(seed)
    edges = np.minimum(np.floor(_log_uniform(rng, lo, hi + 1, (ngrids, 3))), hi).astype(int)
    kinds = rng.choice(["laplacian", "gaussblur3", "wave13pt"], ngrids)
    return [(str(k), tuple(int(v) for v in e)) for k, e in zip(kinds, edges)]


# single-grid entry point of each batch kind, per dtype
SINGLE = {("laplacian", 0): "laplacian", ("laplacian", 1): "laplacian_f32",
          ("wave13pt", 0): "wave13pt", ("wave13pt", 1): "wave13pt_f32",
          ("gaussblur3", 0): "gaussblur3_f64", ("gaussblur3", 1): "gaussblur3_f32"}


def check_sample(grids, seed: int, n: int = 24) -> list[int]:
    """Ids of n grids spread over the kinds (each kind's grids drawn from
    seed), the sample whose batched results are checked."""
    rng = np.random.default_rng(seed + 99)
    ids = []
    for kind in ("laplacian", "wave13pt", "gaussblur3"):
        pool = [i for i, (k, _) in enumerate(grids) if k == kind]
        ids += [int(i) for i in rng.choice(pool, min(len(pool), n // 3), replace=False)]
    return sorted(ids)


def single_grid_replay(kg, ctx, kind: str, dtype: int, dims, levels, steps: int) -> np.ndarray:
    """The grid stepped `steps` times by its kind's single-grid entry point
    from the given initial levels; returns the newest level (ns, ny, nx)."""
    fn = SINGLE[(kind, dtype)]
    bufs = [ctx.array(a.ravel()) for a in levels]
    for _ in range(steps):
        ctx.call(fn, *dims, *COEFS[kind], *bufs)
        bufs = bufs[1:] + bufs[:1]
    ctx.sync()
    # after the rotation the level written last is the one before the input
    # position: laplacian/gaussblur3 [in, out] -> [out, in]: bufs[0];
    # wave13pt [w0, w1, w2] -> [w1, w2, w0]: bufs[1]
    newest = bufs[0] if kind != "wave13pt" else bufs[1]
    out = newest.download().reshape(dims[2], dims[1], dims[0])
    for b in bufs:
        b.free()
    return out


def run_batch(kg, ctx, dtype: int, steps: int = 100, seed: int = 1, ngrids: int = 4096,
              tune_reps: int = 2, seq_steps: int = 3, check: bool = True,
              shapes: dict | None = None) -> dict:
    """check: after the timed steps, a sample of grids (check_sample) is
    compared bit for bit with the same grids stepped as often by the
    single-grid entry points from the same initial levels.  shapes: tile
    shapes {grid id: (lanes per row, planes per tile)} from an earlier tuned
    run (result key "shapes"), applied instead of autotuning -- profiling
    runs use it so that their traces hold the stepping launches only."""
    grids = batch_grids(seed, ngrids)
    b = kg.Batch(ctx)
    checked = None
    try:
        for kind, dims in grids:
            b.add(kind, dtype, dims, COEFS[kind])
        b.fill(seed)
        sample = check_sample(grids, seed) if check else []
        initial = {g: [b.download(g, l) for l in range(3 if grids[g][0] == "wave13pt" else 2)]
                   for g in sample}
        t0 = time.perf_counter()
        if shapes is None:
            ntuned = b.tune(tune_reps)
        else:
            ntuned = 0
            for g, (lpr, ppt) in shapes.items():
                b.set_shape(int(g), int(lpr), int(ppt))
        t_tune = time.perf_counter() - t0
        used = {g: b.shape(g) for g in range(len(grids))}
        info = b.info()
        reserved, in_use = b.memory()
        b.step(1, graph=False)                 # warm
        b.step(steps, graph=True)              # capture + first replay
        ctx.sync()
        t0 = time.perf_counter()
        b.step(steps, graph=True)
        ctx.sync()
        t = time.perf_counter() - t0
        t0 = time.perf_counter()
        if seq_steps > 0:
            b.step_sequential(seq_steps)
        ctx.sync()
        t_seq = time.perf_counter() - t0
        if check:
            bad = []
            for g in sample:
                kind, dims = grids[g]
                got = b.download(g)
                want = single_grid_replay(kg, ctx, kind, dtype, dims, initial[g], b.steps_done(g))
                if not np.array_equal(got.view(np.uint8), want.view(np.uint8)):
                    bad.append(g)
            checked = {"grids": len(sample), "steps": b.steps_done(sample[0]) if sample else 0,
                       "against": "single-grid entry points, same initial levels",
                       "bit_identical": not bad, "mismatched": bad}
    finally:
        b.close()
    pts = info["points_per_step"]
    return {
        "grids": ngrids, "steps": steps, "dtype": "f64" if dtype == kg.F64 else "f32",
        "points_per_step": pts, "hbm_bytes_resident": in_use, "arena_reserved": reserved,
        "work_items": info["work_items"], "workgroups": info["workgroups"],
        "classes_tuned": ntuned, "tune_s": round(t_tune, 3),
        "ms_per_step": 1e3 * t / steps,
        "gpts_s": pts * steps / t / 1e9,
        "gbs": info["bytes_per_step"] * steps / t / 1e9,
        "frac_hbm": info["bytes_per_step"] * steps / t / 1e9 / HBM_PEAK_GBS,
        "sequential_ms_per_step": 1e3 * t_seq / seq_steps if seq_steps > 0 else None,
        "sequential_gpts_s": pts * seq_steps / t_seq / 1e9 if seq_steps > 0 else None,
        "checked": checked,
        "shapes": used,
    }


# ------------------------------------------------------------------ serving
def serve_jobs(n: int, rate: float, seed: int = 7, vmin: float = 32 ** 3,
               vmax: float = 512 * 256 * 256) -> list[dict]:
    """n jobs with exponential inter-arrival times of mean 1/rate seconds."""
    rng = np.random.default_rng(seed)
    kinds = ["laplacian", "wave13pt", "gaussblur3", "jacobi2", "life2"]
    t = np.cumsum(rng.exponential(1.0 / rate, n)) if rate > 0 else np.zeros(n)
    jobs = []
    for i in range(n):
        kind = kinds[int(rng.integers(0, 5))]
        v = float(_log_uniform(rng, vmin, vmax))
        s = (v / (512 * 256 * 256)) ** (1.0 / 3.0)
        nx, ny, ns = max(3, round(512 * s)), max(3, round(256 * s)), max(3, round(256 * s))
        if kind in ("jacobi2", "life2"):       # 2D problem nx x (ny*ns), as the 2D tests run
            ny, ns = ny * ns, 1
        steps = int(round(float(_log_uniform(rng, 10, 1000))))
        dtype = 0 if rng.random() < 0.5 else 1
        jobs.append({"kind": kind, "dtype": dtype, "dims": (nx, ny, ns), "steps": steps,
                     "coef": COEFS[kind], "tol": 1e-8 if kind == "jacobi2" else 0.0,
                     "arrival": float(t[i]), "seed": 100 + i})
    return jobs


def run_serve(kg, ctx, njobs: int = 120, loads=(0.5, 0.8, 1.0, 1.5), seed: int = 7,
              budget: int = 250 * 10 ** 9, max_active: int = 256) -> dict:
    # capacity: the same jobs, all present at t = 0, continuous batching
    base = serve_jobs(njobs, 0.0, seed)
    _, st0, _ = kg.serve(ctx, base, budget, max_active)
    cap = njobs / st0["wall"]
    out = {"jobs": njobs, "budget_bytes": budget, "capacity_jobs_s": cap,
           "capacity_gpts_s": st0["points"] / st0["wall"] / 1e9, "loads": []}
    for load in loads:
        jobs = serve_jobs(njobs, load * cap, seed)
        row = {"offered_load": load, "rate_jobs_s": load * cap}
        sums = {}
        for name, mode in (("continuous", kg.SERVE_CONTINUOUS), ("sequential", kg.SERVE_SEQUENTIAL)):
            res, st, _ = kg.serve(ctx, jobs, budget, max_active, mode)
            ok = [r for r in res if r["status"] == 0]
            sums[name] = [(r["status"], r["steps_run"], r["checksum"]) for r in res]
            row[name] = {"p50_s": st["p50"], "p99_s": st["p99"], "mean_s": st["mean_latency"],
                         "gpts_s": st["points"] / st["wall"] / 1e9, "wall_s": st["wall"],
                         "peak_active": st["peak_active"], "peak_bytes": st["peak_bytes"],
                         "completed": len(ok)}
        # every job completed, and each job's final field (its FP64 sum and
        # sum of squares) is the same in the mix as when run alone
        row["checked"] = {"all_completed": all(row[m]["completed"] == njobs for m in sums),
                          "same_results_in_the_mix": sums["continuous"] == sums["sequential"]}
        out["loads"].append(row)
    return out


# ---------------------------------------------------------------- precision
def run_precision(kg, ctx, n: int = 512, steps: int = 1000, check_every: int = 250,
                  time_steps: int = 50, seed: int = 3) -> dict:
    """FP64 / FP32 / BF16 runs in lockstep from the same FP64 initial field;
    errors of the FP32 and BF16 fields against the FP64 one (max abs and
    relative L2), FP64 sums of squares (and, for wave13pt, the kinetic term
    sum (w_new - w_old)^2) at checkpoints, and time per step per dtype."""
    N = n * n * n
    out = {"grid": [n, n, n], "steps": steps}
    dts = [("f64", kg.F64), ("f32", kg.F32), ("bf16", kg.BF16)]
    for kind, nlev in (("laplacian", 2), ("wave13pt", 3)):
        co = COEFS[kind]
        lv = {}
        for name, code in dts:
            lv[name] = [ctx.empty((N,), {"f64": np.float64, "f32": np.float32,
                                         "bf16": np.uint16}[name]) for _ in range(nlev)]
        for l in range(nlev):
            kg_fill = ctx.lib.kg_fill_uniform
            _c = kg._check
            _c(ctx.lib, kg_fill(ctx.ptr, kg.F64, N, lv["f64"][l].ptr, seed * 10 + l), "fill")
            for name, code in dts[1:]:
                _c(ctx.lib, ctx.lib.kg_convert(ctx.ptr, code, lv[name][l].ptr, kg.F64,
                                               lv["f64"][l].ptr, N), "convert")
        fn = {"f64": kind, "f32": kind + "_f32", "bf16": kind + "_bf16"}

        def step(name, bufs):
            if kind == "wave13pt":
                ctx.call(fn[name], n, n, n, *co, bufs[0], bufs[1], bufs[2])
                return [bufs[1], bufs[2], bufs[0]]
            ctx.call(fn[name], n, n, n, *co, bufs[0], bufs[1])
            return [bufs[1], bufs[0]]

        def newest(bufs):
            return bufs[1] if kind == "wave13pt" else bufs[0]

        def older(bufs):
            return bufs[0] if kind == "wave13pt" else bufs[1]

        checks = []
        for s in range(1, steps + 1):
            for name, _ in dts:
                lv[name] = step(name, lv[name])
            if s % check_every == 0 or s == steps:
                ref = newest(lv["f64"])
                st64 = ctx.field_stats(ref, kg.F64)
                kin64 = ctx.field_stats(ref, kg.F64, older(lv["f64"]), kg.F64)["err_sumsq"]
                row = {"step": s, "f64_sumsq": st64["sumsq"], "f64_kinetic": kin64}
                for name, code in dts[1:]:
                    e = ctx.field_stats(newest(lv[name]), code, ref, kg.F64)
                    kin = ctx.field_stats(newest(lv[name]), code, older(lv[name]), code)["err_sumsq"]
                    row[name] = {"max_abs_err": e["err_max"],
                                 "rel_l2_err": math.sqrt(e["err_sumsq"] / max(st64["sumsq"], 1e-300)),
                                 "sumsq_rel_drift": e["sumsq"] / st64["sumsq"] - 1.0,
                                 "kinetic_rel_drift": (kin / kin64 - 1.0) if kin64 > 0 else 0.0}
                checks.append(row)
        timing = {}
        for name, _ in dts:
            bufs = lv[name]
            for _ in range(3):
                bufs = step(name, bufs)
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(time_steps):
                bufs = step(name, bufs)
            ctx.sync()
            t = (time.perf_counter() - t0) / time_steps
            bpp = {"f64": 8, "f32": 4, "bf16": 2}[name] * (3 if kind == "wave13pt" else 2)
            timing[name] = {"ms_per_step": 1e3 * t, "gpts_s": N / t / 1e9,
                            "gbs": N * bpp / t / 1e9, "frac_hbm": N * bpp / t / 1e9 / HBM_PEAK_GBS}
        for name, _ in dts:
            for buf in lv[name]:
                buf.free()
        out[kind] = {"coef": co, "checks": checks, "timing": timing}
    return out


# ------------------------------------------------------------------ Poisson
def run_poisson(kg, ctx, n: int = 8192, tol: float = 1e-10, seed: int = 11,
                vcycles: int = 2) -> dict:
    """2D Poisson on n x n cells ((n+1)^2 vertices), random right-hand side
    uniform in [-1, 1) (zero on the boundary), FP32 multigrid V-cycles inside
    FP64 defect correction to ||f - A u|| / ||f|| < tol."""
    N = n + 1
    f = ctx.empty((N * N,), np.float64)
    u = ctx.empty((N * N,), np.float64)
    kg._check(ctx.lib, ctx.lib.kg_fill_uniform(ctx.ptr, kg.F64, N * N, f.ptr, seed), "fill")
    # zero the boundary rows/columns of f and the initial guess
    fh = f.download().reshape(N, N)
    fh[0, :] = fh[-1, :] = fh[:, 0] = fh[:, -1] = 0
    f.upload(fh)
    u.upload(np.zeros((N, N)))
    t0 = time.perf_counter()
    rep = kg.poisson_solve(ctx, n, f, u, tol=tol, vcycles=vcycles)
    wall = time.perf_counter() - t0
    f.free()
    u.free()
    rep.update({"n_cells": n, "vertices": N * N, "tol": tol, "wall_s": wall,
                "vcycles_per_correction": vcycles})
    return rep


def run_typed(kg, ctx, n: int = 512, steps: int = 10, seed: int = 3) -> dict:
    """The precision variants alone at n^3: each of laplacian and wave13pt in
    FP32 and BF16 and per-plane gaussblur in FP64 and FP32 for `steps`
    launches (levels rotated, cold caches); mean launch time, grid points/s
    and algorithmic HBM bytes/s per kernel (used for profiling)."""
    N = n * n * n
    out = {"grid": [n, n, n], "steps": steps}
    cases = [("laplacian_f32", np.float32, 2, LAP, 8), ("laplacian_bf16", np.uint16, 2, LAP, 4),
             ("wave13pt_f32", np.float32, 3, WAVE, 12), ("wave13pt_bf16", np.uint16, 3, WAVE, 6),
             ("gaussblur3_f64", np.float64, 2, GAUSS, 16), ("gaussblur3_f32", np.float32, 2, GAUSS, 8)]
    for fn, dt, nlev, co, bpp in cases:
        code = {np.float64: kg.F64, np.float32: kg.F32, np.uint16: kg.BF16}[dt]
        lv = [ctx.empty((N,), dt) for _ in range(nlev)]
        for l, b in enumerate(lv):
            if code == kg.BF16:
                tmp = ctx.empty((N,), np.float32)
                kg._check(ctx.lib, ctx.lib.kg_fill_uniform(ctx.ptr, kg.F32, N, tmp.ptr, seed + l), "fill")
                kg._check(ctx.lib, ctx.lib.kg_convert(ctx.ptr, kg.BF16, b.ptr, kg.F32, tmp.ptr, N),
                          "convert")
                ctx.sync()
                tmp.free()
            else:
                kg._check(ctx.lib, ctx.lib.kg_fill_uniform(ctx.ptr, code, N, b.ptr, seed + l), "fill")
        ctx.call(fn, n, n, n, *co, *lv)
        ctx.sync()
        ctx.drain()
        for _ in range(steps):
            ctx.call(fn, n, n, n, *co, *lv)
            lv = lv[1:] + lv[:1]
        t = float(np.mean(ctx.drain()))
        for b in lv:
            b.free()
        out[fn] = {"us": round(t * 1e6, 2), "gpts_s": round(N / t / 1e9, 1),
                   "gbs": round(N * bpp / t / 1e9, 1), "frac_hbm": round(N * bpp / t / 1e9 / HBM_PEAK_GBS, 4)}
    return out


def _main(argv=None) -> int:
    """python3 kglayers.py {batch,serve,precision,poisson,typed} [--dtype f32|f64]:
    run one layer on device 0 and print its JSON (used for profiling)."""
    import argparse
    import importlib.util
    import json
    from pathlib import Path
    ap = argparse.ArgumentParser()
    ap.add_argument("layer", choices=["batch", "serve", "precision", "poisson", "typed"])
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"])
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--save-shapes", help="batch: write the tuned tile shapes to this JSON file")
    ap.add_argument("--load-shapes", help="batch: apply these tile shapes instead of autotuning")
    ap.add_argument("--no-check", action="store_true", help="batch: skip the sampled result check")
    ap.add_argument("--tune-reps", type=int, default=2, help="batch: launches per autotuner trial")
    ap.add_argument("--seq-steps", type=int, default=3,
                    help="batch: steps of the one-launch-per-grid comparison (0: none)")
    args = ap.parse_args(argv)
    spec = importlib.util.spec_from_file_location("kgstencil", Path(__file__).with_name("kgstencil.py"))
    kg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(kg)
    with kg.Context(0) as ctx:
        if args.layer == "batch":
            shapes = json.load(open(args.load_shapes)) if args.load_shapes else None
            r = run_batch(kg, ctx, kg.F64 if args.dtype == "f64" else kg.F32, steps=args.steps,
                          check=not args.no_check, shapes=shapes, seq_steps=args.seq_steps,
                          tune_reps=args.tune_reps)
            if args.save_shapes:
                with open(args.save_shapes, "w") as fh:
                    json.dump(r["shapes"], fh)
            r.pop("shapes")
        elif args.layer == "serve":
            r = run_serve(kg, ctx)
        elif args.layer == "precision":
            r = run_precision(kg, ctx)
        elif args.layer == "typed":
            r = run_typed(kg, ctx, steps=args.steps)
        else:
            r = run_poisson(kg, ctx)
    print(json.dumps(r), flush=True)
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(_main())
