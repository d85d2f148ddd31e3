"""GPU tests of the serving loop (kg_serve_run) and of the 2D kinds it
serves (jacobi2, life2 in the batch).

* Each job's final field is bit-identical whether it runs alone or in a mix
  with other jobs that arrive while it runs (continuous batching), and in the
  sequential comparison mode.
* Admission respects the HBM budget and the active-job cap; a job larger than
  the budget is refused (KG_ERR_NOMEM) without stopping the others.
* jacobi2 jobs with a tolerance stop once the largest update is below it.
* The batch's 2D kinds equal the FP64 drop-in entry points kg_jacobi and
  kg_gameoflife applied to each xy-plane (bit-exact).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F64, F32 = 0, 1
COEF = {"laplacian": (0.25, 0.125), "wave13pt": (1.296875, 0.125, -0.0078125),
        "gaussblur3": (0.7, 0.31, 0.12, 0.05, 0.02, 0.004), "jacobi2": (0.0, 0.2, 0.05),
        "life2": ()}


def _jobs(rng, n, t_step=0.001):
    """n jobs over every kind and both types: a burst of 8 at t = 0, then one
    every t_step seconds."""
    kinds = list(COEF)
    jobs = []
    for i in range(n):
        kind = kinds[i % len(kinds)]
        dims = [int(v) for v in rng.integers(6, 40, 3)]
        if kind in ("jacobi2", "life2"):
            dims[2] = 1 if i % 2 else int(rng.integers(2, 4))
        jobs.append({"kind": kind, "dtype": F64 if (i // 5) % 2 == 0 else F32,
                     "dims": tuple(dims), "steps": int(rng.integers(3, 40)), "coef": COEF[kind],
                     "arrival": max(0, i - 7) * t_step, "seed": 1000 + i})
    return jobs


def test_serve_result_same_alone_and_in_the_mix(gpu_ctx, kgmod):
    rng = np.random.default_rng(1)
    jobs = _jobs(rng, 20)
    res, st, outs = kgmod.serve(gpu_ctx, jobs, 1 << 30, max_active=8, outputs=True)
    assert all(r["status"] == 0 for r in res)
    assert st["peak_active"] >= 2, "jobs never overlapped on the device"
    seq, _, souts = kgmod.serve(gpu_ctx, jobs, 1 << 30, mode=kgmod.SERVE_SEQUENTIAL, outputs=True)
    for i, j in enumerate(jobs):
        alone = dict(j, arrival=0.0)
        r1, _, o1 = kgmod.serve(gpu_ctx, [alone], 1 << 30, outputs=True)
        assert r1[0]["steps_run"] == res[i]["steps_run"] == j["steps"]
        assert o1[0].tobytes() == outs[i].tobytes(), f"job {i} {j['kind']} differs in the mix"
        assert r1[0]["checksum"] == res[i]["checksum"]
        assert souts[i].tobytes() == outs[i].tobytes(), f"job {i} differs in sequential mode"
        assert res[i]["admitted"] >= j["arrival"] - 1e-9
        assert res[i]["finished"] >= res[i]["admitted"]


def test_serve_budget_and_cap(gpu_ctx, kgmod):
    rng = np.random.default_rng(2)
    jobs = _jobs(rng, 10, t_step=0.0)
    for j in jobs:
        j["dims"] = (40, 40, 40)
        j["dtype"] = F64
    per = 40 * 40 * 40 * 8 * 3   # at most 3 levels of 40^3 FP64 (rows already 16-byte multiples)
    budget = 2 * per + 1000
    big = {"kind": "laplacian", "dtype": F64, "dims": (200, 200, 200), "steps": 2,
           "coef": COEF["laplacian"], "arrival": 0.0}
    res, st, _ = kgmod.serve(gpu_ctx, jobs + [big], budget, max_active=5)
    assert st["peak_bytes"] <= budget
    assert 1 <= st["peak_active"] <= 3
    assert res[-1]["status"] == -3          # KG_ERR_NOMEM: never fits
    assert all(r["status"] == 0 for r in res[:-1])
    res2, st2, _ = kgmod.serve(gpu_ctx, jobs, 1 << 30, max_active=2)
    assert st2["peak_active"] == 2


def test_serve_jacobi_stops_at_tolerance(gpu_ctx, kgmod):
    j = {"kind": "jacobi2", "dtype": F64, "dims": (12, 11, 1), "steps": 20000, "tol": 1e-8,
         "coef": COEF["jacobi2"], "arrival": 0.0, "seed": 5}
    res, _, outs = kgmod.serve(gpu_ctx, [j], 1 << 30, outputs=True)
    n = res[0]["steps_run"]
    assert 10 <= n < 20000 and n % 10 == 0
    # one more sweep moves the field by less than the tolerance
    w = outs[0][0]
    nxt = w.copy()
    c0, c1, c2 = COEF["jacobi2"]
    nxt[1:-1, 1:-1] = (c0 * w[1:-1, 1:-1] + c1 * (w[1:-1, :-2] + w[:-2, 1:-1] + w[1:-1, 2:] + w[2:, 1:-1]) +
                       c2 * (w[:-2, :-2] + w[2:, :-2] + w[:-2, 2:] + w[2:, 2:]))
    assert np.max(np.abs(nxt - w)) < 1e-8

    # the job ends with the field of step n exactly (the check that stops it
    # is read one step later), and the check at step n - 10 had not converged
    def fixed(k):
        r, _, o = kgmod.serve(gpu_ctx, [dict(j, steps=k, tol=0.0)], 1 << 30, outputs=True)
        assert r[0]["steps_run"] == k
        return o[0][0]
    assert np.array_equal(fixed(n), w)
    if n > 10:
        assert np.max(np.abs(fixed(n - 10) - fixed(n - 11))) >= 1e-8


@pytest.mark.parametrize("shape", [None, 64, -1])
@pytest.mark.parametrize("kind", ["jacobi2", "life2"])
def test_batch_2d_kinds_equal_dropin_per_plane(gpu_ctx, kgmod, kind, shape):
    """Default shape (the direct row body), the plane engine with 64 lanes per
    row, and packed rows: each the 2D entry point applied to every plane."""
    rng = np.random.default_rng(3)
    b = kgmod.Batch(gpu_ctx)
    try:
        for dims in [(37, 21, 1), (64, 40, 3), (5, 4, 2), (130, 17, 1), (300, 70, 2), (1030, 9, 1)]:
            nx, ny, ns = dims
            lv = [rng.uniform(-1, 1, (ns, ny, nx)) for _ in range(2)]
            if kind == "life2":
                lv = [(a > 0.2).astype(np.float64) for a in lv]
            # batch levels share their boundary (the x/y frame of every plane)
            edge = np.ones((ns, ny, nx), bool)
            edge[:, 1:ny - 1, 1:nx - 1] = False
            lv[1][edge] = lv[0][edge]
            gid = b.add(kind, F64, dims, COEF[kind])
            if shape is not None:
                b.set_shape(gid, shape, 4)
            for l in range(2):
                b.upload(gid, l, lv[l])
            b.step(3, graph=True)
            got = b.download(gid)
            bufs = [gpu_ctx.array(a) for a in lv]
            for _ in range(3):
                for k in range(ns):
                    off = k * nx * ny * 8
                    if kind == "jacobi2":
                        gpu_ctx.call("jacobi", nx, ny, *COEF[kind], bufs[0].ptr + off, bufs[1].ptr + off)
                    else:
                        gpu_ctx.call("gameoflife", nx, ny, bufs[0].ptr + off, bufs[1].ptr + off)
                bufs = [bufs[1], bufs[0]]
            gpu_ctx.sync()
            want = bufs[0].download()
            assert got.tobytes() == want.tobytes(), (kind, dims)
            b.remove(gid)
    finally:
        b.close()


@pytest.mark.parametrize("kind", ["jacobi2", "life2"])
def test_batch_2d_kinds_fp32_same_in_every_shape(gpu_ctx, kgmod, kind):
    """FP32 2D kinds have no single-grid entry point: the direct row body, the
    plane engine and packed rows must agree bit for bit (same op body)."""
    rng = np.random.default_rng(4)
    dims = (131, 45, 2)
    nx, ny, ns = dims
    lv = [rng.uniform(-1, 1, (ns, ny, nx)).astype(np.float32) for _ in range(2)]
    if kind == "life2":
        lv = [(a > 0.2).astype(np.float32) for a in lv]
    edge = np.ones((ns, ny, nx), bool)
    edge[:, 1:ny - 1, 1:nx - 1] = False
    lv[1][edge] = lv[0][edge]
    outs = []
    for shape in (-2, 64, -1, 16):
        b = kgmod.Batch(gpu_ctx)
        try:
            gid = b.add(kind, F32, dims, COEF[kind])
            b.set_shape(gid, shape, 8)
            for l in range(2):
                b.upload(gid, l, lv[l])
            b.step(3, graph=False)
            outs.append(b.download(gid))
        finally:
            b.close()
    for o in outs[1:]:
        assert o.tobytes() == outs[0].tobytes(), kind
