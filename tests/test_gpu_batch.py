"""GPU tests of the batch layer (kg_batch_*): many grids of mixed kind, type
and size stepped by one persistent launch per step.

The bar is bit-exactness against stepping each grid alone with the
single-grid entry point of its kind and type (kg_laplacian / _f32,
kg_wave13pt / _f32, kg_gaussblur3_f64 / _f32), which the parity tests pin to
the oracle; FP64 grids are also checked against the oracle directly.  Graph
replays, tuning, forced tile shapes and grids joining / leaving mid-run must
not change any result.

A batch grid's levels share their boundary values (kg_batch_fill writes them
so; the uploads here do the same): a batch step writes whole rows of the
output level, its x-boundary points copied from the input level, and with
equal boundaries that is exactly the single-grid result, whose boundary
points are never written.
"""
import ctypes as C

import numpy as np
import pytest

import kgspec as K

pytestmark = pytest.mark.gpu

F64, F32 = 0, 1
COEFS = {
    "laplacian": (0.25, 0.125),
    "wave13pt": (1.296875, 0.125, -0.0078125),
    "gaussblur3": (0.7, 0.31, 0.12, 0.05, 0.02, 0.004),
}
SINGLE = {("laplacian", F64): "laplacian", ("laplacian", F32): "laplacian_f32",
          ("wave13pt", F64): "wave13pt", ("wave13pt", F32): "wave13pt_f32",
          ("gaussblur3", F64): "gaussblur3_f64", ("gaussblur3", F32): "gaussblur3_f32"}
NLEV = {"laplacian": 2, "wave13pt": 3, "gaussblur3": 2}
# interior box of each kind: x and y [lo, n-hi), z [zlo, ns-zhi)
BOX = {"laplacian": (1, 1, 1, 1), "wave13pt": (2, 2, 2, 2), "gaussblur3": (2, 2, 0, 0)}


def _levels(kind, dtype, dims, rng):
    """Random levels whose cells outside the kind's interior box are equal
    in every level."""
    nx, ny, ns = dims
    t = np.float64 if dtype == F64 else np.float32
    lv = [rng.uniform(-1, 1, (ns, ny, nx)).astype(t) for _ in range(NLEV[kind])]
    lo, hi, zlo, zhi = BOX[kind]
    inside = np.zeros((ns, ny, nx), bool)
    inside[zlo:ns - zhi, lo:ny - hi, lo:nx - hi] = True
    for a in lv[1:]:
        a[~inside] = lv[0][~inside]
    return lv


def _single(ctx, kind, dtype, dims, levels, nsteps):
    """Step one grid alone with the single-grid entry point; returns the
    levels in batch order after nsteps (newest per kg_batch_newest_level)."""
    fn = SINGLE[(kind, dtype)]
    bufs = [ctx.array(a) for a in levels]
    co = COEFS[kind]
    for _ in range(nsteps):
        if kind == "wave13pt":
            ctx.call(fn, *dims, *co, bufs[0], bufs[1], bufs[2])
            bufs = [bufs[1], bufs[2], bufs[0]]
        else:
            ctx.call(fn, *dims, *co, bufs[0], bufs[1])
            bufs = [bufs[1], bufs[0]]
    ctx.sync()
    newest = bufs[1] if kind == "wave13pt" else bufs[0]
    out = newest.download()
    for b in bufs:
        b.free()
    return out


def _mixed_specs(rng, n, lo=3, hi=70):
    kinds = list(COEFS)
    specs = []
    for g in range(n):
        kind = kinds[g % 3]
        dtype = F64 if (g // 3) % 2 == 0 else F32
        dims = tuple(int(v) for v in rng.integers(lo, hi, 3))
        specs.append((kind, dtype, dims))
    return specs


@pytest.fixture()
def batch(kgmod, gpu_ctx):
    b = kgmod.Batch(gpu_ctx)
    yield b
    b.close()


def _fill(batch, specs, rng):
    ids, init = [], []
    for kind, dtype, dims in specs:
        gid = batch.add(kind, dtype, dims, COEFS[kind])
        lv = _levels(kind, dtype, dims, rng)
        for l, a in enumerate(lv):
            batch.upload(gid, l, a)
        ids.append(gid)
        init.append(lv)
    return ids, init


def _check_all(gpu_ctx, batch, specs, ids, init, nsteps, what):
    for (kind, dtype, dims), gid, lv in zip(specs, ids, init):
        got = batch.download(gid)
        want = _single(gpu_ctx, kind, dtype, dims, lv, nsteps)
        assert got.tobytes() == want.tobytes(), f"{what}: grid {gid} {kind} {dtype} {dims}"


def test_batch_mixed_matches_single_grid(gpu_ctx, batch):
    rng = np.random.default_rng(1)
    specs = _mixed_specs(rng, 36) + [("laplacian", F64, (3, 3, 3)), ("wave13pt", F32, (4, 5, 4)),
                                     ("gaussblur3", F64, (5, 5, 1)), ("laplacian", F32, (1, 7, 9)),
                                     ("wave13pt", F64, (130, 9, 12)), ("gaussblur3", F32, (260, 33, 5))]
    ids, init = _fill(batch, specs, rng)
    batch.step(2, graph=False)
    _check_all(gpu_ctx, batch, specs, ids, init, 2, "eager 2 steps")
    batch.step(3, graph=True)
    batch.step(3, graph=True)
    _check_all(gpu_ctx, batch, specs, ids, init, 8, "graph replays")
    for gid in ids:
        assert batch.steps_done(gid) == 8


def test_batch_fp64_matches_oracle(gpu_ctx, oracle, batch):
    rng = np.random.default_rng(2)
    specs = [("laplacian", F64, (37, 21, 13)), ("wave13pt", F64, (29, 18, 16)),
             ("gaussblur3", F64, (41, 23, 6))]
    ids, init = _fill(batch, specs, rng)
    batch.step(4, graph=True)
    vp = lambda a: C.c_void_p(a.ctypes.data)
    for (kind, dtype, dims), gid, lv in zip(specs, ids, init):
        bufs = [np.array(a, copy=True) for a in lv]
        co = [C.c_double(c) for c in COEFS[kind]]
        for _ in range(4):
            if kind == "wave13pt":
                oracle.lib.oracle_wave13pt(*dims, *co, vp(bufs[0]), vp(bufs[1]), vp(bufs[2]))
                bufs = [bufs[1], bufs[2], bufs[0]]
            elif kind == "laplacian":
                oracle.lib.oracle_laplacian(*dims, *co, vp(bufs[0]), vp(bufs[1]))
                bufs = [bufs[1], bufs[0]]
            else:
                oracle.lib.oracle_gaussblur3_f64(*dims, *co, vp(bufs[0]), vp(bufs[1]))
                bufs = [bufs[1], bufs[0]]
        want = bufs[1] if kind == "wave13pt" else bufs[0]
        assert batch.download(gid).tobytes() == want.tobytes(), kind


def test_batch_graph_replays_interleaved_with_eager_work(gpu_ctx, batch):
    """A captured step graph replayed several times, with eager launches and
    host synchronisation between replays, gives the eager results."""
    rng = np.random.default_rng(3)
    specs = _mixed_specs(rng, 12)
    ids, init = _fill(batch, specs, rng)
    done = 0
    for r in range(4):
        batch.step(2, graph=True)
        done += 2
        batch.step(1, graph=False)
        done += 1
        gpu_ctx.sync()
        _check_all(gpu_ctx, batch, specs, ids, init, done, f"replay {r}")


def test_batch_shapes_and_tuning_do_not_change_results(gpu_ctx, batch):
    rng = np.random.default_rng(4)
    specs = _mixed_specs(rng, 9, 5, 90)
    ids, init = _fill(batch, specs, rng)
    batch.step(1, graph=False)
    for lpr in (16, 32, 64, 0, -1):    # 0: the flat body, -1: packed rows
        for kc in (1, 7, 64):
            batch.set_shape(ids[0], lpr, kc)
            batch.set_shape(ids[4], lpr, kc)
            assert batch.shape(ids[0]) == (lpr, kc)
            batch.step(1, graph=True)
    _check_all(gpu_ctx, batch, specs, ids, init, 16, "forced shapes")
    assert batch.tune(reps=1) >= 1
    batch.step(2, graph=True)
    _check_all(gpu_ctx, batch, specs, ids, init, 18, "after tuning")


@pytest.mark.parametrize("shape,kc", [(-1, 3), (-1, 16), (-2, 8), (-3, 1), (-3, 5), (-3, 16)])
def test_batch_packed_and_direct_rows_match_single_grid(gpu_ctx, batch, shape, kc):
    """Every grid in packed-row tiles (-1: whole rows back to back across the
    workgroup) or, for gaussblur3, the direct row body (-2) or the column
    march (-3: kc rows per wave; the other kinds then take 64 lanes per row):
    rows of 3..69 points, FP64 and FP32, grids of fewer planes than a tile
    has plane columns."""
    rng = np.random.default_rng(40 + kc)
    specs = _mixed_specs(rng, 24) + [("wave13pt", F64, (5, 70, 6)), ("gaussblur3", F32, (69, 5, 3)),
                                     ("laplacian", F32, (3, 3, 3)), ("gaussblur3", F64, (300, 41, 2)),
                                     ("gaussblur3", F32, (515, 9, 3))]
    ids, init = _fill(batch, specs, rng)
    for g in ids:
        batch.set_shape(g, shape, kc)
    batch.step(3, graph=True)
    _check_all(gpu_ctx, batch, specs, ids, init, 3, f"shape {shape} kc={kc}")


def test_batch_grids_join_and_leave(gpu_ctx, batch):
    rng = np.random.default_rng(5)
    specs = _mixed_specs(rng, 6)
    ids, init = _fill(batch, specs, rng)
    batch.step(3, graph=True)
    batch.remove(ids[2])
    late = [("wave13pt", F32, (33, 17, 20)), ("laplacian", F64, (50, 40, 9))]
    lids, linit = _fill(batch, late, rng)
    r0, u0 = batch.memory()
    batch.step(4, graph=True)
    keep = [i for i in range(6) if i != 2]
    _check_all(gpu_ctx, batch, [specs[i] for i in keep], [ids[i] for i in keep],
               [init[i] for i in keep], 7, "stayed")
    _check_all(gpu_ctx, batch, late, lids, linit, 4, "joined late")
    # the removed grid's memory is reused, not released to the driver
    batch.remove(lids[0])
    lids2, linit2 = _fill(batch, [late[0]], rng)
    assert batch.memory()[0] == r0
    batch.step(2, graph=False)
    _check_all(gpu_ctx, batch, [late[0]], lids2, linit2, 2, "reused slot")


def test_batch_rejects_bad_specs(kgmod, batch):
    with pytest.raises(kgmod.KGError):
        batch.add("laplacian", 2, (8, 8, 8), (0.1, 0.2))      # BF16 is not batched
    with pytest.raises(kgmod.KGError):
        batch.add("laplacian", 0, (0, 8, 8), (0.1, 0.2))
    with pytest.raises(kgmod.KGError):
        batch.remove(12345)


@pytest.mark.parametrize("lpr", [None, -1, -3])
def test_batch_fill_and_sequential_path(gpu_ctx, kgmod, lpr):
    """kg_batch_fill is deterministic in the seed; the per-grid sequential
    launches give the persistent kernel's results (default shapes, every
    grid in packed rows, every gaussblur3 grid in the column march)."""
    rng = np.random.default_rng(6)
    specs = _mixed_specs(rng, 9, 4, 60)
    outs = []
    for mode in ("batch", "sequential"):
        b = kgmod.Batch(gpu_ctx)
        ids = [b.add(k, d, dims, COEFS[k]) for k, d, dims in specs]
        if lpr is not None:
            for g in ids:
                b.set_shape(g, lpr, 8)
        b.fill(1234)
        first = [b.download(g, 0) for g in ids]
        if mode == "batch":
            b.step(5, graph=True)
        else:
            b.step_sequential(5)
        outs.append((first, [b.download(g) for g in ids]))
        b.close()
    for (k, d, dims), f0, f1, o0, o1 in zip(specs, outs[0][0], outs[1][0], outs[0][1], outs[1][1]):
        assert f0.tobytes() == f1.tobytes()
        assert np.all(np.abs(f0) <= 1) and np.std(f0) > 0.3
        assert o0.tobytes() == o1.tobytes(), (k, d, dims)


@pytest.mark.parametrize("dtype", [F64, F32])
def test_batch_column_march_2d_kinds(gpu_ctx, kgmod, dtype):
    """jacobi2 and life2 grids (one 2D stencil per plane) give the same bits
    in the column march (-3) as in the direct row body (-2) and the plane
    engine (64 lanes per row), after 6 steps; odd widths, 1 to 9 planes."""
    coef = {"jacobi2": (0.0, 0.2, 0.05), "life2": ()}
    dims = [(130, 67, 1), (67, 130, 3), (515, 9, 9), (4, 300, 2), (129, 33, 5)]
    outs = []
    for shape, kc in ((-2, 8), (-3, 1), (-3, 7), (-3, 32), (64, 8)):
        b = kgmod.Batch(gpu_ctx)
        ids = []
        for kind in ("jacobi2", "life2"):
            for d in dims:
                g = b.add(kind, dtype, d, coef[kind])
                b.set_shape(g, shape, kc)
                ids.append(g)
        b.fill(77)
        b.step(6, graph=True)
        outs.append([b.download(g).tobytes() for g in ids])
        b.close()
    for o in outs[1:]:
        assert o == outs[0]
