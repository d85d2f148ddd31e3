"""GPU parity of the suite's non-stencil tests against the oracle.

FP64 (the reference's real = double): vecadd and matvec bit-exact (the GPU
keeps matvec's sequential row sum, no contraction), also against the golden
fixtures written from the reference's own build; sincos within 4 ulp of 2
(8.9e-16 absolute) of the host libm result: the device's FP64 sin/cos and
glibc's are each within an ulp of the correctly rounded values, and the sums
have magnitude below 2.  The single-precision _f32 forms: vecadd and matvec
bit-exact, sincos within 4.8e-7 (2 ulp of the sum's magnitude, 2).
"""
import ctypes as C
import json

import numpy as np
import pytest

import kgspec as K

pytestmark = pytest.mark.gpu

vp = lambda a: C.c_void_p(a.ctypes.data)   # callers keep `a` referenced
SINCOS_TOL64 = 4 * 2.0 ** -51
SINCOS_TOL32 = 4.8e-7


def _gpu(ctx, fn, shape_args, ins, out_init):
    bufs = [ctx.array(a) for a in ins] + [ctx.array(out_init)]
    ctx.call(fn, *shape_args, *bufs)
    ctx.sync()
    got = bufs[-1].download()
    for b in bufs:
        b.free()
    return got


def _gpu_shifted(ctx, fn, shape_args, ins, out_init, off):
    """The same with every array `off` elements past its allocation's start
    (8-byte but not 16-byte aligned for odd off)."""
    allocs = [ctx.array(np.concatenate([np.zeros(off, a.dtype), a])) for a in list(ins) + [out_init]]
    es = ins[0].dtype.itemsize
    ptrs = [C.c_void_p(b.ptr + es * off) for b in allocs]
    rc = getattr(ctx.lib, "kg_" + fn)(ctx.ptr, *shape_args, *ptrs)
    assert rc == 0, ctx.lib.kg_last_error()
    ctx.sync()
    got = allocs[-1].download()[off:]
    for b in allocs:
        b.free()
    return got


@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 5, 7), (64, 4, 19), (512, 8, 21), (33, 33, 33)])
def test_vecadd(gpu_ctx, oracle, dims):
    rng = np.random.default_rng(sum(dims))
    n = int(np.prod(dims))
    a, b = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    want = np.zeros(n)
    oracle.lib.oracle_vecadd(C.c_size_t(n), vp(a), vp(b), vp(want))
    got = _gpu(gpu_ctx, "vecadd", dims, [a, b], np.zeros(n))
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))
    got = _gpu_shifted(gpu_ctx, "vecadd", dims, [a, b], np.zeros(n), 1)
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("nx,ny", [(1, 1), (7, 5), (32, 128), (33, 129), (512, 1000), (300, 257)])
def test_matvec(gpu_ctx, oracle, nx, ny):
    rng = np.random.default_rng(nx * 7 + ny)
    A, x = rng.uniform(-1, 1, nx * ny), rng.uniform(-1, 1, nx)
    want = np.zeros(ny)
    oracle.lib.oracle_matvec(nx, ny, vp(A), vp(x), vp(want))
    got = _gpu(gpu_ctx, "matvec", (nx, ny), [A, x], np.zeros(ny))
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))


def test_matvec_full_size(gpu_ctx, oracle):
    nx, ny = 512, 65536
    rng = np.random.default_rng(9)
    A, x = rng.uniform(-1, 1, nx * ny), rng.uniform(-1, 1, nx)
    want = np.zeros(ny)
    oracle.lib.oracle_matvec(nx, ny, vp(A), vp(x), vp(want))
    got = _gpu(gpu_ctx, "matvec", (nx, ny), [A, x], np.zeros(ny))
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))


@pytest.mark.parametrize("dims,scale", [((3, 4, 5), 1.0), ((512, 256, 4), 1.0), ((65, 3, 7), 1.0),
                                        ((256, 16, 16), 100.0)])
def test_sincos(gpu_ctx, oracle, dims, scale):
    rng = np.random.default_rng(11)
    n = int(np.prod(dims))
    x, y = rng.uniform(-scale, scale, n), rng.uniform(-scale, scale, n)
    want = np.zeros(n)
    oracle.lib.oracle_sincos(C.c_size_t(n), vp(x), vp(y), vp(want))
    got = _gpu(gpu_ctx, "sincos", dims, [x, y], np.zeros(n))
    np.testing.assert_allclose(got, want, rtol=0, atol=SINCOS_TOL64)
    got = _gpu_shifted(gpu_ctx, "sincos", dims, [x, y], np.zeros(n), 1)
    np.testing.assert_allclose(got, want, rtol=0, atol=SINCOS_TOL64)


def test_golden_fixtures(gpu_ctx):
    g = np.load(K.ROOT / "tests" / "golden" / "vecmath.npz")
    meta = json.loads(bytes(g["meta"]).decode())
    for ci, case in enumerate(meta["vecadd"]):
        dims = case["dims"]
        for nt in (1, 3):
            st = [gpu_ctx.array(g[f"vecadd_c{ci}_in_{r}"]) for r in ("w0", "w1", "w2")]
            for _ in range(nt):
                gpu_ctx.call("vecadd", *dims, *st)
                st = [st[1], st[2], st[0]]
            gpu_ctx.sync()
            for r, b in zip(("w0", "w1", "w2"), st):
                np.testing.assert_array_equal(b.download().view(np.uint64),
                                              g[f"vecadd_c{ci}_nt{nt}_{r}"].view(np.uint64))
                b.free()
    for ci, case in enumerate(meta["matvec"]):
        nx, ny = case["dims"]
        got = _gpu(gpu_ctx, "matvec", (nx, ny), [g[f"matvec_c{ci}_in_A"], g[f"matvec_c{ci}_in_x"]],
                   np.zeros(ny))
        np.testing.assert_array_equal(got.view(np.uint64), g[f"matvec_c{ci}_y"].view(np.uint64))
    for ci, case in enumerate(meta["sincos"]):
        got = _gpu(gpu_ctx, "sincos", case["dims"],
                   [g[f"sincos_c{ci}_in_x"], g[f"sincos_c{ci}_in_y"]],
                   np.zeros(int(np.prod(case["dims"]))))
        np.testing.assert_allclose(got, g[f"sincos_c{ci}_xy"], rtol=0, atol=SINCOS_TOL64)


# ---- single-precision forms ----

@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 5, 7), (512, 8, 21)])
def test_vecadd_f32(gpu_ctx, oracle, dims):
    rng = np.random.default_rng(sum(dims))
    n = int(np.prod(dims))
    a, b = (rng.uniform(-1, 1, n).astype(np.float32) for _ in range(2))
    want = np.zeros(n, np.float32)
    oracle.lib.oracle_vecadd_f32(C.c_size_t(n), vp(a), vp(b), vp(want))
    got = _gpu(gpu_ctx, "vecadd_f32", dims, [a, b], np.zeros(n, np.float32))
    np.testing.assert_array_equal(got, want)
    got = _gpu_shifted(gpu_ctx, "vecadd_f32", dims, [a, b], np.zeros(n, np.float32), 1)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("nx,ny", [(1, 1), (65, 129), (512, 1000)])
def test_matvec_f32(gpu_ctx, oracle, nx, ny):
    rng = np.random.default_rng(nx * 7 + ny)
    A = rng.uniform(-1, 1, nx * ny).astype(np.float32)
    x = rng.uniform(-1, 1, nx).astype(np.float32)
    want = np.zeros(ny, np.float32)
    oracle.lib.oracle_matvec_f32(nx, ny, vp(A), vp(x), vp(want))
    got = _gpu(gpu_ctx, "matvec_f32", (nx, ny), [A, x], np.zeros(ny, np.float32))
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("dims", [(3, 4, 5), (512, 256, 4)])
def test_sincos_f32(gpu_ctx, oracle, dims):
    rng = np.random.default_rng(11)
    n = int(np.prod(dims))
    x, y = (rng.uniform(-1, 1, n).astype(np.float32) for _ in range(2))
    want = np.zeros(n, np.float32)
    oracle.lib.oracle_sincos_f32(C.c_size_t(n), vp(x), vp(y), vp(want))
    got = _gpu(gpu_ctx, "sincos_f32", dims, [x, y], np.zeros(n, np.float32))
    np.testing.assert_allclose(got, want, rtol=0, atol=SINCOS_TOL32)
