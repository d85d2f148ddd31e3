"""GPU parity of the suite's single-precision tests against the oracle:
vecadd and matvec bit-exact (the GPU keeps matvec's sequential row sum);
sincos within 2 ulp of the sum's magnitude (the device sinf/cosf and the host
libm are each within 1 ulp of the correctly rounded value, so the two sums
differ by at most 2 ulp of each term plus the rounding of the sum: tolerance
4.8e-7 absolute on sums of magnitude <= 2)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

vp = lambda a: C.c_void_p(a.ctypes.data)   # callers keep `a` referenced


@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 5, 7), (64, 4, 19), (512, 8, 21), (33, 33, 33)])
def test_vecadd(gpu_ctx, oracle, dims):
    rng = np.random.default_rng(sum(dims))
    n = int(np.prod(dims))
    a, b = (rng.uniform(-1, 1, n).astype(np.float32) for _ in range(2))
    want = np.zeros(n, np.float32)
    oracle.lib.oracle_vecadd_f32(C.c_size_t(n), vp(a), vp(b), vp(want))
    da, db, dc = gpu_ctx.array(a), gpu_ctx.array(b), gpu_ctx.array(np.zeros(n, np.float32))
    gpu_ctx.call("vecadd", *dims, da, db, dc)
    gpu_ctx.sync()
    np.testing.assert_array_equal(dc.download(), want)


@pytest.mark.parametrize("nx,ny", [(1, 1), (7, 5), (64, 128), (65, 129), (512, 1000), (300, 257)])
def test_matvec(gpu_ctx, oracle, nx, ny):
    rng = np.random.default_rng(nx * 7 + ny)
    A = rng.uniform(-1, 1, nx * ny).astype(np.float32)
    x = rng.uniform(-1, 1, nx).astype(np.float32)
    want = np.zeros(ny, np.float32)
    oracle.lib.oracle_matvec_f32(nx, ny, vp(A), vp(x), vp(want))
    dA, dx, dy = gpu_ctx.array(A), gpu_ctx.array(x), gpu_ctx.array(np.zeros(ny, np.float32))
    gpu_ctx.call("matvec", nx, ny, dA, dx, dy)
    gpu_ctx.sync()
    np.testing.assert_array_equal(dy.download(), want)


def test_matvec_full_size(gpu_ctx, oracle):
    nx, ny = 512, 65536
    rng = np.random.default_rng(9)
    A = rng.uniform(-1, 1, nx * ny).astype(np.float32)
    x = rng.uniform(-1, 1, nx).astype(np.float32)
    want = np.zeros(ny, np.float32)
    oracle.lib.oracle_matvec_f32(nx, ny, vp(A), vp(x), vp(want))
    dA, dx, dy = gpu_ctx.array(A), gpu_ctx.array(x), gpu_ctx.array(np.zeros(ny, np.float32))
    gpu_ctx.call("matvec", nx, ny, dA, dx, dy)
    gpu_ctx.sync()
    np.testing.assert_array_equal(dy.download(), want)


@pytest.mark.parametrize("dims", [(3, 4, 5), (512, 256, 4)])
def test_sincos(gpu_ctx, oracle, dims):
    rng = np.random.default_rng(11)
    n = int(np.prod(dims))
    x, y = (rng.uniform(-1, 1, n).astype(np.float32) for _ in range(2))
    want = np.zeros(n, np.float32)
    oracle.lib.oracle_sincos_f32(C.c_size_t(n), vp(x), vp(y), vp(want))
    dx, dy, dz = gpu_ctx.array(x), gpu_ctx.array(y), gpu_ctx.array(np.zeros(n, np.float32))
    gpu_ctx.call("sincos", *dims, dx, dy, dz)
    gpu_ctx.sync()
    np.testing.assert_allclose(dz.download(), want, rtol=0, atol=4.8e-7)
