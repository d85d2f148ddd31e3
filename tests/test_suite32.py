"""The suite's single-precision tests (vecadd, matvec, sincos; real = float
in the reference).  CPU side: the oracle's restatements pinned against the
reference's own IEEE build (matvec, sincos) and against what fixes them
independently (exact float addition; the forward error bound of a row sum;
float64 sin/cos rounded to float, which libm's faithful sinf/cosf match to
1 ulp).
"""
import ctypes as C

import numpy as np
import pytest

import kgspec as K

vp = lambda a: C.c_void_p(a.ctypes.data)   # callers keep `a` referenced


def test_vecadd_is_float_addition(oracle):
    rng = np.random.default_rng(1)
    n = 10007
    a, b = (rng.uniform(-1, 1, n).astype(np.float32) for _ in range(2))
    out = np.zeros(n, np.float32)
    oracle.lib.oracle_vecadd_f32(C.c_size_t(n), vp(a), vp(b), vp(out))
    np.testing.assert_array_equal(out, a + b)


def test_matvec_within_row_sum_bound(oracle):
    rng = np.random.default_rng(2)
    nx, ny = 300, 70
    A = rng.uniform(-1, 1, (ny, nx)).astype(np.float32)
    x = rng.uniform(-1, 1, nx).astype(np.float32)
    y = np.zeros(ny, np.float32)
    oracle.lib.oracle_matvec_f32(nx, ny, vp(A), vp(x), vp(y))
    exact = A.astype(np.float64) @ x.astype(np.float64)
    bound = 2 * nx * 2.0 ** -24 * (np.abs(A).astype(np.float64) @ np.abs(x).astype(np.float64))
    assert np.all(np.abs(y - exact) <= bound)


def test_sincos_within_an_ulp_of_correct_rounding(oracle):
    rng = np.random.default_rng(3)
    n = 20000
    x, y = (rng.uniform(-1, 1, n).astype(np.float32) for _ in range(2))
    out = np.zeros(n, np.float32)
    oracle.lib.oracle_sincos_f32(C.c_size_t(n), vp(x), vp(y), vp(out))
    s = np.sin(x.astype(np.float64)).astype(np.float32)
    c = np.cos(y.astype(np.float64)).astype(np.float32)
    want = s + c
    # glibc sinf/cosf are faithful (within 1 ulp), equal to the correctly
    # rounded result in most cases; the sum is within 1 ulp of 2 (2.4e-7)
    assert np.mean(out == want) > 0.95
    np.testing.assert_allclose(out, want, rtol=0, atol=2.4e-7)


def test_oracle_matches_reference_ieee_build(oracle, ref):
    lib = ref.lib
    rng = np.random.default_rng(4)
    for nx, ny in [(1, 1), (7, 5), (64, 33), (129, 17)]:
        A = rng.uniform(-1, 1, nx * ny).astype(np.float32)
        x = rng.uniform(-1, 1, nx).astype(np.float32)
        y1, y2 = np.zeros(ny, np.float32), np.zeros(ny, np.float32)
        lib.matvec(nx, ny, vp(A), vp(x), vp(y1))
        oracle.lib.oracle_matvec_f32(nx, ny, vp(A), vp(x), vp(y2))
        np.testing.assert_array_equal(y1, y2)
    for dims in [(3, 4, 5), (17, 9, 2)]:
        n = int(np.prod(dims))
        a, b = (rng.uniform(-1, 1, n).astype(np.float32) for _ in range(2))
        o1, o2 = np.zeros(n, np.float32), np.zeros(n, np.float32)
        d = [C.c_int(v) for v in dims]
        lib.sincos_(C.byref(d[0]), C.byref(d[1]), C.byref(d[2]), vp(a), vp(b), vp(o1))
        oracle.lib.oracle_sincos_f32(C.c_size_t(n), vp(a), vp(b), vp(o2))
        np.testing.assert_array_equal(o1, o2)
