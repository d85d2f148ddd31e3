"""The suite's non-stencil tests (vecadd, matvec, sincos), CPU side.

The reference builds them with real = double (makefile.in:12 overrides the
`real ?= float` of their gcc makefiles; sincos's Fortran kernel gets
-fdefault-real-8).  The oracle's FP64 restatements are pinned against the
reference's own IEEE build (golden fixtures, and the build itself on ragged
sizes when it is present) and against what fixes them independently: exact
float64 addition, the sequential row sum (numpy's cumulative sum adds in
order), and float64 sin/cos.  The single-precision forms beside them (the
library's _f32 entry points; the reference does not build these) are pinned
by arithmetic only.
"""
import ctypes as C
import json

import numpy as np
import pytest

import kgspec as K

vp = lambda a: C.c_void_p(a.ctypes.data)   # callers keep `a` referenced
GOLDEN = K.ROOT / "tests" / "golden" / "vecmath.npz"


def _golden():
    g = np.load(GOLDEN)
    return g, json.loads(bytes(g["meta"]).decode())


def test_oracle_matches_golden(oracle):
    g, meta = _golden()
    for ci, case in enumerate(meta["vecadd"]):
        n = int(np.prod(case["dims"]))
        for nt in (1, 3):
            st = [g[f"vecadd_c{ci}_in_{r}"].copy() for r in ("w0", "w1", "w2")]
            for _ in range(nt):
                oracle.lib.oracle_vecadd(C.c_size_t(n), vp(st[0]), vp(st[1]), vp(st[2]))
                st = [st[1], st[2], st[0]]
            for r, a in zip(("w0", "w1", "w2"), st):
                np.testing.assert_array_equal(a.view(np.uint64),
                                              g[f"vecadd_c{ci}_nt{nt}_{r}"].view(np.uint64))
    for ci, case in enumerate(meta["matvec"]):
        nx, ny = case["dims"]
        y = np.zeros(ny)
        A, x = g[f"matvec_c{ci}_in_A"], g[f"matvec_c{ci}_in_x"]
        oracle.lib.oracle_matvec(nx, ny, vp(A), vp(x), vp(y))
        np.testing.assert_array_equal(y.view(np.uint64), g[f"matvec_c{ci}_y"].view(np.uint64))
    for ci, case in enumerate(meta["sincos"]):
        n = int(np.prod(case["dims"]))
        xy = np.zeros(n)
        x, y = g[f"sincos_c{ci}_in_x"], g[f"sincos_c{ci}_in_y"]
        oracle.lib.oracle_sincos(C.c_size_t(n), vp(x), vp(y), vp(xy))
        np.testing.assert_array_equal(xy.view(np.uint64), g[f"sincos_c{ci}_xy"].view(np.uint64))


def test_vecadd_is_float64_addition(oracle):
    rng = np.random.default_rng(1)
    n = 10007
    a, b = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    out = np.zeros(n)
    oracle.lib.oracle_vecadd(C.c_size_t(n), vp(a), vp(b), vp(out))
    np.testing.assert_array_equal(out, a + b)


def test_matvec_is_the_sequential_row_sum(oracle):
    """y[j] = ((0 + A[0,j] x[0]) + A[1,j] x[1]) + ...: each product rounded,
    then added in column order -- numpy's cumulative sum of the rounded
    products along the row is exactly that."""
    rng = np.random.default_rng(2)
    nx, ny = 300, 70
    A = rng.uniform(-1, 1, (ny, nx))
    x = rng.uniform(-1, 1, nx)
    y = np.zeros(ny)
    oracle.lib.oracle_matvec(nx, ny, vp(A), vp(x), vp(y))
    want = np.cumsum(A * x[None, :], axis=1)[:, -1]
    np.testing.assert_array_equal(y.view(np.uint64), want.view(np.uint64))


def test_sincos_within_an_ulp(oracle):
    rng = np.random.default_rng(3)
    n = 20000
    x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    out = np.zeros(n)
    oracle.lib.oracle_sincos(C.c_size_t(n), vp(x), vp(y), vp(out))
    want = np.sin(x) + np.cos(y)
    # libm's and numpy's sin/cos are each within 1 ulp; the sums are <= 2
    np.testing.assert_allclose(out, want, rtol=0, atol=2 * 2.0 ** -51)
    assert np.mean(out == want) > 0.9


def test_oracle_matches_reference_ieee_build(oracle, ref):
    lib = ref.lib
    rng = np.random.default_rng(4)
    for nx, ny in [(1, 1), (7, 5), (64, 33), (129, 17)]:
        A = rng.uniform(-1, 1, nx * ny)
        x = rng.uniform(-1, 1, nx)
        y1, y2 = np.zeros(ny), np.zeros(ny)
        lib.matvec(nx, ny, vp(A), vp(x), vp(y1))
        oracle.lib.oracle_matvec(nx, ny, vp(A), vp(x), vp(y2))
        np.testing.assert_array_equal(y1.view(np.uint64), y2.view(np.uint64))
    for dims in [(1, 1, 1), (3, 4, 5), (17, 9, 2), (64, 3, 7)]:
        n = int(np.prod(dims))
        a, b = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
        o1, o2 = np.zeros(n), np.zeros(n)
        d = [C.c_int(v) for v in dims]
        lib.sincos_(C.byref(d[0]), C.byref(d[1]), C.byref(d[2]), vp(a), vp(b), vp(o1))
        oracle.lib.oracle_sincos(C.c_size_t(n), vp(a), vp(b), vp(o2))
        np.testing.assert_array_equal(o1.view(np.uint64), o2.view(np.uint64))
        lib.vecadd(*[C.c_int(v) for v in dims], vp(a), vp(b), vp(o1))
        oracle.lib.oracle_vecadd(C.c_size_t(n), vp(a), vp(b), vp(o2))
        np.testing.assert_array_equal(o1.view(np.uint64), o2.view(np.uint64))


# ---- single-precision forms (the library's _f32 entry points) ----

def test_vecadd_f32_is_float_addition(oracle):
    rng = np.random.default_rng(1)
    n = 10007
    a, b = (rng.uniform(-1, 1, n).astype(np.float32) for _ in range(2))
    out = np.zeros(n, np.float32)
    oracle.lib.oracle_vecadd_f32(C.c_size_t(n), vp(a), vp(b), vp(out))
    np.testing.assert_array_equal(out, a + b)


def test_matvec_f32_is_the_sequential_row_sum(oracle):
    rng = np.random.default_rng(2)
    nx, ny = 300, 70
    A = rng.uniform(-1, 1, (ny, nx)).astype(np.float32)
    x = rng.uniform(-1, 1, nx).astype(np.float32)
    y = np.zeros(ny, np.float32)
    oracle.lib.oracle_matvec_f32(nx, ny, vp(A), vp(x), vp(y))
    want = np.cumsum(A * x[None, :], axis=1, dtype=np.float32)[:, -1]
    np.testing.assert_array_equal(y, want)


def test_sincos_f32_within_an_ulp_of_correct_rounding(oracle):
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
