"""Pins the oracle's precision variants (FP32 arithmetic, bfloat16 storage)
and its per-plane gaussblur against what fixes them independently of the
code under test:

* bfloat16 rounding: textbook values, ties-to-even cases, overflow, NaN, and
  a nearest-value search over random floats (distance-based, not the bit
  trick the oracle uses);
* FP32 stencils: on small integers with dyadic coefficients every operation
  is exact, so the FP32 result must equal the (reference-pinned) FP64 oracle
  bit for bit; on random data the FP32 result must lie within the forward
  error bound of the FP64 one (a dropped term or wrong sign fails both);
* gaussblur per xy-plane: plane k of the 3D result equals the reference-
  pinned 2D gaussblur oracle applied to plane k.
"""
import ctypes as C

import numpy as np
import pytest

import kgspec as K

f32, f64 = C.c_float, C.c_double
vp = lambda a: C.c_void_p(a.ctypes.data)   # callers keep `a` referenced


@pytest.fixture(scope="module")
def lib(oracle):
    L = oracle.lib
    L.oracle_f32_to_bf16.restype = C.c_uint16
    L.oracle_f32_to_bf16.argtypes = [f32]
    L.oracle_bf16_to_f32.restype = f32
    L.oracle_bf16_to_f32.argtypes = [C.c_uint16]
    return L


def to_bf16(lib, x):
    return np.array([lib.oracle_f32_to_bf16(float(v)) for v in np.asarray(x, np.float32)],
                    np.uint16)


def test_bf16_known_values(lib):
    cases = {1.0: 0x3F80, -2.0: 0xC000, 0.0: 0x0000, 3.140625: 0x4049,
             1.0 + 2 ** -8: 0x3F80,           # tie -> even (0x3F80)
             1.0 + 3 * 2 ** -8: 0x3F82,       # tie -> even (0x3F82)
             1.0 + 2 ** -8 + 2 ** -20: 0x3F81,
             float(np.finfo(np.float32).max): 0x7F80,   # rounds to +inf
             float("inf"): 0x7F80, float("-inf"): 0xFF80}
    for x, want in cases.items():
        assert lib.oracle_f32_to_bf16(x) == want, (x, hex(lib.oracle_f32_to_bf16(x)), hex(want))
    assert lib.oracle_f32_to_bf16(-0.0) == 0x8000
    h = lib.oracle_f32_to_bf16(float("nan"))
    assert (h & 0x7F80) == 0x7F80 and (h & 0x7F) != 0


def test_bf16_round_trip_is_identity(lib):
    pats = np.arange(1 << 16, dtype=np.uint32)
    vals = (pats << 16).astype(np.uint32).view(np.float32)
    finite = np.isfinite(vals)
    for p, v in zip(pats[finite][::37], vals[finite][::37]):
        assert lib.oracle_f32_to_bf16(float(v)) == p


def test_bf16_rounds_to_nearest(lib):
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(4000) * np.exp(rng.uniform(-30, 30, 4000))).astype(np.float32)
    got = to_bf16(lib, x)
    u = x.view(np.uint32)
    lo = (u & 0xFFFF0000).view(np.float32).astype(np.float64)
    hi = ((u & 0xFFFF0000) + 0x10000).astype(np.uint32).view(np.float32).astype(np.float64)
    xd = x.astype(np.float64)
    dlo, dhi = np.abs(xd - lo), np.abs(hi - xd)
    back = (got.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    best = np.minimum(dlo, dhi)
    np.testing.assert_array_equal(np.abs(back - xd), best)
    tie = dlo == dhi
    assert np.all((got[tie] & 1) == 0)


def _ints(rng, n, lo=-8, hi=8):
    return rng.integers(lo, hi + 1, n).astype(np.float64)


def test_f32_laplacian_exact_on_integers(lib, oracle):
    nx, ny, ns = 13, 9, 7
    rng = np.random.default_rng(1)
    w0 = _ints(rng, nx * ny * ns)
    out64 = _ints(rng, nx * ny * ns)
    out32 = out64.astype(np.float32)
    outbf = to_bf16(lib, out32)
    lib.oracle_laplacian(nx, ny, ns, f64(0.25), f64(0.125), vp(w0), vp(out64))
    w0f, w0b = w0.astype(np.float32), to_bf16(lib, w0)   # keep the inputs alive across the calls
    lib.oracle_laplacian_f32(nx, ny, ns, f32(0.25), f32(0.125), vp(w0f), vp(out32))
    np.testing.assert_array_equal(out32.astype(np.float64), out64)
    lib.oracle_laplacian_bf16(nx, ny, ns, f32(0.25), f32(0.125), vp(w0b), vp(outbf))
    np.testing.assert_array_equal((outbf.astype(np.uint32) << 16).view(np.float32), out64)


def test_f32_wave13pt_exact_on_integers(lib):
    nx, ny, ns = 11, 10, 9
    rng = np.random.default_rng(2)
    w0, w1 = _ints(rng, nx * ny * ns), _ints(rng, nx * ny * ns)
    w2 = _ints(rng, nx * ny * ns)
    w2f = w2.astype(np.float32)
    m = (1.296875, 0.125, -0.0078125)
    lib.oracle_wave13pt(nx, ny, ns, *map(f64, m), vp(w0), vp(w1), vp(w2))
    w0f, w1f = w0.astype(np.float32), w1.astype(np.float32)
    lib.oracle_wave13pt_f32(nx, ny, ns, *map(f32, m), vp(w0f), vp(w1f), vp(w2f))
    np.testing.assert_array_equal(w2f.astype(np.float64), w2)


@pytest.mark.parametrize("which", ["laplacian", "wave13pt"])
def test_f32_within_forward_error_bound(lib, which):
    """|r32 - r64| <= gamma_n * sum |terms| (n = operations on the longest
    path, u = 2^-24), with the terms evaluated in FP64."""
    nx, ny, ns = 17, 12, 10
    rng = np.random.default_rng(5)
    n = nx * ny * ns
    a0, a1 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    a0, a1 = a0.astype(np.float32).astype(np.float64), a1.astype(np.float32).astype(np.float64)
    out64, out32 = np.zeros(n), np.zeros(n, np.float32)
    if which == "laplacian":
        c = (np.float32(0.37), np.float32(-0.21))
        lib.oracle_laplacian(nx, ny, ns, *map(f64, map(float, c)), vp(a0), vp(out64))
        a0f = a0.astype(np.float32)
        lib.oracle_laplacian_f32(nx, ny, ns, *map(f32, c), vp(a0f), vp(out32))
        g = a0.reshape(ns, ny, nx)
        absum = np.zeros_like(g)
        absum[1:-1, 1:-1, 1:-1] = (abs(c[0]) * abs(g[1:-1, 1:-1, 1:-1]) + abs(c[1]) * (
            abs(g[1:-1, 1:-1, 2:]) + abs(g[1:-1, 1:-1, :-2]) + abs(g[1:-1, 2:, 1:-1]) +
            abs(g[1:-1, :-2, 1:-1]) + abs(g[2:, 1:-1, 1:-1]) + abs(g[:-2, 1:-1, 1:-1])))
        nops = 9
    else:
        c = (np.float32(1.2), np.float32(0.11), np.float32(-0.013))
        lib.oracle_wave13pt(nx, ny, ns, *map(f64, map(float, c)), vp(a0), vp(a1), vp(out64))
        a0f, a1f = a0.astype(np.float32), a1.astype(np.float32)
        lib.oracle_wave13pt_f32(nx, ny, ns, *map(f32, c), vp(a0f), vp(a1f), vp(out32))
        g, h = a1.reshape(ns, ny, nx), a0.reshape(ns, ny, nx)
        absum = np.zeros_like(g)
        I = (slice(2, -2),) * 3
        def sh(dz, dy, dx):
            return abs(g[2 + dz:ns - 2 + dz, 2 + dy:ny - 2 + dy, 2 + dx:nx - 2 + dx])
        absum[I] = (abs(c[0]) * sh(0, 0, 0) + abs(h[I]) +
                    abs(c[1]) * sum(sh(*d) for d in [(0, 0, 1), (0, 0, -1), (0, 1, 0), (0, -1, 0), (1, 0, 0), (-1, 0, 0)]) +
                    abs(c[2]) * sum(sh(*d) for d in [(0, 0, 2), (0, 0, -2), (0, 2, 0), (0, -2, 0), (2, 0, 0), (-2, 0, 0)]))
        nops = 17
    u = 2.0 ** -24
    bound = nops * u / (1 - nops * u) * absum.ravel()
    err = np.abs(out32.astype(np.float64) - out64)
    assert np.all(err <= bound + 1e-300), float(np.max(err - bound))
    assert np.any(out64 != 0)


def test_gaussblur3_f64_is_2d_gaussblur_per_plane(lib):
    nx, ny, ns = 14, 11, 4
    rng = np.random.default_rng(9)
    w0 = rng.uniform(-1, 1, nx * ny * ns)
    s = (0.7, 0.31, 0.12, 0.05, 0.02, 0.004)
    out = rng.uniform(-1, 1, nx * ny * ns)
    want = out.copy()
    lib.oracle_gaussblur3_f64(nx, ny, ns, *map(f64, s), vp(w0), vp(out))
    for k in range(ns):
        plane_in = np.ascontiguousarray(w0[k * nx * ny:(k + 1) * nx * ny])
        plane_out = np.ascontiguousarray(want[k * nx * ny:(k + 1) * nx * ny])
        lib.oracle_gaussblur(nx, ny, *map(f64, s), vp(plane_in), vp(plane_out))
        want[k * nx * ny:(k + 1) * nx * ny] = plane_out
    np.testing.assert_array_equal(out, want)


def test_gaussblur3_f32_exact_on_integers(lib):
    nx, ny, ns = 12, 10, 3
    rng = np.random.default_rng(10)
    w0 = _ints(rng, nx * ny * ns)
    s = (8.0, 1.0, 1.0, 1.0, 1.0, 1.0)   # denominator 32: f = 1/32 exact
    o64 = _ints(rng, nx * ny * ns)
    o32 = o64.astype(np.float32)
    lib.oracle_gaussblur3_f64(nx, ny, ns, *map(f64, s), vp(w0), vp(o64))
    w0f = w0.astype(np.float32)
    lib.oracle_gaussblur3_f32(nx, ny, ns, *map(f32, s), vp(w0f), vp(o32))
    np.testing.assert_array_equal(o32.astype(np.float64), o64)
