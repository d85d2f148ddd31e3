"""Pins the oracle (oracle/liboracle.so) before it is trusted as the checker.

1. Golden fixtures written from the reference's own kernel functions
   (tests/golden/, made by gen_golden.py from oracle/_ref): bit-exact.
2. The reference's kernels compiled from its sources (oracle/_ref), on further
   seeded and ragged sizes: bit-exact (skipped where _ref is not built).
3. What the mathematics fixes: exact closed forms on integer-valued fields,
   polynomial reproduction of the tricubic weights, Conway's rule, and the
   in-place whispering sweep against its closed form.
"""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

import kgspec as K
from conftest import assert_bit_equal

GOLDEN = Path(__file__).resolve().parent / "golden"


def load_fixture(name):
    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    meta = json.loads(bytes(z["meta"]).decode())
    return z, meta


# ---------------------------------------------------------------- 1. golden
@pytest.mark.parametrize("name", list(K.SPECS))
def test_oracle_matches_golden(oracle, name):
    spec = K.SPECS[name]
    z, meta = load_fixture(name)
    written = [k.split("_", 2)[2] for k in z.files if k.startswith("c0_nt1_")]
    assert written, "fixture has no outputs"
    for ci, m in enumerate(meta):
        arrays = {r: z[f"c{ci}_in_{r}"] for r in spec.arrays}
        for nt in (1, 3):
            out = oracle.run(spec, tuple(m["dims"]), m["coefs"], arrays, nt)
            for role in written:
                assert_bit_equal(out[role], z[f"c{ci}_nt{nt}_{role}"], f"{name} case {ci} nt={nt} {role}")


# ------------------------------------------------------------------- 2. _ref
RAGGED3 = [(0, 0, 0), (1, 1, 1), (3, 3, 3), (4, 5, 6), (5, 4, 3), (33, 7, 5), (65, 3, 9), (7, 66, 4)]
RAGGED2 = [(0, 0), (1, 1), (3, 3), (4, 5), (5, 4), (70, 9), (9, 70), (129, 6)]


@pytest.mark.parametrize("name", K.STENCILS)
def test_oracle_matches_reference_build(oracle, ref, name):
    spec = K.SPECS[name]
    sizes = RAGGED3 if spec.dim == 3 else RAGGED2
    for si, dims in enumerate(sizes):
        coefs, arrays = K.random_case(spec, dims, 77 + si)
        for nt in (1, 2, 5):
            a = oracle.run(spec, dims, coefs, arrays, nt)
            b = ref.run(spec, dims, coefs, arrays, nt)
            for role in a:
                assert_bit_equal(a[role], b[role], f"{name} {dims} nt={nt} {role}")


def test_oracle_matmul_matches_reference_build(oracle, ref):
    spec = K.SPECS["matmul"]
    for dims in [(1, 1, 1), (16, 4, 16), (20, 13, 7), (64, 33, 9)]:
        coefs, arrays = K.random_case(spec, dims, 5)
        a = oracle.run(spec, dims, coefs, arrays, 2)
        b = ref.run(spec, dims, coefs, arrays, 2)
        assert_bit_equal(a["C"], b["C"], f"matmul {dims}")


def test_reference_fast_math_gap_is_normwise_small(oracle):
    """The reference's own gcc flags (-O3 -ffast-math) contract and reassociate;
    the oracle is the IEEE source-order form.  Record the norm-wise gap."""
    if not K.REF_FAST_LIB.exists():
        pytest.skip("oracle/_ref not built")
    fast = K.RefBackend(K.REF_FAST_LIB)
    for name in ("laplacian", "wave13pt", "tricubic", "gaussblur"):
        spec = K.SPECS[name]
        dims = (40, 20, 12) if spec.dim == 3 else (70, 40)
        coefs, arrays = K.random_case(spec, dims, 3)
        a = oracle.run(spec, dims, coefs, arrays, 3)
        b = fast.run(spec, dims, coefs, arrays, 3)
        for role in spec.pair:
            scale = np.max(np.abs(a[role]))
            err = np.max(np.abs(a[role] - b[role])) / scale
            assert err < 64 * np.finfo(np.float64).eps, (name, role, err)


# ------------------------------------------------------- 3. closed forms
def _grid3(nx, ny, ns):
    k, j, i = np.meshgrid(np.arange(ns), np.arange(ny), np.arange(nx), indexing="ij")
    return i.astype(np.float64).ravel(), j.astype(np.float64).ravel(), k.astype(np.float64).ravel()


def _interior3(nx, ny, ns, lo, hi):
    m = np.zeros((ns, ny, nx), dtype=bool)
    m[lo:ns - hi, lo:ny - hi, lo:nx - hi] = True
    return m.ravel()


def _interior2(nx, ny, lo, hi):
    m = np.zeros((ny, nx), dtype=bool)
    m[lo:ny - hi, lo:nx - hi] = True
    return m.ravel()


def test_laplacian_quadratic_field_exact(oracle):
    nx, ny, ns = 11, 9, 7
    i, j, k = _grid3(nx, ny, ns)
    w = i * i + 2 * j * j + 3 * k * k           # sum of 6 face nbrs = 6w + 2 + 4 + 6
    w1 = np.full_like(w, -7.0)
    st = oracle.run(K.SPECS["laplacian"], (nx, ny, ns), {"alpha": 0.5, "beta": 0.25},
                    {"w0": w, "w1": w1}, 1)
    m = _interior3(nx, ny, ns, 1, 1)
    out = st["w0"]                               # newest level after the swap
    np.testing.assert_array_equal(out[m], 0.5 * w[m] + 0.25 * (6 * w[m] + 12))
    np.testing.assert_array_equal(out[~m], -7.0)  # boundary of w1 untouched


def test_divergence_gradient_linear_fields_exact(oracle):
    nx, ny, ns = 10, 8, 6
    i, j, k = _grid3(nx, ny, ns)
    coefs = {"alpha": 0.5, "beta": 0.25, "gamma": 2.0}
    st = oracle.run(K.SPECS["divergence"], (nx, ny, ns), coefs,
                    {"u": np.zeros_like(i), "ux": 3 * i, "uy": 5 * j * j, "uz": 7 * k}, 1)
    m = _interior3(nx, ny, ns, 1, 1)
    np.testing.assert_array_equal(st["u"][m], 0.5 * 6 + 0.25 * 20 * j[m] + 2.0 * 14)
    st = oracle.run(K.SPECS["gradient"], (nx, ny, ns), coefs,
                    {"u": i + 3 * j * j + 5 * k, "ux": np.zeros_like(i), "uy": np.zeros_like(i),
                     "uz": np.zeros_like(i)}, 1)
    np.testing.assert_array_equal(st["ux"][m], 0.5 * 2)
    np.testing.assert_array_equal(st["uy"][m], 0.25 * 12 * j[m])
    np.testing.assert_array_equal(st["uz"][m], 2.0 * 10)
    assert np.all(st["ux"][~m] == 0)


def test_wave13pt_quadratic_field_exact(oracle):
    nx, ny, ns = 12, 9, 8
    i, j, k = _grid3(nx, ny, ns)
    w1 = i * i + j * j + k * k        # r=1 sum: 6w+6, r=2 sum: 6w+24
    w0 = 2 * i + k
    st = oracle.run(K.SPECS["wave13pt"], (nx, ny, ns), {"m0": 2.0, "m1": 0.5, "m2": 0.25},
                    {"w0": w0, "w1": w1, "w2": np.zeros_like(w0)}, 1)
    m = _interior3(nx, ny, ns, 2, 2)
    want = 2.0 * w1[m] - w0[m] + 0.5 * (6 * w1[m] + 6) + 0.25 * (6 * w1[m] + 24)
    np.testing.assert_array_equal(st["w1"][m], want)    # rotation: w1 <- newest


def test_uxx1_linear_fields_exact(oracle):
    nx, ny, ns = 32, 7, 6
    i, j, k = _grid3(nx, ny, ns)
    st = oracle.run(K.SPECS["uxx1"], (nx, ny, ns), {"c1": 0.5, "c2": 0.25},
                    {"u0": k.copy(), "u1": np.zeros_like(i), "d1": np.full_like(i, 2.0),
                     "xx": i, "xy": 2 * j, "xz": 4 * k}, 1)
    m = _interior3(nx, ny, ns, 2, 1)
    # per axis slope s: c1*s + c2*3s;  (1 + 2 + 4)*(0.5 + 0.75) = 8.75; dth/d = (1/32)/2
    np.testing.assert_array_equal(st["u0"][m], k[m] + 8.75 / 64)


@pytest.mark.parametrize("name", ["tricubic", "tricubic2"])
def test_tricubic_reproduces_cubics(oracle, name):
    """Lagrange-cubic weights at nodes {-1,0,1,2}: the stencil evaluates
    p(i-a) q(j-b) r(k-c) exactly for cubic p, q, r (up to rounding)."""
    nx, ny, ns = 9, 8, 7
    i, j, k = _grid3(nx, ny, ns)
    p = lambda x: 0.5 * x ** 3 - x ** 2 + 2 * x - 3
    q = lambda x: -0.25 * x ** 3 + x + 1
    r = lambda x: 0.125 * x ** 3 + 0.5 * x ** 2 - 2
    rng = np.random.default_rng(0)
    a, b, c = (rng.uniform(-1, 1, i.size) for _ in range(3))
    st = oracle.run(K.SPECS[name], (nx, ny, ns), {},
                    {"u0": p(i) * q(j) * r(k), "u1": np.zeros_like(i), "a": a, "b": b, "c": c}, 1)
    lo, hi = K.SPECS[name].lo, K.SPECS[name].hi
    m = _interior3(nx, ny, ns, lo, hi)
    want = p(i - a) * q(j - b) * r(k - c)
    np.testing.assert_allclose(st["u0"][m], want[m], rtol=1e-11, atol=1e-9)


def test_lapgsrb_quadratic_field_exact(oracle):
    nx, ny, ns = 10, 9, 8
    i, j, k = _grid3(nx, ny, ns)
    w = i * i + j * j + k * k
    st = oracle.run(K.SPECS["lapgsrb"], (nx, ny, ns), {"c0": 2.0, "c1": 0.5, "c2": 0.25, "c3": 0.125},
                    {"w0": w, "w1": np.zeros_like(w)}, 1)
    m = _interior3(nx, ny, ns, 2, 2)
    want = 2.0 * w[m] + 0.5 * (6 * w[m] + 6) + 0.25 * (12 * w[m] + 24) + 0.125 * (6 * w[m] + 24)
    np.testing.assert_array_equal(st["w0"][m], want)


def test_jacobi_quadratic_field_exact(oracle):
    nx, ny = 12, 9
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    i, j = i.ravel().astype(float), j.ravel().astype(float)
    w = i * i + 2 * j * j
    st = oracle.run(K.SPECS["jacobi"], (nx, ny), {"c0": 2.0, "c1": 0.5, "c2": 0.25},
                    {"w0": w, "w1": np.zeros_like(w)}, 1)
    m = _interior2(nx, ny, 1, 1)
    want = 2.0 * w[m] + 0.5 * (4 * w[m] + 6) + 0.25 * (4 * w[m] + 12)
    np.testing.assert_array_equal(st["w0"][m], want)


def test_gaussblur_box_filter(oracle):
    """All six weights equal -> normalised 25-point box filter; on an affine
    field the box mean equals the centre value."""
    nx, ny = 14, 11
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    w = (3 * i - 2 * j + 1).ravel().astype(float)
    s = {k: 0.3 for k in ("s0", "s1", "s2", "s4", "s5", "s8")}
    st = oracle.run(K.SPECS["gaussblur"], (nx, ny), s, {"w0": w, "w1": np.zeros_like(w)}, 1)
    m = _interior2(nx, ny, 2, 2)
    np.testing.assert_allclose(st["w0"][m], w[m], rtol=1e-14, atol=1e-13)


def _conway(board):
    ny, nx = board.shape
    out = board.copy()
    for y in range(1, ny - 1):
        for x in range(1, nx - 1):
            n = int(board[y - 1:y + 2, x - 1:x + 2].sum() - board[y, x])
            out[y, x] = 1 if (n == 3 or (board[y, x] and n == 2)) else 0
    return out


def test_gameoflife_fp64_rule_is_conway_on_01(oracle):
    rng = np.random.default_rng(4)
    nx, ny = 13, 11
    board = (rng.random((ny, nx)) < 0.35).astype(np.float64)
    cur = board.ravel().copy()
    ref_board = board.astype(np.uint8)
    for _ in range(4):
        st = oracle.run(K.SPECS["gameoflife"], (nx, ny), {}, {"u0": cur, "u1": cur.copy()}, 1)
        cur = st["u0"]
        ref_board = _conway(ref_board)
        np.testing.assert_array_equal((cur > 0.5).reshape(ny, nx).astype(np.uint8), ref_board)


def test_life_u8_is_conway(oracle):
    lib = oracle.lib
    rng = np.random.default_rng(5)
    nx, ny = 17, 12
    b = (rng.random((ny, nx)) < 0.3).astype(np.uint8)
    out = b.copy()
    lib.oracle_life_u8(C.c_int(nx), C.c_int(ny), C.c_void_p(b.ctypes.data), C.c_void_p(out.ctypes.data))
    np.testing.assert_array_equal(out, _conway(b))


def test_life_u8_known_patterns(oracle):
    """Textbook answers: a blinker has period 2 (vertical <-> horizontal), a
    block is still, a glider moves (+1, +1) every 4 generations."""
    lib = oracle.lib

    def steps(board, nt):
        ny, nx = board.shape
        cur = board.copy()
        for _ in range(nt):
            nxt = cur.copy()
            lib.oracle_life_u8(C.c_int(nx), C.c_int(ny), C.c_void_p(cur.ctypes.data),
                               C.c_void_p(nxt.ctypes.data))
            cur = nxt
        return cur

    b = np.zeros((9, 9), np.uint8)
    b[3:6, 4] = 1
    h = np.zeros_like(b)
    h[4, 3:6] = 1
    np.testing.assert_array_equal(steps(b, 1), h)
    np.testing.assert_array_equal(steps(b, 2), b)
    blk = np.zeros((6, 6), np.uint8)
    blk[2:4, 2:4] = 1
    np.testing.assert_array_equal(steps(blk, 3), blk)
    g = np.zeros((20, 20), np.uint8)
    for (y, x) in [(1, 2), (2, 3), (3, 1), (3, 2), (3, 3)]:
        g[2 + y, 2 + x] = 1
    np.testing.assert_array_equal(steps(g, 4), np.roll(g, (1, 1), axis=(0, 1)))


def test_whispering_closed_form_equals_serial_sweep(oracle):
    """The GPU formulation (no in-place dependencies) reproduces the serial
    in-place sweep of whispering/whispering.c:87-106 bit for bit."""
    lib = oracle.lib
    for (nx, ny) in [(3, 3), (5, 4), (37, 13), (64, 9)]:
        rng = np.random.default_rng(nx * 100 + ny)
        n = nx * ny
        e = rng.uniform(-1, 1, 2 * n)
        h0, u0, ca, cb, da, db = (rng.uniform(-1, 1, n) for _ in range(6))
        mu, eps = 0.7, -0.3
        e_serial = e.copy()
        h1a, u1a = np.zeros(n), np.zeros(n)
        p = lambda a: C.c_void_p(a.ctypes.data)
        lib.oracle_whispering(C.c_int(nx), C.c_int(ny), C.c_double(mu), C.c_double(eps), p(e_serial),
                              p(h0), p(h1a), p(u0), p(u1a), p(ca), p(cb), p(da), p(db))
        e_out = np.zeros_like(e)
        h1b, u1b = np.zeros(n), np.zeros(n)
        lib.oracle_whispering_closed(C.c_int(nx), C.c_int(ny), C.c_double(mu), C.c_double(eps),
                                     p(e), p(e_out), p(h0), p(h1b), p(u0), p(u1b), p(ca), p(cb),
                                     p(da), p(db))
        assert_bit_equal(e_out, e_serial, f"e {nx}x{ny}")
        assert_bit_equal(h1b, h1a, "h1")
        assert_bit_equal(u1b, u1a, "u_em1")


def test_matmul_against_numpy(oracle):
    spec = K.SPECS["matmul"]
    nx, ny, ns = 24, 17, 9
    coefs, arrays = K.random_case(spec, (nx, ny, ns), 9)
    out = oracle.run(spec, (nx, ny, ns), coefs, arrays, 2)
    A = arrays["A"].reshape(ny, nx).T
    B = arrays["B"].reshape(ns, ny).T
    want = 2 * (A @ B)
    np.testing.assert_allclose(out["C"].reshape(ns, nx).T, want, rtol=1e-12, atol=1e-12)
