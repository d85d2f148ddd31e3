"""GPU parity of the precision variants (FP32 arithmetic with FP32 or
bfloat16 storage, and gaussblur per xy-plane in FP64/FP32) against the
oracle's restatements, and of kg_field_stats against FP64 numpy.

The typed kernels evaluate the same expressions in the same order as the
oracle in IEEE single precision (no contraction), so the bar is bit-exact
(tolerance 0).  Statistics: maxima are order-free and must match exactly;
sums use a different (tree) order than numpy, tolerance 1e-12 relative.
"""
import ctypes as C

import numpy as np
import pytest

import kgspec as K

pytestmark = pytest.mark.gpu

f32, f64 = C.c_float, C.c_double
vp = lambda a: C.c_void_p(a.ctypes.data)   # callers keep `a` referenced

SIZES = [(3, 3, 3), (5, 4, 6), (8, 8, 8), (33, 7, 5), (64, 4, 19), (66, 9, 35), (129, 17, 6),
         (256, 12, 40), (513, 9, 11), (520, 6, 7)]


def to_bf16(x):
    """float32 -> bfloat16 bits, nearest even (inputs here are finite)."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _run_gpu(ctx, fn, dims, coefs, ins, out_init, nt, rotate):
    bufs = [ctx.array(a) for a in ins] + [ctx.array(out_init)]
    for _ in range(nt):
        ctx.call(fn, *dims, *coefs, *bufs)
        bufs = rotate(bufs)
    ctx.sync()
    res = [b.download() for b in bufs]
    for b in bufs:
        b.free()
    return res


def _run_oracle(lib, fn, dims, coefs, ins, out_init, nt, rotate, ctype):
    bufs = [np.array(a, copy=True) for a in ins] + [np.array(out_init, copy=True)]
    for _ in range(nt):
        getattr(lib, "oracle_" + fn)(*dims, *map(ctype, coefs), *map(vp, bufs))
        bufs = rotate(bufs)
    return bufs


def _case(which, dtype, dims, seed):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    mk = lambda: rng.uniform(-1, 1, n).astype(np.float32)
    if which == "laplacian":
        ins, coefs = [mk()], (0.25, 0.125)
        rot = lambda b: [b[1], b[0]]
    elif which == "wave13pt":
        ins, coefs = [mk(), mk()], (1.296875, 0.125, -0.0078125)
        rot = lambda b: [b[1], b[2], b[0]]
    else:
        ins, coefs = [mk()], (0.7, 0.31, 0.12, 0.05, 0.02, 0.004)
        rot = lambda b: [b[1], b[0]]
    out = mk()
    if dtype == "bf16":
        ins, out = [to_bf16(a) for a in ins], to_bf16(out)
    elif dtype == "f64":
        ins, out = [a.astype(np.float64) for a in ins], out.astype(np.float64)
    return ins, out, coefs, rot


@pytest.mark.parametrize("which,dtype", [("laplacian", "f32"), ("laplacian", "bf16"),
                                         ("wave13pt", "f32"), ("wave13pt", "bf16"),
                                         ("gaussblur3", "f32"), ("gaussblur3", "f64")])
def test_typed_matches_oracle(gpu_ctx, oracle, which, dtype):
    fn = f"{which}_{dtype}"
    ctype = f64 if dtype == "f64" else f32
    for si, dims in enumerate(SIZES):
        ins, out, coefs, rot = _case(which, dtype, dims, 100 + si)
        for nt in (1, 3):
            want = _run_oracle(oracle.lib, fn, dims, coefs, ins, out, nt, rot, ctype)
            got = _run_gpu(gpu_ctx, fn, dims, coefs, ins, out, nt, rot)
            for g, w in zip(got, want):
                np.testing.assert_array_equal(g.view(np.uint8), w.view(np.uint8),
                                              err_msg=f"{fn} {dims} nt={nt}")


@pytest.mark.parametrize("which,dtype", [("laplacian", "f32"), ("wave13pt", "f32"),
                                         ("laplacian", "bf16")])
def test_typed_tile_order_tail(gpu_ctx, oracle, which, dtype):
    """600x300x40: 450 (FP32 laplacian), 500 (BF16 laplacian) and 75
    (wave13pt) workgroups, not whole rounds of 8 chunks of two tile rows; the
    last ones keep their hardware ids."""
    dims = (600, 300, 40)
    fn = f"{which}_{dtype}"
    ins, out, coefs, rot = _case(which, dtype, dims, 9)
    want = _run_oracle(oracle.lib, fn, dims, coefs, ins, out, 2, rot, f32)
    got = _run_gpu(gpu_ctx, fn, dims, coefs, ins, out, 2, rot)
    for g, w in zip(got, want):
        assert np.array_equal(g.view(np.uint8), w.view(np.uint8)), fn


@pytest.mark.parametrize("which,dtype", [("laplacian", "f32"), ("wave13pt", "f32"),
                                         ("laplacian", "bf16"), ("gaussblur3", "f32"),
                                         ("gaussblur3", "f64")])
def test_typed_full_size_one_step(gpu_ctx, oracle, which, dtype):
    """512x256x256, one step, bit-exact against the oracle."""
    dims = (512, 256, 256)
    fn = f"{which}_{dtype}"
    ins, out, coefs, rot = _case(which, dtype, dims, 7)
    want = _run_oracle(oracle.lib, fn, dims, coefs, ins, out, 1, rot, f64 if dtype == "f64" else f32)
    got = _run_gpu(gpu_ctx, fn, dims, coefs, ins, out, 1, rot)
    for g, w in zip(got, want):
        assert np.array_equal(g.view(np.uint8), w.view(np.uint8)), fn


@pytest.mark.parametrize("dtype,code,np_t", [("f64", 0, np.float64), ("f32", 1, np.float32),
                                             ("bf16", 2, None)])
def test_field_stats(gpu_ctx, kgmod, dtype, code, np_t):
    rng = np.random.default_rng(31)
    for n in (1, 255, 1000, 1 << 20, 3_000_017):
        x32 = rng.uniform(-2, 2, n).astype(np.float32)
        r64 = rng.uniform(-2, 2, n)
        if dtype == "bf16":
            xh = to_bf16(x32)
            xv = (xh.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        else:
            xh = x32.astype(np_t)
            xv = xh.astype(np.float64)
        dx, dr = gpu_ctx.array(xh), gpu_ctx.array(r64)
        st = gpu_ctx.field_stats(dx, code, dr, kgmod.F64)
        st0 = gpu_ctx.field_stats(dx, code)
        d = xv - r64
        assert st["max"] == np.max(np.abs(xv))
        assert st["err_max"] == np.max(np.abs(d))
        np.testing.assert_allclose(st["sum"], np.sum(xv), rtol=1e-12, atol=1e-12 * np.sum(np.abs(xv)))
        np.testing.assert_allclose(st["sumsq"], np.sum(xv * xv), rtol=1e-12)
        np.testing.assert_allclose(st["err_sumsq"], np.sum(d * d), rtol=1e-12)
        assert st0["err_sumsq"] == 0 and st0["err_max"] == 0 and st0["sum"] == st["sum"]
        # deterministic: a second call returns the same bits
        assert gpu_ctx.field_stats(dx, code, dr, kgmod.F64) == st
        dx.free()
        dr.free()


def test_typed_rejects_aliasing(gpu_ctx, kgmod):
    a = gpu_ctx.empty((4 * 4 * 4,), np.float32)
    with pytest.raises(kgmod.KGError):
        gpu_ctx.call("laplacian_f32", 4, 4, 4, 0.1, 0.2, a, a)
    with pytest.raises(kgmod.KGError):
        gpu_ctx.call("wave13pt_bf16", 4, 4, 4, 0.1, 0.2, 0.3, a, a, a)
    with pytest.raises(kgmod.KGError):
        gpu_ctx.field_stats(a, 7)
