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
                                         ("laplacian", "bf16"), ("wave13pt", "bf16")])
def test_typed_tile_order_tail(gpu_ctx, oracle, which, dtype):
    """600x300x40: 450 (FP32 laplacian), 500 (BF16 laplacian) and 75 / 74
    (FP32 / BF16 wave13pt) workgroups, not whole rounds of 8 chunks (two tile
    rows, or for BF16 wave13pt two whole z-columns); the last ones keep their
    hardware ids."""
    dims = (600, 300, 40)
    fn = f"{which}_{dtype}"
    ins, out, coefs, rot = _case(which, dtype, dims, 9)
    want = _run_oracle(oracle.lib, fn, dims, coefs, ins, out, 2, rot, f32)
    got = _run_gpu(gpu_ctx, fn, dims, coefs, ins, out, 2, rot)
    for g, w in zip(got, want):
        assert np.array_equal(g.view(np.uint8), w.view(np.uint8)), fn


@pytest.mark.parametrize("which,dtype", [("laplacian", "f32"), ("wave13pt", "f32"),
                                         ("laplacian", "bf16"), ("wave13pt", "bf16"),
                                         ("gaussblur3", "f32"), ("gaussblur3", "f64")])
def test_typed_full_size_one_step(gpu_ctx, oracle, which, dtype):
    """512x256x256, one step, bit-exact against the oracle."""
    dims = (512, 256, 256)
    fn = f"{which}_{dtype}"
    ins, out, coefs, rot = _case(which, dtype, dims, 7)
    want = _run_oracle(oracle.lib, fn, dims, coefs, ins, out, 1, rot, f64 if dtype == "f64" else f32)
    got = _run_gpu(gpu_ctx, fn, dims, coefs, ins, out, 1, rot)
    for g, w in zip(got, want):
        assert np.array_equal(g.view(np.uint8), w.view(np.uint8)), fn


@pytest.mark.parametrize("dims", [(264, 150, 3), (520, 71, 2), (128, 200, 1), (256, 300, 40)])
@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_gaussblur3_column_march_pieces_and_chunks(gpu_ctx, oracle, dtype, dims):
    """Per-plane gaussblur on the column march: rows of several pieces (the
    x-neighbours across a piece boundary come from the halo points the end
    lanes load, and a last piece mostly past the row), row chunks of 16 rows
    per wave with a partial last workgroup, one to forty planes (256x300x40
    in FP64: 400 workgroups, 6 whole rounds of the XCD-chunked tile order and
    a tail); three steps, bit for bit against the oracle."""
    fn = f"gaussblur3_{dtype}"
    ins, out, coefs, rot = _case("gaussblur3", dtype, dims, 31)
    want = _run_oracle(oracle.lib, fn, dims, coefs, ins, out, 3, rot, f64 if dtype == "f64" else f32)
    got = _run_gpu(gpu_ctx, fn, dims, coefs, ins, out, 3, rot)
    for g, w in zip(got, want):
        assert np.array_equal(g.view(np.uint8), w.view(np.uint8)), (fn, dims)


@pytest.mark.parametrize("which,dtype", [("laplacian", "f32"), ("laplacian", "bf16"),
                                         ("wave13pt", "f32"), ("gaussblur3", "f32"),
                                         ("gaussblur3", "f64")])
def test_typed_unaligned_fields(gpu_ctx, oracle, which, dtype):
    """Fields one element past a 16-byte boundary (nx a multiple of every
    vector width): the entry points take their one-point-per-lane paths and
    still match the oracle bit for bit over two steps."""
    dims = (264, 21, 9)
    fn = f"{which}_{dtype}"
    ins, out, coefs, rot = _case(which, dtype, dims, 41)
    want = _run_oracle(oracle.lib, fn, dims, coefs, ins, out, 2, rot, f64 if dtype == "f64" else f32)
    host = ins + [out]
    big = [gpu_ctx.empty((a.size + 1,), a.dtype) for a in host]
    try:
        for b, a in zip(big, host):
            b.upload(np.concatenate([np.zeros(1, a.dtype), a]))
        ptrs = [C.c_void_p(b.ptr + a.dtype.itemsize) for b, a in zip(big, host)]
        for _ in range(2):
            gpu_ctx.call(fn, *dims, *coefs, *ptrs)
            ptrs = rot(ptrs)
        gpu_ctx.sync()
        got = [b.download()[1:] for b in big]
    finally:
        for b in big:
            b.free()
    # want[k] is the buffer in position k after two rotations
    order = rot(rot(list(range(len(host)))))
    for i, w in zip(order, want):
        assert np.array_equal(got[i].view(np.uint8), w.view(np.uint8)), (fn, i)


@pytest.mark.parametrize("which,dtype", [("laplacian", "f32"), ("laplacian", "bf16"),
                                         ("wave13pt", "f32"), ("wave13pt", "bf16")])
def test_typed_z_chunk_seams(gpu_ctx, oracle, which, dtype):
    """264x20x150: more z-planes than one chunk of any typed tile shape (the
    longest is 64 planes), so chunks start inside the grid and their first
    planes come from the previous chunk's range; two steps, bit for bit."""
    dims = (264, 20, 150)
    fn = f"{which}_{dtype}"
    ins, out, coefs, rot = _case(which, dtype, dims, 23)
    want = _run_oracle(oracle.lib, fn, dims, coefs, ins, out, 2, rot, f32)
    got = _run_gpu(gpu_ctx, fn, dims, coefs, ins, out, 2, rot)
    for g, w in zip(got, want):
        assert np.array_equal(g.view(np.uint8), w.view(np.uint8)), fn


def test_wave13pt_bf16_z_column_runs(gpu_ctx, oracle):
    """512x200x150 BF16 wave13pt: 25 tile rows of 3 z-chunks each (75
    workgroups), dealt to the XCDs in runs of 2 whole z-columns -- one whole
    round of 8 runs and a tail; two steps, bit for bit."""
    dims = (512, 200, 150)
    ins, out, coefs, rot = _case("wave13pt", "bf16", dims, 29)
    want = _run_oracle(oracle.lib, "wave13pt_bf16", dims, coefs, ins, out, 2, rot, f32)
    got = _run_gpu(gpu_ctx, "wave13pt_bf16", dims, coefs, ins, out, 2, rot)
    for g, w in zip(got, want):
        assert np.array_equal(g.view(np.uint8), w.view(np.uint8))


def _oracle_to_bf16(oracle):
    fn = oracle.lib.oracle_f32_to_bf16
    fn.argtypes, fn.restype = [f32], C.c_uint16
    return fn


def _special_f32_patterns():
    """Every upper half-word (sign, exponent, top 7 mantissa bits) with lower
    halves that round down, tie (to even either way), round up and carry:
    NaN payloads (quiet and signalling), +-inf, FLT_MAX, subnormals, +-0."""
    hi = np.arange(1 << 16, dtype=np.uint32) << 16
    lo = np.array([0x0000, 0x0001, 0x7FFF, 0x8000, 0x8001, 0xC000, 0xFFFF], np.uint32)
    return (hi[:, None] | lo[None, :]).ravel()


def test_bf16_conversion_special_values(gpu_ctx, kgmod, oracle):
    """kg_convert F32 -> BF16 runs the conversion the BF16 stencils store
    through; its bits equal the oracle's round-to-nearest-even rule on
    458752 patterns that cover ties, NaN payloads, overflow to infinity and
    subnormals."""
    to16 = _oracle_to_bf16(oracle)
    pat = _special_f32_patterns()
    x = pat.view(np.float32)
    want = np.array([to16(C.c_float(v)) for v in x.tolist()], np.uint16)
    dx = gpu_ctx.array(x)
    dy = gpu_ctx.empty(x.shape, np.uint16)
    gpu_ctx.call("convert", kgmod.BF16, dy, kgmod.F32, dx, x.size)
    got = dy.download()
    dx.free()
    dy.free()
    nan_w = (want & 0x7FFF) > 0x7F80
    nan_g = (got & 0x7FFF) > 0x7F80
    assert np.array_equal(nan_w, nan_g), "NaN-ness differs"
    bad = np.flatnonzero((got != want) & ~nan_w)
    assert bad.size == 0, [(hex(pat[i]), hex(got[i]), hex(want[i])) for i in bad[:8]]
    # NaNs stay NaNs with the sign kept (payloads are not part of the contract)
    assert np.array_equal(got[nan_w] >> 15, want[nan_w] >> 15)


def test_bf16_stencil_rounding(gpu_ctx, oracle):
    """laplacian_bf16 with alpha = 1 + 2^-8, beta = 0: every product of a
    bfloat16 input with alpha has its lowest bits at the rounding position,
    so the stores round ties to even, carry into the exponent, overflow to
    infinity (inputs near the bfloat16 maximum) and produce subnormals (tiny
    inputs); bit-exact against the oracle.  Inputs near the maximum sit on a
    lattice of pitch 4 in every axis, so no neighbour sum overflows (beta = 0
    would turn an infinite sum into NaN)."""
    rng = np.random.default_rng(5)
    nx, ny, ns = 68, 12, 9
    n = nx * ny * ns
    w0 = rng.integers(0, 1 << 16, n, dtype=np.uint16)
    # magnitudes below 2^97: exponent field at most 0xE0
    exp = (w0 >> 7) & 0xFF
    w0 = np.where(exp > 0xE0, w0 & 0x807F | (0xE0 << 7), w0).astype(np.uint16)
    small = np.array([0x0001, 0x8001, 0x007F, 0x0080, 0x00FF, 0x3F80, 0x3FFF, 0x0000, 0x8000],
                     np.uint16)
    idx = rng.choice(n, 600, replace=False)
    w0[idx] = small[rng.integers(0, small.size, idx.size)]
    large = np.array([0x7F7F, 0xFF7F, 0x7F7E, 0x7F00], np.uint16)
    g = w0.reshape(ns, ny, nx)
    g[::4, ::4, ::4] = large[rng.integers(0, large.size, g[::4, ::4, ::4].shape)]
    w0 = g.ravel().copy()
    out = np.zeros(n, np.uint16)
    coefs = (1.00390625, 0.0)
    rot = lambda b: [b[1], b[0]]
    dims = (nx, ny, ns)
    want = _run_oracle(oracle.lib, "laplacian_bf16", dims, coefs, [w0], out, 1, rot, f32)
    got = _run_gpu(gpu_ctx, "laplacian_bf16", dims, coefs, [w0], out, 1, rot)
    assert np.any(want[0] == 0x7F80) and np.any((want[0] & 0x7F80) == 0)
    for gg, ww in zip(got, want):
        assert np.array_equal(gg, ww), [(i, hex(gg[i]), hex(ww[i])) for i in np.flatnonzero(gg != ww)[:8]]


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
