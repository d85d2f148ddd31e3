"""GPU parity: libkgstencil.so (through the C ABI) against the oracle.

FP64 stencils must be bit-identical to the oracle (which is bit-identical to
the reference's IEEE build): tolerance 0 ulp.  matmul runs on the FP64 matrix
cores and is checked norm-wise (tolerance stated in the test).
"""
import json
from pathlib import Path

import numpy as np
import pytest

import kgspec as K
from conftest import assert_bit_equal

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"

SIZES3 = [(1, 1, 1), (3, 3, 3), (4, 5, 6), (5, 4, 3), (6, 6, 6), (33, 7, 5), (64, 4, 19),
          (65, 3, 9), (66, 9, 35), (7, 66, 4), (130, 17, 40), (512, 8, 21)]
SIZES2 = [(1, 1), (3, 3), (4, 5), (5, 4), (6, 6), (70, 9), (9, 70), (129, 6), (256, 37),
          (257, 65), (512, 100), (1000, 67)]


@pytest.mark.parametrize("name", K.STENCILS)
def test_gpu_matches_golden(gpu_ctx, name):
    spec = K.SPECS[name]
    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    meta = json.loads(bytes(z["meta"]).decode())
    written = [k.split("_", 2)[2] for k in z.files if k.startswith("c0_nt1_")]
    gpu = K.GpuBackend(gpu_ctx)
    for ci, m in enumerate(meta):
        arrays = {r: z[f"c{ci}_in_{r}"] for r in spec.arrays}
        for nt in (1, 3):
            out = gpu.run(spec, tuple(m["dims"]), m["coefs"], arrays, nt)
            for role in written:
                assert_bit_equal(out[role], z[f"c{ci}_nt{nt}_{role}"], f"{name} case {ci} nt={nt} {role}")


@pytest.mark.parametrize("name", K.STENCILS)
def test_gpu_matches_oracle_ragged(gpu_ctx, oracle, name):
    spec = K.SPECS[name]
    gpu = K.GpuBackend(gpu_ctx)
    for si, dims in enumerate(SIZES3 if spec.dim == 3 else SIZES2):
        coefs, arrays = K.random_case(spec, dims, 500 + si)
        for nt in (1, 4):
            want = oracle.run(spec, dims, coefs, arrays, nt)
            got = gpu.run(spec, dims, coefs, arrays, nt)
            for role in want:
                assert_bit_equal(got[role], want[role], f"{name} {dims} nt={nt} {role}")


@pytest.mark.parametrize("name", K.STENCILS)
def test_gpu_matches_oracle_unaligned(gpu_ctx, oracle, name):
    """Fields 8 bytes past a 16-byte boundary (a caller's sub-array): the
    entry points then take their 8-byte-lane paths, with the same results."""
    spec = K.SPECS[name]
    sizes = [(66, 9, 35), (512, 8, 21)] if spec.dim == 3 else [(256, 37), (512, 100)]
    for si, dims in enumerate(sizes):
        coefs, arrays = K.random_case(spec, dims, 700 + si)
        want = oracle.run(spec, dims, coefs, arrays, 2)
        got = K.GpuBackend(gpu_ctx).run(spec, dims, coefs, arrays, 2, offset=1)
        for role in want:
            assert_bit_equal(got[role], want[role], f"{name} {dims} unaligned {role}")


FULL3 = (512, 256, 256)
FULL2 = (512, 65536)


@pytest.mark.parametrize("name", K.STENCILS)
def test_gpu_matches_oracle_full_size_one_step(gpu_ctx, oracle, name):
    """The benchmark configuration (512x256x256 / 512x65536), one step, every
    output bit-identical to the oracle."""
    spec = K.SPECS[name]
    dims = FULL3 if spec.dim == 3 else FULL2
    coefs, arrays = K.random_case(spec, dims, 2024)
    want = oracle.run(spec, dims, coefs, arrays, 1)
    got = K.GpuBackend(gpu_ctx).run(spec, dims, coefs, arrays, 1)
    for role in want:
        assert_bit_equal(got[role], want[role], f"{name} full nt=1 {role}")


@pytest.mark.parametrize("name,dims", [("laplacian", (300, 200, 40)), ("gradient", (300, 200, 40)),
                                       ("wave13pt", (300, 200, 40)), ("jacobi", (600, 1100)),
                                       ("gaussblur", (600, 1100)), ("gameoflife", (600, 1100))])
def test_gpu_tile_order_tails(gpu_ctx, oracle, name, dims):
    """Grids whose workgroup count is not a whole number of tile-order rounds
    (8 chunks of tile rows, or 8 XCD-contiguous parts): the workgroups past
    the last whole round keep their hardware ids.  Launches here: laplacian
    300 workgroups (rounds of 48), gradient 300 (48), wave13pt 150 (48),
    jacobi and gaussblur 175 (160), gameoflife 345 (160)."""
    spec = K.SPECS[name]
    coefs, arrays = K.random_case(spec, dims, 88)
    want = oracle.run(spec, dims, coefs, arrays, 2)
    got = K.GpuBackend(gpu_ctx).run(spec, dims, coefs, arrays, 2)
    for role in want:
        assert_bit_equal(got[role], want[role], f"{name} {dims} {role}")


# Extents at and around the kernels' tile edges: 64- and 128-point row
# pieces (3D), 256- and 512-point pieces (2D), tiles of 4, 8, 10, 12 and 16
# rows, 8-, 32- and 64-plane z-chunks and 8-, 32- and 64-row y-chunks.
EDGE_X3 = [127, 128, 129, 130, 255, 257, 258, 386]
EDGE_Y3 = [9, 10, 11, 12, 13, 14, 18, 19, 22, 23, 26]
EDGE_Z3 = [7, 8, 9, 10, 11, 33, 34, 35, 66]
EDGE_X2 = [127, 128, 129, 255, 256, 257, 511, 512, 513, 514, 1026]
EDGE_Y2 = [7, 8, 9, 10, 33, 34, 35, 63, 64, 65, 66, 130]


@pytest.mark.parametrize("name", K.STENCILS)
def test_gpu_matches_oracle_tile_edges(gpu_ctx, oracle, name):
    """Six seeded shapes per stencil drawn from the tile-edge extents above:
    partial and whole last tiles, chunks and row pieces; two steps, every
    output bit-identical to the oracle."""
    spec = K.SPECS[name]
    rng = np.random.default_rng(K.STENCILS.index(name) + 900)
    gpu = K.GpuBackend(gpu_ctx)
    for si in range(6):
        if spec.dim == 3:
            dims = (int(rng.choice(EDGE_X3)), int(rng.choice(EDGE_Y3)), int(rng.choice(EDGE_Z3)))
        else:
            dims = (int(rng.choice(EDGE_X2)), int(rng.choice(EDGE_Y2)))
        coefs, arrays = K.random_case(spec, dims, 1000 + si)
        if name == "uxx1":
            arrays["d1"] = np.abs(arrays["d1"]) + 1.0
        want = oracle.run(spec, dims, coefs, arrays, 2)
        got = gpu.run(spec, dims, coefs, arrays, 2)
        for role in want:
            assert_bit_equal(got[role], want[role], f"{name} {dims} {role}")


@pytest.mark.parametrize("name,dims", [("laplacian", (512, 512, 1100)), ("jacobi", (512, 600000))])
def test_gpu_fields_above_2GiB(gpu_ctx, oracle, name, dims):
    """Fields larger than 2^31 bytes (2.3 GB and 2.5 GB per array), so byte
    offsets pass 2^31: one step bit-identical to the oracle.  Only a 3D
    grid's xy-plane is limited to 2^31 bytes."""
    spec = K.SPECS[name]
    coefs, arrays = K.random_case(spec, dims, 4242)
    assert all(a.nbytes > 2**31 for a in arrays.values())
    want = oracle.run(spec, dims, coefs, arrays, 1)
    got = K.GpuBackend(gpu_ctx).run(spec, dims, coefs, arrays, 1)
    for role in want:
        assert_bit_equal(got[role], want[role], f"{name} {dims} {role}")


@pytest.mark.parametrize("dims", [(512, 520, 40), (512, 256, 200)])
def test_gpu_lapgsrb_long_z_chunks(gpu_ctx, oracle, dims):
    """lapgsrb lengthens its z-chunks when the launch would exceed about two
    workgroups per compute unit: (512, 520, 40) makes one chunk of every
    plane (520 tiles in xy), (512, 256, 200) two chunks of 98 planes."""
    spec = K.SPECS["lapgsrb"]
    coefs, arrays = K.random_case(spec, dims, 77)
    want = oracle.run(spec, dims, coefs, arrays, 2)
    got = K.GpuBackend(gpu_ctx).run(spec, dims, coefs, arrays, 2)
    for role in want:
        assert_bit_equal(got[role], want[role], f"lapgsrb {dims} {role}")


@pytest.mark.parametrize("name", ["laplacian", "wave13pt", "whispering", "gameoflife", "uxx1"])
def test_gpu_matches_oracle_ten_steps(gpu_ctx, oracle, name):
    """Ten ping-pong steps (the suite's nt) at a mid size."""
    spec = K.SPECS[name]
    dims = (96, 40, 36) if spec.dim == 3 else (300, 500)
    coefs, arrays = K.random_case(spec, dims, 31)
    if name == "uxx1":
        arrays["d1"] = np.abs(arrays["d1"]) + 1.0   # keep dth/d finite over 10 steps
    want = oracle.run(spec, dims, coefs, arrays, 10)
    got = K.GpuBackend(gpu_ctx).run(spec, dims, coefs, arrays, 10)
    for role in want:
        assert_bit_equal(got[role], want[role], f"{name} nt=10 {role}")


def test_gpu_special_values(gpu_ctx, oracle):
    """Infinities, NaNs, signed zeros and subnormals propagate as on the CPU."""
    spec = K.SPECS["gameoflife"]
    dims = (70, 40)
    coefs, arrays = K.random_case(spec, dims, 8)
    u0 = arrays["u0"]
    u0[::7] = 0.0
    u0[3::11] = -0.0
    u0[5::13] = 5e-324
    u0[9::17] = 1.0
    arrays["u0"] = u0
    want = oracle.run(spec, dims, coefs, arrays, 3)
    got = K.GpuBackend(gpu_ctx).run(spec, dims, coefs, arrays, 3)
    for role in want:
        assert_bit_equal(got[role], want[role], f"gameoflife special {role}")
    spec = K.SPECS["laplacian"]
    dims = (20, 10, 8)
    coefs, arrays = K.random_case(spec, dims, 9)
    arrays["w0"][::5] = np.inf
    arrays["w0"][2::9] = -np.inf
    arrays["w0"][4::23] = np.nan
    arrays["w0"][6::31] = 1e-310
    want = oracle.run(spec, dims, coefs, arrays, 2)
    got = K.GpuBackend(gpu_ctx).run(spec, dims, coefs, arrays, 2)
    for role in want:
        assert_bit_equal(got[role], want[role], f"laplacian special {role}")


@pytest.mark.parametrize("dims", [(16, 4, 16), (20, 13, 7), (64, 33, 9), (512, 256, 256)])
def test_gpu_matmul_normwise(gpu_ctx, oracle, dims):
    """FP64 MFMA blocks k by 4 with fused products; the sequential-k reference
    rounds after every product.  Tolerance: |C_gpu - C_ref| <= 4*ny*eps*(|A||B| + |C0|)
    elementwise, nt = 3 accumulations."""
    spec = K.SPECS["matmul"]
    coefs, arrays = K.random_case(spec, dims, 12)
    nt = 3
    want = oracle.run(spec, dims, coefs, arrays, nt)
    got = K.GpuBackend(gpu_ctx).run(spec, dims, coefs, arrays, nt)
    nx, ny, ns = dims
    A = np.abs(arrays["A"].reshape(ny, nx).T)
    B = np.abs(arrays["B"].reshape(ns, ny).T)
    bound = 4 * ny * np.finfo(np.float64).eps * nt * (A @ B) + 1e-300
    diff = np.abs(got["C"] - want["C"]).reshape(ns, nx).T
    assert np.all(diff <= bound), float(np.max(diff / bound))


def test_gpu_timing_and_regcount(gpu_ctx):
    spec = K.SPECS["laplacian"]
    dims = (64, 16, 16)
    coefs, arrays = K.random_case(spec, dims, 1)
    gpu_ctx.drain()
    K.GpuBackend(gpu_ctx).run(spec, dims, coefs, arrays, 3)
    t = gpu_ctx.last_kernel_time()
    assert 0 < t < 1.0
    times = gpu_ctx.drain()
    assert len(times) == 3 and all(x > 0 for x in times)
    assert gpu_ctx.regcount("laplacian") > 0


def test_gpu_rejects_bad_arguments(gpu_ctx, kgmod):
    a = gpu_ctx.empty(64)
    with pytest.raises(kgmod.KGError, match="invalid argument"):
        gpu_ctx.call("laplacian", 4, 4, 4, 0.1, 0.2, a, a)          # aliasing
    with pytest.raises(kgmod.KGError, match="invalid argument"):
        gpu_ctx.call("laplacian", -1, 4, 4, 0.1, 0.2, a, gpu_ctx.empty(64))
    with pytest.raises(kgmod.KGError, match="invalid argument"):
        gpu_ctx.call("gameoflife", 8, 8, None, a)
    with pytest.raises(kgmod.KGError, match="invalid argument"):
        gpu_ctx.call("whispering", 8, 8, 0.5, 0.5, a, a, *[gpu_ctx.empty(64) for _ in range(8)])
