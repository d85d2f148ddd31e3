"""GPU parity of kg_life_u8 (gameoflife on uint8 cells, X mode) against the
oracle's oracle_life_u8, which is itself pinned to a brute-force Conway rule
and to the glider/blinker known answers in test_oracle.py.  Integer work:
tolerance 0 (every byte equal, border bytes of the output untouched)."""
import ctypes as C

import numpy as np
import pytest

import kgspec as K

pytestmark = pytest.mark.gpu

SIZES = [(3, 3), (4, 5), (16, 3), (17, 12), (18, 18), (33, 5), (48, 40), (100, 37),
         (4096, 19), (4113, 9), (8192, 33), (1040, 70),
         # band and segment edges of the vector kernel: 16-row bands, 1024-cell
         # wave segments, 4096-cell workgroups, a last band of one row
         (32, 66), (2064, 100), (1024, 35), (5120, 34), (4112, 67),
         # 150 workgroups: rounds of 8 chunks of 8 bands (128) plus a tail
         (8192, 300), (4096, 600)]


def _oracle_steps(oracle, board, nt, fill=0xAB):
    ny, nx = board.shape
    cur = board.copy()
    for _ in range(nt):
        nxt = np.full_like(cur, fill) if nt == 1 else cur.copy()
        oracle.lib.oracle_life_u8(C.c_int(nx), C.c_int(ny), C.c_void_p(cur.ctypes.data),
                                  C.c_void_p(nxt.ctypes.data))
        cur = nxt
    return cur


def _gpu_steps(ctx, board, nt, fill=0xAB):
    ny, nx = board.shape
    a = ctx.array(board)
    b = ctx.array(np.full_like(board, fill) if nt == 1 else board)
    for _ in range(nt):
        ctx.call("life_u8", nx, ny, a, b)
        a, b = b, a
    ctx.sync()
    out = a.download()
    a.free()
    b.free()
    return out


@pytest.mark.parametrize("dims", SIZES)
def test_life_u8_matches_oracle_ragged(gpu_ctx, oracle, dims):
    nx, ny = dims
    rng = np.random.default_rng(nx * 7 + ny)
    board = (rng.random((ny, nx)) < 0.3).astype(np.uint8)
    # one step into a buffer prefilled with 0xAB: the border must stay 0xAB
    np.testing.assert_array_equal(_gpu_steps(gpu_ctx, board, 1), _oracle_steps(oracle, board, 1),
                                  err_msg=f"life_u8 {dims} one step")
    np.testing.assert_array_equal(_gpu_steps(gpu_ctx, board, 5), _oracle_steps(oracle, board, 5),
                                  err_msg=f"life_u8 {dims} five steps")


def test_life_u8_nonbinary_cells(gpu_ctx, oracle):
    """Any nonzero byte counts as alive (values up to 0xff, including 0x80)."""
    rng = np.random.default_rng(11)
    nx, ny = 4096 + 64, 40
    board = rng.integers(0, 256, (ny, nx), dtype=np.uint8)
    board[rng.random((ny, nx)) < 0.6] = 0
    board[0, :16] = 0x80
    np.testing.assert_array_equal(_gpu_steps(gpu_ctx, board, 1), _oracle_steps(oracle, board, 1))


def test_life_u8_full_size_bernoulli(gpu_ctx, oracle):
    """The X-mode configuration: 8192x8192 cells, Bernoulli p = 0.3, 3 steps."""
    rng = np.random.default_rng(2024)
    board = (rng.random((8192, 8192)) < 0.3).astype(np.uint8)
    np.testing.assert_array_equal(_gpu_steps(gpu_ctx, board, 3), _oracle_steps(oracle, board, 3))


def test_life_u8_glider_translates(gpu_ctx):
    """A glider moves one cell diagonally every 4 generations."""
    nx, ny = 64, 48
    board = np.zeros((ny, nx), np.uint8)
    for (y, x) in [(1, 2), (2, 3), (3, 1), (3, 2), (3, 3)]:
        board[10 + y, 20 + x] = 1
    out = _gpu_steps(gpu_ctx, board, 8)
    np.testing.assert_array_equal(out, np.roll(board, (2, 2), axis=(0, 1)))


def test_life_u8_rejects_bad_arguments(gpu_ctx, kgmod):
    a = gpu_ctx.empty((64,), np.uint8)
    with pytest.raises(kgmod.KGError):
        gpu_ctx.call("life_u8", 8, 8, a, a)
    with pytest.raises(kgmod.KGError):
        gpu_ctx.call("life_u8", 8, 8, None, a)
    b = gpu_ctx.empty((80,), np.uint8)
    with pytest.raises(kgmod.KGError):
        gpu_ctx.call("life_u8", 8, 8, a.ptr + 1, b)
