"""kg_poisson_solve_noalloc against kg_poisson_solve, through the C ABI.

The no-allocation entry point runs the same kernels in the same order on a
workspace reserved inside the library, so the solution and every residual of
the report must be bit-identical to the allocating entry point's.  Two calls
in a row, the second on a smaller grid, show that the shared workspace's
leftover contents from the previous call do not leak into the next one.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rhs(n, seed):
    rng = np.random.default_rng(seed)
    f = rng.uniform(-1, 1, (n + 1, n + 1))
    f[0, :] = f[-1, :] = f[:, 0] = f[:, -1] = 0
    return f


def _solve(ctx, kgmod, f, n, noalloc, u0=None):
    df = ctx.array(f)
    du = ctx.array(np.zeros_like(f) if u0 is None else u0)
    rep = kgmod.poisson_solve(ctx, n, df, du, tol=1e-10, noalloc=noalloc)
    u = du.download()
    df.free()
    du.free()
    return u, rep


def _same(a, b, what):
    ua, ra = a
    ub, rb = b
    assert np.array_equal(ua.view(np.uint64), ub.view(np.uint64)), f"{what}: u differs"
    for k in ("outer", "vcycles", "levels", "rel_residual", "history"):
        assert ra[k] == rb[k], (what, k, ra[k], rb[k])


def test_noalloc_matches_solve_two_calls(gpu_ctx, kgmod):
    big, small = kgmod.POISSON_NOALLOC_MAX_N, 256
    f_big, f_small = _rhs(big, 41), _rhs(small, 42)
    want_big = _solve(gpu_ctx, kgmod, f_big, big, False)
    want_small = _solve(gpu_ctx, kgmod, f_small, small, False)
    got_big = _solve(gpu_ctx, kgmod, f_big, big, True)
    got_small = _solve(gpu_ctx, kgmod, f_small, small, True)
    _same(got_big, want_big, f"n={big}")
    _same(got_small, want_small, f"n={small} after n={big}")
    assert want_big[1]["rel_residual"] < 1e-10 and want_small[1]["rel_residual"] < 1e-10


def test_noalloc_nonzero_initial_guess(gpu_ctx, kgmod):
    n = 128
    f = _rhs(n, 43)
    u0 = np.zeros_like(f)
    u0[1:-1, 1:-1] = np.random.default_rng(44).uniform(-1e-3, 1e-3, (n - 1, n - 1))
    _same(_solve(gpu_ctx, kgmod, f, n, True, u0), _solve(gpu_ctx, kgmod, f, n, False, u0),
          "nonzero initial guess")


def test_noalloc_rejects_sizes_beyond_its_workspace(gpu_ctx, kgmod):
    n = 2 * kgmod.POISSON_NOALLOC_MAX_N
    a = gpu_ctx.empty(((n + 1) * (n + 1),), np.float64)
    b = gpu_ctx.empty(((n + 1) * (n + 1),), np.float64)
    with pytest.raises(kgmod.KGError, match="power of two in"):
        kgmod.poisson_solve(gpu_ctx, n, a, b, noalloc=True)
    with pytest.raises(kgmod.KGError):
        kgmod.poisson_solve(gpu_ctx, 6, a, b, noalloc=True)
    with pytest.raises(kgmod.KGError):
        kgmod.poisson_solve(gpu_ctx, 8, a, a, noalloc=True)
    a.free()
    b.free()
