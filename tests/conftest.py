import sys
from pathlib import Path

import numpy as np
import pytest

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))

import kgspec as K  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libkgstencil.so)")
    config.addinivalue_line("markers", "slow: full-size CPU oracle runs (tens of seconds)")


@pytest.fixture(scope="session")
def oracle():
    if not K.ORACLE_LIB.exists():
        pytest.fail(f"{K.ORACLE_LIB} not built: run `make -C oracle` (or __graft_entry__.build())")
    return K.OracleBackend()


@pytest.fixture(scope="session")
def ref():
    if not K.REF_IEEE_LIB.exists():
        pytest.skip("oracle/_ref not built (needs the reference sources): `make -C oracle ref`")
    return K.RefBackend()


@pytest.fixture(scope="session")
def kgmod():
    return K.load_kgstencil()


@pytest.fixture(scope="session")
def gpu_ctx(kgmod):
    """The product library on device 0.  Fails (does not skip) when the
    library or the device is missing: GPU tests must run the HIP path."""
    ctx = kgmod.Context(0)
    yield ctx
    ctx.close()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def assert_bit_equal(got, want, what=""):
    got = np.ascontiguousarray(got, dtype=np.float64)
    want = np.ascontiguousarray(want, dtype=np.float64)
    assert got.shape == want.shape, f"{what}: shape {got.shape} != {want.shape}"
    gb, wb = got.view(np.uint64), want.view(np.uint64)
    # NaN payloads may differ between CPU and GPU; compare NaN-ness there.
    both_nan = np.isnan(got) & np.isnan(want)
    bad = (gb != wb) & ~both_nan
    if bad.any():
        idx = np.flatnonzero(bad)
        i = idx[0]
        raise AssertionError(f"{what}: {idx.size} of {got.size} values differ; first at flat "
                             f"index {i}: got {got[i]!r} want {want[i]!r}")
