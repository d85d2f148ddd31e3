"""The C ABI library loads and exports every symbol include/kg_stencil.h
declares; host-side error handling works without a GPU."""
import ctypes as C
import re
from pathlib import Path

import pytest

import kgspec as K

HEADER = K.ROOT / "include" / "kg_stencil.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_every_reference_kernel():
    names = declared_functions()
    for k in ["laplacian", "divergence", "gradient", "wave13pt", "uxx1", "tricubic", "tricubic2",
              "lapgsrb", "jacobi", "gaussblur", "whispering", "gameoflife", "matmul"]:
        assert f"kg_{k}" in names


def test_library_exports_every_declared_symbol(kgmod):
    lib = kgmod.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_functions()) <= set(kgmod.EXPORTS)


def test_strerror_and_null_arguments(kgmod):
    lib = kgmod.load()
    assert lib.kg_strerror(0) == b"success"
    assert lib.kg_strerror(-1) == b"invalid argument"
    assert lib.kg_strerror(-4) == b"no such HIP device"
    # Argument checks run before any device call.
    assert lib.kg_laplacian(None, 4, 4, 4, 0.1, 0.2, None, None) == -1
    assert b"null" in lib.kg_last_error()
    assert lib.kg_init(None, 0) == -1
    assert lib.kg_device_count(None) == -1
    assert lib.kg_destroy(None) == 0


def test_library_is_built_for_gfx950():
    so = K.PKG_DIR / "lib" / "libkgstencil.so"
    data = so.read_bytes()
    assert b"gfx950" in data
