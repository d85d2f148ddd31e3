"""ctypes binding of libkgstencil.so (include/kg_stencil.h).

Python host access to the MI355X stencil library: device buffers, transfers,
per-launch timing and one method per reference kernel.  The library is the
only compute path; if it is not built, or no HIP device is present, the calls
raise -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "lib" / "libkgstencil.so"

_dp = C.c_void_p
_i = C.c_int
_d = C.c_double
_f = C.c_float

# kg_dtype codes
F64, F32, BF16 = 0, 1, 2

# name -> argtypes after the context pointer (None marks a device pointer)
SIGNATURES = {
    "kg_laplacian": [_i, _i, _i, _d, _d, _dp, _dp],
    "kg_divergence": [_i, _i, _i, _d, _d, _d, _dp, _dp, _dp, _dp],
    "kg_gradient": [_i, _i, _i, _d, _d, _d, _dp, _dp, _dp, _dp],
    "kg_wave13pt": [_i, _i, _i, _d, _d, _d, _dp, _dp, _dp],
    "kg_uxx1": [_i, _i, _i, _d, _d, _dp, _dp, _dp, _dp, _dp, _dp],
    "kg_tricubic": [_i, _i, _i, _dp, _dp, _dp, _dp, _dp],
    "kg_tricubic2": [_i, _i, _i, _dp, _dp, _dp, _dp, _dp],
    "kg_lapgsrb": [_i, _i, _i, _d, _d, _d, _d, _dp, _dp],
    "kg_jacobi": [_i, _i, _d, _d, _d, _dp, _dp],
    "kg_gaussblur": [_i, _i, _d, _d, _d, _d, _d, _d, _dp, _dp],
    "kg_whispering": [_i, _i, _d, _d] + [_dp] * 10,
    "kg_gameoflife": [_i, _i, _dp, _dp],
    "kg_life_u8": [_i, _i, _dp, _dp],
    "kg_laplacian_f32": [_i, _i, _i, _f, _f, _dp, _dp],
    "kg_laplacian_bf16": [_i, _i, _i, _f, _f, _dp, _dp],
    "kg_wave13pt_f32": [_i, _i, _i, _f, _f, _f, _dp, _dp, _dp],
    "kg_wave13pt_bf16": [_i, _i, _i, _f, _f, _f, _dp, _dp, _dp],
    "kg_gaussblur3_f64": [_i, _i, _i, _d, _d, _d, _d, _d, _d, _dp, _dp],
    "kg_gaussblur3_f32": [_i, _i, _i, _f, _f, _f, _f, _f, _f, _dp, _dp],
    "kg_field_stats": [_i, C.c_size_t, _dp, _i, _dp, C.POINTER(_d)],
    "kg_matmul": [_i, _i, _i, _dp, _dp, _dp],
    "kg_stream_copy": [_dp, _dp, C.c_size_t],
}

# Every symbol include/kg_stencil.h declares.
EXPORTS = [
    "kg_device_count", "kg_init", "kg_destroy", "kg_alloc", "kg_free", "kg_h2d", "kg_d2h",
    "kg_d2d", "kg_sync", "kg_timing_enable", "kg_last_kernel_time", "kg_timing_drain",
    "kg_last_kernel_name", "kg_kernel_regcount", "kg_strerror", "kg_last_error",
] + list(SIGNATURES)


class KGError(RuntimeError):
    pass


_lib = None


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load libkgstencil.so and declare its prototypes (raises if not built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise KGError(f"{p} is not built; run `make -C kernelgen-perf-tests_amd` "
                      "or __graft_entry__.build()")
    lib = C.CDLL(str(p))
    lib.kg_strerror.restype = C.c_char_p
    lib.kg_strerror.argtypes = [_i]
    lib.kg_last_error.restype = C.c_char_p
    lib.kg_last_error.argtypes = []
    lib.kg_last_kernel_name.restype = C.c_char_p
    lib.kg_last_kernel_name.argtypes = [_dp]
    lib.kg_device_count.argtypes = [C.POINTER(_i)]
    lib.kg_init.argtypes = [C.POINTER(_dp), _i]
    lib.kg_destroy.argtypes = [_dp]
    lib.kg_alloc.argtypes = [_dp, C.POINTER(_dp), C.c_size_t]
    lib.kg_free.argtypes = [_dp, _dp]
    for f in ("kg_h2d", "kg_d2h", "kg_d2d"):
        getattr(lib, f).argtypes = [_dp, _dp, _dp, C.c_size_t]
    lib.kg_sync.argtypes = [_dp]
    lib.kg_timing_enable.argtypes = [_dp, _i]
    lib.kg_last_kernel_time.argtypes = [_dp, C.POINTER(_d)]
    lib.kg_timing_drain.argtypes = [_dp, C.POINTER(_d), _i, C.POINTER(_i)]
    lib.kg_kernel_regcount.argtypes = [_dp, C.c_char_p, C.POINTER(_i)]
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = [_dp] + args
        fn.restype = _i
    if path is None:
        _lib = lib
    return lib


def _check(lib, rc: int, what: str) -> None:
    if rc != 0:
        msg = lib.kg_last_error().decode(errors="replace")
        raise KGError(f"{what}: {lib.kg_strerror(rc).decode()} ({rc}): {msg}")


def device_count() -> int:
    lib = load()
    n = _i(0)
    _check(lib, lib.kg_device_count(C.byref(n)), "kg_device_count")
    return n.value


class DeviceArray:
    """A device allocation owned by a Context, with a numpy shape/dtype."""

    def __init__(self, ctx: "Context", shape, dtype=np.float64):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _dp()
        _check(ctx.lib, ctx.lib.kg_alloc(ctx.ptr, C.byref(p), self.nbytes), "kg_alloc")
        self.ptr = p.value or 0

    @property
    def handle(self) -> int:
        return self.ptr

    def upload(self, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"upload of {a.nbytes} bytes into a {self.nbytes}-byte buffer")
        _check(self.ctx.lib, self.ctx.lib.kg_h2d(self.ctx.ptr, self.ptr, a.ctypes.data, a.nbytes),
               "kg_h2d")
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _check(self.ctx.lib, self.ctx.lib.kg_d2h(self.ctx.ptr, out.ctypes.data, self.ptr,
                                                 self.nbytes), "kg_d2h")
        return out

    def copy_from(self, other: "DeviceArray") -> "DeviceArray":
        if other.nbytes != self.nbytes:
            raise ValueError("device copy between buffers of different size")
        _check(self.ctx.lib, self.ctx.lib.kg_d2d(self.ctx.ptr, self.ptr, other.ptr, self.nbytes),
               "kg_d2d")
        return self

    def free(self) -> None:
        if self.ptr:
            _check(self.ctx.lib, self.ctx.lib.kg_free(self.ctx.ptr, self.ptr), "kg_free")
            self.ptr = 0


class Context:
    """One kg_ctx: a device and a stream.  Kernel methods take DeviceArrays."""

    def __init__(self, device: int = 0, lib_path=None):
        self.lib = load(lib_path)
        p = _dp()
        _check(self.lib, self.lib.kg_init(C.byref(p), device), "kg_init")
        self.ptr = p.value
        self._buffers: list[DeviceArray] = []

    def close(self) -> None:
        if self.ptr:
            for b in self._buffers:
                b.free()
            self._buffers.clear()
            self.lib.kg_destroy(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def empty(self, shape, dtype=np.float64) -> DeviceArray:
        b = DeviceArray(self, shape, dtype)
        self._buffers.append(b)
        return b

    def array(self, a: np.ndarray) -> DeviceArray:
        return self.empty(a.shape, a.dtype).upload(a)

    def sync(self) -> None:
        _check(self.lib, self.lib.kg_sync(self.ptr), "kg_sync")

    def timing(self, on: bool) -> None:
        _check(self.lib, self.lib.kg_timing_enable(self.ptr, int(on)), "kg_timing_enable")

    def last_kernel_time(self) -> float:
        s = _d()
        _check(self.lib, self.lib.kg_last_kernel_time(self.ptr, C.byref(s)), "kg_last_kernel_time")
        return s.value

    def drain(self, cap: int = 8192) -> list[float]:
        buf = (_d * cap)()
        n = _i(0)
        _check(self.lib, self.lib.kg_timing_drain(self.ptr, buf, cap, C.byref(n)), "kg_timing_drain")
        return list(buf[: n.value])

    def regcount(self, kernel: str) -> int:
        n = _i(0)
        _check(self.lib, self.lib.kg_kernel_regcount(self.ptr, kernel.encode(), C.byref(n)),
               "kg_kernel_regcount")
        return n.value

    def field_stats(self, x: "DeviceArray", dtype: int, ref: "DeviceArray | None" = None,
                    ref_dtype: int = F64, n: int | None = None) -> dict:
        """FP64 sum, sum of squares, max |x|, and (with ref) sum (x-ref)^2 and
        max |x-ref| of a device field (kg_field_stats)."""
        out = (_d * 5)()
        if n is None:
            n = int(np.prod(x.shape))
        rc = self.lib.kg_field_stats(self.ptr, dtype, n, x.ptr, ref_dtype,
                                     ref.ptr if ref is not None else None, out)
        _check(self.lib, rc, "kg_field_stats")
        return {"sum": out[0], "sumsq": out[1], "max": out[2], "err_sumsq": out[3],
                "err_max": out[4]}

    def call(self, fn: str, *args) -> None:
        """Call kg_<fn>; DeviceArray arguments are passed as device pointers."""
        conv = [a.ptr if isinstance(a, DeviceArray) else a for a in args]
        rc = getattr(self.lib, "kg_" + fn)(self.ptr, *conv)
        _check(self.lib, rc, "kg_" + fn)
