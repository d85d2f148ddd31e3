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
    "kg_vecadd": [_i, _i, _i, _dp, _dp, _dp],
    "kg_matvec": [_i, _i, _dp, _dp, _dp],
    "kg_sincos": [_i, _i, _i, _dp, _dp, _dp],
    "kg_vecadd_f32": [_i, _i, _i, _dp, _dp, _dp],
    "kg_matvec_f32": [_i, _i, _dp, _dp, _dp],
    "kg_sincos_f32": [_i, _i, _i, _dp, _dp, _dp],
    "kg_laplacian_f32": [_i, _i, _i, _f, _f, _dp, _dp],
    "kg_laplacian_bf16": [_i, _i, _i, _f, _f, _dp, _dp],
    "kg_wave13pt_f32": [_i, _i, _i, _f, _f, _f, _dp, _dp, _dp],
    "kg_wave13pt_bf16": [_i, _i, _i, _f, _f, _f, _dp, _dp, _dp],
    "kg_gaussblur3_f64": [_i, _i, _i, _d, _d, _d, _d, _d, _d, _dp, _dp],
    "kg_gaussblur3_f32": [_i, _i, _i, _f, _f, _f, _f, _f, _f, _dp, _dp],
    "kg_field_stats": [_i, C.c_size_t, _dp, _i, _dp, C.POINTER(_d)],
    "kg_fill_uniform": [_i, C.c_size_t, _dp, C.c_ulonglong],
    "kg_convert": [_i, _dp, _i, _dp, C.c_size_t],
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
    _declare_batch(lib)
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


# --------------------------------------------------------------- batches
KIND_LAPLACIAN, KIND_WAVE13PT, KIND_GAUSSBLUR3, KIND_JACOBI2, KIND_LIFE2 = 0, 1, 2, 3, 4
KINDS = {"laplacian": KIND_LAPLACIAN, "wave13pt": KIND_WAVE13PT, "gaussblur3": KIND_GAUSSBLUR3,
         "jacobi2": KIND_JACOBI2, "life2": KIND_LIFE2}
NLEV = {KIND_LAPLACIAN: 2, KIND_WAVE13PT: 3, KIND_GAUSSBLUR3: 2, KIND_JACOBI2: 2, KIND_LIFE2: 2}


class GridSpec(C.Structure):
    _fields_ = [("kind", _i), ("dtype", _i), ("nx", _i), ("ny", _i), ("ns", _i),
                ("coef", _d * 6)]


_BATCH_SIGS = {
    "kg_batch_create": [_dp, C.POINTER(_dp)],
    "kg_batch_destroy": [_dp],
    "kg_batch_add": [_dp, C.POINTER(GridSpec), C.POINTER(_i)],
    "kg_batch_remove": [_dp, _i],
    "kg_batch_upload": [_dp, _i, _i, _dp],
    "kg_batch_download": [_dp, _i, _i, _dp],
    "kg_batch_newest_level": [_dp, _i, C.POINTER(_i)],
    "kg_batch_steps_done": [_dp, _i, C.POINTER(C.c_longlong)],
    "kg_batch_step": [_dp, _i, _i],
    "kg_batch_info": [_dp, C.POINTER(_d), C.POINTER(_d), C.POINTER(_i), C.POINTER(_i)],
    "kg_batch_tune": [_dp, _i, C.POINTER(_i)],
    "kg_batch_grid_shape": [_dp, _i, C.POINTER(_i), C.POINTER(_i)],
    "kg_batch_set_shape": [_dp, _i, _i, _i],
    "kg_batch_memory": [_dp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "kg_batch_fill": [_dp, C.c_ulonglong],
    "kg_batch_fill_grid": [_dp, _i, C.c_ulonglong],
    "kg_batch_step_sequential": [_dp, _i],
}
EXPORTS += list(_BATCH_SIGS) + ["kg_serve_run", "kg_poisson_solve", "kg_poisson_solve_noalloc"]


def _declare_batch(lib):
    if getattr(lib, "_kg_batch_declared", False):
        return
    for name, args in _BATCH_SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _i
    lib._kg_batch_declared = True


class Batch:
    """Many independent 3D grids stepped together (kg_batch_* of the C ABI).

    Host arrays are dense (ns, ny, nx) in the grid's dtype (float64/float32).
    A grid's levels share one boundary: the cells outside the kind's interior
    box must hold the same values in every uploaded level (fill() does so).
    """

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.lib = ctx.lib
        _declare_batch(self.lib)
        p = _dp()
        _check(self.lib, self.lib.kg_batch_create(ctx.ptr, C.byref(p)), "kg_batch_create")
        self.ptr = p.value
        self.specs: dict[int, tuple] = {}

    def close(self) -> None:
        if self.ptr:
            self.lib.kg_batch_destroy(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add(self, kind: str, dtype: int, dims, coefs) -> int:
        s = GridSpec()
        s.kind, s.dtype = KINDS[kind], dtype
        s.nx, s.ny, s.ns = (int(d) for d in dims)
        for q, c in enumerate(coefs):
            s.coef[q] = float(c)
        gid = _i(0)
        _check(self.lib, self.lib.kg_batch_add(self.ptr, C.byref(s), C.byref(gid)), "kg_batch_add")
        self.specs[gid.value] = (kind, dtype, tuple(int(d) for d in dims))
        return gid.value

    def remove(self, gid: int) -> None:
        _check(self.lib, self.lib.kg_batch_remove(self.ptr, gid), "kg_batch_remove")
        del self.specs[gid]

    def _np_dtype(self, gid):
        return np.float64 if self.specs[gid][1] == F64 else np.float32

    def upload(self, gid: int, level: int, a: np.ndarray) -> None:
        kind, dt, (nx, ny, ns) = self.specs[gid]
        a = np.ascontiguousarray(a, dtype=self._np_dtype(gid))
        if a.size != nx * ny * ns:
            raise ValueError("upload: wrong size")
        _check(self.lib, self.lib.kg_batch_upload(self.ptr, gid, level, a.ctypes.data),
               "kg_batch_upload")

    def download(self, gid: int, level: int | None = None) -> np.ndarray:
        kind, dt, (nx, ny, ns) = self.specs[gid]
        if level is None:
            level = self.newest_level(gid)
        out = np.empty((ns, ny, nx), dtype=self._np_dtype(gid))
        _check(self.lib, self.lib.kg_batch_download(self.ptr, gid, level, out.ctypes.data),
               "kg_batch_download")
        return out

    def newest_level(self, gid: int) -> int:
        v = _i(0)
        _check(self.lib, self.lib.kg_batch_newest_level(self.ptr, gid, C.byref(v)),
               "kg_batch_newest_level")
        return v.value

    def steps_done(self, gid: int) -> int:
        v = C.c_longlong(0)
        _check(self.lib, self.lib.kg_batch_steps_done(self.ptr, gid, C.byref(v)),
               "kg_batch_steps_done")
        return v.value

    def step(self, n: int = 1, graph: bool = True) -> None:
        _check(self.lib, self.lib.kg_batch_step(self.ptr, n, int(graph)), "kg_batch_step")

    def info(self) -> dict:
        p, b, it, wg = _d(), _d(), _i(), _i()
        _check(self.lib, self.lib.kg_batch_info(self.ptr, C.byref(p), C.byref(b), C.byref(it),
                                                C.byref(wg)), "kg_batch_info")
        return {"points_per_step": p.value, "bytes_per_step": b.value, "work_items": it.value,
                "workgroups": wg.value}

    def tune(self, reps: int = 3) -> int:
        n = _i(0)
        _check(self.lib, self.lib.kg_batch_tune(self.ptr, reps, C.byref(n)), "kg_batch_tune")
        return n.value

    def shape(self, gid: int) -> tuple[int, int]:
        a, b = _i(), _i()
        _check(self.lib, self.lib.kg_batch_grid_shape(self.ptr, gid, C.byref(a), C.byref(b)),
               "kg_batch_grid_shape")
        return a.value, b.value

    def set_shape(self, gid: int, lanes_per_row: int, planes_per_tile: int) -> None:
        """Forces the tile shape of grid gid's size class: 64/32/16 lanes per
        row, 0 the flat body, -1 packed rows, -2 the direct row body (planar
        kinds: gaussblur3, jacobi2, life2)."""
        _check(self.lib, self.lib.kg_batch_set_shape(self.ptr, gid, lanes_per_row,
                                                     planes_per_tile), "kg_batch_set_shape")

    def fill(self, seed: int) -> None:
        _check(self.lib, self.lib.kg_batch_fill(self.ptr, seed), "kg_batch_fill")

    def step_sequential(self, n: int = 1) -> None:
        _check(self.lib, self.lib.kg_batch_step_sequential(self.ptr, n), "kg_batch_step_sequential")

    def memory(self) -> tuple[int, int]:
        r, u = C.c_size_t(), C.c_size_t()
        _check(self.lib, self.lib.kg_batch_memory(self.ptr, C.byref(r), C.byref(u)),
               "kg_batch_memory")
        return r.value, u.value


# ---------------------------------------------------------------- Poisson
class PoissonReport(C.Structure):
    _fields_ = [("outer", _i), ("vcycles", _i), ("levels", _i), ("rel_residual", _d),
                ("seconds", _d), ("history", _d * 32)]


# largest n of kg_poisson_solve_noalloc (KG_POISSON_NOALLOC_MAX_N)
POISSON_NOALLOC_MAX_N = 1024


def poisson_solve(ctx: Context, n: int, f: DeviceArray, u: DeviceArray, tol: float = 1e-10,
                  max_outer: int = 20, vcycles: int = 2, noalloc: bool = False) -> dict:
    """kg_poisson_solve: FP32 multigrid inside FP64 defect correction.
    noalloc=True calls kg_poisson_solve_noalloc (same solve, workspace
    reserved inside the library, n <= POISSON_NOALLOC_MAX_N)."""
    name = "kg_poisson_solve_noalloc" if noalloc else "kg_poisson_solve"
    fn = getattr(ctx.lib, name)
    fn.argtypes = [_dp, _i, _dp, _dp, _d, _i, _i, C.POINTER(PoissonReport)]
    fn.restype = _i
    rep = PoissonReport()
    _check(ctx.lib, fn(ctx.ptr, n, f.ptr, u.ptr, tol, max_outer, vcycles, C.byref(rep)), name)
    return {"outer": rep.outer, "vcycles": rep.vcycles, "levels": rep.levels,
            "rel_residual": rep.rel_residual, "seconds": rep.seconds,
            "history": list(rep.history[: min(rep.outer + 1, 32)])}


# ---------------------------------------------------------------- serving
SERVE_CONTINUOUS, SERVE_SEQUENTIAL = 0, 1


class Job(C.Structure):
    _fields_ = [("kind", _i), ("dtype", _i), ("nx", _i), ("ny", _i), ("ns", _i), ("steps", _i),
                ("tol", _d), ("coef", _d * 6), ("arrival", _d), ("seed", C.c_ulonglong),
                ("out", _dp)]


class JobResult(C.Structure):
    _fields_ = [("admitted", _d), ("finished", _d), ("steps_run", C.c_longlong),
                ("checksum", _d * 2), ("status", _i)]


class ServeStats(C.Structure):
    _fields_ = [("wall", _d), ("points", _d), ("p50", _d), ("p99", _d), ("mean_latency", _d),
                ("peak_active", _i), ("peak_bytes", C.c_size_t)]


def serve(ctx: Context, jobs: list[dict], hbm_budget: int, max_active: int = 64,
          mode: int = SERVE_CONTINUOUS, outputs: bool = False):
    """Run kg_serve_run over jobs (dicts: kind, dtype, dims, steps, coef,
    arrival, seed, tol).  Returns (results, stats, outs) where outs[i] is the
    final field (ns, ny, nx) when outputs is set."""
    lib = ctx.lib
    fn = lib.kg_serve_run
    fn.argtypes = [_dp, C.POINTER(Job), _i, C.c_size_t, _i, _i, C.POINTER(JobResult),
                   C.POINTER(ServeStats)]
    fn.restype = _i
    n = len(jobs)
    arr = (Job * max(n, 1))()
    outs = []
    for i, j in enumerate(jobs):
        a = arr[i]
        a.kind = KINDS[j["kind"]]
        a.dtype = j["dtype"]
        a.nx, a.ny, a.ns = j["dims"]
        a.steps = j["steps"]
        a.tol = j.get("tol", 0.0)
        for q, c in enumerate(j.get("coef", ())):
            a.coef[q] = c
        a.arrival = j.get("arrival", 0.0)
        a.seed = j.get("seed", i + 1)
        if outputs:
            nx, ny, ns = j["dims"]
            o = np.empty((ns, ny, nx), np.float64 if j["dtype"] == F64 else np.float32)
            outs.append(o)
            a.out = o.ctypes.data
        else:
            outs.append(None)
            a.out = None
    res = (JobResult * max(n, 1))()
    st = ServeStats()
    _check(lib, fn(ctx.ptr, arr, n, hbm_budget, max_active, mode, res, C.byref(st)),
           "kg_serve_run")
    results = [{"admitted": res[i].admitted, "finished": res[i].finished,
                "latency": res[i].finished - jobs[i].get("arrival", 0.0),
                "steps_run": res[i].steps_run, "checksum": (res[i].checksum[0], res[i].checksum[1]),
                "status": res[i].status} for i in range(n)]
    stats = {f: getattr(st, f) for f, _ in ServeStats._fields_}
    return results, stats, outs
