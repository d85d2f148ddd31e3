"""Kernel specs and backends shared by the parity tests.

A spec names a reference kernel's arguments in the reference's own order
(dims, coefficients, arrays), how its main() rotates buffers between steps
(e.g. laplacian/laplacian.c:299-300), and how it fills inputs with glibc rand()
(e.g. laplacian/laplacian.c:141-164).  Three backends run the same spec:

* OracleBackend -- oracle/liboracle.so (the checker; IEEE source order),
* RefBackend    -- oracle/_ref/libref_ieee.so, the reference's own kernel
                   functions compiled from its sources (pins the oracle),
* GpuBackend    -- libkgstencil.so through the C ABI (the product).
"""
from __future__ import annotations

import ctypes as C
import importlib.util
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_LIB = ROOT / "oracle" / "liboracle.so"
REF_IEEE_LIB = ROOT / "oracle" / "_ref" / "libref_ieee.so"
REF_FAST_LIB = ROOT / "oracle" / "_ref" / "libref_fast.so"
PKG_DIR = ROOT / "kernelgen-perf-tests_amd"


def load_kgstencil():
    spec = importlib.util.spec_from_file_location("kgstencil", PKG_DIR / "kgstencil.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@dataclass
class Spec:
    name: str
    dim: int                      # 3: (nx, ny, ns); 2: (nx, ny)
    coefs: tuple                  # coefficient names, in signature order
    arrays: tuple                 # array roles, in signature order
    rotate: str                   # "pp2", "rot3", "none", "whisper"
    pair: tuple = ()              # roles rotated by pp2/rot3
    fill: tuple = ()              # arrays filled by rand(), in the per-index order
    coef_gen: tuple = ()          # divisor applied to real_rand() per coefficient
    fortran: bool = False         # by-reference dims/coefs (jacobi_, matmul_)
    ref_symbol: str = ""
    lo: int = 1                   # interior [lo, n-hi) (for info / invariants)
    hi: int = 1
    bytes_per_point: int = 16     # algorithmic HBM bytes per grid point per step
    sizes: dict = field(default_factory=dict)   # role -> multiple of N (whispering e: 2)


SPECS = {s.name: s for s in [
    Spec("laplacian", 3, ("alpha", "beta"), ("w0", "w1"), "pp2", ("w0", "w1"),
         fill=("w0", "w1"), coef_gen=(1, 1), bytes_per_point=16),
    Spec("divergence", 3, ("alpha", "beta", "gamma"), ("u", "ux", "uy", "uz"), "none",
         fill=("u", "ux", "uy", "uz"), coef_gen=(1, 1, 1), bytes_per_point=32),
    Spec("gradient", 3, ("alpha", "beta", "gamma"), ("u", "ux", "uy", "uz"), "none",
         fill=("u", "ux", "uy", "uz"), coef_gen=(1, 1, 1), bytes_per_point=32),
    Spec("wave13pt", 3, ("m0", "m1", "m2"), ("w0", "w1", "w2"), "rot3", ("w0", "w1", "w2"),
         fill=("w0", "w1", "w2"), coef_gen=(1, 6., 6.), lo=2, hi=2, bytes_per_point=24),
    Spec("uxx1", 3, ("c1", "c2"), ("u0", "u1", "d1", "xx", "xy", "xz"), "pp2", ("u0", "u1"),
         fill=("u0", "u1", "d1", "xx", "xy", "xz"), coef_gen=(1, 1), lo=2, hi=1,
         bytes_per_point=48),
    Spec("tricubic", 3, (), ("u0", "u1", "a", "b", "c"), "pp2", ("u0", "u1"),
         fill=("u0", "u1", "a", "b", "c"), lo=1, hi=2, bytes_per_point=40),
    Spec("tricubic2", 3, (), ("u0", "u1", "a", "b", "c"), "pp2", ("u0", "u1"),
         fill=("u0", "u1", "a", "b", "c"), lo=2, hi=2, bytes_per_point=40),
    Spec("lapgsrb", 3, ("c0", "c1", "c2", "c3"), ("w0", "w1"), "pp2", ("w0", "w1"),
         fill=("w0", "w1"), coef_gen=(1, 6., 12., 6.), lo=2, hi=2, bytes_per_point=16),
    Spec("jacobi", 2, ("c0", "c1", "c2"), ("w0", "w1"), "pp2", ("w0", "w1"),
         fill=("w0", "w1"), coef_gen=(1, 4., 4.), fortran=True, ref_symbol="jacobi_",
         bytes_per_point=16),
    Spec("gaussblur", 2, ("s0", "s1", "s2", "s4", "s5", "s8"), ("w0", "w1"), "pp2",
         ("w0", "w1"), fill=("w0", "w1"), coef_gen=(1, 1, 1, 1, 1, 1), lo=2, hi=2,
         bytes_per_point=16),
    Spec("whispering", 2, ("mu", "epsilon"),
         ("e", "h0", "h1", "u_em0", "u_em1", "ca", "cb", "da", "db"), "whisper",
         fill=("e", "e", "h0", "h1", "u_em0", "u_em1", "ca", "cb", "da", "db"),
         coef_gen=(1, 1), bytes_per_point=96, sizes={"e": 2}),
    Spec("gameoflife", 2, (), ("u0", "u1"), "pp2", ("u0", "u1"), fill=("u0", "u1"),
         bytes_per_point=16),
    Spec("matmul", 3, (), ("A", "B", "C"), "none", fortran=True, ref_symbol="matmul_",
         bytes_per_point=0),
]}

STENCILS = [n for n in SPECS if n != "matmul"]


def npoints(spec: Spec, dims) -> int:
    return int(np.prod(dims))


def array_shapes(spec: Spec, dims) -> dict:
    """Flat sizes per role."""
    if spec.name == "matmul":
        nx, ny, ns = dims
        return {"A": nx * ny, "B": ny * ns, "C": nx * ns}
    n = npoints(spec, dims)
    return {r: n * spec.sizes.get(r, 1) for r in spec.arrays}


def random_case(spec: Spec, dims, seed: int):
    """Seeded numpy inputs, uniform [-1,1) like the reference's real_rand()."""
    rng = np.random.default_rng(seed)
    coefs = {}
    for name, div in zip(spec.coefs, spec.coef_gen or (1,) * len(spec.coefs)):
        coefs[name] = float(rng.uniform(-1, 1)) / div
    arrays = {r: rng.uniform(-1, 1, size=n) for r, n in array_shapes(spec, dims).items()}
    if spec.name == "matmul":
        arrays["C"] = np.zeros_like(arrays["C"])
    return coefs, arrays


# ------------------------------------------------- reference main() replay
# Which buffers each reference program averages, and how (e.g.
# laplacian/laplacian.c:159-165 and :374-377).  "final" names the role after
# the last rotation; the reference's idxs bookkeeping selects the buffer the
# role points to (for two-level ping-pong that is the input of the last step).
MEANS = {
    "laplacian": (("w0", "w1"), 2, ("w1",), 1),
    "divergence": (("u", "ux", "uy", "uz"), 4, ("u",), 1),
    "gradient": (("u", "ux", "uy", "uz"), 4, ("ux", "uy", "uz"), 3),
    "wave13pt": (("w0", "w1", "w2"), 3, ("w1",), 1),
    "uxx1": (("u0", "u1", "d1", "xx", "xy", "xz"), 6, ("u1",), 1),
    "tricubic": (("u0", "u1", "a", "b", "c"), 5, ("u1",), 1),
    "tricubic2": (("u0", "u1", "a", "b", "c"), 5, ("u1",), 1),
    "lapgsrb": (("w0", "w1"), 2, ("w1",), 1),
    "jacobi": (("w0", "w1"), 2, ("w1",), 1),
    "gaussblur": (("w0", "w1"), 2, ("w1",), 1),
    "gameoflife": (("u0", "u1"), 2, ("u1",), 1),
}


def replay_reference_main(oracle, name: str, dims, nt: int):
    """Re-run a reference test program's main() with the oracle's kernels:
    glibc rand() with its default seed, coefficients then arrays drawn in the
    reference's order, nt steps with its rotation.  Returns (initial mean,
    final mean, coefficients)."""
    spec = SPECS[name]
    lib = oracle.lib
    lib.oracle_real_rand.restype = C.c_double
    lib.oracle_srand(C.c_uint(1))
    coefs = {}
    for cname, div in zip(spec.coefs, spec.coef_gen or ()):
        coefs[cname] = lib.oracle_real_rand() / div
    shapes = array_shapes(spec, dims)
    arrays = {r: np.empty(n, dtype=np.float64) for r, n in shapes.items()}
    n = npoints(spec, dims)
    if name == "matmul":
        for r in ("A", "B"):
            a = arrays[r]
            ptrs = (C.c_void_p * 1)(a.ctypes.data)
            lib.oracle_fill_interleaved(C.c_size_t(a.size), 1, ptrs)
        arrays["C"][:] = 0.0
        i_mean = arrays["A"].mean() + arrays["B"].mean()
    elif name == "whispering":
        roles = ("h0", "h1", "u_em0", "u_em1", "ca", "cb", "da", "db")
        bases = [arrays["e"].ctypes.data, arrays["e"].ctypes.data + 8] + \
                [arrays[r].ctypes.data for r in roles]
        strides = [2, 2] + [1] * len(roles)
        lib.oracle_fill_strided(C.c_size_t(n), len(bases), (C.c_void_p * len(bases))(*bases),
                                (C.c_long * len(strides))(*strides))
        i_mean = (arrays["e"].sum() + sum(arrays[r].sum() for r in roles)) / n / 12
    else:
        roles = spec.fill
        ptrs = (C.c_void_p * len(roles))(*[arrays[r].ctypes.data for r in roles])
        lib.oracle_fill_interleaved(C.c_size_t(n), len(roles), ptrs)
        mroles, div, _, _ = MEANS[name]
        i_mean = sum(arrays[r].sum() for r in mroles) / n / div
    st = dict(arrays)
    for _ in range(nt):
        oracle.step(spec, dims, coefs, st)
        rotate(spec, st)
    if name == "matmul":
        f_mean = st["C"].mean()
    elif name == "whispering":
        f_mean = (st["e"].sum() + st["h1"].sum() + st["u_em1"].sum()) / n / 4
    else:
        _, _, froles, fdiv = MEANS[name]
        f_mean = sum(st[r].sum() for r in froles) / n / fdiv
    return float(i_mean), float(f_mean), coefs


def driver_format(value: float) -> str:
    """A mean as the benchmark table shows it: printed %f by the program, then
    squeezed to 8 characters with %.*g (reference benchmark:27-53)."""
    v = float("%f" % value)
    for w in range(6, -1, -1):
        s = "%.*g" % (w, v)
        if len(s) <= 8:
            return s
    raise ValueError(value)


# ------------------------------------------------------------------ rotation
def rotate(spec: Spec, st: dict) -> None:
    if spec.rotate == "pp2":
        a, b = spec.pair
        st[a], st[b] = st[b], st[a]
    elif spec.rotate == "rot3":
        a, b, c = spec.pair
        st[a], st[b], st[c] = st[b], st[c], st[a]
    elif spec.rotate == "whisper":
        st["h0"], st["h1"] = st["h1"], st["h0"]
        st["u_em0"], st["u_em1"] = st["u_em1"], st["u_em0"]
        if "e_next" in st:
            st["e"], st["e_next"] = st["e_next"], st["e"]


# ------------------------------------------------------------------ backends
class _HostBackend:
    """Calls a C library whose kernel functions take the reference signature."""

    def __init__(self, path: Path, prefix: str, fortran_ok: bool):
        self.lib = C.CDLL(str(path))
        self.prefix = prefix
        self.fortran_ok = fortran_ok

    def _fn(self, spec: Spec):
        if self.prefix == "oracle_":
            return getattr(self.lib, "oracle_" + spec.name)
        return getattr(self.lib, spec.ref_symbol or spec.name)

    def step(self, spec: Spec, dims, coefs: dict, st: dict) -> None:
        fn = self._fn(spec)
        args = []
        by_ref = spec.fortran and self.prefix != "oracle_"
        keep = []
        for d in dims:
            if by_ref:
                v = C.c_int(int(d)); keep.append(v); args.append(C.byref(v))
            else:
                args.append(C.c_int(int(d)))
        for cname in spec.coefs:
            if by_ref:
                v = C.c_double(coefs[cname]); keep.append(v); args.append(C.byref(v))
            else:
                args.append(C.c_double(coefs[cname]))
        for role in spec.arrays:
            a = st[role]
            assert a.dtype == np.float64 and a.flags.c_contiguous
            args.append(C.c_void_p(a.ctypes.data))
        fn.restype = None
        fn(*args)

    def run(self, spec: Spec, dims, coefs: dict, arrays: dict, nt: int) -> dict:
        st = {k: np.array(v, dtype=np.float64, copy=True) for k, v in arrays.items()}
        for _ in range(nt):
            self.step(spec, dims, coefs, st)
            rotate(spec, st)
        return st


class OracleBackend(_HostBackend):
    def __init__(self, path: Path = ORACLE_LIB):
        super().__init__(path, "oracle_", True)


class RefBackend(_HostBackend):
    def __init__(self, path: Path = REF_IEEE_LIB):
        super().__init__(path, "", True)


class _Shifted:
    """A device array seen `offset` doubles past its start."""

    def __init__(self, arr, offset, shape):
        self.arr, self.offset, self.shape = arr, offset, shape
        self.ptr = arr.ptr + 8 * offset

    def download(self):
        return self.arr.download()[self.offset:].reshape(self.shape)


class GpuBackend:
    """Runs a spec on the device through the C ABI; buffers stay resident
    across steps, as in the reference's cuda target."""

    def __init__(self, ctx):
        self.ctx = ctx

    def run(self, spec: Spec, dims, coefs: dict, arrays: dict, nt: int, offset: int = 0) -> dict:
        """offset > 0 places every array that many doubles past the start of
        its allocation (8-byte but not 16-byte aligned for odd offsets)."""
        ctx = self.ctx
        src = dict(arrays)
        if spec.name == "whispering":
            src["e_next"] = arrays["e"]
        allocs, st = {}, {}
        for r, a in src.items():
            a = np.ascontiguousarray(a, dtype=np.float64)
            if offset:
                allocs[r] = ctx.array(np.concatenate([np.zeros(offset), a.ravel()]))
                st[r] = _Shifted(allocs[r], offset, a.shape)
            else:
                allocs[r] = st[r] = ctx.array(a)
        for _ in range(nt):
            self.step(spec, dims, coefs, st)
            rotate(spec, st)
        ctx.sync()
        out = {r: st[r].download() for r in arrays}
        for v in allocs.values():
            v.free()
        return out

    def step(self, spec: Spec, dims, coefs: dict, st: dict) -> None:
        cargs = [coefs[c] for c in spec.coefs]
        if spec.name == "whispering":
            arrs = [st["e"], st["e_next"]] + [st[r] for r in spec.arrays[1:]]
        else:
            arrs = [st[r] for r in spec.arrays]
        self.ctx.call(spec.name, *[int(d) for d in dims], *cargs,
                      *[a.ptr if isinstance(a, _Shifted) else a for a in arrs])

