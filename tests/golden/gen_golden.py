"""Write the golden fixtures under tests/golden/ from the reference's own kernels.

Runs only in a checkout where the reference tree is present: it calls the
kernel functions of oracle/_ref/libref_ieee.so, which oracle/build_ref.sh
compiles from the reference's sources (IEEE flags: -O2 -fno-fast-math
-ffp-contract=off).  Each fixture holds seeded uniform[-1,1) inputs, the
coefficients, and every written array after nt = 1 and nt = 3 steps with the
reference main()'s buffer rotation.  The files are data only.

    python tests/golden/gen_golden.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
import kgspec as K  # noqa: E402

CASES = {3: [(19, 7, 6), (66, 5, 5)], 2: [(37, 13), (70, 7)]}
MATMUL_CASES = [(20, 12, 9), (33, 17, 18)]
STEPS = (1, 3)


def written_roles(spec: K.Spec):
    if spec.name == "divergence":
        return ("u",)
    if spec.name == "gradient":
        return ("ux", "uy", "uz")
    if spec.name == "matmul":
        return ("C",)
    if spec.name == "whispering":
        return ("e", "h0", "h1", "u_em0", "u_em1")
    return spec.pair


VEC_CASES = [(19, 7, 6), (66, 5, 5)]
MATVEC_CASES = [(37, 13), (70, 7)]


def vecmath_fixture(ref) -> dict:
    """vecadd (3-level rotation, nt = 1 and 3), matvec and sincos (same
    inputs every step: one call) from the reference's FP64 build."""
    import ctypes as C
    lib = ref.lib
    vp = lambda a: C.c_void_p(a.ctypes.data)
    payload, meta = {}, {"vecadd": [], "matvec": [], "sincos": []}
    for ci, dims in enumerate(VEC_CASES):
        rng = np.random.default_rng(2000 + ci)
        n = int(np.prod(dims))
        w = [rng.uniform(-1, 1, n) for _ in range(3)]
        for r, a in zip(("w0", "w1", "w2"), w):
            payload[f"vecadd_c{ci}_in_{r}"] = a
        for nt in STEPS:
            st = [a.copy() for a in w]
            for _ in range(nt):
                lib.vecadd(C.c_int(dims[0]), C.c_int(dims[1]), C.c_int(dims[2]),
                           vp(st[0]), vp(st[1]), vp(st[2]))
                st = [st[1], st[2], st[0]]
            for r, a in zip(("w0", "w1", "w2"), st):
                payload[f"vecadd_c{ci}_nt{nt}_{r}"] = a
        meta["vecadd"].append({"dims": list(dims)})
        x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
        xy = rng.uniform(-1, 1, n)
        payload[f"sincos_c{ci}_in_x"], payload[f"sincos_c{ci}_in_y"] = x, y
        d = [C.c_int(v) for v in dims]
        lib.sincos_(C.byref(d[0]), C.byref(d[1]), C.byref(d[2]), vp(x), vp(y), vp(xy))
        payload[f"sincos_c{ci}_xy"] = xy
        meta["sincos"].append({"dims": list(dims)})
    for ci, (nx, ny) in enumerate(MATVEC_CASES):
        rng = np.random.default_rng(3000 + ci)
        A, x, y = rng.uniform(-1, 1, nx * ny), rng.uniform(-1, 1, nx), rng.uniform(-1, 1, ny)
        payload[f"matvec_c{ci}_in_A"], payload[f"matvec_c{ci}_in_x"] = A, x
        lib.matvec(C.c_int(nx), C.c_int(ny), vp(A), vp(x), vp(y))
        payload[f"matvec_c{ci}_y"] = y
        meta["matvec"].append({"dims": [nx, ny]})
    payload["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    return payload


def main() -> int:
    if not K.REF_IEEE_LIB.exists():
        print(f"{K.REF_IEEE_LIB} missing: run `make -C oracle ref` first", file=sys.stderr)
        return 1
    ref = K.RefBackend()
    index = {}
    for name, spec in K.SPECS.items():
        cases = MATMUL_CASES if name == "matmul" else CASES[spec.dim]
        payload = {}
        meta = []
        for ci, dims in enumerate(cases):
            seed = 1000 + 17 * ci + sum(map(ord, name))
            coefs, arrays = K.random_case(spec, dims, seed)
            meta.append({"dims": list(dims), "seed": seed, "coefs": coefs})
            for role, a in arrays.items():
                payload[f"c{ci}_in_{role}"] = a
            for nt in STEPS:
                out = ref.run(spec, dims, coefs, arrays, nt)
                for role in written_roles(spec):
                    payload[f"c{ci}_nt{nt}_{role}"] = out[role]
        payload["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
        np.savez_compressed(HERE / f"{name}.npz", **payload)
        index[name] = {"cases": meta, "steps": list(STEPS), "written": list(written_roles(spec))}
    vm = vecmath_fixture(ref)
    np.savez_compressed(HERE / "vecmath.npz", **vm)
    index["vecmath"] = json.loads(bytes(vm["meta"]).decode())
    (HERE / "index.json").write_text(json.dumps(
        {"source": "oracle/_ref/libref_ieee.so (reference sources, -O2 -fno-fast-math "
                   "-ffp-contract=off; jacobi/matmul/sincos via amdflang; real = double)",
         "kernels": index}, indent=1))
    print(f"wrote {len(index)} fixtures to {HERE}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
