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
    (HERE / "index.json").write_text(json.dumps(
        {"source": "oracle/_ref/libref_ieee.so (reference sources, -O2 -fno-fast-math "
                   "-ffp-contract=off; jacobi/matmul via amdflang)",
         "kernels": index}, indent=1))
    print(f"wrote {len(index)} fixtures to {HERE}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
