"""The `hip` target programs on the device, through libkgstencil.so.

Each prints the same lines as the `gcc` target for the same arguments; the
final means may differ only where the gcc target's -ffast-math reassociation
shows through %f (tolerance: 1e-12 relative plus the %f quantum).  With
PROFILING_FNAME set they print one kernel-time and one regcount line per
launch, and the driver reports t_krn / nreg_krn for them.
"""
import os
import re
import subprocess

import pytest

import kgspec as K
from test_targets import ALL, TESTS2, args_for, build, run

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def built():
    for t in ("gcc", "hip"):
        p = K.ROOT / "laplacian" / t / "laplacian"
        if not p.exists():
            build(t)


def means(out):
    return {k: float(v) for k, v in re.findall(r"^(initial mean|final mean) = (\S+)", out, re.M)}


@pytest.mark.parametrize("name", ALL)
def test_hip_program_matches_gcc_program(built, name):
    print(f"hip target {name} {args_for(name)}", flush=True)
    h = run(K.ROOT / name / "hip" / name, args_for(name), {"NO_TIMING": "1"})
    g = run(K.ROOT / name / "gcc" / name, args_for(name), {"NO_TIMING": "1"})
    assert h.returncode == 0, h.stdout + h.stderr
    hl = [l for l in h.stdout.splitlines() if "mean" not in l]
    gl = [l for l in g.stdout.splitlines() if "mean" not in l]
    assert hl == gl
    mh, mg = means(h.stdout), means(g.stdout)
    assert mh.keys() == mg.keys()
    for k in mh:
        assert abs(mh[k] - mg[k]) <= 1e-12 * abs(mg[k]) + 1e-6, (k, mh[k], mg[k])


@pytest.mark.parametrize("name", ["laplacian", "jacobi", "whispering", "matmul"])
def test_hip_program_profiling_lines(built, name):
    kernel = (K.ROOT / name / "hip" / "kernel").read_text().strip()
    nt = args_for(name)[-1]
    r = run(K.ROOT / name / "hip" / name, args_for(name), {"PROFILING_FNAME": kernel})
    assert r.returncode == 0, r.stderr
    times = re.findall(rf"^{kernel} kernel time = (\S+)$", r.stdout, re.M)
    regs = re.findall(rf"^{kernel} regcount = (\d+)$", r.stdout, re.M)
    assert len(times) == nt and len(regs) == nt
    assert all(float(t) > 0 for t in times) and int(regs[0]) > 0
    for line in ("init time", "device buffer alloc time", "data load time", "data save time",
                 "device buffer free time"):
        assert line in r.stdout


def test_benchmark_driver_hip_gcc(built):
    out = subprocess.run([str(K.ROOT / "benchmark"), "64", "32", "32", "2", "1", "hip", "gcc"],
                         capture_output=True, text=True, cwd=K.ROOT, timeout=600).stdout
    rows = {}
    for m in re.finditer(r"^\|\s+(\w+) \|\s+(hip|gcc) \|(.*)$", out, re.M):
        rows[(m.group(1), m.group(2))] = [c.strip() for c in m.group(3).strip("|").split("|")]
    for name in ALL:
        h, g = rows[(name, "hip")], rows[(name, "gcc")]
        assert "FAIL" not in h, (name, h)
        float(h[5])              # t_krn from the per-launch kernel time lines
        int(float(h[6]))         # nreg_krn
        assert h[0] == g[0]      # i_mean
    assert "Times for chart >>" in out
