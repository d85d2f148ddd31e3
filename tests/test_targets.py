"""The host programs of the suite (<test>/<test>.c) and the benchmark driver.

CPU-side: the `gcc` target (CPU backend of the C ABI, reference flags) prints
what the reference's own gcc-target programs print (oracle/_ref/bin) for the
same arguments, keeps the reference's argument checks and stdout contract, and
the driver turns that output into the reference's table.
"""
import os
import re
import subprocess

import pytest

import kgspec as K

TESTS3 = ["laplacian", "divergence", "gradient", "wave13pt", "uxx1", "tricubic", "tricubic2",
          "lapgsrb", "matmul", "sincos", "vecadd"]
TESTS2 = ["jacobi", "gaussblur", "whispering", "gameoflife", "matvec"]
ALL = TESTS3 + TESTS2


def args_for(name, small=True):
    if name in TESTS2:
        return [66, 34, 3] if small else [512, 65536, 10]
    return [34, 18, 12, 3] if small else [512, 256, 256, 10]


def build(target):
    subprocess.run(["make", "-s", "-C", str(K.ROOT), target], check=True)


@pytest.fixture(scope="module")
def gcc_built():
    build("gcc")


def run(path, args, env=None):
    e = dict(os.environ)
    e.pop("NO_TIMING", None)
    e.pop("PROFILING_FNAME", None)
    if env:
        e.update(env)
    return subprocess.run([str(path), *map(str, args)], capture_output=True, text=True,
                          env=e, timeout=300)


@pytest.mark.parametrize("name", ALL)
def test_gcc_target_prints_what_the_reference_program_prints(gcc_built, name):
    ref = K.ROOT / "oracle" / "_ref" / "bin" / name
    if not ref.exists():
        pytest.skip("oracle/_ref not built")
    ours = run(K.ROOT / name / "gcc" / name, args_for(name), {"NO_TIMING": "1"})
    theirs = run(ref, args_for(name), {"NO_TIMING": "1"})
    assert ours.returncode == 0 and theirs.returncode == 0
    assert ours.stdout == theirs.stdout


@pytest.mark.parametrize("name", ALL)
def test_gcc_omp_target_prints_what_the_gcc_target_prints(gcc_built, name):
    """The OpenMP build of the CPU target (a second, multi-core baseline)
    computes every point as the single-threaded build does."""
    build(f"{name}.gcc-omp")
    omp = run(K.ROOT / name / "gcc-omp" / name, args_for(name), {"NO_TIMING": "1",
                                                                 "OMP_NUM_THREADS": "4"})
    one = run(K.ROOT / name / "gcc" / name, args_for(name), {"NO_TIMING": "1"})
    assert omp.returncode == 0 and one.returncode == 0
    assert omp.stdout == one.stdout


@pytest.mark.parametrize("name", ["laplacian", "gameoflife", "matmul", "vecadd"])
def test_stdout_contract(gcc_built, name):
    out = run(K.ROOT / name / "gcc" / name, args_for(name)).stdout
    assert re.search(r"^initial mean = -?\d+\.\d{6}$", out, re.M)
    assert re.search(r"^compute time = \d+\.\d{6} sec$", out, re.M)
    assert re.search(r"^final mean = -?\d+\.\d{6}$", out, re.M)
    # the gcc target has no device phases
    assert "init time" not in out and "data load time" not in out


def test_argument_errors(gcc_built):
    p = K.ROOT / "laplacian" / "gcc" / "laplacian"
    r = run(p, [4, 4, 4])
    assert r.returncode == 1 and r.stdout.startswith("Usage: ")
    r = run(p, [4, -3, 4, 1])
    assert r.returncode == 1 and "Value for ny is invalid: -3" in r.stdout
    r = run(K.ROOT / "gameoflife" / "gcc" / "gameoflife", [8, 8, 1, 1])
    assert r.returncode == 1 and "<nx> <ny> <nt>" in r.stdout


def test_benchmark_driver_table(gcc_built):
    out = subprocess.run([str(K.ROOT / "benchmark"), "34", "18", "12", "2", "1", "gcc"],
                         capture_output=True, text=True, cwd=K.ROOT, timeout=600).stdout
    rows = {m.group(1): m.group(0) for m in re.finditer(r"^\|\s+(\w+) \|\s+gcc \|.*$", out, re.M)}
    assert sorted(rows) == sorted(ALL)
    for name, row in rows.items():
        cells = [c.strip() for c in row.strip("|").split("|")]
        assert "FAIL" not in cells, row
        float(cells[6])          # t_comp
        float(cells[7])          # t_krn = t_comp / niters
    assert "Data for table >>" in out and "<< Data for table" in out


def test_benchmark_driver_gcc_omp_rows(gcc_built):
    build("gcc-omp")
    out = subprocess.run([str(K.ROOT / "benchmark"), "34", "18", "12", "2", "1", "gcc-omp"],
                         capture_output=True, text=True, cwd=K.ROOT, timeout=600).stdout
    rows = {m.group(1): m.group(0) for m in re.finditer(r"^\|\s+(\w+) \|\s+gcc-omp \|.*$", out, re.M)}
    assert sorted(rows) == sorted(ALL)
    for row in rows.values():
        cells = [c.strip() for c in row.strip("|").split("|")]
        assert "FAIL" not in cells, row
        float(cells[7])          # t_krn = t_comp / niters
    assert "gcc-omp target host:" in out


@pytest.mark.parametrize("name", ["laplacian", "whispering", "tricubic", "matvec"])
def test_gcc_check_rule_runs_clean_under_sanitizers(gcc_built, name):
    """`make check` builds the program and its CPU backend with ASan and UBSan
    and runs it; any report fails the run."""
    dims = ["NX=34", "NY=18", "NS=12", "NT=2"]
    r = subprocess.run(["make", "-s", "-C", str(K.ROOT / name / "gcc"), "check", *dims],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert "final mean" in r.stdout


def test_benchmark_usage():
    out = subprocess.run([str(K.ROOT / "benchmark")], capture_output=True, text=True).stdout
    assert out.startswith("Usage: benchmark <nx> <ny> <ns> <niters> <nruns> <target>")
