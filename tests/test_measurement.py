"""The measurement tooling on the CPU: bench.py's one-GPU contract and the
rocprofv3 summariser (kgprof.py) that turns FETCH_SIZE / WRITE_SIZE counters
into calibrated HBM bytes per launch."""
import csv
import importlib.util
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _kgprof():
    spec = importlib.util.spec_from_file_location("kgprof", ROOT / "kernelgen-perf-tests_amd" / "kgprof.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "single device" in r.stderr and not r.stdout.strip()


def _write_csv(path, header, rows):
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(header)
        w.writerows(rows)


def test_kgprof_scales_counters_by_the_copy_kernel(tmp_path):
    """Two stream_copy launches read and write 1 GiB each; their FETCH_SIZE
    shows half the reads (as the counter reads on MI355X) and WRITE_SIZE all
    the writes, so the scales come out 2.0 and 1.0, and a laplacian launch's
    traffic follows from its own counters times those scales."""
    copy = "(anonymous namespace)::stream_copy_k(double*, double const*, long long)"
    lap = "void kgp::plane_kernel<kgops::LaplacianOp<kgp::Cfg<2> > >(int)"
    trace = tmp_path / "t_kernel_trace.csv"
    _write_csv(trace, ["Kernel_Name", "Start_Timestamp", "End_Timestamp"],
               [[copy, 0, 200_000], [copy, 300_000, 500_000], [lap, 600_000, 700_000],
                [lap, 800_000, 902_000]])
    gib = 1 << 30
    head = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]
    # FETCH_SIZE: two dimensions per dispatch, summed by the reader
    _write_csv(tmp_path / "f.csv", head,
               [[1, copy, "FETCH_SIZE", gib / 1024 / 4], [1, copy, "FETCH_SIZE", gib / 1024 / 4],
                [2, copy, "FETCH_SIZE", gib / 1024 / 2], [3, lap, "FETCH_SIZE", 150_000.0],
                [4, lap, "FETCH_SIZE", 250_000.0]])
    _write_csv(tmp_path / "w.csv", head,
               [[1, copy, "WRITE_SIZE", gib / 1024], [2, copy, "WRITE_SIZE", gib / 1024],
                [3, lap, "WRITE_SIZE", 260_000.0], [4, lap, "WRITE_SIZE", 264_000.0]])
    kp = _kgprof()
    out_json = tmp_path / "traffic.json"
    assert kp.main(["--trace", str(trace), "--fetch", str(tmp_path / "f.csv"), "--write",
                    str(tmp_path / "w.csv"), "--copy-bytes", str(gib), "--traffic-json",
                    str(out_json), "--out", str(tmp_path / "s.txt")]) == 0
    text = (tmp_path / "s.txt").read_text()
    assert "FETCH_SIZE x 2.0000, WRITE_SIZE x 1.0000" in text
    t = json.loads(out_json.read_text())
    rd = 200_000.0 * 1024 * 2.0        # mean FETCH_SIZE x 1 KiB x scale
    wr = 262_000.0 * 1024 * 1.0
    assert t["laplacian"]["read_bytes"] == round(rd)
    assert t["laplacian"]["write_bytes"] == round(wr)
    assert t["laplacian"]["traffic_bytes"] == round(rd + wr)
    assert t["laplacian"]["algorithmic_bytes"] == 16 * 512 * 256 * 256
    assert t["laplacian"]["mean_launch_us"] == 101.0
    assert t["stream_copy"]["traffic_bytes"] == 2 * gib
