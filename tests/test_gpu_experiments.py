"""The device benchmark programs (the reference's local GPU sort and merge
programs, host-link transfers, the in-place exchange, allocation) and the
experiment runner over them, on the GPU: each program's CSV line has the
reference's columns and the program's own output check passes."""
import csv
import pathlib
import subprocess
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG = ROOT / "multi-gpu-sorting_amd"
BIN = PKG / "bin"
sys.path.insert(0, str(PKG / "experiments"))

pytestmark = pytest.mark.gpu


def _run(args, timeout=90):
    path = BIN / args[0]
    if not path.exists():
        pytest.fail("%s is not built" % args[0])
    print("run:", " ".join(str(a) for a in args), flush=True)
    r = subprocess.run([str(path)] + [str(a) for a in args[1:]], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1, r.stdout
    return next(csv.reader(lines))


@pytest.mark.parametrize("alg", ["mgs_sort", "mgs_sort_lsd", "hipcub::DeviceRadixSort::SortKeys", "thrust::sort",
                                 "hipcub::DeviceMergeSort::SortKeys", "cub::DeviceRadixSort::SortKeys",
                                 "mgpu::mergesort"])
@pytest.mark.parametrize("n", [0, 1, 1000003])
def test_local_gpu_sort_program(alg, n):
    row = _run(["local_gpu_sort_algorithm_benchmark", n, alg, 0])
    assert row[:3] == [str(n), alg, "0"] and float(row[3]) >= 0


@pytest.mark.parametrize("dist", ["uniform", "zipf", "sorted"])
def test_local_gpu_sort_program_long_keys(dist):
    for alg in ("mgs_sort", "hipcub::DeviceRadixSort::SortKeys"):
        row = _run(["local_gpu_sort_algorithm_benchmark", 300001, alg, 0, "long", dist])
        assert row[1] == alg


@pytest.mark.parametrize("alg", ["mgs_merge", "mgpu::merge", "mgs_multiway_merge", "thrust::merge"])
@pytest.mark.parametrize("n", [1, 2, 2000001])
def test_local_gpu_merge_program(alg, n):
    row = _run(["local_gpu_merge_algorithm_benchmark", n, alg, 0])
    assert row[:3] == [str(n), alg, "0"]


def test_local_gpu_multiway_program():
    row = _run(["local_gpu_multiway_merge_benchmark", "100000,3,0,250000,17", 0])
    assert row[:2] == ["100000,3,0,250000,17", "0"] and float(row[2]) > 0
    row = _run(["local_gpu_multiway_merge_benchmark", ",".join(["4099"] * 16), 0])
    assert len(row) == 3


def test_data_transfer_program():
    row = _run(["data_transfer_benchmark", 64 << 20, "0", 3, "serial"])
    assert len(row) == 12 and row[1] == "0" and row[6] == "" and float(row[7]) == 0
    h2d, d2h, both = float(row[5]), float(row[9]), float(row[11])
    assert h2d > 0 and d2h > 0 and both > 0


@pytest.mark.parametrize("block", ["1048576", "3333333.3333", "10000000", "999"])
def test_inplace_memcpy_program(block):
    row = _run(["inplace_memcpy_benchmark", 10000000, "0", block])
    assert row[:3] == ["10000000", "0", str(min(int(float(block)), 10000000))] and float(row[3]) > 0


def test_memory_program():
    row = _run(["memory_benchmark", 1 << 30, 1, "parallel"])
    assert row[:4] and row[1] == "1" and float(row[4]) > 0


def test_runner_over_device_programs(tmp_path):
    import plot_experiments
    import run_experiments
    import settings

    out = tmp_path / "run"
    ids = ["local_gpu_sort_algorithm_to_sort_duration", "number_sorted_sublists_to_gpu_multiway_merge_duration",
           "gpus_to_h_to_d_total_duration_and_d_to_h_total_duration", "num_gpus_to_memory_allocation_total_duration"]
    print("runner:", ids, flush=True)
    rc = run_experiments.main(["-e", ",".join(ids), "--scale", "0.0001", "--repetitions", "1", "--output", str(out),
                               "--timeout", "60"])
    assert rc == 0
    exps = {x.identifier: x for x in settings.experiments()}
    for ident in ids:
        x = exps[ident]
        rows = list(csv.reader((out / ("%s_%s.csv" % (x.executable, ident))).open()))
        assert rows[0] == x.columns and len(rows) == 1 + len(x.arguments)
    assert plot_experiments.main([str(out)]) == 0


def test_runner_over_sort_benchmark(tmp_path):
    import run_experiments
    import settings

    out = tmp_path / "run"
    ident = "sort_algorithm_to_sort_duration"
    print("runner:", ident, flush=True)
    assert run_experiments.main(["-e", ident, "--scale", "0.00005", "--repetitions", "1", "--output", str(out),
                                 "--timeout", "60"]) == 0
    x = {e.identifier: e for e in settings.experiments()}[ident]
    rows = list(csv.DictReader((out / ("sort_benchmark_%s.csv" % ident)).open()))
    assert len(rows) == len(x.arguments)
    assert {r["algorithm"] for r in rows} == {"gnu_parallel", "paradis", "thrust", "tanasic", "gowanlock", "hip"}
    assert all(r["verified"] == "1" for r in rows)
