"""The experiment grid, its runner and its plots (the reference's
scripts/settings.py, run_experiments.py and plot_experiments.py), and the
host-only benchmark programs they drive, on the CPU."""
import csv
import os
import pathlib
import subprocess
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG = ROOT / "multi-gpu-sorting_amd"
EXP = PKG / "experiments"
BIN = PKG / "bin"
sys.path.insert(0, str(EXP))

import settings  # noqa: E402
import run_experiments  # noqa: E402

# Every experiment identifier of the reference's grid (scripts/settings.py:52-770).
REFERENCE_IDENTIFIERS = [
    "num_elements_to_gowanlock_sort_duration_for_gpus", "num_elements_to_tanasic_sort_duration_for_gpus",
    "sort_algorithm_to_sort_duration", "data_type_to_gowanlock_sort_duration", "data_type_to_tanasic_sort_duration",
    "distribution_type_to_sort_duration_for_algorithm", "num_elements_to_sort_duration_profilers_nsys",
    "gpus_to_tanasic_sort_duration_profilers_nsys", "gpus_to_gowanlock_sort_duration_profilers_nsys",
    "num_bytes_to_h_to_d_total_duration_and_d_to_h_total_duration", "num_bytes_to_p_to_p_total_duration",
    "gpus_to_h_to_d_total_duration_and_d_to_h_total_duration", "gpus_to_p_to_p_total_duration",
    "num_gpus_to_memory_allocation_total_duration", "local_gpu_sort_algorithm_to_sort_duration",
    "local_gpu_merge_algorithm_to_merge_duration", "local_cpu_sort_algorithm_to_sort_duration",
    "local_cpu_merge_algorithm_to_merge_duration", "number_sorted_sublists_to_multiway_merge_duration",
    "large_data_to_gowanlock_sort_duration_for_gpus", "large_data_to_gowanlock_vs_cpu_sort_duration_for_gpus",
    "gpus_to_sort_duration_profilers_nsys", "block_size_to_inplace_data_transfer_duration",
]


def _need_bin(name):
    path = BIN / name
    if not path.exists():
        pytest.skip("%s not built" % name)
    return str(path)


def test_grid_covers_the_reference():
    exps = settings.experiments()
    ids = [x.identifier for x in exps]
    assert len(ids) == len(set(ids))
    for ident in REFERENCE_IDENTIFIERS:
        assert ident in ids or ident in settings.OMITTED, ident
    for x in exps:
        assert x.arguments, x.identifier
        for args in x.arguments:
            assert len(args) == len(x.parameters), (x.identifier, args)
        # one GPU only
        if "--gpus" in x.parameters:
            assert {a[x.parameters.index("--gpus")] for a in x.arguments} == {"0"}


def test_grid_programs_are_built_by_the_makefile():
    makefile = (PKG / "Makefile").read_text()
    for x in settings.experiments():
        assert x.executable in makefile, x.executable
        assert (PKG / "host" / "main" / (x.executable + ".cpp")).exists(), x.executable


def test_sort_columns_match_the_program():
    # the reference's fourteen columns first, in its order
    # (scripts/settings.py:42-46), then this build's seven
    assert settings.SORT_COLUMNS[:14] == [
        "num_elements", "algorithm", "gpus", "data_type", "distribution_type", "distribution_seed", "num_threads",
        "chunk_size", "merge_group_size", "num_buffers", "memory_allocation_duration", "sort_duration",
        "memory_deallocation_duration", "total_duration"]
    src = (PKG / "host" / "main" / "sort_benchmark.cpp").read_text()
    for col in settings.SORT_COLUMNS[14:]:
        assert col in src


def test_scaling_and_commands():
    exps = {x.identifier: x for x in settings.experiments()}
    x = exps["large_data_to_gowanlock_sort_duration_for_gpus"]
    args = x.arguments[-1]
    s = run_experiments.scaled_arguments(x, args, 0.5)
    assert s[0] == args[0] // 2 and s[3] == args[3] // 2 and s[4:] == args[4:]
    x = exps["block_size_to_inplace_data_transfer_duration"]
    s = run_experiments.scaled_arguments(x, (10 ** 9, 0, 10 ** 7), 0.001)
    assert s == (10 ** 6, 0, 10 ** 4)
    x = exps["number_sorted_sublists_to_multiway_merge_duration"]
    s = run_experiments.scaled_arguments(x, ("1000,2000,3",), 0.5)
    assert s == ("500,1000,2",) and run_experiments.num_keys(x, s) == 1502
    cmd = run_experiments.command(exps["sort_algorithm_to_sort_duration"],
                                  (1000, "paradis", "0", "int", "uniform", 7), BIN, False, None)
    assert cmd[1:] == ["1000", "--algorithm", "paradis", "--gpus", "0", "--data_type", "int", "--distribution_type",
                       "uniform", "--distribution_seed", "7"]
    prof = run_experiments.command(exps["gpus_to_sort_duration_profilers_nsys"],
                                   exps["gpus_to_sort_duration_profilers_nsys"].arguments[0], BIN, False,
                                   pathlib.Path("/tmp/p"))
    assert prof[0] == "rocprofv3" and "--pmc" not in prof and prof[prof.index("--") + 1].endswith("sort_benchmark")
    assert run_experiments.is_host_run(exps["sort_algorithm_to_sort_duration"],
                                       (1000, "paradis", "0", "int", "uniform", 7))


def test_csv_line_check():
    assert run_experiments.csv_lines('1,"a,b",3\n', 3, True) == ['1,"a,b",3']
    assert run_experiments.csv_lines('1,2\n', 3, True) is None
    assert run_experiments.csv_lines('profiler says hi\n1,2,3\n', 3, False) == ["1,2,3"]
    assert run_experiments.csv_lines('', 3, False) is None


def test_dry_run_of_the_whole_grid(capsys):
    assert run_experiments.main(["--dry-run", "--scale", "0.001"]) == 0
    out = capsys.readouterr().out.splitlines()
    assert len(out) >= sum(len(x.arguments) for x in settings.experiments())
    assert run_experiments.main(["-e", "no_such_experiment"]) == 2


def test_runner_and_plots_on_host_programs(tmp_path):
    for name in ("local_cpu_sort_algorithm_benchmark", "local_cpu_merge_algorithm_benchmark",
                 "local_cpu_multiway_merge_benchmark"):
        _need_bin(name)
    out = tmp_path / "run"
    ids = ("local_cpu_sort_algorithm_to_sort_duration,local_cpu_merge_algorithm_to_merge_duration,"
           "number_sorted_sublists_to_multiway_merge_duration")
    rc = run_experiments.main(["-e", ids, "--scale", "0.00005", "--max-elements", "1e6", "--repetitions", "2",
                               "--output", str(out), "--timeout", "120"])
    assert rc == 0
    exps = {x.identifier: x for x in settings.experiments()}
    for ident in ids.split(","):
        x = exps[ident]
        path = out / ("%s_%s.csv" % (x.executable, ident))
        rows = list(csv.reader(path.open()))
        assert rows[0] == x.columns
        assert len(rows) > 2 and all(len(r) == len(x.columns) for r in rows)
    import plot_experiments
    assert plot_experiments.main([str(out)]) == 0
    summary = (out / "summary.md").read_text()
    for ident in ids.split(","):
        assert "## " + ident in summary
        assert (out / (ident + ".png")).stat().st_size > 1000


def test_runner_reports_failures(tmp_path, capsys):
    rc = run_experiments.main(["-e", "local_cpu_sort_algorithm_to_sort_duration", "--scale", "0.000001",
                               "--repetitions", "1", "--bin", str(tmp_path / "nothing"), "--output",
                               str(tmp_path / "run")])
    assert rc == 1
    assert "FAILED" in capsys.readouterr().out
    assert not list((tmp_path / "run").glob("*.csv"))


@pytest.mark.parametrize("alg", ["gnu_parallel::sort", "paradis::sort", "std::sort"])
def test_local_cpu_sort_program(alg):
    r = subprocess.run([_need_bin("local_cpu_sort_algorithm_benchmark"), "123457", alg, "3"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    row = next(csv.reader([r.stdout]))
    assert row[:3] == ["123457", alg, "3"] and float(row[3]) > 0


def test_local_cpu_programs_reject_bad_arguments():
    sort = _need_bin("local_cpu_sort_algorithm_benchmark")
    for args in (["10"], ["x", "paradis::sort"], ["10", "quick::sort"], ["10", "paradis::sort", "0"],
                 ["-5", "paradis::sort"]):
        r = subprocess.run([sort] + args, capture_output=True, text=True, timeout=30)
        assert r.returncode == 1 and "Usage" not in r.stdout and "./local_cpu_sort" in r.stdout
    r = subprocess.run([sort, "10", "boost::sort::block_indirect_sort"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and "Boost" in r.stdout
    mw = _need_bin("local_cpu_multiway_merge_benchmark")
    for bad in ("1,,2", "1,x", ""):
        r = subprocess.run([mw, bad], capture_output=True, text=True, timeout=30)
        assert r.returncode == 1


def test_local_cpu_merge_programs():
    r = subprocess.run([_need_bin("local_cpu_merge_algorithm_benchmark"), "54321", "gnu_parallel::multiway_merge",
                        "2"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.startswith('54321,"gnu_parallel::multiway_merge",2,')
    r = subprocess.run([_need_bin("local_cpu_multiway_merge_benchmark"), "1000,0,77,5000", "2"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.startswith('"1000,0,77,5000",2,')


def test_device_programs_refuse_more_than_one_gpu():
    r = subprocess.run([_need_bin("memory_benchmark"), "1000", "2", "parallel"], capture_output=True, text=True,
                       timeout=30)
    assert r.returncode == 1 and "one GPU" in r.stdout
    r = subprocess.run([_need_bin("inplace_memcpy_benchmark"), "1000", "0", "1.5x"], capture_output=True, text=True,
                       timeout=30)
    assert r.returncode == 1
