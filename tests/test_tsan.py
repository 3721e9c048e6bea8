"""ThreadSanitizer run of the threaded host code that does not touch the
device: PARADIS and the generator under OpenMP (with the OpenMP runtime's
sanitizer support, libarcher, so its barriers are seen), and the HET sort's
merge thread (ChunkMerger) fed by producer threads, including its exception
path.  Built by `make tsan` with ROCm's clang; every result is also checked
against std::sort."""
import os
import subprocess

import pytest

from conftest import ROOT

PKG = os.path.join(ROOT, "multi-gpu-sorting_amd")
ARCHER = "/opt/rocm/llvm/lib/libarcher.so"


def test_threaded_host_code_is_race_free():
    if not os.path.exists("/opt/rocm/llvm/bin/clang++"):
        pytest.skip("ROCm clang not installed")
    b = subprocess.run(["make", "-s", "-C", PKG, "tsan"], capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-2000:]
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = f"halt_on_error=0 exitcode=66 suppressions={os.path.join(PKG, 'host', 'main', 'tsan.supp')}"
    if os.path.exists(ARCHER):
        env["OMP_TOOL_LIBRARIES"] = ARCHER
    env["OMP_NUM_THREADS"] = "8"
    r = subprocess.run([os.path.join(PKG, "bin", "tsan_check")], capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert "WARNING: ThreadSanitizer" not in out, out[-4000:]
    assert r.returncode == 0 and "all passed" in r.stdout, out[-4000:]


@pytest.mark.parametrize("extra", [[], ["--solo_min_log2", "12", "--sort_streams", "3"]], ids=["batched", "solo+3streams"])
def test_serving_layer_threads_are_race_free(extra):
    """The serving benchmark (SortServer: client, scheduler, completion
    threads; pinned staging ring; pool) built against a host emulation of
    the HIP stream API whose streams are worker threads (host/main/tsan_hip),
    under ThreadSanitizer; every output equals the host sort."""
    if not os.path.exists("/opt/rocm/llvm/bin/clang++"):
        pytest.skip("ROCm clang not installed")
    b = subprocess.run(["make", "-s", "-C", PKG, "tsan_serve_host"], capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr[-2000:]
    env = dict(os.environ)
    env["TSAN_OPTIONS"] = f"halt_on_error=0 exitcode=66 suppressions={os.path.join(PKG, 'host', 'main', 'tsan.supp')}"
    if os.path.exists(ARCHER):
        env["OMP_TOOL_LIBRARIES"] = ARCHER
    env["OMP_NUM_THREADS"] = "8"
    cmd = [os.path.join(PKG, "bin", "serve_tsan"), "--requests", "96", "--max_log2", "15", "--budget_gib", "1",
           "--data", "device,host", "--loads", "inf,2,0.5"] + extra
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert "WARNING: ThreadSanitizer" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    rows = [line for line in r.stdout.splitlines() if '"mode"' in line]
    assert len(rows) == 18 and all('"matches_host_sort": true' in x for x in rows)
