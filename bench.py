#!/usr/bin/env python3
"""Headline benchmark: sort 2^30 uniform u32 keys in HBM on one MI355X.

One step = one call of mgs_sort (libmgsort, HIP kernels) over 2^30 keys
already resident in HBM, written out of place so every step sorts the same
unsorted input.  Keys come from the reference's DataGenerator (uniform,
seed 2147483647; libmgsort_host).  Prints one JSON line:

  value          whole-job GKeys/s (keys / wall time of the K timed steps)
  roofline       the dominant kernel (the LSD pass kernel): algorithmic bytes
                 per launch (8 B per key: one read and one write of the keys)
                 over its average launch time, from HIP events recorded around
                 each launch on the library's stream during the timed steps;
                 `traffic` is the HBM bytes per launch measured by rocprofv3
                 (FETCH_SIZE x2 + WRITE_SIZE x1, calibrated on a plain copy),
                 read from profiles/ when present
  cpu_baseline   the oracle's sort (oracle/oracle_sort.c, one core) on a
                 bounded sample of the same workload
  cpu_parallel   PARADIS and GNU parallel sort on all host cores (sample)

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N]
"""
import argparse
import ctypes
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "multi-gpu-sorting_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

SEED = 2147483647
KEY_BYTES = 4
ALG_BYTES_PER_KEY = KEY_BYTES + KEY_BYTES * 2 * KEY_BYTES  # 36: histogram read + 4 x (read + write)
PASS_BYTES_PER_KEY = 2 * KEY_BYTES  # one LSD pass: read + write
PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_lib():
    path = os.path.join(ROOT, "multi-gpu-sorting_amd", "lib", "libmgsort_host.so")
    L = ctypes.CDLL(path)
    c, vp, sz = ctypes.c_char_p, ctypes.c_void_p, ctypes.c_size_t
    L.mgh_generate.argtypes = [c, c, ctypes.c_uint32, vp, sz]
    L.mgh_generate.restype = ctypes.c_int
    L.mgh_paradis_sort.argtypes = [c, vp, sz, ctypes.c_int]
    L.mgh_gnu_parallel_sort.argtypes = [c, vp, sz, ctypes.c_int]
    L.mgh_event_create.restype = vp
    L.mgh_event_destroy.argtypes = [vp]
    L.mgh_event_elapsed_ms.argtypes = [vp, vp]
    L.mgh_event_elapsed_ms.restype = ctypes.c_float
    return L


def generate(H, n, np):
    keys = np.empty(n, dtype=np.uint32)
    rc = H.mgh_generate(b"u32", b"uniform", SEED, keys.ctypes.data, n)
    if rc != 0:
        raise RuntimeError(f"mgh_generate failed ({rc})")
    return keys


def traffic_from_profiles():
    """HBM bytes per launch of the pass kernel, if a calibrated rocprofv3
    summary is committed (profiles/*pass_kernel*2e30*.txt: 'traffic_bytes: X')."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    for f in sorted(os.listdir(d)):
        if "pass_kernel" in f and f.endswith(".txt"):
            m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(os.path.join(d, f)).read())
            if m:
                return float(m.group(1))
    return None


def cpu_baseline(H, np):
    """Oracle sort (one core) and the parallel CPU sorts on bounded samples."""
    import pyoracle
    out = {}
    n = 1 << 25
    keys = generate(H, n, np)
    reps, t_total = 0, 0.0
    while reps < 8 and t_total < 12.0:
        t0 = time.perf_counter()
        k, _ = pyoracle.sort(pyoracle.K_U32, keys)
        t_total += time.perf_counter() - t0
        reps += 1
    assert np.all(k[1:] >= k[:-1])
    out["cpu_baseline"] = {"value": round(n * reps / t_total / 1e9, 4), "unit": "GKeys/s", "cores": 1,
                           "kind": "port",
                           "sample": f"{reps} sorts of 2^25 uniform u32 keys (reference DataGenerator, seed {SEED}) "
                                     f"by the oracle's byte-wise LSD counting sort, {t_total:.1f} s"}
    threads = len(os.sched_getaffinity(0))
    n2 = 1 << 28
    keys = generate(H, n2, np)
    par = {}
    for name, fn in (("paradis", H.mgh_paradis_sort), ("gnu_parallel_sort", H.mgh_gnu_parallel_sort)):
        k = keys.copy()
        t0 = time.perf_counter()
        fn(b"u32", k.ctypes.data, n2, threads)
        dt = time.perf_counter() - t0
        assert np.all(k[1:] >= k[:-1])
        par[name] = round(n2 / dt / 1e9, 4)
    out["cpu_parallel"] = {**par, "unit": "GKeys/s", "threads": threads,
                           "sample": "2^28 uniform u32 keys, one sort each"}
    return out


def batched_section():
    """4096 uneven arrays (log-uniform lengths in [2^8, 2^22]) of u64 keys with
    u32 payloads: one graph-replayed mgs_segmented_sort vs one mgs_sort per
    array (eager loop and the same loop captured in a graph)."""
    from mgsort import batched
    out = {"workload": "4096 arrays, lengths round(2^(8+14U)) (mt19937_64 seed 2024), u64 keys + u32 payload",
           "bytes_per_key": 200}
    for dist in ("uniform", "zipf-1.2"):
        r = batched.run(dist=dist, log=log)
        out[dist] = {"total_keys": r.total_keys, "batched_ms": round(r.batched_ms, 3),
                     "loop_ms": round(r.loop_ms, 3),
                     "loop_graph_ms": None if r.loop_graph_ms is None else round(r.loop_graph_ms, 3),
                     "speedup_vs_loop": round(r.loop_ms / r.batched_ms, 3),
                     "speedup_vs_loop_graph": None if r.loop_graph_ms is None else round(r.loop_graph_ms / r.batched_ms, 3),
                     "batched_gkeys_s": round(r.total_keys / r.batched_ms / 1e6, 3),
                     "batched_roofline_frac": round(r.total_keys * 200 / (r.batched_ms / 1e3) / (PEAK_GBPS * 1e9), 4),
                     "tuning": r.tuning, "verified": r.verified}
        if not r.verified:
            raise RuntimeError(f"batched {dist}: output failed the check")
    return out


def serving_section(requests=192):
    """Poisson stream of mixed requests (log-uniform 2^10..2^26 keys; 50% u32,
    30% u64, 20% u32+u32 pairs; uniform / nearly sorted) through the serving
    layer: continuous batching vs the plain per-request path."""
    import subprocess
    exe = os.path.join(ROOT, "multi-gpu-sorting_amd", "bin", "serve_benchmark")
    p = subprocess.Popen([exe, "--requests", str(requests), "--budget_gib", "64", "--loads", "inf,2,1,0.5"],
                         stdout=subprocess.PIPE, text=True)
    rows = []
    for line in p.stdout:  # progress as it comes
        log("serve:", line.rstrip())
        if line.startswith("{"):
            rows.append(json.loads(line))
    rc = p.wait(timeout=900)
    if rc != 0:
        raise RuntimeError(f"serve_benchmark failed ({rc})")
    runs = [x for x in rows if "mode" in x]
    out = {"workload": f"{requests} requests, sizes round(2^(10+16U)), mt19937_64 seed 7; 64 GiB HBM pool",
           "runs": runs}
    for data in ("device", "host"):
        sat = {x["mode"]: x for x in runs if x["data"] == data and x["load"] == "inf"}
        if "batched" in sat and "plain" in sat:
            out[f"{data}_saturated_speedup_vs_plain"] = round(sat["batched"]["gkeys_s"] / sat["plain"]["gkeys_s"], 3)
            out[f"{data}_saturated_speedup_vs_plain_overlap"] = round(
                sat["batched"]["gkeys_s"] / sat["plain_overlap"]["gkeys_s"], 3)
    out["same_as_alone"] = all(x["same_as_alone"] for x in runs)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 30)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched segmented-sort section")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop section")
    ap.add_argument("--segmented-only", action="store_true",
                    help="only sort the batched workload (4096 arrays, u64 keys + u32 payload, uniform) with "
                         "mgs_segmented_sort --steps times (a profiling target)")
    ap.add_argument("--copy-only", action="store_true",
                    help="only copy n u32 keys device to device (calibrates the profiler's byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: the sort path of this build runs on one GPU "
              "(multi-GPU exchange is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import mgsort

    if args.segmented_only:
        from mgsort import batched
        lengths = batched.segment_lengths(4096, seed=2024)
        offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
        total = int(offsets[-1])
        keys = batched.make_keys(total, "uniform", 2024, torch.device("cuda"))
        vals = batched.local_index(lengths, torch.device("cuda"))
        ok, ov = torch.empty_like(keys), torch.empty_like(vals)
        d_off = torch.from_numpy(offsets).cuda()
        tn = mgsort.Tuning(0, 0, 32768, 0)  # the autotuner's usual choice on this workload
        temp = mgsort.temp_buffer(mgsort.lib.mgs_segmented_sort_temp_bytes(mgsort.KEY_U64, mgsort.VAL_U32, total, 4096),
                                  "cuda")
        ms = []
        for k in range(args.warmup + args.steps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            mgsort.segmented_sort(mgsort.KEY_U64, keys, d_off, vals, keys_out=ok, vals_out=ov, tuning=tn, temp=temp)
            e1.record()
            torch.cuda.synchronize()
            if k >= args.warmup:
                ms.append(e0.elapsed_time(e1))
        if not batched.check_segments(ok, ov, keys, vals, offsets):
            raise RuntimeError("segmented output failed the check")
        print(json.dumps({"segmented_only": True, "total_keys": total, "ms_min": round(min(ms), 3),
                          "gkeys_s": round(total / min(ms) / 1e6, 3)}), flush=True)
        return

    if args.copy_only:
        a = torch.empty(args.n, dtype=torch.int32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1)
        for _ in range(args.warmup + args.steps):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_only": True, "n": args.n, "bytes_per_copy": 8 * args.n}), flush=True)
        return

    H = host_lib()
    n = args.n
    log(f"generating {n} keys")
    keys = generate(H, n, np)
    src = torch.from_numpy(keys.view(np.int32)).cuda()
    out = torch.empty_like(src)
    temp = mgsort.temp_buffer(mgsort.lib.mgs_sort_temp_bytes(mgsort.KEY_U32, mgsort.VAL_NONE, n), "cuda")
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    L = mgsort.lib
    marks = ctypes.c_int(0)
    rc = L.mgs_sort_timed(0, 0, src.data_ptr(), out.data_ptr(), None, None, n, temp.data_ptr(), temp.numel(), sp,
                          None, 0, ctypes.byref(marks))
    if rc != 0:
        raise RuntimeError(f"mgs_sort_timed: {L.mgs_status_str(rc)}")
    nm = marks.value
    stage_names = L.mgs_sort_timed_stages().decode().split(",")
    events = [[H.mgh_event_create() for _ in range(nm)] for _ in range(args.steps)]
    ev_arrays = [(ctypes.c_void_p * nm)(*e) for e in events]

    for _ in range(args.warmup):
        mgsort.sort("u32", src, keys_out=out, temp=temp)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        rc = L.mgs_sort_timed(0, 0, src.data_ptr(), out.data_ptr(), None, None, n, temp.data_ptr(), temp.numel(), sp,
                              ev_arrays[k], nm, ctypes.byref(marks))
        if rc != 0:
            raise RuntimeError(f"mgs_sort_timed: {L.mgs_status_str(rc)}")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps

    # per-stage averages from the events
    stage_ms = {}
    for i in range(nm - 1):
        vals = [H.mgh_event_elapsed_ms(events[k][i], events[k][i + 1]) for k in range(args.steps)]
        name = stage_names[i] if i < len(stage_names) else f"stage{i}"
        stage_ms[name] = sum(vals) / len(vals)
    for e in events:
        for x in e:
            H.mgh_event_destroy(x)

    # correctness of the last step's output: sorted, and the same multiset
    o64 = out.to(torch.int64) & 0xFFFFFFFF
    s64 = src.to(torch.int64) & 0xFFFFFFFF
    ok = bool((o64[1:] >= o64[:-1]).all().item()) and int(o64.sum().item()) == int(s64.sum().item()) and \
        int((o64 * o64 % 1000003).sum().item()) == int((s64 * s64 % 1000003).sum().item())
    del o64, s64
    if not ok:
        raise RuntimeError("sorted output failed the check")

    # ceiling reference: a plain device-to-device copy of the same 4 GiB
    # (HIP's copy kernel), which reads and writes the bytes a pass moves
    copy_ms = []
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        out.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        copy_ms.append(e0.elapsed_time(e1))
    copy_gbps = n * PASS_BYTES_PER_KEY / (min(copy_ms) / 1e3) / 1e9

    # the dominant kernel: the slowest of the passes that read and write
    # every key once (the three-pass path's passes, or the LSD passes)
    passes = [k for k in stage_ms if k.startswith("msd_pass") or k == "msd_local"]
    if not passes or min(stage_ms[k] for k in passes) < 0.05:
        passes = [f"pass{p}" for p in range(KEY_BYTES)]
    dom = max(passes, key=lambda k: stage_ms[k])
    avg_pass_ms = stage_ms[dom]
    achieved = n * PASS_BYTES_PER_KEY / (avg_pass_ms / 1e3) / 1e9
    value = n / (ms_per_step / 1e3) / 1e9
    result = {
        "metric": "GKeys/s & HBM GB/s sorting 2^30 u32 on 1 MI355X; batched p99 latency vs plain",
        "value": round(value, 3),
        "unit": "GKeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic: reference DataGenerator uniform u32, seed 2147483647, resident in HBM",
        "config": {"workload": f"in-HBM sort of {n} uniform u32 keys (out of place, stable LSD radix)",
                   "n": n, "key": "u32", "values": "none", "distribution": "uniform", "seed": SEED},
        "hbm_gbps_alg": round(n * ALG_BYTES_PER_KEY / (ms_per_step / 1e3) / 1e9, 1),
        "roofline_frac_whole_sort": round(n * ALG_BYTES_PER_KEY / (ms_per_step / 1e3) / 1e9 / PEAK_GBPS, 4),
        "roofline": {"bound": "hbm", "kernel": dom,
                     "achieved": round(achieved, 1), "peak": PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_GBPS, 4), "traffic": traffic_from_profiles(),
                     "alg_bytes_per_launch": n * PASS_BYTES_PER_KEY, "avg_launch_ms": round(avg_pass_ms, 4)},
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "copy_ceiling": {"what": "device-to-device copy of the 4 GiB key array (hipMemcpy's kernel), same bytes as one pass",
                         "gbps": round(copy_gbps, 1), "frac_of_peak": round(copy_gbps / PEAK_GBPS, 4),
                         "pass_vs_copy": round(achieved / copy_gbps, 4)},
    }
    del src, out, temp
    torch.cuda.empty_cache()
    if not args.no_batched:
        result["batched"] = batched_section()
        torch.cuda.empty_cache()
    if not args.no_serving:
        result["serving"] = serving_section()
    if not args.no_cpu:
        result.update(cpu_baseline(H, np))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
