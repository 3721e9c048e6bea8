#!/usr/bin/env python3
"""Headline benchmark: sort 2^30 uniform u32 keys in HBM on one MI355X.

One step = one call of mgs_sort (libmgsort, HIP kernels) over 2^30 keys
already resident in HBM, written out of place so every step sorts the same
unsorted input.  Keys come from the reference's DataGenerator (uniform,
seed 2147483647; libmgsort_host).  Prints one JSON line:

  value          whole-job GKeys/s (keys / wall time of the K timed steps)
  roofline       the dominant kernel (the slowest launch that reads and
                 writes every key once): algorithmic bytes per launch (8 B per
                 key) over its average launch time, from HIP events recorded
                 around each launch on the library's stream during the timed
                 steps; `traffic` is the HBM bytes per launch measured by
                 rocprofv3 (FETCH_SIZE x2 + WRITE_SIZE x1, calibrated on a
                 plain copy), read from the profiles/ summary of that kernel
  cpu_baseline   the reference's CPU path: PARADIS radix sort (this build's
                 implementation from Cho et al. 2015; the reference's
                 third_party/simple_paradis is absent) on all host cores, on
                 the same 2^30 input, 10 repetitions (median, min)
  configs        the other configurations: 2^24 u32 (GPU, PARADIS, GNU
                 parallel sort), 2^30 u32 under six distributions and u64 /
                 f32 uniform, the merge of two sorted 2^29 runs (pivot,
                 in-place swap, two merge-path merges)
  het            host link (H2D, D2H, both, in-place block sweep) and the
                 heterogeneous sort of 2^33 u32 keys from pinned host memory
                 (8 chunks of 2^30, 3 buffers, merge groups of 8; CPU and GPU
                 final merge) through the sort_benchmark driver
  batched        4096 uneven arrays: one segmented sort vs one sort per array
  serving        the serving loop: batched vs plain per-request path
  cpu_parallel   GNU parallel sort and __gnu_parallel::multiway_merge

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-...]
"""
import argparse
import ctypes
import json
import os
import re
import statistics
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "multi-gpu-sorting_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

SEED = 2147483647
KEY_BYTES = 4
ALG_BYTES_PER_KEY = KEY_BYTES + KEY_BYTES * 2 * KEY_BYTES  # 36: histogram read + 4 x (read + write)
PASS_BYTES_PER_KEY = 2 * KEY_BYTES  # one pass: read + write
PEAK_GBPS = 8000.0
SORT_BENCH = os.path.join(ROOT, "multi-gpu-sorting_amd", "bin", "sort_benchmark")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def alg_bytes(kb, vb=0):
    """Fixed roofline formula: one histogram read plus one read and one write
    of keys (and payloads) per byte of the key."""
    return kb + kb * 2 * (kb + vb)


def host_threads():
    return len(os.sched_getaffinity(0))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_lib():
    path = os.path.join(ROOT, "multi-gpu-sorting_amd", "lib", "libmgsort_host.so")
    L = ctypes.CDLL(path)
    c, vp, sz = ctypes.c_char_p, ctypes.c_void_p, ctypes.c_size_t
    L.mgh_generate.argtypes = [c, c, ctypes.c_uint32, vp, sz]
    L.mgh_generate.restype = ctypes.c_int
    L.mgh_paradis_sort.argtypes = [c, vp, sz, ctypes.c_int]
    L.mgh_gnu_parallel_sort.argtypes = [c, vp, sz, ctypes.c_int]
    L.mgh_multiway_merge.argtypes = [c, vp, sz, sz, vp, ctypes.c_int]
    L.mgh_link_probe.argtypes = [ctypes.c_int, ctypes.c_int, sz, sz, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.mgh_last_error.restype = c
    L.mgh_event_create.restype = vp
    L.mgh_event_destroy.argtypes = [vp]
    L.mgh_event_elapsed_ms.argtypes = [vp, vp]
    L.mgh_event_elapsed_ms.restype = ctypes.c_float
    return L


def generate(H, np, n, dtype="u32", dist="uniform", seed=SEED):
    t = {"u32": np.uint32, "u64": np.uint64, "f32": np.float32}[dtype]
    keys = np.empty(n, dtype=t)
    rc = H.mgh_generate(dtype.encode(), dist.encode(), seed, keys.ctypes.data, n)
    if rc != 0:
        raise RuntimeError(f"mgh_generate {dtype} {dist} failed ({rc})")
    return keys


def to_dev(a, np):
    import torch
    return torch.from_numpy(a.view(np.int32 if a.itemsize == 4 else np.int64)).cuda()


def traffic_from_profiles(kernel):
    """HBM bytes per launch of `kernel` from a calibrated rocprofv3 summary in
    profiles/ ('kernel: <name>' and 'traffic_bytes_per_launch: X')."""
    d = os.path.join(ROOT, "profiles")
    if not os.path.isdir(d):
        return None
    for f in sorted(os.listdir(d)):
        if not f.endswith(".txt"):
            continue
        text = open(os.path.join(d, f)).read()
        if re.search(rf"^kernel:\s*{re.escape(kernel)}\s*$", text, re.M):
            m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
            if m:
                return float(m.group(1))
    return None


# ---------------------------------------------------------------------------
# device-side checks
# ---------------------------------------------------------------------------
def enc_dev(kt, t):
    """enc() as an int64 tensor whose signed order is the sort order."""
    import torch
    import mgsort
    if t.element_size() == 4:
        x = t.to(torch.int64) & 0xFFFFFFFF
        if kt == mgsort.KEY_F32:
            x = torch.where(x == 0x80000000, torch.zeros_like(x), x)
            x = torch.where((x & 0x80000000) != 0, x ^ 0xFFFFFFFF, x | 0x80000000)
        return x
    return t ^ (-(1 << 63))


def check_sorted_permutation(kt, out, src):
    """Sorted by enc and exactly the input's multiset (both sorted by torch)."""
    import torch
    e = enc_dev(kt, out)
    ok = bool((e[1:] >= e[:-1]).all().item())
    del e
    a = out.to(torch.int64) if out.element_size() == 4 else out
    b = src.to(torch.int64) if src.element_size() == 4 else src
    ok = ok and torch.equal(torch.sort(a)[0], torch.sort(b)[0])
    torch.cuda.empty_cache()
    return ok


def time_sorts(mgsort, kt, src, out, temp, reps, warm=1):
    """ms of each of `reps` out-of-place sorts (HIP events on the stream)."""
    import torch
    s = torch.cuda.current_stream()
    for _ in range(warm):
        mgsort.sort(kt, src, keys_out=out, temp=temp, stream=s)
    ms = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        mgsort.sort(kt, src, keys_out=out, temp=temp, stream=s)
        e1.record(s)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    return ms


PATH_NAMES = {1: "small", 2: "lsd", 3: "three_pass"}


# ---------------------------------------------------------------------------
# sections
# ---------------------------------------------------------------------------
def configs_section(H, np, headline_keys):
    import torch
    import mgsort
    out = {}
    threads = host_threads()

    # configuration 1: 2^24 uniform u32, seed 2147483647
    n = 1 << 24
    a = generate(H, np, n)
    exp = np.sort(a)  # std::sort order of unsigned keys
    d = to_dev(a, np)
    o = torch.empty_like(d)
    temp = mgsort.temp_buffer(mgsort.lib.mgs_sort_temp_bytes(mgsort.KEY_U32, mgsort.VAL_NONE, n), "cuda")
    ms = time_sorts(mgsort, "u32", d, o, temp, 10, warm=2)
    gpu_exact = bool(np.array_equal(o.cpu().numpy().view(np.uint32), exp))
    c1 = {"n": n, "gpu_ms_median": round(statistics.median(ms), 4), "gpu_ms_min": round(min(ms), 4),
          "gpu_gkeys_s": round(n / statistics.median(ms) / 1e6, 2),
          "gpu_path": PATH_NAMES.get(mgsort.lib.mgs_sort_path(mgsort.KEY_U32, mgsort.VAL_NONE, n, temp.data_ptr(),
                                                              temp.numel(), None)),
          "gpu_bit_exact_vs_std_sort": gpu_exact}
    for name, fn in (("paradis", H.mgh_paradis_sort), ("gnu_parallel_sort", H.mgh_gnu_parallel_sort)):
        ts, exact = [], True
        for _ in range(10):
            k = a.copy()
            t0 = time.perf_counter()
            fn(b"u32", k.ctypes.data, n, threads)
            ts.append(time.perf_counter() - t0)
            exact = exact and bool(np.array_equal(k, exp))
        c1[name] = {"s_median": round(statistics.median(ts), 5), "s_min": round(min(ts), 5),
                    "gkeys_s": round(n / statistics.median(ts) / 1e9, 4), "threads": threads,
                    "bit_exact_vs_std_sort": exact}
    out["config1_2e24_u32"] = c1
    log("configs: config 1 done", json.dumps(c1))
    if not (gpu_exact and c1["paradis"]["bit_exact_vs_std_sort"] and c1["gnu_parallel_sort"]["bit_exact_vs_std_sort"]):
        raise RuntimeError("configuration 1: an output differs from std::sort")
    del d, o, temp
    torch.cuda.empty_cache()

    # configuration 2: 2^30 keys, distributions and key types
    n = 1 << 30
    c2 = {}
    cases = [("u32", dd) for dd in ("uniform", "normal-clamped", "zipf", "sorted", "reverse-sorted",
                                    "nearly-sorted-swaps")] + [("u64", "uniform"), ("f32", "uniform")]
    for dtype, dist in cases:
        kt = mgsort.key_type(dtype)
        kb = mgsort.KEY_BYTES[kt]
        if dtype == "u32" and dist == "uniform":
            a = headline_keys
        else:
            t0 = time.perf_counter()
            a = generate(H, np, n, dtype, dist)
            log(f"configs: generated 2^30 {dtype} {dist} in {time.perf_counter() - t0:.1f} s")
        src = to_dev(a, np)
        del a
        o = torch.empty_like(src)
        temp = mgsort.temp_buffer(mgsort.lib.mgs_sort_temp_bytes(kt, mgsort.VAL_NONE, n), "cuda")
        ms = time_sorts(mgsort, dtype, src, o, temp, 5)
        path = mgsort.lib.mgs_sort_path(kt, mgsort.VAL_NONE, n, temp.data_ptr(), temp.numel(), None)
        status = mgsort.temp_status(temp)
        del temp
        torch.cuda.empty_cache()
        ok = status == 0 and check_sorted_permutation(kt, o, src)
        med = statistics.median(ms)
        rec = {"ms_median": round(med, 4), "ms_min": round(min(ms), 4), "gkeys_s": round(n / med / 1e6, 2),
               "roofline_frac": round(n * alg_bytes(kb) / (med / 1e3) / 1e9 / PEAK_GBPS, 4),
               "bytes_per_key": alg_bytes(kb), "path": PATH_NAMES.get(path), "verified": ok}
        c2[f"{dtype}_{dist}"] = rec
        log(f"configs: 2^30 {dtype} {dist}: {rec}")
        if not ok:
            raise RuntimeError(f"2^30 {dtype} {dist}: output failed the check")
        del src, o
        torch.cuda.empty_cache()
    out["config2_2e30"] = c2

    # merge of two sorted 2^29 runs on one device
    m = 1 << 29
    src = to_dev(headline_keys, np)
    halves = src.clone()
    mgsort.sort("u32", halves[:m])
    mgsort.sort("u32", halves[m:])
    torch.cuda.synchronize()
    d = torch.empty_like(halves)
    res = torch.zeros(1, dtype=torch.int64, device="cuda")
    o = torch.empty_like(halves)
    mtemp = mgsort.temp_buffer(mgsort.lib.mgs_merge_temp_bytes(mgsort.KEY_U32, mgsort.VAL_NONE, m, m), "cuda")
    t_piv, t_swap, t_merge, p = [], [], [], 0
    for rep in range(6):
        d.copy_(halves)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        mgsort.select_pivot("u32", [d[:m]], [d[m:]], m, res)
        e1.record()
        torch.cuda.synchronize()  # the swap and merge sizes depend on the pivot
        p = int(res.item())
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        ev[0].record()
        if p > 0:
            mgsort.swap_ranges("u32", d[m - p:m], d[m:m + p])
        ev[1].record()
        if p > 0:
            mgsort.merge("u32", d[:m - p], d[m - p:m], o[:m], temp=mtemp)
            mgsort.merge("u32", d[m:m + p], d[m + p:], o[m:], temp=mtemp)
        else:
            o.copy_(d)
        ev[2].record()
        torch.cuda.synchronize()
        if rep:
            t_piv.append(e0.elapsed_time(e1))
            t_swap.append(ev[0].elapsed_time(ev[1]))
            t_merge.append(ev[1].elapsed_time(ev[2]))
    ok = check_sorted_permutation(mgsort.KEY_U32, o, src)
    N = 2 * m
    total = statistics.median(t_piv) + statistics.median(t_swap) + statistics.median(t_merge)
    bytes_moved = 8 * N + 16 * p
    out["config2_merge_2x2e29"] = {
        "p": p, "pivot_ms": round(statistics.median(t_piv), 4), "swap_ms": round(statistics.median(t_swap), 4),
        "merges_ms": round(statistics.median(t_merge), 4), "total_ms": round(total, 4),
        "alg_bytes": bytes_moved, "alg_bytes_formula": "8 N + 16 p (u32): merge-path read + write, swap of p each side",
        "hbm_gbps_alg": round(bytes_moved / (total / 1e3) / 1e9, 1),
        "roofline_frac": round(bytes_moved / (total / 1e3) / 1e9 / PEAK_GBPS, 4),
        "merge_gbps": round(8 * N / (statistics.median(t_merge) / 1e3) / 1e9, 1),
        "verified": ok}
    log("configs: merge", json.dumps(out["config2_merge_2x2e29"]))
    if not ok:
        raise RuntimeError("merge of two sorted 2^29 runs failed the check")
    del src, halves, d, o, mtemp
    torch.cuda.empty_cache()
    return out


def het_section(H, het_n):
    out = {}
    link = {}
    gb = 1 << 30
    for mode, name in ((0, "h2d"), (1, "d2h"), (2, "bidirectional")):
        r = ctypes.c_double(0)
        rc = H.mgh_link_probe(0, mode, gb, 0, 5, ctypes.byref(r))
        link[f"{name}_gbps"] = round(r.value, 2) if rc == 0 else None
    sweep = {}
    for blk in (1 << 20, 4 << 20, 16 << 20, 41 << 20, 64 << 20, 256 << 20):
        r = ctypes.c_double(0)
        rc = H.mgh_link_probe(0, 3, gb, blk, 3, ctypes.byref(r))
        sweep[f"{blk >> 20}MiB"] = round(r.value, 2) if rc == 0 else None
    link["inplace_1GiB_gbps_by_block"] = sweep
    link["what"] = ("pinned host memory, 1 GiB per transfer, best of 5; bidirectional and in-place count both "
                    "directions; in-place = D2H of a buffer's blocks with each block refilled by H2D after it left")
    out["host_link"] = link
    log("het: link", json.dumps(link))

    runs = {}
    chunk = 1 << 30
    for backend in ("cpu", "gpu"):
        n = het_n
        row = None
        while n >= chunk:
            cmd = [SORT_BENCH, str(n), "--algorithm", "gowanlock", "--data_type", "u32", "--chunk_size", str(chunk),
                   "--merge_group_size", "8", "--num_buffers", "3", "--num_threads", str(host_threads()),
                   "--merge_backend", backend]
            log("het:", " ".join(cmd[1:]))
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
            line = r.stdout.strip().splitlines()[-1] if r.stdout.strip() else ""
            if r.returncode == 0 and line and not line.startswith("Error"):
                row = line.split(",")
                break
            log(f"het: {n} keys failed ({r.returncode}): {line} {r.stderr[-300:]}")
            n //= 2
        if row is None:
            runs[backend] = None
            continue
        sort_s = float(row[11])
        merge_s = float(row[18])
        rec = {"n": n, "chunks": n // chunk, "sort_phase_s": round(sort_s, 4), "merge_phase_s": round(merge_s, 4),
               "gkeys_s": round(n / sort_s / 1e9, 3), "verified": row[-1] == "1"}
        if link.get("bidirectional_gbps"):
            # link roofline: every key crosses the link once each way
            t_link = 8 * n / (link["bidirectional_gbps"] * 1e9)
            rec["link_bound_s"] = round(t_link, 4)
            rec["frac_of_link_bound"] = round(t_link / sort_s, 4)
        runs[backend] = rec
        log("het:", backend, json.dumps(rec))
        if not rec["verified"]:
            raise RuntimeError(f"HET {backend}: output failed the check")
    out["gowanlock_u32"] = {"config": "chunk 2^30, 3 buffers (in-place transfers, blocks of chunk/100), merge groups of 8",
                            "runs": runs}
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


def run_serve(args, tag):
    exe = os.path.join(ROOT, "multi-gpu-sorting_amd", "bin", "serve_benchmark")
    p = subprocess.Popen([exe] + args, stdout=subprocess.PIPE, text=True)
    rows = []
    for line in p.stdout:  # progress as it comes
        log(f"serve{tag}:", line.rstrip())
        if line.startswith("{"):
            rows.append(json.loads(line))
    rc = p.wait(timeout=900)
    if rc != 0:
        raise RuntimeError(f"serve_benchmark {' '.join(args)} failed ({rc})")
    runs = [x for x in rows if "mode" in x]
    out = {"runs": runs}
    for x in rows:
        if x.get("event") in ("generated", "ceiling"):
            out[x["event"]] = x
    for data in ("device", "host"):
        sat = {x["mode"]: x for x in runs if x["data"] == data and x["load"] == "inf"}
        if "batched" in sat and "plain" in sat:
            out[f"{data}_saturated_speedup_vs_plain"] = round(sat["batched"]["gkeys_s"] / sat["plain"]["gkeys_s"], 3)
            out[f"{data}_saturated_speedup_vs_plain_overlap"] = round(
                sat["batched"]["gkeys_s"] / sat["plain_overlap"]["gkeys_s"], 3)
    out["matches_host_sort"] = all(x["matches_host_sort"] for x in runs) and out["ceiling"]["matches_host_sort"]
    return out


def serving_section(requests=192):
    """Poisson streams of mixed requests (50% u32, 30% u64, 20% u32+u32
    pairs; uniform / nearly sorted) through the serving layer: continuous
    batching vs the plain per-request path, with every output checked against
    a host sort.  The mixed stream has sizes log-uniform in 2^10..2^26 keys
    (most keys sit in requests of 2^21 and more); the small-request stream
    2^10..2^20 (1024 requests)."""
    out = {"mixed": run_serve(["--requests", str(requests), "--budget_gib", "64", "--loads", "inf,2,1,0.5"], ""),
           "small": run_serve(["--requests", "1024", "--max_log2", "20", "--budget_gib", "64", "--data", "device",
                               "--loads", "inf,1"], " small")}
    out["mixed"]["workload"] = (f"{requests} requests, sizes round(2^(10+16U)), mt19937_64 seed 7; 64 GiB HBM pool; "
                                "requests of 2^21 keys and more sorted alone, two sort streams")
    out["small"]["workload"] = "1024 requests, sizes round(2^(10+10U)), mt19937_64 seed 7; device data"
    return out


def cpu_section(H, np, keys, gpu_sorted):
    """The reference CPU path on the GPU box's host: PARADIS on the headline
    input (10 repetitions), the GNU parallel sort (one run on 2^28 keys) and
    __gnu_parallel::multiway_merge of 8 sorted runs of 2^27 keys, with the
    device k-way merge of the same runs beside it."""
    threads = host_threads()
    n = keys.size
    ts, exact = [], True
    work = np.empty_like(keys)
    for rep in range(10):
        np.copyto(work, keys)
        t0 = time.perf_counter()
        H.mgh_paradis_sort(b"u32", work.ctypes.data, n, threads)
        ts.append(time.perf_counter() - t0)
        if rep == 0:
            exact = bool(np.array_equal(work, gpu_sorted))
    med = statistics.median(ts)
    base = {"value": round(n / med / 1e9, 4), "unit": "GKeys/s", "cores": threads, "kind": "port",
            "sample": (f"PARADIS (in-place parallel radix sort, Cho et al. 2015; this build's implementation, the "
                       f"reference's simple_paradis submodule being absent) on the same {n} uniform u32 keys, "
                       f"10 repetitions: median {med:.3f} s, min {min(ts):.3f} s; output identical to the GPU's: "
                       f"{exact}"),
            "s_median": round(med, 4), "s_min": round(min(ts), 4), "gkeys_s_best": round(n / min(ts) / 1e9, 4),
            "cpu_model": cpu_model(), "numa": "not bound (one process, default policy)",
            "bit_exact_vs_gpu": exact}
    par = {"threads": threads, "cpu_model": base["cpu_model"]}
    n2 = 1 << 28
    k = keys[:n2].copy()
    t0 = time.perf_counter()
    H.mgh_gnu_parallel_sort(b"u32", k.ctypes.data, n2, threads)
    par["gnu_parallel_sort_2e28_gkeys_s"] = round(n2 / (time.perf_counter() - t0) / 1e9, 4)
    # multiway merge of 8 sorted runs (the reference's final HET merge)
    runs = 8
    rl = n // runs
    src = keys.copy()
    for r in range(runs):
        H.mgh_paradis_sort(b"u32", src[r * rl:].ctypes.data, rl, threads)
    dst = np.empty_like(src)
    tm = []
    for _ in range(3):
        t0 = time.perf_counter()
        H.mgh_multiway_merge(b"u32", src.ctypes.data, runs, rl, dst.ctypes.data, threads)
        tm.append(time.perf_counter() - t0)
    ok = bool(np.array_equal(dst, gpu_sorted))
    par["multiway_merge_8_runs"] = {"keys": n, "s_median": round(statistics.median(tm), 4),
                                    "gkeys_s": round(n / statistics.median(tm) / 1e9, 4),
                                    "host_gbps": round(8 * n / statistics.median(tm) / 1e9, 2),
                                    "bit_exact": ok}
    del dst
    # the same 8 runs merged on the device (mgs_multiway_merge, the HET sort's
    # device merge backend)
    import torch
    import mgsort
    d_src = torch.from_numpy(src.view(np.int32)).cuda()
    del src
    d_out = torch.empty_like(d_src)
    druns = [d_src[r * rl:(r + 1) * rl] for r in range(runs)]
    temp = mgsort.temp_buffer(mgsort.lib.mgs_multiway_merge_temp_bytes(mgsort.KEY_U32, runs, n), "cuda")
    gm = []
    for rep in range(6):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        mgsort.multiway_merge("u32", druns, d_out, temp=temp)
        e1.record()
        torch.cuda.synchronize()
        if rep > 0:
            gm.append(e0.elapsed_time(e1))
    gok = mgsort.temp_status(temp) == 0 and bool(np.array_equal(d_out.cpu().numpy().view(np.uint32), gpu_sorted))
    ms = statistics.median(gm)
    par["gpu_multiway_merge_8_runs"] = {"keys": n, "ms_median": round(ms, 3), "gkeys_s": round(n / ms / 1e6, 2),
                                        "hbm_gbps_alg": round(8 * n / ms / 1e6, 1),
                                        "roofline_frac": round(8 * n / ms / 1e6 / PEAK_GBPS, 4),
                                        "speedup_vs_cpu": round(statistics.median(tm) * 1e3 / ms, 2),
                                        "bit_exact": gok}
    del d_src, d_out, druns, temp
    torch.cuda.empty_cache()
    return {"cpu_baseline": base, "cpu_parallel": par}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 30)
    ap.add_argument("--het-n", type=int, default=1 << 33, help="keys of the heterogeneous sort (halved on failure)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-configs", action="store_true", help="skip the other configurations")
    ap.add_argument("--no-het", action="store_true", help="skip the host link and the heterogeneous sort")
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
        tn = mgsort.Tuning(0, 0, 32768, int(lengths.max()))
        temp = mgsort.temp_buffer(mgsort.lib.mgs_segmented_sort_temp_bytes_tuned(
            mgsort.KEY_U64, mgsort.VAL_U32, total, 4096, ctypes.byref(tn)), "cuda")
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
    keys = generate(H, np, n)
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
    path = L.mgs_sort_path(mgsort.KEY_U32, mgsort.VAL_NONE, n, temp.data_ptr(), temp.numel(), None)
    status = mgsort.temp_status(temp)
    if status != 0:
        raise RuntimeError(f"sort status: {L.mgs_status_str(status)}")
    del temp
    torch.cuda.empty_cache()

    # correctness of the last step's output: sorted, and exactly the input's
    # multiset
    if not check_sorted_permutation(mgsort.KEY_U32, out, src):
        raise RuntimeError("sorted output failed the check")
    gpu_sorted = out.cpu().numpy().view(np.uint32)

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

    # the dominant kernel: the slowest of the launches that read and write
    # every key once (the three-pass path's passes, or the LSD passes)
    passes = [k for k in stage_ms if k.startswith("msd_pass") or k == "msd_local"]
    if not passes or min(stage_ms[k] for k in passes) < 0.05:
        passes = [f"pass{p}" for p in range(KEY_BYTES)]
    dom = max(passes, key=lambda k: stage_ms[k])
    dom_kernel = {"msd_pass1": "msd_pass_kernel (pass 1)", "msd_pass2": "msd_pass_kernel (pass 2)",
                  "msd_local": "msd_local_kernel"}.get(dom, "pass_kernel")
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
        "config": {"workload": f"in-HBM sort of {n} uniform u32 keys (out of place, stable radix sort)",
                   "n": n, "key": "u32", "values": "none", "distribution": "uniform", "seed": SEED,
                   "path": PATH_NAMES.get(path)},
        "hbm_gbps_alg": round(n * ALG_BYTES_PER_KEY / (ms_per_step / 1e3) / 1e9, 1),
        "roofline_frac_whole_sort": round(n * ALG_BYTES_PER_KEY / (ms_per_step / 1e3) / 1e9 / PEAK_GBPS, 4),
        "roofline": {"bound": "hbm", "kernel": dom_kernel, "stage": dom,
                     "achieved": round(achieved, 1), "peak": PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_GBPS, 4), "traffic": traffic_from_profiles(dom_kernel),
                     "alg_bytes_per_launch": n * PASS_BYTES_PER_KEY, "avg_launch_ms": round(avg_pass_ms, 4)},
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "copy_ceiling": {"what": "device-to-device copy of the 4 GiB key array (hipMemcpy's kernel), same bytes as one pass",
                         "gbps": round(copy_gbps, 1), "frac_of_peak": round(copy_gbps / PEAK_GBPS, 4),
                         "dominant_vs_copy": round(achieved / copy_gbps, 4)},
    }
    del src, out
    torch.cuda.empty_cache()
    if not args.no_configs:
        result["configs"] = configs_section(H, np, keys)
        torch.cuda.empty_cache()
    if not args.no_het:
        result["het"] = het_section(H, args.het_n)
    if not args.no_batched:
        result["batched"] = batched_section()
        torch.cuda.empty_cache()
    if not args.no_serving:
        result["serving"] = serving_section()
    if not args.no_cpu:
        result.update(cpu_section(H, np, keys, gpu_sorted))
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
