"""The five benchmark configurations' own workloads on the GPU, checked
against the oracle.

The reference itself only checks std::is_sorted (src/sort_benchmark.cu:127),
so these tests are what shows that each configuration's output is the
reference's: bit-exact against the oracle where the oracle finishes in
seconds, and at full size through size-independent properties computed on
the device (sorted by enc, the exact multiset of raw keys, and for pairs with
an index payload the gather identity out_keys == in_keys[out_vals] with a
permutation check, which together pin the stable sort).

  1  2^24 uniform u32 keys, seed 2147483647: mgs_sort, PARADIS and the GNU
     parallel sort, each against the oracle
  2  u32 / u64 / f32 x {uniform, normal-clamped, zipf, sorted, reverse-sorted,
     nearly-sorted-swaps} (+ f32 float-specials) at 2^24 against the oracle;
     one 2^30 run per key type with device-side properties; the merge of two
     sorted 2^29 runs (pivot, in-place swap, two merge-path merges) at full
     size with properties and at 2^24 against the oracle's Tanasic steps
  3  GowanlockSort with 9 chunks, 3 buffers, merge groups of 8, both merge
     backends, against the oracle (2^23 + 777 keys: the chunk arithmetic of
     the 2^33 configuration at a size the oracle sorts in a second)
  4  4096 arrays with log-uniform lengths in [2^8, 2^22] (1.76e9 u64 keys +
     u32 payloads), uniform and zipf(1.2): per-segment properties at full
     shape and the oracle on a seeded sample of 64 segments with the largest
  5  SortServer, plain and batched, host and device data, two loads, the
     configuration's type mix: every request's output against the oracle
"""
import ctypes
import os
import time

import numpy as np
import pytest

import pyoracle as O
from conftest import ROOT

pytestmark = pytest.mark.gpu

HOST_LIB = os.path.join(ROOT, "multi-gpu-sorting_amd", "lib", "libmgsort_host.so")
SEED = 2147483647  # the reference's default seed (src/argument_limits.cuh:156)
TYPES = {"u32": (O.K_U32, np.uint32), "u64": (O.K_U64, np.uint64), "f32": (O.K_F32, np.float32)}
INT64_MIN = -(1 << 63)


def log(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def mg():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mgsort
    return mgsort


@pytest.fixture(scope="module")
def H(mg):
    L = ctypes.CDLL(HOST_LIB)
    c, vp, sz = ctypes.c_char_p, ctypes.c_void_p, ctypes.c_size_t
    L.mgh_generate.argtypes = [c, c, ctypes.c_uint32, vp, sz]
    L.mgh_paradis_sort.argtypes = [c, vp, sz, ctypes.c_int]
    L.mgh_gnu_parallel_sort.argtypes = [c, vp, sz, ctypes.c_int]
    L.mgh_gowanlock_sort.argtypes = [c, vp, sz, ctypes.c_int, sz, sz, sz, c, ctypes.c_int, sz]
    L.mgh_last_error.restype = c
    L.mgh_serve.argtypes = [vp, sz, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                            ctypes.POINTER(ctypes.c_double)]
    return L


def generate(H, dtype, dist, seed, n):
    _, t = TYPES[dtype]
    a = np.empty(n, dtype=t)
    rc = H.mgh_generate(dtype.encode(), dist.encode(), seed, a.ctypes.data, n)
    assert rc == 0, (dtype, dist, rc)
    return a


def bits(a):
    return a.view(np.uint32 if a.itemsize == 4 else np.uint64)


def to_dev(a):
    import torch
    return torch.from_numpy(a.view(np.int32 if a.itemsize == 4 else np.int64)).cuda()


def host_threads():
    return min(16, len(os.sched_getaffinity(0)))


# ---------------------------------------------------------------------------
# device-side properties
# ---------------------------------------------------------------------------
def enc_dev(kt, t):
    """enc() on the device as an int64 tensor whose signed order is the sort
    order (32-bit keys: the unsigned image; 64-bit keys: image ^ 2^63)."""
    import torch
    if t.element_size() == 4:
        x = t.to(torch.int64) & 0xFFFFFFFF
        if kt == O.K_F32:
            x = torch.where(x == 0x80000000, torch.zeros_like(x), x)
            neg = (x & 0x80000000) != 0
            x = torch.where(neg, x ^ 0xFFFFFFFF, x | 0x80000000)
        return x
    x = t.clone()
    if kt == O.K_U64:
        return x ^ INT64_MIN
    raise ValueError(kt)


def assert_sorted_dev(kt, t):
    e = enc_dev(kt, t)
    assert bool((e[1:] >= e[:-1]).all().item()), "not sorted by enc"


def assert_same_multiset_dev(a, b):
    import torch
    sa = torch.sort(a.to(torch.int64) if a.element_size() == 4 else a)[0]
    sb = torch.sort(b.to(torch.int64) if b.element_size() == 4 else b)[0]
    assert torch.equal(sa, sb), "output is not a permutation of the input"


def assert_stable_pairs_dev(kt, in_k, out_k, out_v):
    """out is the stable sort of (in_k, index): strictly increasing in
    (enc(key), index), every index once, and every key where its index says."""
    import torch
    n = in_k.numel()
    idx = out_v.to(torch.int64) & 0xFFFFFFFF
    e = enc_dev(kt, out_k)
    inc = (e[1:] > e[:-1]) | ((e[1:] == e[:-1]) & (idx[1:] > idx[:-1]))
    assert bool(inc.all().item()), "not stable-sorted by (enc, index)"
    seen = torch.zeros(n, dtype=torch.bool, device=in_k.device)
    seen[idx] = True
    assert bool(seen.all().item()), "payload is not a permutation"
    assert torch.equal(in_k[idx], out_k), "a key is not the input key of its payload index"


# ---------------------------------------------------------------------------
# configuration 1
# ---------------------------------------------------------------------------
def test_config1_plain_2e24_gpu_and_cpu_paths(mg, H):
    import torch
    n = 1 << 24
    a = generate(H, "u32", "uniform", SEED, n)
    exp, _ = O.sort(O.K_U32, a)
    assert np.array_equal(exp, np.sort(a))  # std::sort order for unsigned keys
    d = to_dev(a)
    mg.sort("u32", d)
    torch.cuda.synchronize()
    assert np.array_equal(bits(d.cpu().numpy()), exp)
    for name, fn in (("paradis", H.mgh_paradis_sort), ("gnu_parallel_sort", H.mgh_gnu_parallel_sort)):
        k = a.copy()
        assert fn(b"u32", k.ctypes.data, n, host_threads()) == 0, name
        assert np.array_equal(k, exp), name


# ---------------------------------------------------------------------------
# configuration 2
# ---------------------------------------------------------------------------
C2_DISTS = ["uniform", "normal-clamped", "zipf", "sorted", "reverse-sorted", "nearly-sorted-swaps"]
C2_CASES = [(t, d) for t in ("u32", "u64", "f32") for d in C2_DISTS] + [("f32", "float-specials")]


@pytest.mark.parametrize("dtype,dist", C2_CASES, ids=[f"{t}-{d}" for t, d in C2_CASES])
def test_config2_distributions_2e24(mg, H, dtype, dist):
    """Keys alone and keys with an index payload, bit-exact against the
    oracle (stability is visible in the payload)."""
    import torch
    kt, _ = TYPES[dtype]
    n = 1 << 24
    a = generate(H, dtype, dist, 4242, n)
    idx = np.arange(n, dtype=np.uint32)
    ek, ev = O.sort(kt, bits(a), idx)
    dk = to_dev(a)
    mg.sort(dtype, dk)
    torch.cuda.synchronize()
    assert np.array_equal(bits(dk.cpu().numpy()), ek), "keys-only sort differs from the oracle"
    dk = to_dev(a)
    dv = torch.from_numpy(idx.view(np.int32)).cuda()
    mg.sort(dtype, dk, dv)
    torch.cuda.synchronize()
    assert np.array_equal(bits(dk.cpu().numpy()), ek), "pair sort keys differ from the oracle"
    assert np.array_equal(dv.cpu().numpy().view(np.uint32), ev), "pair sort payloads differ from the oracle"


C2_FULL = [("u32", "uniform"), ("u64", "zipf"), ("f32", "float-specials")]


@pytest.mark.parametrize("dtype,dist", C2_FULL, ids=[f"{t}-{d}" for t, d in C2_FULL])
def test_config2_full_size_2e30_properties(mg, H, dtype, dist):
    import torch
    kt, _ = TYPES[dtype]
    n = 1 << 30
    t0 = time.time()
    a = generate(H, dtype, dist, SEED, n)
    log(f"  generated 2^30 {dtype} {dist} in {time.time() - t0:.1f} s")
    src = to_dev(a)
    del a
    out = torch.empty_like(src)
    temp = mg.temp_buffer(mg.lib.mgs_sort_temp_bytes(mg.key_type(dtype), mg.VAL_NONE, n), "cuda")
    mg.sort(dtype, src, keys_out=out, temp=temp)
    mg.check_temp(temp, "keys-only 2^30")
    assert_sorted_dev(kt, out)
    assert_same_multiset_dev(src, out)
    log("  keys-only: sorted, same multiset")
    del out, temp
    torch.cuda.empty_cache()
    idx = torch.arange(n, dtype=torch.int32, device="cuda")
    ok, ov = torch.empty_like(src), torch.empty_like(idx)
    mg.sort(dtype, src, idx, keys_out=ok, vals_out=ov)
    torch.cuda.synchronize()
    del idx
    assert_stable_pairs_dev(kt, src, ok, ov)
    log("  pairs: stable sort of (key, index)")


def test_config2_merge_of_two_sorted_2e29_runs(mg, H):
    """One device: sort both halves, find the pivot, swap L's top p with R's
    bottom p in place, merge each half (src/tanasic_sort.cuh:89-174), at the
    configuration's full size; checked by the pivot's defining inequality,
    global order and the multiset."""
    import torch
    m = 1 << 29
    a = generate(H, "u32", "uniform", SEED, 2 * m)
    inp = to_dev(a)
    del a
    d = inp.clone()
    mg.sort("u32", d[:m])
    mg.sort("u32", d[m:])
    res = torch.zeros(1, dtype=torch.int64, device="cuda")
    mg.select_pivot("u32", [d[:m]], [d[m:]], m, res)
    torch.cuda.synchronize()
    p = int(res.item())
    e = enc_dev(O.K_U32, d)
    L, R = e[:m], e[m:]
    # p = max k with L[M-k] > R[k-1]  (k == 0 when no such k)
    assert 0 <= p <= m
    if p > 0:
        assert int(L[m - p]) > int(R[p - 1])
    if p < m:
        assert int(L[m - p - 1]) <= int(R[p])
    log(f"  pivot p = {p} (of {m})")
    del e, L, R
    out = torch.empty_like(d)
    if p > 0:
        mg.swap_ranges("u32", d[m - p:m], d[m:m + p])
        mg.merge("u32", d[:m - p], d[m - p:m], out[:m])
        mg.merge("u32", d[m:m + p], d[m + p:], out[m:])
    else:
        out.copy_(d)
    torch.cuda.synchronize()
    assert_sorted_dev(O.K_U32, out)
    assert_same_multiset_dev(inp, out)


def test_config2_merge_2e24_against_oracle(mg, H):
    import torch
    m = 1 << 23
    a = generate(H, "u32", "uniform", 77, 2 * m)
    idx = np.arange(2 * m, dtype=np.uint32)
    hk, hv = a.copy(), idx.copy()
    for h in range(2):
        s = slice(h * m, (h + 1) * m)
        hk[s], hv[s] = O.sort(O.K_U32, hk[s], hv[s])
    ek, ev, ep = O.tanasic_1dev(O.K_U32, hk, hv)
    dk, dv = to_dev(hk), torch.from_numpy(hv.view(np.int32)).cuda()
    res = torch.zeros(1, dtype=torch.int64, device="cuda")
    mg.select_pivot("u32", [dk[:m]], [dk[m:]], m, res)
    torch.cuda.synchronize()
    p = int(res.item())
    assert p == ep
    ok, ov = torch.empty_like(dk), torch.empty_like(dv)
    mg.swap_ranges("u32", dk[m - p:m], dk[m:m + p], dv[m - p:m], dv[m:m + p])
    mg.merge("u32", dk[:m - p], dk[m - p:m], ok[:m], dv[:m - p], dv[m - p:m], ov[:m])
    mg.merge("u32", dk[m:m + p], dk[m + p:], ok[m:], dv[m:m + p], dv[m + p:], ov[m:])
    torch.cuda.synchronize()
    assert np.array_equal(bits(ok.cpu().numpy()), ek)
    assert np.array_equal(ov.cpu().numpy().view(np.uint32), ev)


# ---------------------------------------------------------------------------
# configuration 3
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("backend", ["cpu", "gpu"])
def test_config3_het_nine_chunks_three_buffers(H, backend):
    chunk = 1 << 20
    n = 8 * chunk + 777  # 9 chunks, the last one short
    a = generate(H, "u32", "uniform", SEED, n)
    k = a.copy()
    rc = H.mgh_gowanlock_sort(b"u32", k.ctypes.data, n, 0, chunk, 8, 3, backend.encode(), host_threads(), 0)
    assert rc == 0, H.mgh_last_error()
    exp, _ = O.sort(O.K_U32, a)
    assert np.array_equal(k, exp)


# ---------------------------------------------------------------------------
# configuration 4
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dist", ["uniform", "zipf-1.2"])
def test_config4_batched_4096_arrays(mg, dist):
    import torch
    from mgsort import batched
    lengths = batched.segment_lengths(4096, seed=2024)
    offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    total = int(offsets[-1])
    keys = batched.make_keys(total, dist, 2024, torch.device("cuda"))
    vals = batched.local_index(lengths, torch.device("cuda"))
    ok, ov = torch.empty_like(keys), torch.empty_like(vals)
    d_off = torch.from_numpy(offsets).cuda()
    tn = mg.Tuning(0, 0, 0, int(lengths.max()))
    temp = mg.temp_buffer(mg.lib.mgs_segmented_sort_temp_bytes_tuned(mg.KEY_U64, mg.VAL_U32, total, 4096,
                                                                      ctypes.byref(tn)), "cuda")
    mg.segmented_sort(mg.KEY_U64, keys, d_off, vals, keys_out=ok, vals_out=ov, tuning=tn, temp=temp)
    mg.check_temp(temp, "segmented sort")
    log(f"  {total} keys in 4096 arrays sorted")
    assert batched.check_segments(ok, ov, keys, vals, offsets)
    # oracle on a seeded sample of 64 segments, the largest among them
    rng = np.random.default_rng(64)
    pick = set(rng.choice(4096, size=63, replace=False).tolist()) | {int(np.argmax(lengths))}
    hk = keys.cpu().numpy().view(np.uint64)
    hv = vals.cpu().numpy().view(np.uint32)
    gk = ok.cpu().numpy().view(np.uint64)
    gv = ov.cpu().numpy().view(np.uint32)
    for s in sorted(pick):
        lo, hi = int(offsets[s]), int(offsets[s + 1])
        ek, ev = O.sort(O.K_U64, hk[lo:hi], hv[lo:hi])
        assert np.array_equal(gk[lo:hi], ek), s
        assert np.array_equal(gv[lo:hi], ev), s


# ---------------------------------------------------------------------------
# configuration 5
# ---------------------------------------------------------------------------
class ServeRequest(ctypes.Structure):
    _fields_ = [("key_type", ctypes.c_int), ("val_type", ctypes.c_int), ("n", ctypes.c_size_t),
                ("keys_in", ctypes.c_void_p), ("vals_in", ctypes.c_void_p), ("keys_out", ctypes.c_void_p),
                ("vals_out", ctypes.c_void_p), ("arrival_s", ctypes.c_double), ("latency_s", ctypes.c_double),
                ("status", ctypes.c_int)]


def serving_mix(H, count, lo, hi, seed):
    """The configuration's mix: sizes round(2^(lo + (hi-lo) U)); 50% u32
    keys, 30% u64 keys, 20% u32 keys with u32 payloads; uniform or
    nearly-sorted-swaps."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        n = int(round(2.0 ** (lo + (hi - lo) * rng.random())))
        t = rng.random()
        dtype, pairs = ("u32", False) if t < 0.5 else (("u64", False) if t < 0.8 else ("u32", True))
        dist = "nearly-sorted-swaps" if rng.random() < 0.5 else "uniform"
        keys = generate(H, dtype, dist, int(rng.integers(1, 2**31)), n)
        vals = np.arange(n, dtype=np.uint32) if pairs else None
        reqs.append((dtype, keys, vals))
    return reqs


@pytest.mark.parametrize("batched", [1, 0], ids=["batched", "plain"])
@pytest.mark.parametrize("host_data", [1, 0], ids=["host", "device"])
@pytest.mark.parametrize("load", ["burst", "paced"])
def test_config5_serving_outputs_match_oracle(mg, H, batched, host_data, load):
    reqs = serving_mix(H, 48, 10, 22, seed=5 + batched + 2 * host_data)
    arr = (ServeRequest * len(reqs))()
    outs = []
    gap = 0.0 if load == "burst" else 0.0005
    for i, (dtype, keys, vals) in enumerate(reqs):
        ko = np.empty_like(keys)
        vo = None if vals is None else np.empty_like(vals)
        outs.append((ko, vo))
        arr[i].key_type = mg.key_type(dtype)
        arr[i].val_type = mg.VAL_NONE if vals is None else mg.VAL_U32
        arr[i].n = keys.size
        arr[i].keys_in = keys.ctypes.data
        arr[i].vals_in = None if vals is None else vals.ctypes.data
        arr[i].keys_out = ko.ctypes.data
        arr[i].vals_out = None if vo is None else vo.ctypes.data
        arr[i].arrival_s = i * gap
    rate = ctypes.c_double(0)
    rc = H.mgh_serve(ctypes.cast(arr, ctypes.c_void_p), len(reqs), 0, batched, 3, host_data, 8.0, ctypes.byref(rate))
    assert rc == 0, H.mgh_last_error()
    lat = sorted(arr[i].latency_s for i in range(len(reqs)))
    log(f"  {len(reqs)} requests, {rate.value:.2f} GKeys/s, p50 {lat[len(lat) // 2] * 1e3:.2f} ms")
    for i, (dtype, keys, vals) in enumerate(reqs):
        kt, _ = TYPES[dtype]
        assert arr[i].status == 0, i
        ek, ev = O.sort(kt, bits(keys), vals)
        ko, vo = outs[i]
        assert np.array_equal(bits(ko), ek), (i, dtype, keys.size)
        if vals is None:
            # keys alone: the output is a permutation of the input
            assert np.array_equal(np.sort(bits(ko)), np.sort(bits(keys)))
        else:
            assert np.array_equal(vo, ev), (i, "payload")
