"""Batched segmented sort workload: thousands of uneven arrays sorted by one
fixed launch sequence (mgs_segmented_sort, captured once in a HIP graph and
replayed) against a loop of one plain mgs_sort per array.

Workload (4096 arrays): lengths round(2^(8 + 14 U)) with U from a seeded
mt19937_64 (log-uniform in [2^8, 2^22]), u64 keys (uniform, or zipf(s=1.2)
over 2^20 distinct values), u32 payload = index of the key inside its array,
so stability is observable.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import (KEY_U64, VAL_U32, lib, segmented_sort, temp_buffer, tuning)

MASK64 = (1 << 64) - 1


class MT19937_64:
    """The 64-bit Mersenne Twister (Matsumoto & Nishimura 2000), the engine
    std::mt19937_64 names; used for the array lengths."""

    def __init__(self, seed: int):
        self.mt = [0] * 312
        self.mt[0] = seed & MASK64
        for i in range(1, 312):
            self.mt[i] = (6364136223846793005 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 62)) + i) & MASK64
        self.i = 312

    def __call__(self) -> int:
        if self.i >= 312:
            for k in range(312):
                x = (self.mt[k] & 0xFFFFFFFF80000000) | (self.mt[(k + 1) % 312] & 0x7FFFFFFF)
                xa = x >> 1
                if x & 1:
                    xa ^= 0xB5026F5AA96619E9
                self.mt[k] = self.mt[(k + 156) % 312] ^ xa
            self.i = 0
        y = self.mt[self.i]
        self.i += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        return y & MASK64


def segment_lengths(nseg: int = 4096, lo_log2: float = 8.0, hi_log2: float = 22.0, seed: int = 2024) -> np.ndarray:
    """round(2^(lo + (hi-lo) U)), U = 53-bit canonical double of mt19937_64."""
    g = MT19937_64(seed)
    out = np.empty(nseg, dtype=np.int64)
    for s in range(nseg):
        u = (g() >> 11) * (1.0 / 9007199254740992.0)
        out[s] = int(round(2.0 ** (lo_log2 + (hi_log2 - lo_log2) * u)))
    return out


def zipf_table(s: float, k: int, device) -> torch.Tensor:
    """CDF of zipf(s) over ranks 1..k (float64, on device)."""
    w = 1.0 / torch.arange(1, k + 1, dtype=torch.float64, device=device) ** s
    c = torch.cumsum(w, 0)
    return c / c[-1]


def make_keys(total: int, dist: str, seed: int, device) -> torch.Tensor:
    """u64 keys as an int64 tensor (bit patterns)."""
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    if dist == "uniform":
        hi = torch.randint(-(2 ** 31), 2 ** 31 - 1, (total,), dtype=torch.int64, device=device, generator=g)
        lo = torch.randint(0, 2 ** 32 - 1, (total,), dtype=torch.int64, device=device, generator=g)
        return (hi << 32) | lo
    if dist.startswith("zipf"):
        s = float(dist.split("-")[1]) if "-" in dist else 1.2
        k = 1 << 20
        cdf = zipf_table(s, k, device)
        hi = torch.randint(-(2 ** 31), 2 ** 31 - 1, (k,), dtype=torch.int64, device=device, generator=g)
        lo = torch.randint(0, 2 ** 32 - 1, (k,), dtype=torch.int64, device=device, generator=g)
        values = (hi << 32) | lo  # seeded value map: rank -> key
        out = torch.empty(total, dtype=torch.int64, device=device)
        step = 1 << 26
        for a in range(0, total, step):
            b = min(total, a + step)
            u = torch.rand(b - a, dtype=torch.float64, device=device, generator=g)
            r = torch.searchsorted(cdf, u).clamp_(max=k - 1)
            out[a:b] = values[r]
        return out
    raise ValueError(dist)


def local_index(lengths: np.ndarray, device) -> torch.Tensor:
    """u32 payload: position of each key inside its array."""
    total = int(lengths.sum())
    offs = np.concatenate([[0], np.cumsum(lengths)[:-1]])
    idx = torch.arange(total, dtype=torch.int64, device=device)
    starts = torch.repeat_interleave(torch.from_numpy(offs).to(device), torch.from_numpy(lengths).to(device))
    return (idx - starts).to(torch.int32)


def check_segments(keys: torch.Tensor, vals: torch.Tensor, in_keys: torch.Tensor, in_vals: torch.Tensor,
                   offsets: np.ndarray) -> bool:
    """Per segment: ascending by unsigned key, ties in payload order (stable),
    and the same multiset of (key, payload) pairs as the input."""
    n = keys.numel()
    seg_id = torch.repeat_interleave(torch.arange(len(offsets) - 1, device=keys.device),
                                     torch.from_numpy(np.diff(offsets)).to(keys.device))
    same = seg_id[1:] == seg_id[:-1]
    ku = keys ^ (-(2 ** 63))  # unsigned order as signed
    a, b = ku[:-1], ku[1:]
    order_ok = bool(((b > a) | ((b == a) & (vals[1:] > vals[:-1])) | ~same).all().item())
    if not order_ok:
        return False
    # multiset check with a mixing hash of (segment, key, payload)
    def h(k, v):
        x = (k * 0x9E3779B97F4A7C15) ^ (v.to(torch.int64) * 0xC2B2AE3D27D4EB4F) ^ (seg_id * 0x165667B19E3779F9)
        x = x ^ (x >> 29)
        return x
    return bool(h(keys, vals).sum().item() == h(in_keys, in_vals).sum().item()) and n == in_keys.numel()


@dataclass
class BatchedResult:
    total_keys: int
    nseg: int
    batched_ms: float
    loop_ms: float
    loop_graph_ms: Optional[float]
    tuning: Dict[str, object]
    verified: bool


def run(dist: str = "uniform", nseg: int = 4096, seed: int = 2024, reps: int = 3, loop_graph: bool = True,
        tune_now: bool = False, log=print) -> BatchedResult:
    """The batch as one segmented sort (captured in a graph and replayed)
    against one mgs_sort per array.  The segmented sort's tuning comes from
    the persisted table (mgsort.tuning) for u64 keys + u32 payloads at this
    batch's size, or, without a table entry (or with tune_now), from timing
    every point of the search space on this batch."""
    dev = torch.device("cuda")
    lengths = segment_lengths(nseg, seed=seed)
    offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    total = int(offsets[-1])
    log(f"batched: {nseg} arrays, {total} keys ({dist})")
    keys = make_keys(total, dist, seed, dev)
    vals = local_index(lengths, dev)
    out_k, out_v = torch.empty_like(keys), torch.empty_like(vals)
    d_off = torch.from_numpy(offsets).to(dev)
    stream = torch.cuda.Stream()
    found = None if tune_now else tuning.lookup(KEY_U64, VAL_U32, total)
    if found is not None:
        tn, entry = found
        point = {"large_variant": tn.large_variant, "small_max": tn.small_max, "medium_max": tn.medium_max}
        tuned = {"point": point, "source": f"table ({entry['key']}+{entry['val']}, 2^{entry['size_log2']} keys)",
                 "table_ms": entry.get("ms"), "table_default_ms": entry.get("default_ms")}
    else:
        with torch.cuda.stream(stream):
            point, best_ms, default_ms, timings = tuning.tune(KEY_U64, keys, d_off, vals, out_k, out_v, stream)
        tn = tuning.as_tuning(point)
        tuned = {"point": point, "source": f"timed on this batch ({len(timings)} points)",
                 "ms": round(best_ms, 3), "default_ms": round(default_ms, 3)}
    # the longest array bounds the segment length: the large-segment passes
    # then use 32-bit look-back descriptors although the batch holds more
    # than 2^30 keys
    tn.max_segment_keys = int(lengths.max())
    tuned["max_segment_keys"] = tn.max_segment_keys
    blk, ipt = tuning.geometry(KEY_U64, VAL_U32, tn.large_variant)
    tuned["geometry"] = {"block": blk, "items_per_thread": ipt}
    need = lib.mgs_segmented_sort_temp_bytes_tuned(KEY_U64, VAL_U32, total, nseg, ctypes.byref(tn))
    temp = temp_buffer(need, dev)

    # the whole batch as one graph, replayed
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        segmented_sort(KEY_U64, keys, d_off, vals, keys_out=out_k, vals_out=out_v, tuning=tn, temp=temp,
                       stream=torch.cuda.current_stream())
    g.replay()
    torch.cuda.synchronize()
    verified = check_segments(out_k, out_v, keys, vals, offsets)
    t0 = time.perf_counter()
    for _ in range(reps):
        g.replay()
    torch.cuda.synchronize()
    batched_ms = (time.perf_counter() - t0) * 1e3 / reps
    del g

    # plain: one mgs_sort per array, eagerly from the host
    out_k.zero_()
    ptr_k, ptr_v, ptr_ok, ptr_ov = keys.data_ptr(), vals.data_ptr(), out_k.data_ptr(), out_v.data_ptr()
    sp = stream.cuda_stream
    tb = lib.mgs_sort_temp_bytes(KEY_U64, VAL_U32, int(lengths.max()))
    ptemp = temp_buffer(tb, dev)

    def loop():
        for s in range(nseg):
            o, ln = int(offsets[s]), int(lengths[s])
            st = lib.mgs_sort(KEY_U64, VAL_U32, ptr_k + 8 * o, ptr_ok + 8 * o, ptr_v + 4 * o, ptr_ov + 4 * o, ln,
                              ptemp.data_ptr(), tb, sp)
            if st != 0:
                raise RuntimeError(f"mgs_sort: {lib.mgs_status_str(st)}")

    loop()
    torch.cuda.synchronize()
    verified = verified and check_segments(out_k, out_v, keys, vals, offsets)
    t0 = time.perf_counter()
    for _ in range(reps):
        loop()
    torch.cuda.synchronize()
    loop_ms = (time.perf_counter() - t0) * 1e3 / reps

    loop_graph_ms = None
    if loop_graph:
        g2 = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g2, stream=stream):
            sp = torch.cuda.current_stream().cuda_stream
            loop()
        g2.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            g2.replay()
        torch.cuda.synchronize()
        loop_graph_ms = (time.perf_counter() - t0) * 1e3 / reps
        del g2
    return BatchedResult(total, nseg, batched_ms, loop_ms, loop_graph_ms, tuned, verified)
