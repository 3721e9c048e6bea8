"""Autotuner of the batched segmented sort (no reference counterpart: the
batching layer is this library's).

The search space per (key type, payload type, size bin):
  large_variant  the tile geometry of the large segments' passes
                 (mgs_segmented_variant: threads x items per thread)
  medium_max     longest segment one workgroup streams alone
  small_max      longest segment sorted inside one workgroup's LDS
Every point is timed on a synthetic batch of that size (segment lengths
log-uniform in [2^8, 2^22], uniform keys, u32 payload = index inside the
segment).  The winners are written to a JSON table, which the library loads
once (first lookup) from seg_tuning.json beside this file, or from the path
in MGS_SEG_TUNING.

  python -m mgsort.tuning --out table.json [--sizes 24,27,30] [--keys u32,u64,f32] [--vals none,u32]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
from typing import Dict, List, Optional, Sequence, Tuple

from . import (KEY_NAMES, VAL_NAMES, Tuning, key_type, lib, segmented_sort, temp_buffer, val_type)

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "seg_tuning.json")
KEY_STR = {v: k for k, v in KEY_NAMES.items()}
VAL_STR = {0: "none", 1: "u32", 2: "u64"}
MEDIUM_MAX = (0, 1 << 15, 1 << 16, 1 << 18)  # 0 = the library default (2^17)
SMALL_MAX = (0, 1024)                          # 0 = the library default (one LDS tile)

_table: Optional[dict] = None


def variant_count() -> int:
    return lib.mgs_segmented_variant_count()


def geometry(kt, vt, variant: int) -> Tuple[int, int]:
    """(threads per workgroup, items per thread) of a large-segment variant."""
    b, i = ctypes.c_int32(0), ctypes.c_int32(0)
    st = lib.mgs_segmented_variant(key_type(kt), val_type(vt), int(variant), ctypes.byref(b), ctypes.byref(i))
    if st != 0:
        raise ValueError(f"mgs_segmented_variant({kt}, {vt}, {variant}): {lib.mgs_status_str(st).decode()}")
    return b.value, i.value


def candidates() -> List[Dict[str, int]]:
    return [{"large_variant": v, "small_max": sm, "medium_max": mm}
            for v in range(variant_count()) for mm in MEDIUM_MAX for sm in SMALL_MAX]


def as_tuning(point: Dict[str, int]) -> Tuning:
    return Tuning(int(point.get("large_variant", 0)), int(point.get("small_max", 0)),
                  int(point.get("medium_max", 0)), int(point.get("max_segment_keys", 0)))


def size_bin(total: int) -> int:
    return int(round(math.log2(max(int(total), 1))))


def load_table(path: Optional[str] = None, reload: bool = False) -> dict:
    """The persisted table (read once; an absent file is an empty table)."""
    global _table
    if _table is not None and not reload and path is None:
        return _table
    p = path or os.environ.get("MGS_SEG_TUNING") or TABLE_PATH
    t = {"format": 1, "entries": []}
    if os.path.exists(p):
        with open(p) as f:
            t = json.load(f)
        if t.get("format") != 1 or not isinstance(t.get("entries"), list):
            raise ValueError(f"{p}: not a segmented-sort tuning table")
    if path is None:
        _table = t
    return t


def lookup(kt, vt, total: int, table: Optional[dict] = None) -> Optional[Tuple[Tuning, dict]]:
    """Tuning of the nearest size bin measured for (kt, vt), with its entry;
    None when the table has no entry for the pair."""
    t = table if table is not None else load_table()
    k, v = KEY_STR[key_type(kt)], VAL_STR[val_type(vt)]
    rows = [e for e in t["entries"] if e["key"] == k and e["val"] == v]
    if not rows:
        return None
    b = size_bin(total)
    e = min(rows, key=lambda r: (abs(r["size_log2"] - b), r["size_log2"]))
    return as_tuning(e["tuning"]), e


def temp_bytes_for(kt, vt, max_total: int, nseg: int, points: Sequence[Dict[str, int]]) -> int:
    """Scratch that serves every point (the large tiles, and so the
    descriptor count, differ between variants)."""
    need = 0
    for p in points:
        tn = as_tuning(p)
        need = max(need, lib.mgs_segmented_sort_temp_bytes_tuned(key_type(kt), val_type(vt), max_total, nseg,
                                                                 ctypes.byref(tn)))
    return need


def time_point(kt, keys, d_off, vals, out_k, out_v, point, temp, stream, reps: int = 2) -> float:
    import torch
    tn = as_tuning(point)
    segmented_sort(kt, keys, d_off, vals, keys_out=out_k, vals_out=out_v, tuning=tn, temp=temp, stream=stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        segmented_sort(kt, keys, d_off, vals, keys_out=out_k, vals_out=out_v, tuning=tn, temp=temp, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(kt, keys, d_off, vals, out_k, out_v, stream, points: Optional[Sequence[Dict[str, int]]] = None,
         reps: int = 2) -> Tuple[Dict[str, int], float, float, List[Tuple[Dict[str, int], float]]]:
    """Times every point on this batch; returns (best point, its ms, the
    default point's ms, all timings)."""
    points = list(points) if points is not None else candidates()
    vt = "none" if vals is None else ("u32" if vals.element_size() == 4 else "u64")
    temp = temp_buffer(temp_bytes_for(kt, vt, keys.numel(), d_off.numel() - 1, points), keys.device, stream)
    timings = [(p, time_point(kt, keys, d_off, vals, out_k, out_v, p, temp, stream, reps)) for p in points]
    best, best_ms = min(timings, key=lambda x: x[1])
    default_ms = next((ms for p, ms in timings
                       if p.get("large_variant", 0) == 0 and p.get("small_max", 0) == 0 and p.get("medium_max", 0) == 0),
                      float("nan"))
    return dict(best), best_ms, default_ms, timings


def synthetic_batch(kt, vt, total: int, seed: int = 2024):
    """Segment lengths round(2^(8 + 14 U)) (U from mt19937_64) until about
    `total` keys, uniform keys, payload = index inside the segment."""
    import numpy as np
    import torch
    from .batched import local_index, segment_lengths
    dev = torch.device("cuda")
    est = max(1, int(total / 600_000 * 1.5) + 16)  # mean length of the distribution is about 4.3e5
    lengths = segment_lengths(est, seed=seed)
    while lengths.sum() < total:
        est *= 2
        lengths = segment_lengths(est, seed=seed)
    cut = int(np.searchsorted(np.cumsum(lengths), total)) + 1
    lengths = lengths[:cut]
    offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    n = int(offsets[-1])
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    if KEY_NAMES[kt] in (KEY_NAMES["u32"], KEY_NAMES["i32"], KEY_NAMES["f32"]):
        keys = torch.randint(-2**31, 2**31 - 1, (n,), dtype=torch.int32, device=dev, generator=g)
    else:
        keys = torch.randint(-2**63, 2**63 - 1, (n,), dtype=torch.int64, device=dev, generator=g)
    vals = None
    if vt == "u32":
        vals = local_index(lengths, dev)
    elif vt == "u64":
        vals = local_index(lengths, dev).to(torch.int64)
    return keys, vals, torch.from_numpy(offsets).to(dev), lengths


def main(argv=None) -> int:
    import torch
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--sizes", default="24,27,30", help="log2 of the batch sizes")
    ap.add_argument("--keys", default="u32,u64,f32")
    ap.add_argument("--vals", default="none,u32")
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--append", action="store_true", help="add to (or replace entries of) an existing --out table")
    args = ap.parse_args(argv)
    props = torch.cuda.get_device_properties(0)
    table = {"format": 1, "device": f"{props.name} ({getattr(props, 'gcnArchName', '?')})",
             "workload": "segment lengths round(2^(8+14U)), U from mt19937_64 seed 2024, until the bin's size; "
                         "uniform keys; u32 payload = index inside the segment",
             "search": {"large_variant": list(range(variant_count())), "medium_max": list(MEDIUM_MAX),
                        "small_max": list(SMALL_MAX)},
             "entries": []}
    stream = torch.cuda.Stream()
    for k in args.keys.split(","):
        for v in args.vals.split(","):
            for lg in (int(x) for x in args.sizes.split(",")):
                t0 = time.time()
                keys, vals, d_off, lengths = synthetic_batch(k, v, 1 << lg)
                out_k = torch.empty_like(keys)
                out_v = None if vals is None else torch.empty_like(vals)
                with torch.cuda.stream(stream):
                    best, ms, dms, timings = tune(k, keys, d_off, vals, out_k, out_v, stream, reps=args.reps)
                blk, ipt = geometry(k, v, best["large_variant"])
                e = {"key": k, "val": v, "size_log2": lg, "total_keys": int(keys.numel()), "nseg": int(len(lengths)),
                     "tuning": best, "geometry": {"block": blk, "items_per_thread": ipt}, "ms": round(ms, 4),
                     "default_ms": round(dms, 4), "speedup_vs_default": round(dms / ms, 4),
                     "points": len(timings)}
                table["entries"].append(e)
                print(json.dumps(e), flush=True)
                print(f"tuned {k}+{v} 2^{lg} in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
                del keys, vals, d_off, out_k, out_v
                torch.cuda.empty_cache()
    if args.append and os.path.exists(args.out):
        old = load_table(args.out)
        new_keys = {(e["key"], e["val"], e["size_log2"]) for e in table["entries"]}
        table["entries"] = [e for e in old["entries"] if (e["key"], e["val"], e["size_log2"]) not in new_keys] + \
            table["entries"]
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
