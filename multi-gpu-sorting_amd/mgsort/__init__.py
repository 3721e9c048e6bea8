"""ctypes binding of libmgsort (include/mgsort.h) for PyTorch device tensors.

Device memory and streams come from PyTorch; every computation is a HIP
kernel of libmgsort.so.  Importing this module fails loudly when the library
has not been built (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(os.path.dirname(_HERE), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libmgsort.so")
HOST_LIB_PATH = os.path.join(LIB_DIR, "libmgsort_host.so")

# mgs_key / mgs_val
KEY_U32, KEY_U64, KEY_I32, KEY_I64, KEY_F32, KEY_F64 = range(6)
VAL_NONE, VAL_U32, VAL_U64 = range(3)
KEY_NAMES = {"u32": KEY_U32, "u64": KEY_U64, "i32": KEY_I32, "i64": KEY_I64, "f32": KEY_F32, "f64": KEY_F64}
VAL_NAMES = {None: VAL_NONE, "none": VAL_NONE, "u32": VAL_U32, "u64": VAL_U64}
KEY_BYTES = {KEY_U32: 4, KEY_I32: 4, KEY_F32: 4, KEY_U64: 8, KEY_I64: 8, KEY_F64: 8}
VAL_BYTES = {VAL_NONE: 0, VAL_U32: 4, VAL_U64: 8}

MGS_OK = 0
PATH_SMALL, PATH_LSD, PATH_THREE_PASS = 1, 2, 3

SYMBOLS = [
    "mgs_sort_temp_bytes", "mgs_sort", "mgs_sort_noalloc", "mgs_sort_timed", "mgs_sort_timed_stages", "mgs_sort_path", "mgs_merge_temp_bytes", "mgs_merge", "mgs_select_pivot",
    "mgs_swap_ranges", "mgs_fill_max", "mgs_segmented_sort_temp_bytes", "mgs_segmented_sort_temp_bytes_tuned",
    "mgs_segmented_sort", "mgs_segmented_sort_ranges", "mgs_segmented_variant_count", "mgs_segmented_variant",
    "mgs_multiway_merge_temp_bytes", "mgs_multiway_merge", "mgs_multiway_merge_portions",
    "mgs_temp_status", "mgs_status_str", "mgs_last_hip_error", "mgs_version",
]

MAX_KEYS = 0xFFFF0000  # MGS_MAX_KEYS


class MgsError(RuntimeError):
    pass


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libmgsort.so not built at {LIB_PATH}; run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
    lib.mgs_sort_temp_bytes.argtypes = [i, i, sz]
    lib.mgs_sort_temp_bytes.restype = sz
    lib.mgs_sort.argtypes = [i, i, vp, vp, vp, vp, sz, vp, sz, vp]
    lib.mgs_sort.restype = i
    lib.mgs_sort_noalloc.argtypes = [i, i, vp, vp, vp, vp, sz, vp]
    lib.mgs_sort_noalloc.restype = i
    lib.mgs_sort_timed.argtypes = [i, i, vp, vp, vp, vp, sz, vp, sz, vp, ctypes.POINTER(vp), i, ctypes.POINTER(i)]
    lib.mgs_sort_timed.restype = i
    lib.mgs_sort_path.argtypes = [i, i, sz, vp, sz, vp]
    lib.mgs_sort_path.restype = i
    lib.mgs_sort_timed_stages.argtypes = []
    lib.mgs_sort_timed_stages.restype = ctypes.c_char_p
    lib.mgs_merge_temp_bytes.argtypes = [i, i, sz, sz]
    lib.mgs_merge_temp_bytes.restype = sz
    lib.mgs_merge.argtypes = [i, i, vp, vp, sz, vp, vp, sz, vp, vp, vp, sz, vp]
    lib.mgs_merge.restype = i
    lib.mgs_select_pivot.argtypes = [i, ctypes.POINTER(vp), ctypes.POINTER(vp), sz, sz, vp, vp]
    lib.mgs_select_pivot.restype = i
    lib.mgs_swap_ranges.argtypes = [i, i, vp, vp, vp, vp, sz, vp]
    lib.mgs_swap_ranges.restype = i
    lib.mgs_fill_max.argtypes = [i, vp, sz, vp]
    lib.mgs_fill_max.restype = i
    lib.mgs_segmented_sort_temp_bytes.argtypes = [i, i, sz, sz]
    lib.mgs_segmented_sort_temp_bytes.restype = sz
    lib.mgs_multiway_merge_temp_bytes.argtypes = [i, sz, sz]
    lib.mgs_multiway_merge_temp_bytes.restype = sz
    lib.mgs_multiway_merge.argtypes = [i, ctypes.POINTER(vp), ctypes.POINTER(sz), sz, vp, vp, sz, vp]
    lib.mgs_multiway_merge.restype = i
    lib.mgs_multiway_merge_portions.argtypes = [i, ctypes.POINTER(vp), ctypes.POINTER(sz), sz, vp, vp, sz, vp, sz,
                                                ctypes.POINTER(sz), ctypes.POINTER(vp)]
    lib.mgs_multiway_merge_portions.restype = i
    lib.mgs_segmented_variant_count.argtypes = []
    lib.mgs_segmented_variant_count.restype = i
    lib.mgs_segmented_variant.argtypes = [i, i, i, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    lib.mgs_segmented_variant.restype = i
    lib.mgs_segmented_sort_temp_bytes_tuned.argtypes = [i, i, sz, sz, vp]
    lib.mgs_segmented_sort_temp_bytes_tuned.restype = sz
    lib.mgs_temp_status.argtypes = [vp, sz, vp]
    lib.mgs_temp_status.restype = i
    lib.mgs_segmented_sort.argtypes = [i, i, vp, vp, vp, vp, vp, sz, sz, vp, vp, sz, vp]
    lib.mgs_segmented_sort.restype = i
    lib.mgs_segmented_sort_ranges.argtypes = [i, i, vp, vp, vp, vp, sz, sz, vp, vp, sz, vp]
    lib.mgs_segmented_sort_ranges.restype = i
    lib.mgs_status_str.argtypes = [i]
    lib.mgs_status_str.restype = ctypes.c_char_p
    lib.mgs_last_hip_error.argtypes = []
    lib.mgs_last_hip_error.restype = i
    lib.mgs_version.argtypes = []
    lib.mgs_version.restype = ctypes.c_char_p
    return lib


lib = _load()


class Tuning(ctypes.Structure):
    _fields_ = [("large_variant", ctypes.c_int32), ("small_max", ctypes.c_int32),
                ("medium_max", ctypes.c_int32), ("max_segment_keys", ctypes.c_int32)]


def _check(st: int, what: str) -> None:
    if st != MGS_OK:
        raise MgsError(f"{what}: {lib.mgs_status_str(st).decode()} (status {st}, hip {lib.mgs_last_hip_error()})")


def _ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    return None if t is None else t.data_ptr()


def _stream(stream) -> Optional[int]:
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream


def key_type(kt) -> int:
    return KEY_NAMES[kt] if isinstance(kt, str) else int(kt)


def val_type(vt) -> int:
    return VAL_NAMES[vt] if (vt is None or isinstance(vt, str)) else int(vt)


def temp_buffer(nbytes: int, device, stream=None) -> torch.Tensor:
    """Scratch for one call.  With `stream`, the buffer belongs to that
    stream in torch's caching allocator, so it is not handed to other work
    before the call's kernels on `stream` have finished with it."""
    if stream is None:
        return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
    with torch.cuda.stream(stream):
        return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def temp_status(temp: Optional[torch.Tensor], stream=None) -> int:
    """mgs_temp_status: waits for `stream` and returns the device-side status
    (MGS_OK, MGS_ERR_DEVICE, MGS_ERR_SKIPPED) of the last call that used temp."""
    if temp is None:
        return MGS_OK
    return lib.mgs_temp_status(_ptr(temp), temp.numel(), _stream(stream))


def check_temp(temp: Optional[torch.Tensor], what: str, stream=None) -> None:
    _check(temp_status(temp, stream), what)


def sort(kt, keys: torch.Tensor, vals: Optional[torch.Tensor] = None, keys_out: Optional[torch.Tensor] = None,
         vals_out: Optional[torch.Tensor] = None, temp: Optional[torch.Tensor] = None, stream=None) -> None:
    """Stable sort by enc(); in place unless keys_out/vals_out are given."""
    kt = key_type(kt)
    vt = VAL_NONE if vals is None else (VAL_U32 if vals.element_size() == 4 else VAL_U64)
    n = keys.numel()
    if keys_out is None:
        keys_out = keys
    if vals is not None and vals_out is None:
        vals_out = vals
    need = lib.mgs_sort_temp_bytes(kt, vt, n)
    if temp is None:
        temp = temp_buffer(need, keys.device, stream)
    st = lib.mgs_sort(kt, vt, _ptr(keys), _ptr(keys_out), _ptr(vals), _ptr(vals_out), n, _ptr(temp),
                      temp.numel(), _stream(stream))
    _check(st, "mgs_sort")


def merge(kt, a: torch.Tensor, b: torch.Tensor, out: torch.Tensor, av=None, bv=None, out_vals=None,
          temp=None, stream=None) -> None:
    kt = key_type(kt)
    vt = VAL_NONE if av is None else (VAL_U32 if av.element_size() == 4 else VAL_U64)
    need = lib.mgs_merge_temp_bytes(kt, vt, a.numel(), b.numel())
    if temp is None:
        temp = temp_buffer(need, out.device, stream)
    st = lib.mgs_merge(kt, vt, _ptr(a), _ptr(av), a.numel(), _ptr(b), _ptr(bv), b.numel(), _ptr(out),
                       _ptr(out_vals), _ptr(temp), temp.numel(), _stream(stream))
    _check(st, "mgs_merge")


def multiway_merge(kt, runs: Sequence[torch.Tensor], out: torch.Tensor, temp=None, stream=None,
                   portions: int = 0) -> Optional[List[int]]:
    """Stable k-way merge of sorted device runs into out (keys only).  With
    portions > 0 the output is written in that many stretches; returns their
    end indices."""
    kt = key_type(kt)
    k = len(runs)
    ptrs = (ctypes.c_void_p * max(k, 1))(*[r.data_ptr() for r in runs])
    lens = (ctypes.c_size_t * max(k, 1))(*[r.numel() for r in runs])
    total = sum(r.numel() for r in runs)
    need = lib.mgs_multiway_merge_temp_bytes(kt, k, total)
    if temp is None:
        temp = temp_buffer(need, out.device, stream)
    if portions <= 0:
        st = lib.mgs_multiway_merge(kt, ptrs, lens, k, _ptr(out), _ptr(temp), temp.numel(), _stream(stream))
        _check(st, "mgs_multiway_merge")
        return None
    ends = (ctypes.c_size_t * portions)()
    st = lib.mgs_multiway_merge_portions(kt, ptrs, lens, k, _ptr(out), _ptr(temp), temp.numel(), _stream(stream),
                                         portions, ends, None)
    _check(st, "mgs_multiway_merge_portions")
    return list(ends)


def select_pivot(kt, left: Sequence[torch.Tensor], right: Sequence[torch.Tensor], part_size: int,
                 result: torch.Tensor, stream=None) -> None:
    """Writes the pivot (uint64) into result[0] (a device or pinned host tensor)."""
    kt = key_type(kt)
    n = len(left)
    L = (ctypes.c_void_p * n)(*[t.data_ptr() for t in left])
    R = (ctypes.c_void_p * n)(*[t.data_ptr() for t in right])
    st = lib.mgs_select_pivot(kt, L, R, part_size, n, result.data_ptr(), _stream(stream))
    _check(st, "mgs_select_pivot")


def swap_ranges(kt, a: torch.Tensor, b: torch.Tensor, av=None, bv=None, stream=None) -> None:
    kt = key_type(kt)
    vt = VAL_NONE if av is None else (VAL_U32 if av.element_size() == 4 else VAL_U64)
    st = lib.mgs_swap_ranges(kt, vt, _ptr(a), _ptr(av), _ptr(b), _ptr(bv), a.numel(), _stream(stream))
    _check(st, "mgs_swap_ranges")


def fill_max(kt, keys: torch.Tensor, stream=None) -> None:
    st = lib.mgs_fill_max(key_type(kt), _ptr(keys), keys.numel(), _stream(stream))
    _check(st, "mgs_fill_max")


def segmented_sort(kt, keys: torch.Tensor, offsets: torch.Tensor, vals: Optional[torch.Tensor] = None,
                   keys_out=None, vals_out=None, max_total: Optional[int] = None, tuning: Optional[Tuning] = None,
                   temp=None, stream=None) -> None:
    """offsets: device int64/uint64 tensor of nseg+1 entries."""
    kt = key_type(kt)
    vt = VAL_NONE if vals is None else (VAL_U32 if vals.element_size() == 4 else VAL_U64)
    nseg = offsets.numel() - 1
    if max_total is None:
        max_total = keys.numel()
    if keys_out is None:
        keys_out = keys
    if vals is not None and vals_out is None:
        vals_out = vals
    tp = ctypes.byref(tuning) if tuning is not None else None
    need = lib.mgs_segmented_sort_temp_bytes_tuned(kt, vt, max_total, nseg, tp)
    if temp is None:
        temp = temp_buffer(need, keys.device, stream)
    st = lib.mgs_segmented_sort(kt, vt, _ptr(keys), _ptr(keys_out), _ptr(vals), _ptr(vals_out), _ptr(offsets),
                                nseg, max_total, tp, _ptr(temp), temp.numel(), _stream(stream))
    _check(st, "mgs_segmented_sort")


def segmented_sort_ranges(kt, keys: torch.Tensor, ranges: torch.Tensor, vals: Optional[torch.Tensor] = None,
                          val_begin: Optional[torch.Tensor] = None, max_total: Optional[int] = None,
                          tuning: Optional[Tuning] = None, temp=None, stream=None) -> None:
    """In-place sort of segments keys[ranges[2s]:ranges[2s+1]] (device int64
    tensor of 2*nseg bounds; segments disjoint, anywhere in keys); payloads
    at vals[val_begin[s]:...] (default: the keys' positions)."""
    kt = key_type(kt)
    vt = VAL_NONE if vals is None else (VAL_U32 if vals.element_size() == 4 else VAL_U64)
    nseg = ranges.numel() // 2
    if max_total is None:
        max_total = keys.numel()
    tp = ctypes.byref(tuning) if tuning is not None else None
    need = lib.mgs_segmented_sort_temp_bytes_tuned(kt, vt, max_total, nseg, tp)
    if temp is None:
        temp = temp_buffer(need, keys.device, stream)
    st = lib.mgs_segmented_sort_ranges(kt, vt, _ptr(keys), _ptr(vals), _ptr(ranges), _ptr(val_begin), nseg, max_total, tp,
                                       _ptr(temp), temp.numel(), _stream(stream))
    _check(st, "mgs_segmented_sort_ranges")
