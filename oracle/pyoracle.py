"""ctypes binding of the CPU oracle (oracle/liboracle.so) — test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, as the checker.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

K_U32, K_U64, K_I32, K_I64, K_F32, K_F64 = range(6)
V_NONE, V_U32, V_U64 = range(3)
KEY_DTYPE = {K_U32: np.uint32, K_U64: np.uint64, K_I32: np.int32, K_I64: np.int64, K_F32: np.float32,
             K_F64: np.float64}
BITS_DTYPE = {K_U32: np.uint32, K_I32: np.uint32, K_F32: np.uint32, K_U64: np.uint64, K_I64: np.uint64,
              K_F64: np.uint64}

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, i = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        L.oracle_sort.argtypes = [i, i, vp, vp, sz]
        L.oracle_sort.restype = i
        L.oracle_encode.argtypes = [i, vp, sz, vp]
        L.oracle_encode.restype = None
        L.oracle_merge.argtypes = [i, i, vp, vp, sz, vp, vp, sz, vp, vp]
        L.oracle_merge.restype = None
        L.oracle_select_pivot.argtypes = [i, ctypes.POINTER(vp), ctypes.POINTER(vp), sz, sz]
        L.oracle_select_pivot.restype = ctypes.c_uint64
        L.oracle_tanasic_1dev.argtypes = [i, i, vp, vp, sz]
        L.oracle_tanasic_1dev.restype = ctypes.c_uint64
        L.oracle_segmented_sort.argtypes = [i, i, vp, vp, vp, sz]
        L.oracle_segmented_sort.restype = i
        if hasattr(L, "oracle_generate"):
            L.oracle_generate.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint32, vp, sz, i]
            L.oracle_generate.restype = i
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def val_type(vals) -> int:
    if vals is None:
        return V_NONE
    return V_U32 if vals.dtype.itemsize == 4 else V_U64


def encode(kt: int, keys: np.ndarray) -> np.ndarray:
    keys = np.ascontiguousarray(keys)
    out = np.empty(keys.size, dtype=np.uint64)
    lib().oracle_encode(kt, _p(keys), keys.size, _p(out))
    return out


def sort(kt: int, keys: np.ndarray, vals: np.ndarray | None = None):
    """Returns sorted copies (keys, vals)."""
    k = np.array(keys, copy=True, order="C")
    v = None if vals is None else np.array(vals, copy=True, order="C")
    if lib().oracle_sort(kt, val_type(v), _p(k), _p(v), k.size) != 0:
        raise MemoryError("oracle_sort")
    return k, v


def merge(kt: int, a, b, av=None, bv=None):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    out = np.empty(a.size + b.size, dtype=a.dtype)
    ov = None if av is None else np.empty(a.size + b.size, dtype=av.dtype)
    lib().oracle_merge(kt, val_type(av), _p(a), _p(av), a.size, _p(b), _p(bv), b.size, _p(out), _p(ov))
    return out, ov


def select_pivot(kt: int, left, right, part_size: int) -> int:
    left = [np.ascontiguousarray(x) for x in left]
    right = [np.ascontiguousarray(x) for x in right]
    n = len(left)
    L = (ctypes.c_void_p * n)(*[x.ctypes.data for x in left])
    R = (ctypes.c_void_p * n)(*[x.ctypes.data for x in right])
    return int(lib().oracle_select_pivot(kt, L, R, part_size, n))


def tanasic_1dev(kt: int, keys, vals=None):
    k = np.array(keys, copy=True, order="C")
    v = None if vals is None else np.array(vals, copy=True, order="C")
    p = lib().oracle_tanasic_1dev(kt, val_type(v), _p(k), _p(v), k.size // 2)
    return k, v, int(p)


def segmented_sort(kt: int, keys, offsets, vals=None):
    k = np.array(keys, copy=True, order="C")
    v = None if vals is None else np.array(vals, copy=True, order="C")
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    if lib().oracle_segmented_sort(kt, val_type(v), _p(k), _p(v), _p(off), off.size - 1) != 0:
        raise MemoryError("oracle_segmented_sort")
    return k, v
