"""The C ABI without a GPU: every function declared in include/*.h is
exported by the built libraries, argument checks reject bad calls before any
device work, and the host library's plan / generator entry points answer."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

from conftest import ROOT

LIB = os.path.join(ROOT, "multi-gpu-sorting_amd", "lib", "libmgsort.so")
HOST_LIB = os.path.join(ROOT, "multi-gpu-sorting_amd", "lib", "libmgsort_host.so")


def declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(mg[sh]_[a-z0-9_]+)\s*\(", text)))


def test_headers_declare_entry_points():
    names = [n for h in glob.glob(os.path.join(ROOT, "include", "*.h")) for n in declared(h)]
    assert "mgs_sort" in names and "mgs_segmented_sort" in names and "mgs_merge" in names


@pytest.mark.parametrize("header", sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))))
def test_library_exports_every_declared_symbol(header):
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared(header) if not hasattr(lib, n)]
    assert not missing, missing


def test_host_library_exports_its_api():
    lib = ctypes.CDLL(HOST_LIB)
    hdr = os.path.join(ROOT, "multi-gpu-sorting_amd", "host", "host_api.h")
    missing = [n for n in declared(hdr) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_symbol_list_matches_header():
    import mgsort
    hdr = declared(os.path.join(ROOT, "include", "mgsort.h"))
    assert sorted(mgsort.SYMBOLS) == hdr


def test_status_strings_and_version():
    import mgsort
    L = mgsort.lib
    assert L.mgs_status_str(0) == b"ok"
    for st in (-1, -2, -3, -4, -5, -6, -7):
        assert L.mgs_status_str(st) and L.mgs_status_str(st) != b"ok"
    assert L.mgs_version()


def test_temp_bytes_monotone_and_typed():
    import mgsort
    L = mgsort.lib
    a = L.mgs_sort_temp_bytes(mgsort.KEY_U32, mgsort.VAL_NONE, 1 << 20)
    b = L.mgs_sort_temp_bytes(mgsort.KEY_U32, mgsort.VAL_NONE, 1 << 24)
    c = L.mgs_sort_temp_bytes(mgsort.KEY_U64, mgsort.VAL_U64, 1 << 24)
    assert 0 < a < b < c
    # the out-of-place sort needs at least one alternate key buffer
    assert b >= 4 << 24


PAIRS = [(k, v) for k in range(6) for v in range(3)]


@pytest.mark.parametrize("kt,vt", PAIRS)
def test_sort_temp_bytes_never_decrease(kt, vt):
    """A buffer sized for n serves every smaller sort: the scratch query does
    not decrease across the small-sort limit, the switch to large tiles at
    2^21 keys for records wider than 4 bytes, the size ranges of the
    three-pass path (2^23, 3 * 2^24, 2^29, 2^30 + 2^27) and up to
    MGS_MAX_KEYS."""
    import mgsort
    L = mgsort.lib
    ns = sorted({1, 2047, 2048, 4095, 4096, 4097, 8191, 8192, 8193, (1 << 21) - 10000, (1 << 21) - 1, 1 << 21,
                 (1 << 21) + 1, (1 << 21) + 10000, (1 << 23) - 1, 1 << 23, 1 << 24, (3 << 24) - 1, 3 << 24,
                 (1 << 26) - 1, 1 << 26, (1 << 29) - 1, 1 << 29, 1 << 30, (1 << 30) + 1,
                 (1 << 30) + (1 << 27), (1 << 30) + (1 << 27) + 1, mgsort.MAX_KEYS - 1, mgsort.MAX_KEYS})
    sizes = [L.mgs_sort_temp_bytes(kt, vt, n) for n in ns]
    assert all(b >= a for a, b in zip(sizes, sizes[1:])), list(zip(ns, sizes))
    # the alternate buffers are in the scratch of every sort that needs scratch
    kb, vb = mgsort.KEY_BYTES[kt], mgsort.VAL_BYTES[vt]
    for n, b in zip(ns, sizes):
        assert b == 0 or b >= n * (kb + vb)


@pytest.mark.parametrize("kt,vt", PAIRS)
def test_sort_scratch_fits_the_reference_buffer_plan(kt, vt):
    """mgs_sort's scratch stays within the reference's device buffer plan,
    N * sizeof(T) + 128,000,000 bytes (src/device_buffers.cuh:30-44), for
    every key and payload type up to MGS_MAX_KEYS, including the reference's
    HET chunk of 4.125e9 keys (scripts/settings.py:614): above the
    descriptor budget the LSD passes run in rounds."""
    import mgsort
    L = mgsort.lib
    kb, vb = mgsort.KEY_BYTES[kt], mgsort.VAL_BYTES[vt]
    for n in (1 << 24, 1 << 30, (1 << 30) + 1, 1 << 31, 4_125_000_000, mgsort.MAX_KEYS):
        extra = L.mgs_sort_temp_bytes(kt, vt, n) - n * (kb + vb)
        assert extra <= 128_000_000, (n, extra)


def test_too_many_keys_is_rejected_before_any_device_work():
    """n above MGS_MAX_KEYS (the last tile must end below 2^32: the kernels
    index keys with 32 bits) returns MGS_ERR_TOO_LARGE before touching the
    (bogus) pointers; MGS_MAX_KEYS itself passes the size check and fails on
    the NULL pointer."""
    import mgsort
    L = mgsort.lib
    bogus = ctypes.c_void_p(0x1000)
    for n in (mgsort.MAX_KEYS + 1, (1 << 32) - 1, 1 << 32, 1 << 40):
        assert L.mgs_sort(0, 0, bogus, bogus, None, None, n, bogus, 1 << 40, None) == -5
        assert L.mgs_segmented_sort(0, 0, bogus, bogus, None, None, bogus, 1, n, None, bogus, 1 << 40, None) == -5
    assert L.mgs_sort(0, 0, None, None, None, None, mgsort.MAX_KEYS, None, 0, None) == -4


def test_temp_status_without_scratch_is_ok():
    import mgsort
    assert mgsort.lib.mgs_temp_status(None, 0, None) == 0


def test_argument_checks_need_no_device():
    """Invalid calls are rejected on the host, before any HIP call."""
    import mgsort
    L = mgsort.lib
    assert L.mgs_sort(17, 0, None, None, None, None, 10, None, 0, None) == -1          # bad key type
    assert L.mgs_sort(0, 9, None, None, None, None, 10, None, 0, None) == -1           # bad value type
    assert L.mgs_sort(0, 0, None, None, None, None, 10, None, 0, None) == -4           # NULL keys
    assert L.mgs_sort(0, 0, None, None, None, None, 0, None, 0, None) == 0             # empty: nothing to do
    assert L.mgs_merge(0, 0, None, None, 0, None, None, 0, None, None, None, 0, None) == 0
    assert L.mgs_fill_max(0, None, 5, None) == -4


def test_gowanlock_plan():
    """PlanGowanlock: chunk and merge-group arithmetic of the heterogeneous
    sort (reference src/gowanlock_sort.cu:280-311)."""
    H = ctypes.CDLL(HOST_LIB)
    out = (ctypes.c_size_t * 6)()
    H.mgh_gowanlock_plan.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t * 6]
    for n, chunk, group in [(1000, 100, 3), (1000, 300, 2), (1 << 20, 1 << 17, 4), (10, 100, 2), (1000, 250, 4)]:
        assert H.mgh_gowanlock_plan(n, chunk, group, out) == 0
        cs, gs, nchunks, groups, chunks_merged, runs = list(out)
        assert cs == chunk
        assert nchunks == -(-n // cs)
        assert chunks_merged == groups * gs
        assert chunks_merged <= nchunks
        # every chunk is in exactly one run of the final multiway merge
        assert runs == groups + (nchunks - chunks_merged)


@pytest.mark.parametrize("dtype,np_t", [("u32", np.uint32), ("i64", np.int64), ("f32", np.float32)])
def test_cpu_baselines_sort(dtype, np_t):
    """PARADIS and the GNU parallel sort produce the oracle's order."""
    import pyoracle as O
    kt = {"u32": O.K_U32, "i64": O.K_I64, "f32": O.K_F32}[dtype]
    H = ctypes.CDLL(HOST_LIB)
    vp = ctypes.c_void_p
    H.mgh_generate.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_uint32, vp, ctypes.c_size_t]
    H.mgh_paradis_sort.argtypes = [ctypes.c_char_p, vp, ctypes.c_size_t, ctypes.c_int]
    H.mgh_gnu_parallel_sort.argtypes = [ctypes.c_char_p, vp, ctypes.c_size_t, ctypes.c_int]
    n = 300_001
    keys = np.empty(n, dtype=np_t)
    assert H.mgh_generate(dtype.encode(), b"uniform", 12345, keys.ctypes.data, n) == 0
    exp, _ = O.sort(kt, keys.view(O.BITS_DTYPE[kt]))
    for fn in (H.mgh_paradis_sort, H.mgh_gnu_parallel_sort):
        for threads in (1, 4):
            k = keys.copy()
            assert fn(dtype.encode(), k.ctypes.data, n, threads) == 0
            # equal keys are indistinguishable in a keys-only sort, so compare encodings
            assert np.array_equal(O.encode(kt, k.view(O.BITS_DTYPE[kt])), O.encode(kt, exp))


def test_multiway_merge_argument_checks():
    """Rejected before any device work: unknown type, too many runs, NULL run
    with keys, NULL output; no runs at all is a no-op."""
    lib = ctypes.CDLL(LIB)
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    lib.mgs_multiway_merge.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(sz), sz, vp, vp, sz, vp]
    lib.mgs_multiway_merge_temp_bytes.argtypes = [ctypes.c_int, sz, sz]
    lib.mgs_multiway_merge_temp_bytes.restype = sz
    runs = (vp * 2)(16, None)
    lens = (sz * 2)(10, 5)
    assert lib.mgs_multiway_merge(9, runs, lens, 2, 16, 16, 1 << 20, None) == -1
    assert lib.mgs_multiway_merge(0, runs, lens, 257, 16, 16, 1 << 20, None) == -4
    assert lib.mgs_multiway_merge(0, runs, lens, 2, 16, 16, 1 << 20, None) == -4   # run 1 NULL with 5 keys
    lens2 = (sz * 2)(10, 0)
    assert lib.mgs_multiway_merge(0, runs, lens2, 2, None, 16, 1 << 20, None) == -4  # no output
    assert lib.mgs_multiway_merge(0, None, None, 0, None, None, 0, None) == 0
    # scratch grows with the keys and the run count, and 0 for too many runs
    a = lib.mgs_multiway_merge_temp_bytes(0, 8, 1 << 30)
    assert a > lib.mgs_multiway_merge_temp_bytes(0, 8, 1 << 20) > 0
    assert lib.mgs_multiway_merge_temp_bytes(0, 257, 1 << 20) == 0
    assert a < (1 << 30) * 4 // 8  # well below the keys themselves
