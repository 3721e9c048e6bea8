"""The segmented sort's autotuner without a GPU: the search space the library
exposes, scratch sizing across it, and the persisted table's lookup."""
import json
import os

import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def tn():
    from mgsort import tuning
    return tuning


KEYS = ["u32", "i32", "f32", "u64", "i64", "f64"]
VALS = ["none", "u32", "u64"]


def test_variant_space(tn):
    assert tn.variant_count() == 3
    pts = tn.candidates()
    assert len(pts) == 3 * len(tn.MEDIUM_MAX) * len(tn.SMALL_MAX)
    assert {"large_variant": 0, "small_max": 0, "medium_max": 0} in pts
    assert len({json.dumps(p, sort_keys=True) for p in pts}) == len(pts)


@pytest.mark.parametrize("k", KEYS)
@pytest.mark.parametrize("v", VALS)
def test_geometries(tn, k, v):
    """Each variant is a distinct (threads, items) pair whose tile fits the
    LDS of one CU (160 KiB) with room for the counters."""
    import mgsort
    rec = mgsort.KEY_BYTES[mgsort.KEY_NAMES[k]] + mgsort.VAL_BYTES[mgsort.VAL_NAMES[v]]
    geoms = [tn.geometry(k, v, i) for i in range(tn.variant_count())]
    assert len(set(geoms)) == len(geoms)
    for b, i in geoms:
        assert b in (256, 512, 1024) and 1 <= i <= 32
        assert b * i * rec <= 128 * 1024
    assert geoms[0] == {4: (512, 16), 8: (512, 20), 12: (512, 20), 16: (512, 14)}[rec]


def test_variant_out_of_range_is_an_error(tn):
    import mgsort
    with pytest.raises(ValueError):
        tn.geometry("u32", "none", 3)
    with pytest.raises(ValueError):
        tn.geometry("u32", "none", -1)
    # the sort rejects it before touching the device
    t = tn.as_tuning({"large_variant": 7})
    st = mgsort.lib.mgs_segmented_sort(0, 0, 16, 16, None, None, 16, 1, 100, __import__("ctypes").byref(t), 16, 1 << 30, None)
    assert st == -4  # MGS_ERR_INVALID


def test_temp_bytes_cover_every_point(tn):
    """Smaller tiles mean more look-back descriptors: the scratch for the
    search grows with the smallest tile, and temp_bytes_for is the max."""
    import ctypes
    import mgsort
    total, nseg = 1 << 30, 4096
    per = []
    for var in range(tn.variant_count()):
        t = tn.as_tuning({"large_variant": var})
        per.append(mgsort.lib.mgs_segmented_sort_temp_bytes_tuned(mgsort.KEY_U64, mgsort.VAL_U32, total, nseg,
                                                                 ctypes.byref(t)))
    assert per[0] == mgsort.lib.mgs_segmented_sort_temp_bytes(mgsort.KEY_U64, mgsort.VAL_U32, total, nseg)
    assert tn.temp_bytes_for("u64", "u32", total, nseg, tn.candidates()) == max(per)
    # variant 2 (512 x 12) has the smallest tile of u64+u32, so the most descriptors
    assert per[2] == max(per) and per[2] > per[0]


def test_lookup_nearest_bin(tn):
    table = {"format": 1, "entries": [
        {"key": "u64", "val": "u32", "size_log2": 24, "tuning": {"large_variant": 2, "small_max": 0, "medium_max": 0}},
        {"key": "u64", "val": "u32", "size_log2": 30, "tuning": {"large_variant": 1, "small_max": 1024, "medium_max": 65536}},
        {"key": "u32", "val": "none", "size_log2": 27, "tuning": {"large_variant": 0, "small_max": 0, "medium_max": 32768}},
    ]}
    t, e = tn.lookup("u64", "u32", 1_758_065_964, table)  # 2^30.7 -> bin 31 -> nearest 30
    assert (t.large_variant, t.small_max, t.medium_max) == (1, 1024, 65536) and e["size_log2"] == 30
    t, e = tn.lookup("u64", "u32", 1 << 26, table)  # bin 26: 24 is nearer than 30
    assert t.large_variant == 2
    t, e = tn.lookup("u64", "u32", 1 << 27, table)  # tie 24 / 30: the smaller bin
    assert e["size_log2"] == 24
    assert tn.lookup("u64", "u64", 1 << 27, table) is None
    t, _ = tn.lookup("u32", None, 1 << 20, table)
    assert t.medium_max == 32768


def test_committed_table_is_well_formed(tn):
    """The table measured on an MI355X that ships with the package: one entry
    per (key, payload, size bin), each naming a point of the search space."""
    path = tn.TABLE_PATH
    if not os.path.exists(path):
        pytest.skip("no committed tuning table")
    t = tn.load_table(path)
    assert t["format"] == 1 and t["entries"]
    seen = set()
    space = {json.dumps(p, sort_keys=True) for p in tn.candidates()}
    for e in t["entries"]:
        k = (e["key"], e["val"], e["size_log2"])
        assert k not in seen
        seen.add(k)
        assert json.dumps({x: e["tuning"][x] for x in ("large_variant", "medium_max", "small_max")},
                          sort_keys=True) in space
        assert e["ms"] <= e["default_ms"] * 1.0001
        assert tuple(e["geometry"].values()) == tn.geometry(e["key"], e["val"], e["tuning"]["large_variant"])


def test_host_table_loader_agrees(tn, tmp_path):
    """The C++ loader of the serving layer (mgh_seg_tuning_lookup) picks the
    same entry as the Python one, for every pair and size bin, and rejects a
    malformed file."""
    import ctypes
    import mgsort
    H = ctypes.CDLL(os.path.join(ROOT, "multi-gpu-sorting_amd", "lib", "libmgsort_host.so"))
    H.mgh_seg_tuning_lookup.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p]
    table = {"format": 1, "entries": [
        {"key": "u64", "val": "u32", "size_log2": 24, "tuning": {"large_variant": 2, "small_max": 0, "medium_max": 0}},
        {"key": "u64", "val": "u32", "size_log2": 30, "tuning": {"large_variant": 1, "small_max": 1024, "medium_max": 65536}},
        {"key": "f32", "val": "none", "size_log2": 27, "tuning": {"large_variant": 0, "small_max": 7, "medium_max": 32768}},
    ]}
    p = tmp_path / "t.json"
    p.write_text(json.dumps(table, indent=1))
    for k in KEYS:
        for v in VALS:
            for lg in (10, 24, 26, 27, 28, 31):
                out = mgsort.Tuning(-9, -9, -9, -9)
                rc = H.mgh_seg_tuning_lookup(str(p).encode(), mgsort.KEY_NAMES[k], mgsort.VAL_NAMES[v], 1 << lg,
                                             ctypes.byref(out))
                want = tn.lookup(k, v, 1 << lg, table)
                if want is None:
                    assert rc == 1 and out.large_variant == -9
                else:
                    t, _ = want
                    assert rc == 0
                    assert (out.large_variant, out.small_max, out.medium_max, out.max_segment_keys) == \
                        (t.large_variant, t.small_max, t.medium_max, t.max_segment_keys)
    bad = tmp_path / "bad.json"
    bad.write_text('{"format": 1, "entries": [{"key": "u64", "val": "u32", "size_log2": 24, "tuning": {"large_variant": 9,'
                   ' "small_max": 0, "medium_max": 0}}]}')
    out = mgsort.Tuning()
    assert H.mgh_seg_tuning_lookup(str(bad).encode(), 1, 1, 1 << 24, ctypes.byref(out)) == -4
    bad.write_text('{"format": 1, "entries": [')
    assert H.mgh_seg_tuning_lookup(str(bad).encode(), 1, 1, 1 << 24, ctypes.byref(out)) == -4
    assert H.mgh_seg_tuning_lookup(str(tmp_path / "absent.json").encode(), 1, 1, 1 << 24, ctypes.byref(out)) == -4
    if os.path.exists(tn.TABLE_PATH):
        out = mgsort.Tuning()
        assert H.mgh_seg_tuning_lookup(tn.TABLE_PATH.encode(), 1, 1, 1_758_065_964, ctypes.byref(out)) == 0
        t, _ = tn.lookup("u64", "u32", 1_758_065_964)
        assert (out.large_variant, out.small_max, out.medium_max) == (t.large_variant, t.small_max, t.medium_max)
