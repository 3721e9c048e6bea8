"""GPU parity of the batched segmented sort against per-segment oracle sorts."""
import numpy as np
import pytest

import pyoracle as O
from helpers import KEY_NAME, from_dev, random_keys, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mg():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mgsort
    return mgsort


def make_offsets(lengths, base=0):
    off = np.zeros(len(lengths) + 1, dtype=np.uint64)
    off[1:] = np.cumsum(lengths)
    return off + np.uint64(base)


@pytest.mark.parametrize("kt", [O.K_U32, O.K_U64, O.K_F32], ids=lambda k: KEY_NAME[k])
@pytest.mark.parametrize("vdt", [None, np.uint32], ids=["keys", "pairs"])
@pytest.mark.parametrize("case", ["tiny", "mixed", "loguniform"])
def test_segmented(mg, kt, vdt, case):
    import torch
    rng = np.random.default_rng(hash(case) % 1000)
    if case == "tiny":
        lengths = rng.integers(0, 40, size=300)
    elif case == "mixed":
        lengths = np.array([0, 1, 2, 4096, 4097, 8192, 8193, 100_000, 131_072, 131_073, 300_001, 5, 0, 2048, 2049])
    else:
        lengths = np.round(2.0 ** (8 + 12 * rng.random(200))).astype(np.int64)
    total = int(lengths.sum())
    keys = random_keys(kt, total, rng, "few" if case == "mixed" else "uniform")
    vals = None if vdt is None else np.arange(total, dtype=vdt)
    off = make_offsets(lengths)
    ek, ev = O.segmented_sort(kt, keys, off, vals)
    dk = to_dev(keys)
    dv = to_dev(vals)
    doff = torch.from_numpy(off.view(np.int64)).cuda()
    mg.segmented_sort(KEY_NAME[kt], dk, doff, dv)
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(dk, keys.dtype), ek)
    if vdt is not None:
        assert np.array_equal(from_dev(dv, vdt), ev)


@pytest.mark.parametrize("kt,vdt", [(O.K_U32, None), (O.K_I32, np.uint32), (O.K_U64, None), (O.K_U64, np.uint32),
                                    (O.K_F64, np.uint64)], ids=["u32", "i32+v32", "u64", "u64+v32", "f64+v64"])
@pytest.mark.parametrize("variant", [0, 1, 2])
def test_segmented_large_variants(mg, kt, vdt, variant):
    """Every tile geometry of the large segments' passes (the autotuner's
    search space) gives the oracle's result; segments from one key to a few
    tiles of the largest geometry, all of them large (medium_max = 1)."""
    import torch
    rng = np.random.default_rng(100 + variant + 7 * kt)
    lengths = np.concatenate([[0, 1, 257, 4095, 4096, 4097, 18431, 18432, 18433],
                              np.round(2.0 ** (9 + 8 * rng.random(40))).astype(np.int64)])
    total = int(lengths.sum())
    keys = random_keys(kt, total, rng, "few" if variant == 1 else "uniform")
    vals = None if vdt is None else rng.integers(0, 2**31, size=total).astype(vdt)
    off = make_offsets(lengths)
    ek, ev = O.segmented_sort(kt, keys, off, vals)
    dk, dv = to_dev(keys), to_dev(vals)
    doff = torch.from_numpy(off.view(np.int64)).cuda()
    t = mg.Tuning(variant, 1, 1, 0)
    mg.segmented_sort(KEY_NAME[kt], dk, doff, dv, tuning=t)
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(dk, keys.dtype), ek)
    if vdt is not None:
        assert np.array_equal(from_dev(dv, vdt), ev)


@pytest.mark.parametrize("small_max,medium_max", [(1, 1), (64, 1 << 14), (4096, 1 << 20)])
def test_segmented_bins_agree(mg, small_max, medium_max):
    """Every bin assignment gives the same result (routing is invisible)."""
    import torch
    rng = np.random.default_rng(11)
    lengths = np.round(2.0 ** (4 + 14 * rng.random(120))).astype(np.int64)
    total = int(lengths.sum())
    keys = random_keys(O.K_U64, total, rng)
    vals = np.arange(total, dtype=np.uint32)
    off = make_offsets(lengths)
    ek, ev = O.segmented_sort(O.K_U64, keys, off, vals)
    dk, dv = to_dev(keys), to_dev(vals)
    doff = torch.from_numpy(off.view(np.int64)).cuda()
    t = mg.Tuning(0, small_max, medium_max, 0)
    mg.segmented_sort("u64", dk, doff, dv, tuning=t)
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(dk, np.uint64), ek)
    assert np.array_equal(from_dev(dv, np.uint32), ev)


def test_segmented_out_of_place_with_base_offset(mg):
    import torch
    rng = np.random.default_rng(12)
    lengths = np.array([10, 5000, 20000, 200_000, 3])
    base = 1000
    total = int(lengths.sum())
    buf = random_keys(O.K_U32, base + total + 50, rng)
    off = make_offsets(lengths, base)
    ek, _ = O.segmented_sort(O.K_U32, buf, off)
    din = to_dev(buf)
    dout = to_dev(buf.copy())
    doff = torch.from_numpy(off.view(np.int64)).cuda()
    mg.segmented_sort("u32", din, doff, keys_out=dout, max_total=total)
    torch.cuda.synchronize()
    got = from_dev(dout, np.uint32)
    assert np.array_equal(got[base:base + total], ek[base:base + total])
    assert np.array_equal(from_dev(din, np.uint32), buf)  # input untouched


@pytest.mark.parametrize("kt", [O.K_U32, O.K_U64], ids=lambda k: KEY_NAME[k])
@pytest.mark.parametrize("vdt", [None, np.uint32], ids=["keys", "pairs"])
def test_segmented_ranges_scattered(mg, kt, vdt):
    """Ranges form: disjoint segments at arbitrary, unordered places of one
    array, sorted in place; the gaps between them stay untouched."""
    import torch
    rng = np.random.default_rng(21)
    lengths = np.concatenate([[0, 1, 7, 2048, 2049, 70_000, 140_000, 300_000],
                              np.round(2.0 ** (6 + 10 * rng.random(60))).astype(np.int64)])
    gaps = rng.integers(0, 3000, size=len(lengths))
    starts, pos = [], 0
    for ln, g in zip(lengths, gaps):
        pos += int(g)
        starts.append(pos)
        pos += int(ln)
    n = pos + 100
    buf = random_keys(kt, n, rng)
    vals = None if vdt is None else np.arange(n, dtype=vdt)
    order = rng.permutation(len(lengths))  # segments listed in shuffled order
    ranges = np.array([[starts[i], starts[i] + lengths[i]] for i in order], dtype=np.uint64).reshape(-1)
    ek, ev = buf.copy(), None if vals is None else vals.copy()
    for i in range(len(lengths)):
        a, b = starts[i], starts[i] + int(lengths[i])
        k, v = O.sort(kt, buf[a:b], None if vals is None else vals[a:b])
        ek[a:b] = k
        if vals is not None:
            ev[a:b] = v
    dk, dv = to_dev(buf), to_dev(vals)
    dr = torch.from_numpy(ranges.view(np.int64)).cuda()
    mg.segmented_sort_ranges(KEY_NAME[kt], dk, dr, dv, max_total=int(lengths.sum()))
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(dk, buf.dtype), ek)
    if vals is not None:
        assert np.array_equal(from_dev(dv, vdt), ev)


def test_segmented_malformed_offsets_are_skipped(mg):
    """A decreasing offset leaves the segments it would make overlap
    unsorted, and no segment reaches past max_total; the rest sort."""
    import torch
    rng = np.random.default_rng(22)
    n = 60_000
    buf = random_keys(O.K_U32, n, rng)
    # segments: [0,10000) ok, [10000,5000) bad, [5000,30000) starts before an
    # earlier offset -> skipped, [30000,50000) ok, [50000,60000) beyond max_total
    off = np.array([0, 10_000, 5_000, 30_000, 50_000, 60_000], dtype=np.uint64)
    din = to_dev(buf)
    temp = mg.temp_buffer(mg.lib.mgs_segmented_sort_temp_bytes(mg.KEY_U32, mg.VAL_NONE, 50_000, 5), "cuda")
    mg.segmented_sort("u32", din, torch.from_numpy(off.view(np.int64)).cuda(), max_total=50_000, temp=temp)
    # the call reports the skipped segments in its status word
    assert mg.temp_status(temp) == -7
    got = from_dev(din, np.uint32)
    exp = buf.copy()
    for a, b in [(0, 10_000), (30_000, 50_000)]:
        exp[a:b] = O.sort(O.K_U32, buf[a:b])[0]
    assert np.array_equal(got, exp)


def test_segmented_ranges_separate_payload_places(mg):
    """Payloads of a segment may sit elsewhere than its keys (val_begin)."""
    import torch
    rng = np.random.default_rng(23)
    lengths = np.array([5, 3000, 150_000, 0, 70_000, 1])
    n = int(lengths.sum())
    kstart = np.concatenate([[0], np.cumsum(lengths)[:-1]])
    vperm = rng.permutation(len(lengths))  # payload blocks stored in another order
    vstart = np.zeros(len(lengths), dtype=np.int64)
    pos = 0
    for i in vperm:
        vstart[i] = pos + 17
        pos += int(lengths[i]) + 17
    keys = random_keys(O.K_U64, n, rng, "few")
    vals = np.zeros(pos + 10, dtype=np.uint32)
    for i in range(len(lengths)):
        vals[vstart[i]:vstart[i] + lengths[i]] = np.arange(lengths[i], dtype=np.uint32) + 1000 * i
    ek, ev = keys.copy(), vals.copy()
    for i in range(len(lengths)):
        a, b = kstart[i], kstart[i] + lengths[i]
        k, v = O.sort(O.K_U64, keys[a:b], vals[vstart[i]:vstart[i] + lengths[i]])
        ek[a:b] = k
        ev[vstart[i]:vstart[i] + lengths[i]] = v
    ranges = np.stack([kstart, kstart + lengths], axis=1).reshape(-1).astype(np.int64)
    dk, dv = to_dev(keys), to_dev(vals)
    mg.segmented_sort_ranges("u64", dk, torch.from_numpy(ranges).cuda(), dv,
                             val_begin=torch.from_numpy(vstart.astype(np.int64)).cuda(), max_total=n)
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(dk, np.uint64), ek)
    assert np.array_equal(from_dev(dv, np.uint32), ev)


def test_segmented_max_segment_hint(mg):
    """mgs_tuning.max_segment_keys: segments up to the hint sort (with the
    smaller scratch of 4-byte descriptors); a longer one is left unsorted and
    reported as skipped; a well-formed call reports ok."""
    import ctypes
    import torch
    rng = np.random.default_rng(24)
    lengths = np.array([1000, 300_000, 50_000, 0, 7])
    n = int(lengths.sum())
    keys = random_keys(O.K_U64, n, rng)
    vals = np.arange(n, dtype=np.uint32)
    off = make_offsets(lengths)
    doff = torch.from_numpy(off.view(np.int64)).cuda()
    for hint, skipped in ((300_000, []), (100_000, [1])):
        tn = mg.Tuning(0, 0, 0, hint)
        need = mg.lib.mgs_segmented_sort_temp_bytes_tuned(mg.KEY_U64, mg.VAL_U32, n, len(lengths), ctypes.byref(tn))
        assert need <= mg.lib.mgs_segmented_sort_temp_bytes(mg.KEY_U64, mg.VAL_U32, n, len(lengths))
        temp = mg.temp_buffer(need, "cuda")
        dk, dv = to_dev(keys), to_dev(vals)
        mg.segmented_sort("u64", dk, doff, dv, tuning=tn, temp=temp)
        assert mg.temp_status(temp) == (-7 if skipped else 0)
        ek, ev = keys.copy(), vals.copy()
        for s in range(len(lengths)):
            a, b = int(off[s]), int(off[s + 1])
            if s not in skipped:
                ek[a:b], ev[a:b] = O.sort(O.K_U64, keys[a:b], vals[a:b])
        assert np.array_equal(from_dev(dk, np.uint64), ek), hint
        assert np.array_equal(from_dev(dv, np.uint32), ev), hint
