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
