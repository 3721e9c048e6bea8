"""GPU parity of merge path, Tanasic pivot, range swap, filler and the
one-device Tanasic merge (pivot + swap + two merges) against the oracle.

Reference: SelectPivot/FindPivot src/tanasic_sort.cuh:21-87, SwapPartitions
:89-146, MergeLocalPartitions :148-174 (thrust::merge :163-167), fillers
:255-260.
"""
import numpy as np
import pytest

import pyoracle as O
from helpers import KEY_NAME, from_dev, random_keys, random_vals, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mg():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mgsort
    return mgsort


def sorted_keys(kt, n, rng, dist="uniform"):
    k = random_keys(kt, n, rng, dist)
    return O.sort(kt, k)[0]


@pytest.mark.parametrize("kt", [O.K_U32, O.K_U64, O.K_I32, O.K_F32, O.K_F64], ids=lambda k: KEY_NAME[k])
@pytest.mark.parametrize("na,nb", [(0, 5), (5, 0), (1, 1), (100, 3), (3839, 3841), (100_000, 77_777),
                                   (1 << 20, 1 << 20)])
@pytest.mark.parametrize("dist", ["uniform", "few"])
def test_merge(mg, kt, na, nb, dist):
    import torch
    rng = np.random.default_rng(na * 31 + nb)
    a = sorted_keys(kt, na, rng, dist)
    b = sorted_keys(kt, nb, rng, dist)
    av = np.arange(na, dtype=np.uint32)
    bv = np.arange(na, na + nb, dtype=np.uint32)
    ek, ev = O.merge(kt, a, b, av, bv)
    out = torch.empty(na + nb, dtype=torch.int32 if a.itemsize == 4 else torch.int64, device="cuda")
    ov = torch.empty(na + nb, dtype=torch.int32, device="cuda")
    mg.merge(KEY_NAME[kt], to_dev(a) if na else out[:0], to_dev(b) if nb else out[:0], out,
             av=to_dev(av) if na else ov[:0], bv=to_dev(bv) if nb else ov[:0], out_vals=ov)
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(out, a.dtype), ek)
    assert np.array_equal(from_dev(ov, np.uint32), ev)


@pytest.mark.parametrize("kt", [O.K_U32, O.K_U64, O.K_F32], ids=lambda k: KEY_NAME[k])
@pytest.mark.parametrize("parts,part_size", [(1, 1), (1, 10), (1, 4096), (2, 1000), (4, 777), (8, 3), (1, 1 << 20)])
@pytest.mark.parametrize("dist", ["uniform", "few", "equal", "sorted"])
def test_pivot(mg, kt, parts, part_size, dist):
    import torch
    rng = np.random.default_rng(parts * 1000 + part_size)
    n = parts * part_size
    if dist == "sorted":  # L entirely below R: p = 0
        allk = sorted_keys(kt, 2 * n, rng)
        Lcat, Rcat = allk[:n], allk[n:]
    else:
        Lcat = sorted_keys(kt, n, rng, dist)
        Rcat = sorted_keys(kt, n, rng, dist)
    L = [Lcat[i * part_size:(i + 1) * part_size] for i in range(parts)]
    R = [Rcat[i * part_size:(i + 1) * part_size] for i in range(parts)]
    expect = O.select_pivot(kt, L, R, part_size)
    res = torch.zeros(1, dtype=torch.int64, device="cuda")
    mg.select_pivot(KEY_NAME[kt], [to_dev(x) for x in L], [to_dev(x) for x in R], part_size, res)
    torch.cuda.synchronize()
    assert int(res.item()) == expect


def test_pivot_into_pinned_host(mg):
    import torch
    rng = np.random.default_rng(1)
    L = sorted_keys(O.K_U32, 5000, rng)
    R = sorted_keys(O.K_U32, 5000, rng)
    res = torch.zeros(1, dtype=torch.int64).pin_memory()
    mg.select_pivot("u32", [to_dev(L)], [to_dev(R)], 5000, res)
    torch.cuda.synchronize()
    assert int(res.item()) == O.select_pivot(O.K_U32, [L], [R], 5000)


@pytest.mark.parametrize("kt", [O.K_U32, O.K_U64, O.K_F32], ids=lambda k: KEY_NAME[k])
@pytest.mark.parametrize("m", [1, 7, 4096, 100_001, 1 << 20])
def test_tanasic_one_device(mg, kt, m):
    """sort halves -> pivot -> swap -> merge both halves == full sort."""
    import torch
    rng = np.random.default_rng(m)
    keys = random_keys(kt, 2 * m, rng)
    vals = random_vals(np.uint32, 2 * m, rng)
    # oracle: halves sorted, then the Tanasic steps
    hk = keys.copy()
    hv = vals.copy()
    for h in range(2):
        s = slice(h * m, (h + 1) * m)
        hk[s], hv[s] = O.sort(kt, hk[s], hv[s])
    ek, ev, ep = O.tanasic_1dev(kt, hk, hv)
    # GPU
    dk, dv = to_dev(hk), to_dev(hv)
    res = torch.zeros(1, dtype=torch.int64, device="cuda")
    mg.select_pivot(KEY_NAME[kt], [dk[:m]], [dk[m:]], m, res)
    torch.cuda.synchronize()
    p = int(res.item())
    assert p == ep
    if p > 0:
        mg.swap_ranges(KEY_NAME[kt], dk[m - p:m], dk[m:m + p], dv[m - p:m], dv[m:m + p])
        ok = torch.empty_like(dk)
        ov = torch.empty_like(dv)
        mg.merge(KEY_NAME[kt], dk[:m - p], dk[m - p:m], ok[:m], dv[:m - p], dv[m - p:m], ov[:m])
        mg.merge(KEY_NAME[kt], dk[m:m + p], dk[m + p:], ok[m:], dv[m:m + p], dv[m + p:], ov[m:])
        dk, dv = ok, ov
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(dk, keys.dtype), ek)
    assert np.array_equal(from_dev(dv, np.uint32), ev)
    # and the whole thing equals a full stable sort of the sorted halves
    fk, _ = O.sort(kt, hk)
    assert np.array_equal(ek, fk)


@pytest.mark.parametrize("kt", [O.K_U32, O.K_U64, O.K_I32, O.K_I64, O.K_F32, O.K_F64], ids=lambda k: KEY_NAME[k])
def test_fill_max_sorts_last(mg, kt):
    import torch
    rng = np.random.default_rng(2)
    keys = random_keys(kt, 1000, rng)
    dk = to_dev(keys)
    full = torch.empty(1010, dtype=dk.dtype, device="cuda")
    full[:1000] = dk
    mg.fill_max(KEY_NAME[kt], full[1000:])
    mg.sort(KEY_NAME[kt], full)
    torch.cuda.synchronize()
    out = from_dev(full, keys.dtype)
    enc = O.encode(kt, out[1000:])
    assert np.all(enc == np.iinfo(np.uint64).max >> (64 - 8 * keys.itemsize))
    assert np.array_equal(out[:1000], O.sort(kt, keys)[0])


@pytest.mark.parametrize("n", [0, 1, 5, 4096, 100_003])
def test_swap(mg, n):
    import torch
    rng = np.random.default_rng(n)
    a = random_keys(O.K_U64, n, rng)
    b = random_keys(O.K_U64, n, rng)
    da, db = to_dev(a), to_dev(b)
    if n:
        mg.swap_ranges("u64", da, db)
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(da, np.uint64), b)
    assert np.array_equal(from_dev(db, np.uint64), a)


# ---------------------------------------------------------------------------
# k-way merge (mgs_multiway_merge): the device counterpart of
# __gnu_parallel::multiway_merge at src/gowanlock_sort.cuh:91-92.  Expected
# output: the stable sort of the runs laid end to end (equal keys in run
# order), from the oracle's stable sort.
# ---------------------------------------------------------------------------
def kway_lengths(k, rng, scale):
    lens = rng.integers(0, scale, size=k)
    if k >= 3:
        lens[1] = 0  # an empty run
        lens[-1] = scale * 3  # one long run
    return lens


@pytest.mark.parametrize("kt", [O.K_U32, O.K_U64, O.K_I32, O.K_F32, O.K_F64], ids=lambda k: KEY_NAME[k])
@pytest.mark.parametrize("k", [1, 2, 3, 8, 17, 64, 256])
@pytest.mark.parametrize("dist", ["uniform", "few", "equal"])
def test_multiway_merge(mg, kt, k, dist):
    import torch
    rng = np.random.default_rng(1000 * k + 7 * kt + len(dist))
    scale = {1: 300_000, 2: 200_000, 3: 100_000, 8: 60_000, 17: 20_000, 64: 5000, 256: 1500}[k]
    lens = kway_lengths(k, rng, scale)
    runs = [sorted_keys(kt, int(n), rng, dist) for n in lens]
    expect = O.sort(kt, np.concatenate(runs) if runs else np.zeros(0, dtype=np.uint32))[0]
    druns = [to_dev(r) for r in runs]
    out = torch.empty(int(lens.sum()), dtype=druns[0].dtype, device="cuda")
    temp = mg.temp_buffer(mg.lib.mgs_multiway_merge_temp_bytes(kt, k, int(lens.sum())), "cuda")
    mg.multiway_merge(KEY_NAME[kt], druns, out, temp=temp)
    assert mg.temp_status(temp) == 0
    assert np.array_equal(from_dev(out, expect.dtype), expect)


@pytest.mark.parametrize("portions", [1, 3, 16])
def test_multiway_merge_portions(mg, portions):
    """Stretches: ends increase to the total, and the output equals the
    one-call merge."""
    import torch
    rng = np.random.default_rng(portions)
    runs = [sorted_keys(O.K_U64, n, rng) for n in (1 << 18, 77_777, 0, 300_001, 5, 1 << 17, 4096, 65_537)]
    total = sum(r.size for r in runs)
    expect = O.sort(O.K_U64, np.concatenate(runs))[0]
    druns = [to_dev(r) for r in runs]
    out = torch.empty(total, dtype=torch.int64, device="cuda")
    ends = mg.multiway_merge("u64", druns, out, portions=portions)
    torch.cuda.synchronize()
    assert len(ends) == portions and ends[-1] == total
    assert all(a <= b for a, b in zip(ends, ends[1:]))
    assert np.array_equal(from_dev(out, np.uint64), expect)


def test_multiway_merge_unsorted_runs_report(mg):
    """Runs that are not sorted: every access stays inside the runs and the
    output; the call either reports MGS_ERR_UNSORTED or writes some order."""
    import torch
    rng = np.random.default_rng(5)
    runs = [random_keys(O.K_U32, n, rng) for n in (100_000, 50_000, 120_000)]
    runs[1] = np.sort(runs[1])[::-1].copy()
    druns = [to_dev(r) for r in runs]
    total = sum(r.size for r in runs)
    out = torch.zeros(total + 4096, dtype=torch.int32, device="cuda")
    temp = mg.temp_buffer(mg.lib.mgs_multiway_merge_temp_bytes(O.K_U32, 3, total), "cuda")
    mg.multiway_merge("u32", druns, out[:total], temp=temp)
    assert mg.temp_status(temp) in (0, -8)
    assert int(out[total:].abs().sum().item()) == 0  # nothing written past the output
