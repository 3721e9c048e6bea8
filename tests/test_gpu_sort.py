"""GPU parity of the radix sort (mgs_sort) against the CPU oracle.

Reference call sites replaced: thrust::sort at src/tanasic_sort.cuh:264-266,
src/gowanlock_sort.cuh:188-190/:337-340, src/sort_benchmark.cu:89; the
reference's own check is std::is_sorted (src/sort_benchmark.cu:127) plus the
max element (src/local_gpu_sort_algorithm_benchmark.cu:91-94).  Here every
output is compared bit-exactly (keys and payloads) with the oracle.
"""
import numpy as np
import pytest

import pyoracle as O
from helpers import KEY_KINDS, KEY_NAME, bits_dtype, from_dev, random_keys, random_vals, to_dev

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mg():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mgsort
    return mgsort


def run_sort(mg, kt, keys, vals, inplace=True):
    import torch
    dk = to_dev(keys)
    dv = to_dev(vals)
    if inplace:
        mg.sort(KEY_NAME[kt], dk, dv)
        ok, ov = dk, dv
    else:
        ok = torch.empty_like(dk)
        ov = None if dv is None else torch.empty_like(dv)
        mg.sort(KEY_NAME[kt], dk, dv, keys_out=ok, vals_out=ov)
        torch.cuda.synchronize()
        # input untouched
        assert np.array_equal(from_dev(dk, keys.dtype), keys)
    torch.cuda.synchronize()
    rk = from_dev(ok, keys.dtype)
    rv = None if vals is None else from_dev(ov, vals.dtype)
    return rk, rv


SIZES = [1, 2, 63, 64, 65, 1000, 2048, 2049, 4096, 4097, 8191, 8192, 8193, 100_003, 1 << 20]


@pytest.mark.parametrize("kt", KEY_KINDS, ids=lambda k: KEY_NAME[k])
@pytest.mark.parametrize("n", SIZES)
def test_sort_keys(mg, kt, n):
    rng = np.random.default_rng(1000 + n + 17 * kt)
    keys = random_keys(kt, n, rng)
    ek, _ = O.sort(kt, keys)
    rk, _ = run_sort(mg, kt, keys, None)
    assert np.array_equal(rk, ek)


@pytest.mark.parametrize("kt", [O.K_U32, O.K_U64, O.K_F32], ids=lambda k: KEY_NAME[k])
@pytest.mark.parametrize("vdt", [np.uint32, np.uint64], ids=["v32", "v64"])
@pytest.mark.parametrize("n", [5, 3000, 4097, 70_001, 1 << 20])
def test_sort_pairs_stable(mg, kt, vdt, n):
    rng = np.random.default_rng(7 + n)
    keys = random_keys(kt, n, rng, "few")  # many ties: stability is visible
    vals = random_vals(vdt, n, rng)
    ek, ev = O.sort(kt, keys, vals)
    rk, rv = run_sort(mg, kt, keys, vals)
    assert np.array_equal(rk, ek)
    assert np.array_equal(rv, ev)


@pytest.mark.parametrize("dist", ["equal", "few", "sorted", "reverse", "lowbits"])
@pytest.mark.parametrize("kt", [O.K_U32, O.K_U64, O.K_F32], ids=lambda k: KEY_NAME[k])
@pytest.mark.parametrize("n", [3000, 300_000])
def test_sort_distributions(mg, dist, kt, n):
    rng = np.random.default_rng(99)
    keys = random_keys(kt, n, rng, dist)
    vals = random_vals(np.uint32, n, rng)
    ek, ev = O.sort(kt, keys, vals)
    rk, rv = run_sort(mg, kt, keys, vals)
    assert np.array_equal(rk, ek)
    assert np.array_equal(rv, ev)


@pytest.mark.parametrize("kt", [O.K_U32, O.K_U64], ids=lambda k: KEY_NAME[k])
@pytest.mark.parametrize("dist", ["uniform", "equal", "lowbits"])
@pytest.mark.parametrize("n", [4000, 500_000])
def test_sort_out_of_place(mg, kt, dist, n):
    rng = np.random.default_rng(5)
    keys = random_keys(kt, n, rng, dist)
    vals = random_vals(np.uint32, n, rng)
    ek, ev = O.sort(kt, keys, vals)
    rk, rv = run_sort(mg, kt, keys, vals, inplace=False)
    assert np.array_equal(rk, ek)
    assert np.array_equal(rv, ev)


def test_sort_repeat_reuses_temp(mg):
    """Same temp buffer across calls (descriptor state must not leak)."""
    import torch
    rng = np.random.default_rng(3)
    n = 1_000_003
    temp = mg.temp_buffer(mg.lib.mgs_sort_temp_bytes(mg.KEY_U32, mg.VAL_NONE, n), "cuda")
    for it in range(3):
        keys = random_keys(O.K_U32, n, rng, ["uniform", "lowbits", "sorted"][it])
        dk = to_dev(keys)
        mg.sort("u32", dk, temp=temp)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(dk, np.uint32), O.sort(O.K_U32, keys)[0])


def test_temp_too_small_is_an_error(mg):
    import torch
    n = 100_000
    dk = torch.zeros(n, dtype=torch.int32, device="cuda")
    small = torch.empty(16, dtype=torch.uint8, device="cuda")
    with pytest.raises(mg.MgsError):
        mg.sort("u32", dk, temp=small)


def test_empty_is_noop(mg):
    import torch
    dk = torch.zeros(0, dtype=torch.int32, device="cuda")
    mg.sort("u32", dk)


@pytest.mark.parametrize("kt,vdt", [(O.K_U32, None), (O.K_U64, np.uint32), (O.K_F32, np.uint32)],
                         ids=["u32", "u64+v32", "f32+v32"])
def test_xcd_local_switch_same_result(mg, kt, vdt, monkeypatch):
    """Both descriptor modes of the LSD passes give the oracle's result on one
    temp buffer used alternately: XCD-local chains (workgroup-scope stores,
    chains served per XCD, the default) and MGS_XCD_LOCAL=0 (agent scope)."""
    import torch
    rng = np.random.default_rng(31 + kt)
    n = 3_000_017  # enough tiles for 16 chains
    keys = random_keys(kt, n, rng, "few")
    vals = None if vdt is None else random_vals(vdt, n, rng)
    ek, ev = O.sort(kt, keys, vals)
    vt = mg.VAL_NONE if vdt is None else mg.VAL_U32
    temp = mg.temp_buffer(mg.lib.mgs_sort_temp_bytes(kt, vt, n), "cuda")
    for mode in ("0", "1", "0", "1"):
        monkeypatch.setenv("MGS_XCD_LOCAL", mode)
        dk, dv = to_dev(keys), to_dev(vals)
        mg.sort(KEY_NAME[kt], dk, dv, temp=temp)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(dk, keys.dtype), ek), mode
        if vdt is not None:
            assert np.array_equal(from_dev(dv, vals.dtype), ev), mode


@pytest.mark.parametrize("dist", ["skewed", "few", "uniform", "byte1"])
@pytest.mark.parametrize("kt,vdt", [(O.K_U32, None), (O.K_F32, np.uint32), (O.K_I32, np.uint64), (O.K_U64, None),
                                    (O.K_F64, np.uint32), (O.K_I64, np.uint64)],
                         ids=["u32", "f32+v32", "i32+v64", "u64", "f64+v32", "i64+v64"])
def test_skew_pair_same_result(mg, kt, vdt, dist, monkeypatch):
    """Large sorts launch every pass as a pair of pass kernels, one-atomic-
    per-key ranking and uniform-digit ranking, and the one that does not suit
    the byte's digit counts exits at once.  MGS_SKEW_PAIR_MIN=0 forces the
    pair on an array below the default size: with bytes whose top digit holds
    more than 1/64 of the keys (bytes 0 and 2 of "skewed", every byte of
    "few") and without (uniform), and with a constant byte 0 below a skewed
    byte 1 (counted from the input by the first kernel of the pair only),
    both settings give the oracle's result on one temp buffer, payloads
    stable."""
    import torch
    rng = np.random.default_rng(131 + kt)
    n = (1 << 21) + 54_321  # large tiles for every record width
    keys = random_keys(kt, n, rng, dist)
    vals = None if vdt is None else random_vals(vdt, n, rng)
    ek, ev = O.sort(kt, keys, vals)
    vt = mg.VAL_NONE if vdt is None else (mg.VAL_U32 if vdt == np.uint32 else mg.VAL_U64)
    temp = mg.temp_buffer(mg.lib.mgs_sort_temp_bytes(kt, vt, n), "cuda")
    for pair_min in ("0", str(1 << 26), "0"):
        monkeypatch.setenv("MGS_SKEW_PAIR_MIN", pair_min)
        dk, dv = to_dev(keys), to_dev(vals)
        mg.sort(KEY_NAME[kt], dk, dv, temp=temp)
        torch.cuda.synchronize()
        assert np.array_equal(from_dev(dk, keys.dtype), ek), pair_min
        if vdt is not None:
            assert np.array_equal(from_dev(dv, vals.dtype), ev), pair_min


@pytest.mark.parametrize("kt,vdt", [(O.K_U64, None), (O.K_I64, np.uint64), (O.K_U32, np.uint32), (O.K_F64, np.uint32),
                                    (O.K_I32, np.uint64)],
                         ids=["u64", "i64+v64", "u32+v32", "f64+v32", "i32+v64"])
def test_sort_large_tile_geometry(mg, kt, vdt):
    """Arrays of 2^21 keys and more with records wider than 4 bytes use the
    large tiles (one workgroup per CU); parity with the oracle, ragged last
    tile, payloads stable."""
    rng = np.random.default_rng(77 + kt)
    n = (1 << 21) + 12_345
    keys = random_keys(kt, n, rng, "few")
    vals = None if vdt is None else random_vals(vdt, n, rng)
    ek, ev = O.sort(kt, keys, vals)
    rk, rv = run_sort(mg, kt, keys, vals, inplace=False)
    assert np.array_equal(rk, ek)
    if vdt is not None:
        assert np.array_equal(rv, ev)


@pytest.mark.parametrize("kt", [O.K_U32, O.K_F32, O.K_I32], ids=lambda k: KEY_NAME[k])
@pytest.mark.parametrize("n", [100_003, (1 << 21) + 999])
@pytest.mark.parametrize("dist", ["uniform", "few", "sorted"])
def test_rank_peers_variant(mg, kt, n, dist, monkeypatch):
    """MGS_RANK_PEERS=1 ranks digits with peer masks (lanes with equal
    digits find each other through an LDS OR; stability relies only on the
    in-order execution of a wavefront's DS instructions) instead of one
    ds_add_rtn_u32 per key: same result as the oracle."""
    import torch
    rng = np.random.default_rng(n + kt)
    keys = random_keys(kt, n, rng, dist)
    monkeypatch.setenv("MGS_RANK_PEERS", "1")
    dk = to_dev(keys)
    temp = mg.temp_buffer(mg.lib.mgs_sort_temp_bytes(kt, mg.VAL_NONE, n), "cuda")
    mg.sort(KEY_NAME[kt], dk, temp=temp)
    assert mg.temp_status(temp) == 0
    assert np.array_equal(from_dev(dk, keys.dtype), O.sort(kt, keys)[0])


@pytest.mark.parametrize("kt,vdt", [(O.K_U32, None), (O.K_U64, np.uint32)], ids=["u32", "u64+v32"])
def test_lookback_bound_reports_status(mg, kt, vdt, monkeypatch):
    """The look-back waits are bounded.  With the bound at 0 sleeps
    (MGS_LB_SPIN_LIMIT=0) a tile whose predecessor has not published yet gives
    up at once; every kernel still ends, and the call's status word says
    MGS_ERR_DEVICE.  A call whose status is ok must have the oracle's output;
    with the default bound the status is ok."""
    import torch
    rng = np.random.default_rng(5 + kt)
    n = 1 << 24
    keys = random_keys(kt, n, rng)
    vals = None if vdt is None else random_vals(vdt, n, rng)
    ek, ev = O.sort(kt, keys, vals)
    vt = mg.VAL_NONE if vdt is None else mg.VAL_U32
    temp = mg.temp_buffer(mg.lib.mgs_sort_temp_bytes(kt, vt, n), "cuda")
    seen = []
    for limit in ("0", "0", None):
        if limit is None:
            monkeypatch.delenv("MGS_LB_SPIN_LIMIT", raising=False)
        else:
            monkeypatch.setenv("MGS_LB_SPIN_LIMIT", limit)
        dk, dv = to_dev(keys), to_dev(vals)
        mg.sort(KEY_NAME[kt], dk, dv, temp=temp)
        st = mg.temp_status(temp)
        seen.append(st)
        assert st in (0, -6)
        if st == 0:
            assert np.array_equal(from_dev(dk, keys.dtype), ek), limit
            if vdt is not None:
                assert np.array_equal(from_dev(dv, vals.dtype), ev), limit
    assert seen[-1] == 0


MSD_KEYS = [O.K_U32, O.K_I32, O.K_F32]


def msd_default_size(n):
    """Sizes the three-pass path is tried for without MGS_MSD_MIN:
    [2^23, 3 * 2^24) and [2^29, 2^30 + 2^27]."""
    return (1 << 23) <= n < (3 << 24) or (1 << 29) <= n <= (1 << 30) + (1 << 27)


def expected_path(kt, keys, forced=True):
    """The three-pass path hands over to the LSD passes when a 16-bit
    sub-bucket exceeds the 20480 keys of the LDS sort, a byte-3 bucket
    holds more than twice its share plus two 18432-key tiles, or a key
    differs from the leading bits that a sample of 1024 keys (positions
    i (n - 1) / 1023) shares, or at once when a bucket holds more than 20
    of the sampled keys.  The sub-buckets are the 16 bits below those
    shared leading bits (at most 16 of them).  forced: the path is tried at
    any size (MGS_MSD_MIN=0)."""
    n = keys.size
    if n <= 4096:
        return 1
    if not forced and not msd_default_size(n):
        return 2
    enc = O.encode(kt, keys).astype(np.uint64)
    sample = enc[(np.arange(1024, dtype=np.uint64) * np.uint64(n - 1)) // np.uint64(1023)]
    same, any_ = np.bitwise_and.reduce(sample), np.bitwise_or.reduce(sample)
    common = 32 - int(same ^ any_).bit_length()
    low = 16 - min(common, 16)
    # a bucket (the window's high byte) holding more than 20 sampled keys
    # hands over before the read
    if np.bincount(((sample >> np.uint64(low + 8)) & np.uint64(0xFF)).astype(np.int64), minlength=256).max() > 20:
        return 2
    if low < 16:
        hi_mask = np.uint64((0xFFFFFFFF << (low + 16)) & 0xFFFFFFFF)
        if np.any((enc & hi_mask) != (same & hi_mask)):
            return 2
    top = ((enc >> np.uint64(low)) & np.uint64(0xFFFF)).astype(np.int64)
    sub = np.bincount(top, minlength=1 << 16)
    bucket = sub.reshape(256, 256).sum(axis=1)
    off = sub.max() > 20480 or bucket.max() > 2 * (n // 256) + 2 * 18432
    return 2 if off else 3


@pytest.mark.parametrize("kt", MSD_KEYS, ids=lambda k: KEY_NAME[k])
@pytest.mark.parametrize("n", [4097, 100_003, 1 << 20, (1 << 21) + 777])
@pytest.mark.parametrize("dist", ["uniform", "sorted", "reverse", "few", "equal", "lowbits", "runs", "nonneg", "low28",
                                  "stray"])
def test_three_pass_path(mg, kt, n, dist, monkeypatch):
    """4-byte keys alone try the three-pass path (MSD pass on byte 3, pass on
    byte 2 inside each bucket, LDS sort of each batch of 16-bit sub-buckets);
    MGS_MSD_MIN=0 tries it at any size.  Keys whose top 16 bits crowd into
    a few sub-buckets (few, equal, lowbits at the larger sizes) hand over to
    the LSD passes.  Runs of keys sharing their top 16 bits (runs, sorted at
    2^26) switch the passes to the uniform-digit ranking.  Keys whose
    leading bits agree (nonneg, low28) sort by the 16 bits below them; one
    key that breaks the agreement a sample saw (stray) hands over to the LSD
    passes.  Either way the output is the oracle's, in place and out of
    place."""
    import torch
    monkeypatch.setenv("MGS_MSD_MIN", "0")
    rng = np.random.default_rng(n + 3 * kt)
    keys = random_keys(kt, n, rng, dist)
    ek, _ = O.sort(kt, keys)
    need = mg.lib.mgs_sort_temp_bytes(kt, mg.VAL_NONE, n)
    temp = mg.temp_buffer(need, "cuda")
    for inplace in (True, False):
        dk = to_dev(keys)
        out = dk if inplace else torch.empty_like(dk)
        mg.sort(KEY_NAME[kt], dk, keys_out=out, temp=temp)
        assert mg.temp_status(temp) == 0
        got = mg.lib.mgs_sort_path(kt, mg.VAL_NONE, n, mg._ptr(temp), temp.numel(), None)
        assert got == expected_path(kt, keys), (inplace, got)
        assert np.array_equal(from_dev(out, keys.dtype), ek), inplace
        if not inplace:
            assert np.array_equal(from_dev(dk, keys.dtype), keys)


@pytest.mark.parametrize("kt,dist,n", [(O.K_U32, "uniform", (1 << 24) + 12_345), (O.K_U32, "sorted", (1 << 24) + 12_345),
                                       (O.K_U32, "few", (1 << 24) + 12_345), (O.K_U32, "crowd", (1 << 24) + 12_345),
                                       (O.K_U32, "reverse", (1 << 24) + 12_345), (O.K_U32, "skewed", (1 << 24) + 12_345),
                                       (O.K_U32, "uniform", (1 << 26) + 12_345), (O.K_U32, "nonneg", (1 << 24) + 12_345),
                                       (O.K_U32, "low28", (1 << 24) + 12_345), (O.K_U32, "stray", (1 << 24) + 12_345),
                                       (O.K_I32, "nonneg", (1 << 24) + 12_345), (O.K_F32, "nonneg", (1 << 24) + 12_345)],
                         ids=lambda v: KEY_NAME.get(v, str(v)) if isinstance(v, int) and v < 8 else str(v))
def test_three_pass_default_size(mg, kt, dist, n):
    """Without environment the path is tried in [2^23, 3 * 2^24) and from
    2^29 keys (2^26 takes the LSD passes); the output is the oracle's either
    way.  Non-negative int keys (the generator's int inputs) and keys below
    2^28 take the three-pass path with the window under their shared leading
    bits."""
    import torch
    rng = np.random.default_rng(26)
    keys = random_keys(kt, n, rng, dist)
    ek, _ = O.sort(kt, keys)
    temp = mg.temp_buffer(mg.lib.mgs_sort_temp_bytes(kt, mg.VAL_NONE, n), "cuda")
    dk = to_dev(keys)
    mg.sort(KEY_NAME[kt], dk, temp=temp)
    assert mg.temp_status(temp) == 0
    path = mg.lib.mgs_sort_path(kt, mg.VAL_NONE, n, mg._ptr(temp), temp.numel(), None)
    assert path == expected_path(kt, keys, forced=False)
    if dist in ("nonneg", "low28") and kt != O.K_F32:  # (the float inputs carry specials that crowd sub-buckets)
        assert path == 3
    if dist == "stray":
        assert path == 2
    assert np.array_equal(from_dev(dk, np.uint32), ek)


@pytest.mark.parametrize("kt,vdt", [(O.K_U32, None), (O.K_F32, np.uint32), (O.K_U64, np.uint32), (O.K_I64, np.uint64)],
                         ids=["u32", "f32+v32", "u64+v32", "i64+v64"])
@pytest.mark.parametrize("budget", ["16384", "40000"])
def test_lsd_rounds(mg, kt, vdt, budget, monkeypatch):
    """A sort whose look-back descriptors exceed the budget (MGS_DESC_BUDGET,
    forced small here) runs every pass in rounds over a window of each
    chain's tiles, carrying the chains' digit counts from round to round:
    same result as the oracle, with the pair of ranking launches and with
    agent-scope descriptors too, on one temp buffer reused."""
    import torch
    monkeypatch.setenv("MGS_DESC_BUDGET", budget)
    monkeypatch.setenv("MGS_MSD", "0")
    rng = np.random.default_rng(int(budget) + kt)
    n = (1 << 21) + 7777
    keys = random_keys(kt, n, rng, "few")
    vals = None if vdt is None else random_vals(vdt, n, rng)
    ek, ev = O.sort(kt, keys, vals)
    vt = mg.VAL_NONE if vdt is None else (mg.VAL_U32 if vdt == np.uint32 else mg.VAL_U64)
    temp = mg.temp_buffer(mg.lib.mgs_sort_temp_bytes(kt, vt, n), "cuda")
    for pair_min, xcd in (("0", "1"), (str(1 << 26), "0"), (str(1 << 26), "1")):
        monkeypatch.setenv("MGS_SKEW_PAIR_MIN", pair_min)
        monkeypatch.setenv("MGS_XCD_LOCAL", xcd)
        dk, dv = to_dev(keys), to_dev(vals)
        mg.sort(KEY_NAME[kt], dk, dv, temp=temp)
        assert mg.temp_status(temp) == 0
        assert np.array_equal(from_dev(dk, keys.dtype), ek), (pair_min, xcd)
        if vdt is not None:
            assert np.array_equal(from_dev(dv, vals.dtype), ev), (pair_min, xcd)
