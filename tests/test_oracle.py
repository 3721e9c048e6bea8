"""Pins the CPU oracle (oracle/oracle_sort.c) to independent references, so
that the GPU parity tests check against something known to be right.

* oracle_sort: numpy's stable sort on the values themselves (integers; floats
  without NaN, where -0.0 == +0.0 keep input order exactly as a stable sort
  on enc() does), numpy's stable argsort for the payload permutation, a
  brute-force insertion sort on tiny inputs, and the NaN / infinity placement
  the key order prescribes (-NaN < -inf < ... < +inf < +NaN).
* oracle_merge: a stable sort of the concatenation A ++ B (A wins ties).
* oracle_select_pivot: the definition by brute force over every k, and the
  invariant that after the swap max(L') <= min(R').
* oracle_tanasic_1dev: on two sorted halves, the sorted whole; its pivot is
  the brute-force pivot of the halves.
* oracle_segmented_sort: numpy per segment.
"""
import numpy as np
import pytest

import pyoracle as O

INT_TYPES = {O.K_U32: np.uint32, O.K_I32: np.int32, O.K_U64: np.uint64, O.K_I64: np.int64}
FLT_TYPES = {O.K_F32: np.float32, O.K_F64: np.float64}


def rand_values(kt, n, rng, dup=False):
    t = {**INT_TYPES, **FLT_TYPES}[kt]
    if kt in FLT_TYPES:
        span = 30 if t == np.float32 else 300
        x = (rng.standard_normal(n) * 10.0 ** rng.integers(-span, span, size=n)).astype(t)
        if n >= 8:  # signed zeros, infinities, denormals (no NaN here)
            x[rng.integers(0, n, 4)] = t(0.0)
            x[rng.integers(0, n, 4)] = t(-0.0)
            x[rng.integers(0, n, 2)] = t(np.inf)
            x[rng.integers(0, n, 2)] = t(-np.inf)
            tiny = np.finfo(t).tiny
            x[rng.integers(0, n, 4)] = t(tiny / 8)
            x[rng.integers(0, n, 4)] = t(-tiny / 3)
    else:
        info = np.iinfo(t)
        hi = 50 if dup else int(info.max)
        lo = -50 if (dup and info.min < 0) else (0 if dup else int(info.min))
        x = rng.integers(lo, hi, size=n, dtype=np.int64 if t != np.uint64 else np.uint64,
                         endpoint=True).astype(t)
    return np.ascontiguousarray(x)


def bits(kt, x):
    return x.view(O.BITS_DTYPE[kt])


@pytest.mark.parametrize("kt", list(INT_TYPES) + list(FLT_TYPES))
@pytest.mark.parametrize("n", [0, 1, 2, 17, 1000, 100_003])
@pytest.mark.parametrize("dup", [False, True])
def test_sort_equals_numpy_stable(kt, n, dup):
    rng = np.random.default_rng(n * 7 + kt)
    x = rand_values(kt, n, rng, dup)
    got, _ = O.sort(kt, bits(kt, x))
    exp = x[np.argsort(x, kind="stable")]
    assert np.array_equal(got, bits(kt, exp))  # bitwise: -0.0/+0.0 in input order


@pytest.mark.parametrize("kt", [O.K_U32, O.K_I64, O.K_F32, O.K_F64])
def test_pairs_permutation_is_stable_argsort(kt):
    rng = np.random.default_rng(3)
    n = 50_000
    x = rand_values(kt, n, rng, dup=True)
    v = np.arange(n, dtype=np.uint32)
    _, gv = O.sort(kt, bits(kt, x), v)
    assert np.array_equal(gv, np.argsort(x, kind="stable").astype(np.uint32))


def test_sort_brute_force_tiny():
    """Insertion sort on enc order written from the key-order definition."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        n = int(rng.integers(0, 12))
        x = rng.integers(0, 6, size=n).astype(np.uint32)
        v = np.arange(n, dtype=np.uint32)
        order = []
        for i in range(n):  # stable insertion
            j = len(order)
            while j > 0 and x[order[j - 1]] > x[i]:
                j -= 1
            order.insert(j, i)
        k, gv = O.sort(O.K_U32, x, v)
        assert list(gv) == order and np.array_equal(k, x[order])


@pytest.mark.parametrize("kt,t,sign,expo", [(O.K_F32, np.uint32, 0x80000000, 0x7F800000),
                                             (O.K_F64, np.uint64, 0x8000000000000000, 0x7FF0000000000000)])
def test_nan_and_infinity_placement(kt, t, sign, expo):
    qnan = expo | (expo >> 1 & ~expo)  # quiet bit set
    vals = np.array([expo | 1, sign | expo | 5, expo, sign | expo, 0, sign, qnan, sign | qnan, 1, sign | 1],
                    dtype=t)
    k, _ = O.sort(kt, vals)
    # -NaNs (largest payload first), -inf, -denormal, zeros, +denormal, +inf, +NaNs
    exp = np.array([sign | qnan, sign | expo | 5, sign | expo, sign | 1, 0, sign, 1, expo, expo | 1, qnan], dtype=t)
    assert np.array_equal(k, exp)


@pytest.mark.parametrize("kt", [O.K_U32, O.K_F64])
@pytest.mark.parametrize("na,nb", [(0, 0), (0, 5), (7, 0), (1000, 333), (5000, 5000)])
def test_merge_equals_stable_sort_of_concatenation(kt, na, nb):
    rng = np.random.default_rng(na + nb)
    a = np.sort(rand_values(kt, na, rng, dup=True), kind="stable")
    b = np.sort(rand_values(kt, nb, rng, dup=True), kind="stable")
    av = np.arange(na, dtype=np.uint32)
    bv = np.arange(na, na + nb, dtype=np.uint32)
    k, v = O.merge(kt, bits(kt, a), bits(kt, b), av, bv)
    cat = np.concatenate([a, b])
    order = np.argsort(cat, kind="stable")
    assert np.array_equal(k, bits(kt, cat[order]))
    assert np.array_equal(v, order.astype(np.uint32))


def pivot_brute(L, R):
    """max k in [0, M] with k == 0 or L[M-k] > R[k-1] (the definition)."""
    M = len(L)
    best = 0
    for k in range(1, M + 1):
        if L[M - k] > R[k - 1]:
            best = k
    return best


@pytest.mark.parametrize("parts", [1, 2, 4])
def test_select_pivot_brute_force_and_invariant(parts):
    rng = np.random.default_rng(parts)
    for trial in range(150):
        m = int(rng.integers(1, 9))
        M = m * parts
        L = np.sort(rng.integers(0, 20, size=M).astype(np.uint32))
        R = np.sort(rng.integers(0, 20, size=M).astype(np.uint32))
        p = O.select_pivot(O.K_U32, np.split(L, parts), np.split(R, parts), m)
        assert p == pivot_brute(L, R)
        L2 = np.concatenate([L[:M - p], R[:p]])
        R2 = np.concatenate([L[M - p:], R[p:]])
        assert L2.max() <= R2.min()


@pytest.mark.parametrize("kt", [O.K_U32, O.K_F32])
def test_tanasic_one_device_merges_sorted_halves(kt):
    rng = np.random.default_rng(9)
    n = 20_000
    x = rand_values(kt, n, rng, dup=True)
    h = n // 2
    left = O.sort(kt, bits(kt, x[:h]))[0]
    right = O.sort(kt, bits(kt, x[h:]))[0]
    k, _, p = O.tanasic_1dev(kt, np.concatenate([left, right]))  # the merge step on two sorted halves
    assert np.array_equal(O.encode(kt, k), np.sort(O.encode(kt, bits(kt, x))))
    assert p == pivot_brute(list(O.encode(kt, left)), list(O.encode(kt, right)))


def test_segmented_sort_equals_numpy_per_segment():
    rng = np.random.default_rng(13)
    lengths = rng.integers(0, 300, size=60)
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.uint64)
    x = rand_values(O.K_I32, int(off[-1]), rng, dup=True)
    v = np.arange(x.size, dtype=np.uint32)
    k, gv = O.segmented_sort(O.K_I32, bits(O.K_I32, x), off, v)
    for s in range(len(lengths)):
        a, b = int(off[s]), int(off[s + 1])
        order = np.argsort(x[a:b], kind="stable")
        assert np.array_equal(k[a:b], bits(O.K_I32, x[a:b][order]))
        assert np.array_equal(gv[a:b], (order + a).astype(np.uint32))
