"""mgs_sort_noalloc (scratch in the library's static device array) against
mgs_sort (caller's temp buffer): bit-identical keys and payloads, two calls
in a row on one stream, the second smaller (the array keeps the first call's
contents), and the size limit returns MGS_ERR_TOO_LARGE."""
import ctypes

import numpy as np
import pytest

import pyoracle as O
from helpers import KEY_NAME, from_dev, random_keys, random_vals, to_dev

pytestmark = pytest.mark.gpu

NOALLOC_BYTES = 64 << 20


@pytest.fixture(scope="module")
def mg():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import mgsort
    return mgsort


def noalloc(mg, kt, vt, keys, vals, n):
    import torch
    st = mg.lib.mgs_sort_noalloc(kt, vt, keys.data_ptr(), keys.data_ptr(), None if vals is None else vals.data_ptr(),
                                 None if vals is None else vals.data_ptr(), n,
                                 torch.cuda.current_stream().cuda_stream)
    return st


@pytest.mark.parametrize("kt,vdt", [(O.K_U32, None), (O.K_F32, np.uint32), (O.K_U64, np.uint32), (O.K_I64, np.uint64)],
                         ids=["u32", "f32+v32", "u64+v32", "i64+v64"])
def test_noalloc_matches_sort_twice(mg, kt, vdt):
    import torch
    vt = mg.VAL_NONE if vdt is None else (mg.VAL_U32 if vdt == np.uint32 else mg.VAL_U64)
    rng = np.random.default_rng(90 + kt)
    # the largest size whose scratch fits the array, then a smaller one
    kb = 4 if kt in (O.K_U32, O.K_I32, O.K_F32) else 8
    vb = 0 if vdt is None else np.dtype(vdt).itemsize
    big = NOALLOC_BYTES // (kb + vb) - 100_000
    while mg.lib.mgs_sort_temp_bytes(kt, vt, big) > NOALLOC_BYTES:
        big -= 10_000
    for n in (big, 123_457):
        keys = random_keys(kt, n, rng, "few")
        vals = random_vals(vdt, n, rng)
        a_k, a_v = to_dev(keys), to_dev(vals)
        assert noalloc(mg, kt, vt, a_k, a_v, n) == 0
        b_k, b_v = to_dev(keys), to_dev(vals)
        mg.sort(KEY_NAME[kt], b_k, b_v)
        torch.cuda.synchronize()
        assert torch.equal(a_k, b_k), n
        if vdt is not None:
            assert torch.equal(a_v, b_v), n
        ek, ev = O.sort(kt, keys, vals)
        assert np.array_equal(from_dev(a_k, keys.dtype), ek), n


def test_noalloc_size_limit(mg):
    import torch
    n = NOALLOC_BYTES // 4
    assert mg.lib.mgs_sort_temp_bytes(mg.KEY_U32, mg.VAL_NONE, n) > NOALLOC_BYTES
    k = torch.zeros(16, dtype=torch.int32, device="cuda")
    # the size check comes before any access to the (short) buffer
    assert mg.lib.mgs_sort_noalloc(mg.KEY_U32, mg.VAL_NONE, k.data_ptr(), k.data_ptr(), None, None, n,
                                   torch.cuda.current_stream().cuda_stream) == -5
