"""Shared test helpers: seeded inputs and numpy <-> device tensor plumbing."""
from __future__ import annotations

import numpy as np

import pyoracle as O

KEY_KINDS = [O.K_U32, O.K_U64, O.K_I32, O.K_I64, O.K_F32, O.K_F64]
KEY_NAME = {O.K_U32: "u32", O.K_U64: "u64", O.K_I32: "i32", O.K_I64: "i64", O.K_F32: "f32", O.K_F64: "f64"}
VAL_KINDS = [None, np.uint32, np.uint64]


def bits_dtype(kt):
    return O.BITS_DTYPE[kt]


def random_keys(kt: int, n: int, rng: np.random.Generator, dist: str = "uniform") -> np.ndarray:
    """Keys as raw bits (uint32/uint64 arrays)."""
    bd = bits_dtype(kt)
    nbits = 32 if bd == np.uint32 else 64
    if dist == "uniform":
        k = rng.integers(0, 2**nbits, size=n, dtype=np.uint64 if nbits == 64 else np.uint64).astype(bd)
    elif dist == "equal":
        k = np.full(n, rng.integers(0, 2**31), dtype=bd)
    elif dist == "few":
        pool = rng.integers(0, 2**nbits, size=7, dtype=np.uint64).astype(bd)
        k = pool[rng.integers(0, 7, size=n)]
    elif dist == "sorted":
        k = random_keys(kt, n, rng)
        k = k[np.argsort(O.encode(kt, k), kind="stable")]
    elif dist == "reverse":
        k = random_keys(kt, n, rng, "sorted")[::-1].copy()
    elif dist == "skewed":  # bytes 0 and 2 have a digit with > 1/64 of the keys, the others are uniform
        k = random_keys(kt, n, rng).astype(np.uint64)
        k = (k & ~np.uint64(0xFF)) | rng.integers(0, 4, size=n).astype(np.uint64)
        half = rng.random(n) < 0.5
        k[half] = (k[half] & ~np.uint64(0xFF0000)) | np.uint64(0x550000)
        k = k.astype(bd)
    elif dist == "byte1":  # byte 0 constant, byte 1 takes 3 values, the higher bytes are uniform
        k = random_keys(kt, n, rng).astype(np.uint64)
        k = (k & ~np.uint64(0xFFFF)) | np.uint64(0x34) | (rng.integers(0, 3, size=n).astype(np.uint64) << np.uint64(8))
        k = k.astype(bd)
    elif dist == "runs":  # runs of 4096 keys sharing their top 16 bits, the runs spread over the top byte
        run = (np.arange(n, dtype=np.uint64) // np.uint64(4096)) * np.uint64(257) & np.uint64(0xFFFF)
        k = ((run << np.uint64(nbits - 16)) | rng.integers(0, 1 << 16, size=n).astype(np.uint64)).astype(bd)
    elif dist == "lowbits":  # only the low byte varies
        k = (rng.integers(0, 256, size=n).astype(np.uint64) | np.uint64(0x1200)).astype(bd)
    elif dist == "nonneg":  # the top bit of the raw keys clear: non-negative ints and floats, u32 below 2^31
        k = random_keys(kt, n, rng) & ~bd(1 << (nbits - 1))
    elif dist == "low28":  # raw keys below 2^28: the top 4 bits of every encoded key agree
        k = random_keys(kt, n, rng) & bd((1 << 28) - 1)
    elif dist == "crowd":  # a third of the keys share their top byte (0x5A), the rest uniform
        k = random_keys(kt, n, rng).astype(np.uint64)
        m = rng.random(n) < 1 / 3
        k[m] = (k[m] & np.uint64((1 << (nbits - 8)) - 1)) | np.uint64(0x5A << (nbits - 8))
        k = k.astype(bd)
    elif dist == "stray":  # low28 with one key of all ones, at a position a sparse sample misses
        k = random_keys(kt, n, rng, "low28")
        k[min(1, n - 1)] = bd((1 << nbits) - 1)
    else:
        raise ValueError(dist)
    if kt in (O.K_F32, O.K_F64) and dist in ("uniform", "few") and n >= 16:
        k = inject_float_specials(kt, k, rng)
    return np.ascontiguousarray(k)


def inject_float_specials(kt: int, k: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    """Plants +-0, +-inf, NaNs with payloads, denormals of both signs."""
    k = k.copy()
    if kt == O.K_F32:
        specials = np.array([0x00000000, 0x80000000, 0x7F800000, 0xFF800000, 0x7FC00000, 0xFFC00000,
                             0x7F800001, 0xFFBFFFFF, 0x00000001, 0x807FFFFF, 0x00400000, 0x80000001,
                             0x7F7FFFFF, 0xFF7FFFFF], dtype=np.uint32)
    else:
        specials = np.array([0x0, 0x8000000000000000, 0x7FF0000000000000, 0xFFF0000000000000,
                             0x7FF8000000000000, 0xFFF8000000000000, 0x7FF0000000000001, 0x0000000000000001,
                             0x800FFFFFFFFFFFFF, 0x7FEFFFFFFFFFFFFF], dtype=np.uint64)
    m = max(len(specials), k.size // 50)
    pos = rng.choice(k.size, size=min(m, k.size), replace=False)
    k[pos] = specials[rng.integers(0, len(specials), size=pos.size)]
    return k


def random_vals(vdt, n: int, rng: np.random.Generator):
    if vdt is None:
        return None
    # the original index makes stability observable
    return np.arange(n, dtype=vdt) ^ vdt(rng.integers(0, 2**31))


def to_dev(a: np.ndarray):
    import torch
    if a is None:
        return None
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    elif a.dtype == np.uint64:
        a = a.view(np.int64)
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def from_dev(t, dtype) -> np.ndarray:
    return t.cpu().numpy().view(dtype)
