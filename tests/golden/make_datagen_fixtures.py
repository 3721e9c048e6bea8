"""Writes tests/golden/datagen_ref.json: digests of the reference's own
DataGenerator (src/data_generator.cuh, compiled by oracle/Makefile into
oracle/_ref/ref_datagen) over a grid of (dtype, distribution, seed, n).

Each entry holds the SHA-256 digest (first 128 bits) of the raw output bytes and the first
and last four values as hex bit patterns.  nearly-sorted is digested with the
positions the reference computes racily (the last iteration of each of the
128 static blocks of n-1 iterations) zeroed.

Run in the container that mounts the reference:  make -C oracle ref &&
python tests/golden/make_datagen_fixtures.py
"""
import hashlib
import itertools
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_datagen")
DT = {"int": np.int32, "long": np.int64, "float": np.float32, "double": np.float64, "u32": np.uint32,
      "u64": np.uint64}
DISTS = ["uniform", "normal", "zero", "staggered", "sorted", "reverse-sorted", "nearly-sorted", "bucket-sorted"]
SEEDS = [2147483647, 0, 4294967295, 12345]
SIZES = [1, 9, 10, 19, 127, 128, 129, 1000, 100003, 1 << 20]


def sha256(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()[:32]


def racy_positions(n: int):
    if n < 2:
        return []
    out = []
    it = n - 1
    q, r = divmod(it, 128)
    for t in range(128):
        qq, rr = (q + 1, 0) if t < r else (q, r)
        lo = qq * t + rr
        hi = lo + qq
        if hi > lo and hi < n - 1:
            out.append(hi - 1)
    return out


def digest(arr: np.ndarray, dist: str) -> dict:
    a = arr.copy()
    if dist == "nearly-sorted":
        a[racy_positions(a.size)] = 0
    bits = a.view(np.uint32 if a.itemsize == 4 else np.uint64)
    return {"sha256": sha256(a.tobytes()), "head": [f"{int(x):x}" for x in bits[:4]],
            "tail": [f"{int(x):x}" for x in bits[-4:]]}


def main():
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: run make -C oracle ref")
    cases = []
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "v.bin")
        for dt, dist, seed, n in itertools.product(DT, DISTS, SEEDS, SIZES):
            if dt in ("u32", "u64") and dist == "normal":
                continue  # the reference's double->unsigned cast is undefined there
            if n < 10 and dist in ("staggered", "bucket-sorted"):
                continue  # the reference divides by zero
            if n > 100003 and seed not in (2147483647,):
                continue
            subprocess.run([REF, dt, dist, str(seed), str(n), out], check=True)
            arr = np.fromfile(out, dtype=DT[dt])
            cases.append({"dtype": dt, "dist": dist, "seed": seed, "n": n, **digest(arr, dist)})
    with open(os.path.join(HERE, "datagen_ref.json"), "w") as f:
        f.write('{"source": "src/data_generator.cuh compiled by oracle/Makefile (oracle/_ref/ref_datagen)",\n')
        f.write(' "digest": "sha256 (first 128 bits) of raw bytes; nearly-sorted with racy positions zeroed",\n')
        f.write(' "cases": [\n')
        f.write(",\n".join(json.dumps(c, separators=(",", ":")) for c in cases))
        f.write("\n]}\n")
    print(len(cases), "cases")


if __name__ == "__main__":
    main()
