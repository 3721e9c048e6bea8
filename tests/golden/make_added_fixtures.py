"""Writes tests/golden/datagen_added.json: digests of the distributions added
for the MI355X benchmark configs (normal-clamped, zipf, zipf-1.2,
nearly-sorted-swaps, float-specials), computed by the oracle's C
restatement (oracle/oracle_datagen.c).  The product generator is checked
against these digests and against the oracle itself (tests/test_datagen.py).

Run:  make -C oracle && python tests/golden/make_added_fixtures.py
"""
import hashlib
import itertools
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402

DT = {"u32": np.uint32, "u64": np.uint64, "f32": np.float32, "f64": np.float64, "i32": np.int32, "i64": np.int64}
ADDED = ["normal-clamped", "zipf", "zipf-1.2", "nearly-sorted-swaps", "float-specials"]
SEEDS = [2147483647, 7]
SIZES = [1, 100, 4097, 100003, 1 << 20]


def generate(dt, dist, seed, n):
    a = np.zeros(n, dtype=DT[dt])
    r = pyoracle.lib().oracle_generate(dt.encode(), dist.encode(), seed, a.ctypes.data, n, 0)
    return a if r == 0 else None


def main():
    cases = []
    for dt, dist, seed, n in itertools.product(DT, ADDED, SEEDS, SIZES):
        a = generate(dt, dist, seed, n)
        if a is None:
            continue  # float-specials exists for floating types only
        bits = a.view(np.uint32 if a.itemsize == 4 else np.uint64)
        cases.append({"dtype": dt, "dist": dist, "seed": seed, "n": n,
                      "sha256": hashlib.sha256(a.tobytes()).hexdigest()[:32],
                      "head": [f"{int(x):x}" for x in bits[:4]], "tail": [f"{int(x):x}" for x in bits[-4:]]})
    with open(os.path.join(HERE, "datagen_added.json"), "w") as f:
        f.write('{"source": "oracle/oracle_datagen.c",\n "cases": [\n')
        f.write(",\n".join(json.dumps(c, separators=(",", ":")) for c in cases))
        f.write("\n]}\n")
    print(len(cases), "cases")


if __name__ == "__main__":
    main()
