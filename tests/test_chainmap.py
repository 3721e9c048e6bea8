"""The chain map of the LSD passes (csrc/radix_sort.hpp), compiled for the
host from the library's own header and checked against plain division.

A pass splits its tiles into chains; chain_of() maps a key position to its
chain with one multiply by a precomputed reciprocal and one correction step.
It must equal pos / chain_keys for every position below 2^32, for every tile
size the kernels use (power-of-two tiles and the 9216-key u32 tile), at the
chain boundaries and at random positions, and every chain must be non-empty.
No counterpart in the reference (its thrust::sort has no chains).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"

PROGRAM = r"""
#include "radix_sort.hpp"
#include <cstdio>
#include <random>
int main() {
  std::mt19937_64 rng(1);
  long bad = 0, checks = 0;
  const uint32_t tiles[] = {2048, 4096, 4608, 8192, 9216, 12288};
  for (uint32_t tile : tiles)
    for (int rep = 0; rep < 600; ++rep) {
      const uint64_t n = rep < 200 ? (rng() % (1ull << 20)) + 1
                       : rep < 400 ? (rng() % ((1ull << 32) - 1)) + 1
                                   : (1ull << 32) - 1 - (rng() % 100000);
      for (int maxc : {1, 8, 16, 32}) {
        const mgs::ChainMap m = mgs::make_chain_map(n, tile, maxc);
        if (m.chains == 0 || m.chains > (uint32_t)maxc) { printf("chain count n=%llu\n", (unsigned long long)n); return 1; }
        if ((uint64_t)(m.chains - 1) * m.tpc >= m.tiles) { printf("empty chain n=%llu\n", (unsigned long long)n); return 1; }
        auto check = [&](uint64_t o) {
          const uint32_t c = mgs::chain_of((uint32_t)o, m);
          ++checks;
          if (c != o / m.chain_keys || c >= m.chains) ++bad;
        };
        for (uint32_t c = 0; c <= m.chains; ++c)
          for (int64_t d = -3; d <= 3; ++d) {
            const int64_t o = (int64_t)((uint64_t)c * m.chain_keys) + d;
            if (o >= 0 && (uint64_t)o < n) check((uint64_t)o);
          }
        for (int k = 0; k < 40; ++k) check(rng() % n);
        check(0);
        check(n - 1);
      }
    }
  printf("checks %ld bad %ld\n", checks, bad);
  return bad != 0;
}
"""


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_chain_of_equals_division(tmp_path):
    src = tmp_path / "chainmap.cpp"
    exe = tmp_path / "chainmap"
    src.write_text(PROGRAM)
    inc = [f"-I{ROOT}/include", f"-I{ROOT}/multi-gpu-sorting_amd/csrc"]
    subprocess.run([HIPCC, "-O2", "-std=c++17", "--offload-host-only", *inc, str(src), "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " bad 0" in r.stdout
