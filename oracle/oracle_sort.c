/*
 * CPU oracle for the sorting hot path — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of what the reference computes on this path, used by
 * tests/ , __graft_entry__.smoke() and bench.py's CPU baseline as the
 * checker.  Nothing in the product links, loads or calls this file.
 *
 *   oracle_enc_*          key order of thrust::sort on arithmetic keys
 *                         (reference src/tanasic_sort.cuh:264-266): the
 *                         unsigned image enc() with -0.0 folded onto +0.0
 *   oracle_sort           stable ascending sort by enc() (keys or pairs):
 *                         byte-wise LSD counting sort
 *   oracle_merge          stable two-way merge, A first on ties
 *                         (thrust::merge at reference src/tanasic_sort.cuh:163)
 *   oracle_select_pivot   SelectPivot's binary search, restated step for step
 *                         (reference src/tanasic_sort.cuh:21-43)
 *   oracle_tanasic_1dev   FindPivot + SwapPartitions + MergeLocalPartitions
 *                         for two halves on one device (reference
 *                         src/tanasic_sort.cuh:89-174)
 *   oracle_segmented_sort per-segment oracle_sort
 *
 * Parity pins (tests/test_oracle.py): oracle_sort against numpy's stable
 * sort of the values themselves (integers, and floats with signed zeros,
 * infinities and denormals, where -0.0 == +0.0 keep input order), numpy's
 * stable argsort for payloads, a brute-force insertion sort, and the
 * prescribed NaN / infinity placement; oracle_merge against a stable sort of
 * the concatenation; the pivot against a brute-force scan over all k and the
 * invariant max(L') <= min(R'); tanasic_1dev and segmented_sort against the
 * sorted whole / numpy per segment.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { K_U32 = 0, K_U64 = 1, K_I32 = 2, K_I64 = 3, K_F32 = 4, K_F64 = 5 };
enum { V_NONE = 0, V_U32 = 1, V_U64 = 2 };

static int key_bytes(int kt) { return (kt == K_U64 || kt == K_I64 || kt == K_F64) ? 8 : 4; }
static int val_bytes(int vt) { return vt == V_U32 ? 4 : (vt == V_U64 ? 8 : 0); }

uint32_t oracle_enc32(int kt, uint32_t x) {
  switch (kt) {
    case K_I32: return x ^ 0x80000000u;
    case K_F32:
      if (x == 0x80000000u) x = 0u; /* -0.0 == +0.0 */
      return (x & 0x80000000u) ? ~x : (x | 0x80000000u);
    default: return x;
  }
}

uint64_t oracle_enc64(int kt, uint64_t x) {
  switch (kt) {
    case K_I64: return x ^ 0x8000000000000000ull;
    case K_F64:
      if (x == 0x8000000000000000ull) x = 0ull;
      return (x & 0x8000000000000000ull) ? ~x : (x | 0x8000000000000000ull);
    default: return x;
  }
}

static uint64_t enc_at(int kt, const void* keys, size_t i) {
  if (key_bytes(kt) == 4) return oracle_enc32(kt, ((const uint32_t*)keys)[i]);
  return oracle_enc64(kt, ((const uint64_t*)keys)[i]);
}

/* Encodes n keys into out (uint64 images). */
void oracle_encode(int kt, const void* keys, size_t n, uint64_t* out) {
  for (size_t i = 0; i < n; ++i) out[i] = enc_at(kt, keys, i);
}

/* Stable ascending sort of keys (and vals) by enc(): LSD counting sort over
 * the bytes of the encoded image.  Returns 0, or -1 on allocation failure. */
int oracle_sort(int kt, int vt, void* keys, void* vals, size_t n) {
  const int kb = key_bytes(kt), vb = val_bytes(vt);
  if (n < 2) return 0;
  uint64_t* e = (uint64_t*)malloc(n * 8);
  uint64_t* e2 = (uint64_t*)malloc(n * 8);
  size_t* idx = (size_t*)malloc(n * sizeof(size_t));
  size_t* idx2 = (size_t*)malloc(n * sizeof(size_t));
  if (!e || !e2 || !idx || !idx2) {
    free(e); free(e2); free(idx); free(idx2);
    return -1;
  }
  oracle_encode(kt, keys, n, e);
  for (size_t i = 0; i < n; ++i) idx[i] = i;
  for (int byte = 0; byte < kb; ++byte) {
    size_t count[257];
    memset(count, 0, sizeof(count));
    const int sh = 8 * byte;
    for (size_t i = 0; i < n; ++i) count[((e[i] >> sh) & 0xFF) + 1]++;
    for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
    for (size_t i = 0; i < n; ++i) {
      const size_t pos = count[(e[i] >> sh) & 0xFF]++;
      e2[pos] = e[i];
      idx2[pos] = idx[i];
    }
    uint64_t* te = e; e = e2; e2 = te;
    size_t* ti = idx; idx = idx2; idx2 = ti;
  }
  /* apply the permutation */
  char* kcopy = (char*)malloc(n * kb);
  char* vcopy = vb ? (char*)malloc(n * vb) : NULL;
  if (!kcopy || (vb && !vcopy)) {
    free(e); free(e2); free(idx); free(idx2); free(kcopy); free(vcopy);
    return -1;
  }
  memcpy(kcopy, keys, n * kb);
  if (vb) memcpy(vcopy, vals, n * vb);
  for (size_t i = 0; i < n; ++i) {
    memcpy((char*)keys + i * kb, kcopy + idx[i] * kb, kb);
    if (vb) memcpy((char*)vals + i * vb, vcopy + idx[i] * vb, vb);
  }
  free(kcopy); free(vcopy); free(e); free(e2); free(idx); free(idx2);
  return 0;
}

/* Stable merge of sorted A and B (A first on equal keys). */
void oracle_merge(int kt, int vt, const void* ak, const void* av, size_t na, const void* bk, const void* bv,
                  size_t nb, void* ok, void* ov) {
  const int kb = key_bytes(kt), vb = val_bytes(vt);
  size_t i = 0, j = 0, o = 0;
  while (i < na || j < nb) {
    int take_a;
    if (i >= na) take_a = 0;
    else if (j >= nb) take_a = 1;
    else take_a = enc_at(kt, ak, i) <= enc_at(kt, bk, j);
    if (take_a) {
      memcpy((char*)ok + o * kb, (const char*)ak + i * kb, kb);
      if (vb) memcpy((char*)ov + o * vb, (const char*)av + i * vb, vb);
      ++i;
    } else {
      memcpy((char*)ok + o * kb, (const char*)bk + j * kb, kb);
      if (vb) memcpy((char*)ov + o * vb, (const char*)bv + j * vb, vb);
      ++j;
    }
    ++o;
  }
}

/* SelectPivot (reference src/tanasic_sort.cuh:21-43) restated: L and R are
 * the virtual concatenations of num_parts partitions of part_size keys each
 * (reference GetValueFromVirtualPartition, :15-19); comparisons use enc(). */
uint64_t oracle_select_pivot(int kt, const void* const* left, const void* const* right, size_t part_size,
                             size_t num_parts) {
  const int kb = key_bytes(kt);
  uint64_t low = 0;
  uint64_t high = (uint64_t)part_size * num_parts;
  const uint64_t total = high;
  while (low < high) {
    const uint64_t mid = high - (high - low) / 2;
    const uint64_t ia = total - mid, ib = mid - 1;
    const void* pa = (const char*)left[ia / part_size] + (ia % part_size) * kb;
    const void* pb = (const char*)right[ib / part_size] + (ib % part_size) * kb;
    const uint64_t a = enc_at(kt, pa, 0), b = enc_at(kt, pb, 0);
    if (a <= b)
      high = mid - 1;
    else
      low = mid;
  }
  return low;
}

/* One-device Tanasic merge of two sorted halves L = keys[0, m) and
 * R = keys[m, 2m): pivot p, swap L's last p with R's first p (reference
 * SwapPartitions, src/tanasic_sort.cuh:89-146), then merge each half at
 * offsets m - p and p (MergeLocalPartitions, :148-174).  Result: keys[0, 2m)
 * sorted.  Returns p. */
uint64_t oracle_tanasic_1dev(int kt, int vt, void* keys, void* vals, size_t m) {
  const int kb = key_bytes(kt), vb = val_bytes(vt);
  const void* lp[1] = {keys};
  const void* rp[1] = {(const char*)keys + m * kb};
  const uint64_t p = oracle_select_pivot(kt, lp, rp, m, 1);
  char* L = (char*)keys;
  char* R = (char*)keys + m * kb;
  char* Lv = vb ? (char*)vals : NULL;
  char* Rv = vb ? (char*)vals + m * vb : NULL;
  char tmp[8];
  for (uint64_t i = 0; i < p; ++i) {
    memcpy(tmp, L + (m - p + i) * kb, kb);
    memcpy(L + (m - p + i) * kb, R + i * kb, kb);
    memcpy(R + i * kb, tmp, kb);
    if (vb) {
      memcpy(tmp, Lv + (m - p + i) * vb, vb);
      memcpy(Lv + (m - p + i) * vb, Rv + i * vb, vb);
      memcpy(Rv + i * vb, tmp, vb);
    }
  }
  if (p > 0) {
    char* ok = (char*)malloc(m * kb + 1);
    char* ov = vb ? (char*)malloc(m * vb + 1) : NULL;
    /* left: merge [0, m-p) and [m-p, m) */
    oracle_merge(kt, vt, L, Lv, m - p, L + (m - p) * kb, vb ? Lv + (m - p) * vb : NULL, p, ok, ov);
    memcpy(L, ok, m * kb);
    if (vb) memcpy(Lv, ov, m * vb);
    /* right: merge [0, p) and [p, m) */
    oracle_merge(kt, vt, R, Rv, p, R + p * kb, vb ? Rv + p * vb : NULL, m - p, ok, ov);
    memcpy(R, ok, m * kb);
    if (vb) memcpy(Rv, ov, m * vb);
    free(ok);
    free(ov);
  }
  return p;
}

/* Sorts each segment [off[s], off[s+1]) independently. */
int oracle_segmented_sort(int kt, int vt, void* keys, void* vals, const uint64_t* off, size_t nseg) {
  const int kb = key_bytes(kt), vb = val_bytes(vt);
  for (size_t s = 0; s < nseg; ++s) {
    const uint64_t b = off[s], e = off[s + 1];
    if (oracle_sort(kt, vt, (char*)keys + b * kb, vb ? (char*)vals + b * vb : NULL, (size_t)(e - b)) != 0)
      return -1;
  }
  return 0;
}
