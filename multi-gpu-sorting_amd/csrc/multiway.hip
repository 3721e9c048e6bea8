// k-way merge of sorted runs on the device: the GPU counterpart of the
// reference's __gnu_parallel::multiway_merge of sorted chunks
// (src/gowanlock_sort.cuh:91-92), used by the heterogeneous sort's device
// merge backend in one pass instead of log2(k) pairwise merge rounds.
//
// Stable order: the output is the stable sort of the runs laid end to end in
// run order, i.e. the lexicographic order of (enc(key), run, index).
//
//   sample    every S-th key of every run (S = kCap / (step + k), step = 5k
//             up to 64 runs, 3k above, at least 8) with its sample id, ids
//             numbered run by run
//   sort      the samples by key, stably (mgs_sort with the ids as payload):
//             (key, run, index) order
//   split     every step-th sorted sample (v, j, i) bounds a piece: run j'
//             contributes the keys that precede (v, j, i) -- upper bound of v
//             for j' < j, lower bound for j' > j, exactly i for j' = j
//   base      each boundary's output position (the sum of its splits)
//   piece     one workgroup per piece: its slices of the k runs into LDS,
//             merged by a tree of LDS merge-path rounds (adjacent slices,
//             earlier runs first on ties), and the piece written out with
//             consecutive lanes on consecutive addresses
//
// A run contributes to a piece only the keys between two of its consecutive
// samples per sample of it inside the piece, plus one more stretch, so a
// piece holds at most (step + k) S <= kCap keys and fits the LDS.
#include <vector>

#include "common.hpp"
#include "internal.hpp"
#include "radix_sort.hpp"

namespace mgs {

namespace {

constexpr int kMwBlock = 1024;
constexpr uint32_t kMwCap = 8192;  // keys per piece at most
// sorted samples per piece: 5k up to 64 runs (at least 8), so that a piece
// holds on average step / (step + k) = 5/6 of the keys its LDS bound allows;
// 3k above 64 runs (3/4), where the samples, one per kCap / 4k keys, would
// cost more to sort than the fuller pieces save (measured: 8 runs 4% faster
// with 5k than with 3k, 256 runs 3% slower)
inline uint32_t piece_step(size_t k) { return (uint32_t)std::max<size_t>(8, (k <= 64 ? 5 : 3) * k); }
constexpr size_t kMwMaxRuns = 256;

inline uint32_t sample_stride(size_t k) { return (uint32_t)(kMwCap / (piece_step(k) + k)); }

size_t align_up(size_t x) { return (x + 255) / 256 * 256; }

struct MwLayout {
  size_t status, ptrs, lens, sbase, skeys, skeys2, sid, sid2, sort_temp, splits, bases, total;
  uint64_t ns, pieces;
};

// ns: sample count, exact when lens are known, else the bound for total keys.
MwLayout mw_layout(int kb, size_t k, uint64_t ns) {
  MwLayout L{};
  L.ns = ns;
  L.pieces = ns == 0 ? 0 : (ns - 1) / piece_step(k) + 1;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = align_up(o + bytes);
    return at;
  };
  L.status = take(4);  // first: the call's status word
  L.ptrs = take(k * 8);
  L.lens = take(k * 8);
  L.sbase = take((k + 1) * 8);
  L.skeys = take(ns * kb);
  L.skeys2 = take(ns * kb);
  L.sid = take(ns * 4);
  L.sid2 = take(ns * 4);
  L.sort_temp = take(sort_temp_bytes(kb == 4 ? MGS_KEY_U32 : MGS_KEY_U64, MGS_VAL_U32, ns));
  L.splits = take((L.pieces + 1) * k * 8);
  L.bases = take((L.pieces + 1) * 8);
  L.total = o;
  return L;
}

uint64_t sample_count(const size_t* lens, size_t k, uint32_t S) {
  uint64_t ns = 0;
  for (size_t j = 0; j < k; ++j) ns += (lens[j] + S - 1) / S;
  return ns;
}

// Sample q: run j with sbase[j] <= q < sbase[j+1], index (q - sbase[j]) S.
template <typename Bits>
__global__ void mw_sample_kernel(const Bits* const* __restrict__ runs, const uint64_t* __restrict__ sbase, uint32_t k,
                                 uint32_t S, uint64_t ns, Bits* __restrict__ skeys, uint32_t* __restrict__ sid,
                                 uint32_t* __restrict__ status) {
  const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q == 0) *status = 0;  // the kernels after this one only OR into it
  if (q >= ns) return;
  uint32_t lo = 0, hi = k;  // last j with sbase[j] <= q
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (sbase[mid] <= q)
      lo = mid;
    else
      hi = mid;
  }
  skeys[q] = runs[lo][(q - sbase[lo]) * S];
  sid[q] = (uint32_t)q;
}

// Keys of run r (length n) that precede value e in encoded order: strictly
// smaller (UPPER false) or smaller or equal (UPPER true).
template <int KT, bool UPPER, typename Bits>
__device__ __forceinline__ uint64_t bound_global(const Bits* r, uint64_t n, Bits e) {
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    const Bits m = Codec<KT>::enc(r[mid]);
    if (UPPER ? (m <= e) : (m < e))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Thread (b, j'): split of run j' at boundary b (1 <= b < pieces); boundary
// 0 and `pieces` are the runs' ends.
template <int KT, typename Bits>
__global__ void mw_split_kernel(const Bits* const* __restrict__ runs, const uint64_t* __restrict__ lens,
                                const uint64_t* __restrict__ sbase, uint32_t k, uint32_t S,
                                const Bits* __restrict__ skeys, const uint32_t* __restrict__ sid, uint64_t pieces,
                                uint32_t step, uint64_t* __restrict__ splits) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (pieces + 1) * k) return;
  const uint64_t b = t / k;
  const uint32_t jj = (uint32_t)(t % k);
  if (b == 0) {
    splits[t] = 0;
    return;
  }
  if (b == pieces) {
    splits[t] = lens[jj];
    return;
  }
  const uint64_t q = b * step;
  const Bits e = Codec<KT>::enc(skeys[q]);
  const uint32_t id = sid[q];
  uint32_t lo = 0, hi = k;  // run of sample id
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (sbase[mid] <= id)
      lo = mid;
    else
      hi = mid;
  }
  const uint32_t j = lo;
  uint64_t s;
  if (jj == j)
    s = (id - sbase[j]) * (uint64_t)S;
  else if (jj < j)
    s = bound_global<KT, true>(runs[jj], lens[jj], e);
  else
    s = bound_global<KT, false>(runs[jj], lens[jj], e);
  splits[t] = s;
}

__global__ void mw_base_kernel(const uint64_t* __restrict__ splits, uint32_t k, uint64_t pieces,
                               uint64_t* __restrict__ bases, const uint32_t* __restrict__ sort_status,
                               uint32_t* __restrict__ status) {
  const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b == 0 && *sort_status) atomicOr(status, *sort_status);  // the sample sort's own status word
  if (b > pieces) return;
  uint64_t s = 0;
  for (uint32_t j = 0; j < k; ++j) s += splits[b * k + j];
  bases[b] = s;
}

// Keys of `a` taken among the first `diag` outputs of the stable merge of
// a (na keys, taken first on ties) and b (nb keys), both in LDS.
template <int KT, typename Bits>
__device__ __forceinline__ uint32_t corank_lds(const Bits* a, uint32_t na, const Bits* b, uint32_t nb, uint32_t diag) {
  uint32_t lo = diag > nb ? diag - nb : 0;
  uint32_t hi = diag < na ? diag : na;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (Codec<KT>::enc(a[mid]) <= Codec<KT>::enc(b[diag - mid - 1]))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// One piece: its slices of the k runs are loaded into LDS back to back (run
// order), then merged in ceil(log2 k) rounds, each merging adjacent slice
// pairs (2i, 2i+1) into the other LDS buffer (earlier runs first on ties, so
// the order stays stable).  In a round every thread writes q consecutive
// outputs (q odd, q * kMwBlock >= the piece's keys, q <= kCap / kMwBlock + 1):
// one co-rank search in its pair, then a serial merge.
template <int KT, typename Bits>
__global__ __launch_bounds__(kMwBlock) void mw_piece_kernel(const Bits* const* __restrict__ runs, uint32_t k,
                                                            const uint64_t* __restrict__ splits,
                                                            const uint64_t* __restrict__ bases, uint64_t first_piece,
                                                            Bits* __restrict__ out, uint32_t* __restrict__ status) {
  // loads per thread: the piece's keys at most
  constexpr uint32_t kLoads = kMwCap / kMwBlock;
  __shared__ Bits buf[2][kMwCap];
  __shared__ uint32_t bnd[kMwMaxRuns + 1];
  __shared__ uint64_t sbeg[kMwMaxRuns];
  __shared__ const Bits* sptr[kMwMaxRuns];
  __shared__ uint32_t sums[4];
  const int tid = threadIdx.x;
  const uint64_t b = first_piece + blockIdx.x;
  // slice j of the piece: runs[j][splits[b][j], splits[b+1][j]); slices that
  // go backwards or more than kCap keys in all (only unsorted runs break the
  // sample bound) leave the piece unmerged and the call reports it;
  // otherwise the piece's keys are exactly [bases[b], bases[b+1]) of the
  // output
  uint32_t len = 0;
  bool bad = false;
  if ((uint32_t)tid < k) {
    const uint64_t lo = splits[b * k + tid], hi = splits[(b + 1) * k + tid];
    bad = hi < lo;
    len = bad ? 0u : (uint32_t)std::min<uint64_t>(hi - lo, kMwCap + 1);
    sbeg[tid] = lo;
    sptr[tid] = runs[tid];
  }
  const uint32_t excl = block_excl_scan_256((uint32_t)tid < k ? len : 0u, sums, tid);
  if ((uint32_t)tid < k) {
    bnd[tid] = min(excl, kMwCap);
    if ((uint32_t)tid == k - 1) {
      bad = bad || excl + len > kMwCap;
      bnd[k] = min(excl + len, kMwCap);
    }
  }
  if (__syncthreads_or(bad ? 1 : 0)) {
    if (tid == 0) atomicOr(status, kStatusUnsorted);
    return;
  }
  const uint32_t cnt = bnd[k];
  // outputs per thread in a merge round, odd so that the lanes' serial
  // merges start in different LDS banks: with several rounds the fewest that
  // cover the piece (a piece of 6800 keys, the average for 8 runs, takes 7
  // per thread);
  // a single round (k <= 2) keeps kCap / kMwBlock + 1, which leaves fewer
  // co-rank searches (measured faster there, slower with more rounds)
  const uint32_t q = k <= 2 ? kMwCap / kMwBlock + 1 : ((cnt + kMwBlock - 1) / kMwBlock) | 1u;
  {
    // every load of the piece issued before the first LDS store
    Bits v[kLoads];
#pragma unroll
    for (uint32_t i = 0; i < kLoads; ++i) {
      const uint32_t e = (uint32_t)tid + i * kMwBlock;
      v[i] = Bits(0);
      if (e < cnt) {
        uint32_t lo = 0, hi = k;  // the slice of slot e: last j with bnd[j] <= e
        while (hi - lo > 1) {
          const uint32_t mid = (lo + hi) >> 1;
          if (bnd[mid] <= e)
            lo = mid;
          else
            hi = mid;
        }
        v[i] = sptr[lo][sbeg[lo] + (e - bnd[lo])];
      }
    }
#pragma unroll
    for (uint32_t i = 0; i < kLoads; ++i) {
      const uint32_t e = (uint32_t)tid + i * kMwBlock;
      if (e < cnt) buf[0][e] = v[i];
    }
  }
  int cur = 0;
  for (uint32_t slices = k; slices > 1; slices = (slices + 1) >> 1) {
    __syncthreads();
    const Bits* src = buf[cur];
    Bits* dst = buf[cur ^ 1];
    const uint32_t pairs = (slices + 1) >> 1;
    uint32_t o = (uint32_t)tid * q;
    const uint32_t end = min(o + q, cnt);
    while (o < end) {
      // the last pair i with bnd[2i] <= o: it ends beyond o (a later pair
      // starts beyond o, or it is the last pair, which ends at cnt)
      uint32_t lo = 0, hi = pairs;
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (bnd[2 * mid] <= o)
          lo = mid;
        else
          hi = mid;
      }
      const uint32_t a0 = bnd[2 * lo], a1 = bnd[min(2 * lo + 1, slices)], b1 = bnd[min(2 * lo + 2, slices)];
      const uint32_t na = a1 - a0, nb = b1 - a1;
      const uint32_t pend = min(end, b1);
      uint32_t ia = corank_lds<KT>(src + a0, na, src + a1, nb, o - a0);
      uint32_t ib = (o - a0) - ia;
      // the heads of both sides in registers: one LDS load per output
      Bits ka = ia < na ? src[a0 + ia] : Bits(0);
      Bits kb = ib < nb ? src[a1 + ib] : Bits(0);
      for (; o < pend; ++o) {  // branch-free: one LDS store and one load per output
        const bool take_a = ib >= nb || (ia < na && Codec<KT>::enc(ka) <= Codec<KT>::enc(kb));
        dst[o] = take_a ? ka : kb;
        ia += take_a ? 1u : 0u;
        ib += take_a ? 0u : 1u;
        const uint32_t at = take_a ? a0 + ia : a1 + ib;  // the taken side's next key (past its end: ignored)
        const Bits nk = src[min(at, kMwCap - 1)];
        ka = take_a ? nk : ka;
        kb = take_a ? kb : nk;
      }
    }
    __syncthreads();
    // boundaries of the merged slices: slice i starts where pair i did
    uint32_t nb_val = 0;
    const uint32_t nslices = pairs;
    if ((uint32_t)tid <= nslices) nb_val = bnd[min(2 * (uint32_t)tid, slices)];
    __syncthreads();
    if ((uint32_t)tid <= nslices) bnd[tid] = nb_val;
    cur ^= 1;
  }
  __syncthreads();
  Bits* dst = out + bases[b];
  for (uint32_t e = tid; e < cnt; e += kMwBlock) dst[e] = buf[cur][e];
}

uint32_t grid_of(uint64_t threads, int block) { return (uint32_t)((threads + block - 1) / block); }

template <int KT>
int mw_launch(const void* const* runs, const size_t* lens, size_t k, void* out, void* temp, size_t temp_bytes,
              hipStream_t s, size_t portions, size_t* portion_end, hipEvent_t* portion_done) {
  using Bits = typename Codec<KT>::Bits;
  constexpr int KB = sizeof(Bits);
  const uint32_t S = sample_stride(k);
  const uint64_t ns = sample_count(lens, k, S);
  if (ns > (uint64_t)MGS_MAX_KEYS) return MGS_ERR_TOO_LARGE;  // sample ids and the sample sort
  const MwLayout L = mw_layout(KB, k, ns);
  if (temp == nullptr || temp_bytes < L.total) return MGS_ERR_TEMP_TOO_SMALL;
  if (ns == 0) {  // no keys: nothing to merge, the status word says ok
    hipError_t z = hipMemsetAsync(temp, 0, 4, s);
    for (size_t p = 0; z == hipSuccess && p < portions; ++p) {
      portion_end[p] = 0;
      if (portion_done) z = hipEventRecord(portion_done[p], s);
    }
    return record_hip(z);
  }
  char* t = static_cast<char*>(temp);
  // run table: host arrays copied into temp (pageable sources; the copy
  // has read them when hipMemcpyAsync returns)
  std::vector<uint64_t> h(3 * k + 1);
  uint64_t acc = 0;
  for (size_t j = 0; j < k; ++j) {
    h[j] = reinterpret_cast<uint64_t>(runs[j]);
    h[k + j] = lens[j];
    h[2 * k + j] = acc;
    acc += (lens[j] + S - 1) / S;
  }
  h[3 * k] = acc;
  hipError_t e = hipMemcpyAsync(t + L.ptrs, h.data(), k * 8, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(t + L.lens, h.data() + k, k * 8, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(t + L.sbase, h.data() + 2 * k, (k + 1) * 8, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return record_hip(e);
  const Bits* const* d_runs = reinterpret_cast<const Bits* const*>(t + L.ptrs);
  const uint64_t* d_lens = reinterpret_cast<const uint64_t*>(t + L.lens);
  const uint64_t* d_sbase = reinterpret_cast<const uint64_t*>(t + L.sbase);
  Bits* sk = reinterpret_cast<Bits*>(t + L.skeys);
  Bits* sk2 = reinterpret_cast<Bits*>(t + L.skeys2);
  uint32_t* sid = reinterpret_cast<uint32_t*>(t + L.sid);
  uint32_t* sid2 = reinterpret_cast<uint32_t*>(t + L.sid2);
  uint64_t* splits = reinterpret_cast<uint64_t*>(t + L.splits);
  uint64_t* bases = reinterpret_cast<uint64_t*>(t + L.bases);
  uint32_t* status = reinterpret_cast<uint32_t*>(t + L.status);
  const uint32_t* sort_status = reinterpret_cast<const uint32_t*>(t + L.sort_temp);  // first word of its scratch

  mw_sample_kernel<Bits><<<grid_of(ns, 256), 256, 0, s>>>(d_runs, d_sbase, (uint32_t)k, S, ns, sk, sid, status);
  int rc = check_launch();
  if (rc != MGS_OK) return rc;
  rc = sort_dispatch((mgs_key)KT, MGS_VAL_U32, sk, sk2, sid, sid2, ns, t + L.sort_temp,
                     sort_temp_bytes((mgs_key)KT, MGS_VAL_U32, ns), s);
  if (rc != MGS_OK) return rc;
  mw_split_kernel<KT, Bits><<<grid_of((L.pieces + 1) * k, 256), 256, 0, s>>>(d_runs, d_lens, d_sbase, (uint32_t)k, S,
                                                                            sk2, sid2, L.pieces, piece_step(k),
                                                                            splits);
  mw_base_kernel<<<grid_of(L.pieces + 1, 256), 256, 0, s>>>(splits, (uint32_t)k, L.pieces, bases, sort_status,
                                                            status);
  rc = check_launch();
  if (rc != MGS_OK) return rc;

  Bits* o = static_cast<Bits*>(out);
  if (portions <= 1) {
    mw_piece_kernel<KT, Bits><<<(uint32_t)L.pieces, kMwBlock, 0, s>>>(d_runs, (uint32_t)k, splits, bases, 0, o, status);
    rc = check_launch();
    if (rc == MGS_OK && portions == 1) {
      size_t n = 0;
      for (size_t j = 0; j < k; ++j) n += lens[j];
      portion_end[0] = n;
      if (portion_done) rc = record_hip(hipEventRecord(portion_done[0], s));
    }
    return rc;
  }
  // portions: pieces [P p / portions, P (p+1) / portions); their output ends
  // come from the bases, read back once
  std::vector<uint64_t> first(portions + 1);
  for (size_t p = 0; p <= portions; ++p) first[p] = L.pieces * p / portions;
  std::vector<uint64_t> hb(L.pieces + 1);
  e = hipMemcpyAsync(hb.data(), bases, (L.pieces + 1) * 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return record_hip(e);
  for (size_t p = 0; p < portions; ++p) {
    const uint64_t np = first[p + 1] - first[p];
    if (np > 0)
      mw_piece_kernel<KT, Bits><<<(uint32_t)np, kMwBlock, 0, s>>>(d_runs, (uint32_t)k, splits, bases, first[p], o,
                                                                  status);
    rc = check_launch();
    if (rc != MGS_OK) return rc;
    portion_end[p] = (size_t)hb[first[p + 1]];
    if (portion_done) {
      rc = record_hip(hipEventRecord(portion_done[p], s));
      if (rc != MGS_OK) return rc;
    }
  }
  return MGS_OK;
}

template <int KB>
size_t mw_bytes(size_t k, size_t total) {
  if (k == 0) return 256;
  const uint32_t S = sample_stride(k);
  return mw_layout(KB, k, total / S + k).total;
}

}  // namespace

size_t multiway_temp_bytes(mgs_key kt, size_t k, size_t total) {
  const int kb = key_bytes(kt);
  if (kb < 0 || k > kMwMaxRuns) return 0;
  return kb == 4 ? mw_bytes<4>(k, total) : mw_bytes<8>(k, total);
}

int multiway_dispatch(mgs_key kt, const void* const* runs, const size_t* lens, size_t k, void* out, void* temp,
                      size_t temp_bytes, hipStream_t s, size_t portions, size_t* portion_end,
                      hipEvent_t* portion_done) {
  if (k > kMwMaxRuns) return MGS_ERR_INVALID;
  switch (kt) {
    case MGS_KEY_U32: return mw_launch<MGS_KEY_U32>(runs, lens, k, out, temp, temp_bytes, s, portions, portion_end, portion_done);
    case MGS_KEY_U64: return mw_launch<MGS_KEY_U64>(runs, lens, k, out, temp, temp_bytes, s, portions, portion_end, portion_done);
    case MGS_KEY_I32: return mw_launch<MGS_KEY_I32>(runs, lens, k, out, temp, temp_bytes, s, portions, portion_end, portion_done);
    case MGS_KEY_I64: return mw_launch<MGS_KEY_I64>(runs, lens, k, out, temp, temp_bytes, s, portions, portion_end, portion_done);
    case MGS_KEY_F32: return mw_launch<MGS_KEY_F32>(runs, lens, k, out, temp, temp_bytes, s, portions, portion_end, portion_done);
    case MGS_KEY_F64: return mw_launch<MGS_KEY_F64>(runs, lens, k, out, temp, temp_bytes, s, portions, portion_end, portion_done);
  }
  return MGS_ERR_BAD_TYPE;
}

}  // namespace mgs
