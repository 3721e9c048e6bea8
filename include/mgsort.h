/*
 * mgsort.h — C ABI of the MI355X (gfx950) sorting library libmgsort.so.
 *
 * This is the drop-in boundary that replaces the device primitives the
 * reference calls through thrust/CUB/moderngpu and its own SelectPivot kernel.
 * Every entry point is stream-ordered and asynchronous: it enqueues HIP kernels
 * on `stream` (a hipStream_t passed as void*, NULL = the legacy default stream)
 * and returns.  The caller selects the device with hipSetDevice() beforehand
 * and owns every buffer; the library never allocates device memory on these
 * paths — scratch comes from the caller's `temp` buffer, whose size is given by
 * the matching *_temp_bytes() query (the two-call pattern of
 * cub::DeviceRadixSort::SortKeys, reference
 * src/local_gpu_sort_algorithm_benchmark.cu:53-65).
 *
 * Key order ("encoded order").  Keys are ordered by an unsigned image enc(x):
 *   unsigned: x;  signed: x with the sign bit flipped;
 *   float/double: bits | sign-bit if the sign is clear, ~bits if it is set,
 *   after mapping -0.0 onto +0.0.
 * Sorts are stable ascending in enc order.  For NaN-free input this is the
 * order of std::stable_sort(less) and of thrust::sort on arithmetic keys
 * (reference src/tanasic_sort.cuh:264-266); -NaN sorts first, +NaN last,
 * -0.0/+0.0 keep their input order, denormals sort by value.
 *
 * Status codes: 0 on success, a negative MGS_ERR_* otherwise (no exceptions,
 * no exit(); mgs_status_str() gives the text).  Unlike the reference, whose
 * CheckCudaError is a no-op in release builds (src/cuda_error.cuh:5-10), every
 * HIP launch is checked.
 */
#ifndef MGSORT_H_
#define MGSORT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Key types (the reference instantiates int, long, float, double:
 * src/sort_benchmark.cu:209-221; u32/u64 are added for the MI355X configs). */
typedef enum mgs_key {
  MGS_KEY_U32 = 0,
  MGS_KEY_U64 = 1,
  MGS_KEY_I32 = 2,
  MGS_KEY_I64 = 3,
  MGS_KEY_F32 = 4,
  MGS_KEY_F64 = 5
} mgs_key;

/* Payload types for key-value sorts (none in the reference). */
typedef enum mgs_val {
  MGS_VAL_NONE = 0,
  MGS_VAL_U32 = 1,
  MGS_VAL_U64 = 2
} mgs_val;

enum {
  MGS_OK = 0,
  MGS_ERR_BAD_TYPE = -1,       /* unknown mgs_key / mgs_val                   */
  MGS_ERR_TEMP_TOO_SMALL = -2, /* temp_bytes below the *_temp_bytes() answer  */
  MGS_ERR_HIP = -3,            /* a HIP call failed; see mgs_last_hip_error() */
  MGS_ERR_INVALID = -4,        /* bad argument (NULL pointer, size, overlap)  */
  MGS_ERR_TOO_LARGE = -5,      /* n above MGS_MAX_KEYS                        */
  MGS_ERR_DEVICE = -6,         /* a kernel gave up a bounded wait or found a  */
                               /* look-back chain not fully served: the       */
                               /* output is not valid (mgs_temp_status)       */
  MGS_ERR_SKIPPED = -7,        /* segments left unsorted: malformed bounds,  */
                               /* beyond max_total_keys or max_segment_keys   */
                               /* (mgs_temp_status)                           */
  MGS_ERR_UNSORTED = -8        /* mgs_multiway_merge: input runs not sorted   */
                               /* (mgs_temp_status)                           */
};

/* Largest key count of one call: the kernels index keys with 32 bits and
 * the last tile (at most 65536 keys) must end below 2^32. */
#define MGS_MAX_KEYS ((size_t)0xFFFF0000u)

/* Tuning of the segmented sort; a zeroed struct (or NULL) selects the
 * defaults chosen by measurement on MI355X. */
typedef struct mgs_tuning {
  int32_t large_variant;    /* tile geometry of the large segments' passes: */
                            /* 0 = default, 0 <= v <                        */
                            /* mgs_segmented_variant_count(), see           */
                            /* mgs_segmented_variant(); others are          */
                            /* MGS_ERR_INVALID                              */
  int32_t small_max;        /* segments up to this length sort inside one   */
                            /* workgroup's LDS (0 = default)                */
  int32_t medium_max;       /* segments up to this length sort with a single*/
                            /* workgroup streaming HBM (0 = default)        */
  int32_t max_segment_keys; /* longest segment the call accepts (0 = any up */
                            /* to max_total_keys); below 2^30 the look-back */
                            /* descriptors take 4 bytes instead of 8        */
} mgs_tuning;

/* --------------------------------------------------------------------------
 * Radix sort.  Replaces thrust::sort(par(alloc).on(stream), b, e) at
 * reference src/tanasic_sort.cuh:264-266, src/gowanlock_sort.cuh:188-190 and
 * :337-340, src/sort_benchmark.cu:89, and cub::DeviceRadixSort::SortKeys at
 * src/local_gpu_sort_algorithm_benchmark.cu:55-65.
 *
 * Sorts n keys (and, if vt != MGS_VAL_NONE, their payloads) stably by enc().
 * keys_in may equal keys_out (in place, thrust::sort semantics) and vals_in
 * may equal vals_out; otherwise the inputs are only read.  n <= MGS_MAX_KEYS.
 *
 * Scratch: mgs_sort_temp_bytes(kt, vt, n) <= n * (sizeof key + sizeof payload)
 * + 128,000,000 for every n <= MGS_MAX_KEYS (the reference's device buffer
 * plan, src/device_buffers.cuh:30-44), and it does not decrease with n, so a
 * buffer sized for n also serves every smaller sort.
 * ------------------------------------------------------------------------ */
size_t mgs_sort_temp_bytes(mgs_key kt, mgs_val vt, size_t n);

int mgs_sort(mgs_key kt, mgs_val vt,
             const void* keys_in, void* keys_out,
             const void* vals_in, void* vals_out,
             size_t n, void* temp, size_t temp_bytes, void* stream);

/* mgs_sort without a temp buffer (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops).  Same kernels as mgs_sort; the scratch is a 64 MiB
 * array in the library's device memory, one per GPU, held until the process
 * ends, so n is limited to mgs_sort_temp_bytes(kt, vt, n) <= 64 MiB (2^23 - 1
 * u32 keys, about 8.3M u64 keys or u32 pairs, 5.5M u64 keys with u32
 * payloads); larger n returns MGS_ERR_TOO_LARGE.  Allocates nothing.  All
 * calls on one GPU share its array: calls that overlap on one GPU (on
 * different streams or host threads) race and must be ordered by the
 * caller. */
int mgs_sort_noalloc(mgs_key kt, mgs_val vt,
                     const void* keys_in, void* keys_out,
                     const void* vals_in, void* vals_out,
                     size_t n, void* stream);

/* Measurement variant of mgs_sort (no reference counterpart; used by the
 * benchmark to time each kernel on the stream it runs on): enqueues the same
 * kernels and records events[i] (hipEvent_t handles created by the caller)
 * before the i-th launch and one after the last, up to num_events.
 * *num_marks receives the number of marks the sequence has (stages + 1), so
 * a first call with num_events = 0 sizes the array. */
int mgs_sort_timed(mgs_key kt, mgs_val vt,
                   const void* keys_in, void* keys_out,
                   const void* vals_in, void* vals_out,
                   size_t n, void* temp, size_t temp_bytes, void* stream,
                   void* const* events, int num_events, int* num_marks);

/* Which algorithm the last mgs_sort / mgs_sort_timed call with these
 * arguments and this temp buffer used (no reference counterpart): waits for
 * `stream` and returns MGS_PATH_SMALL (one workgroup in LDS, no scratch),
 * MGS_PATH_LSD (onesweep LSD passes) or MGS_PATH_THREE_PASS (the MSD pass,
 * the bucket pass and the LDS sort of sub-buckets, tried for 4-byte keys
 * alone of 2^23 to 3 * 2^24 keys and of 2^29 keys and more; it hands over to
 * the LSD passes when the key distribution does not suit it), or a negative
 * status. */
enum { MGS_PATH_SMALL = 1, MGS_PATH_LSD = 2, MGS_PATH_THREE_PASS = 3 };
int mgs_sort_path(mgs_key kt, mgs_val vt, size_t n, const void* temp, size_t temp_bytes, void* stream);

/* Comma-separated names of the stages of this thread's last mgs_sort_timed
 * call, one per mark but the last (no reference counterpart). */
const char* mgs_sort_timed_stages(void);

/* --------------------------------------------------------------------------
 * Merge-path merge.  Replaces thrust::merge at reference
 * src/tanasic_sort.cuh:163-167 and src/local_gpu_merge_algorithm_benchmark.cu
 * :157-158 (and mgpu::merge at :165-167).  Stable: on equal keys the element
 * of A comes first.  Output must not overlap the inputs.
 * ------------------------------------------------------------------------ */
size_t mgs_merge_temp_bytes(mgs_key kt, mgs_val vt, size_t na, size_t nb);

int mgs_merge(mgs_key kt, mgs_val vt,
              const void* a_keys, const void* a_vals, size_t na,
              const void* b_keys, const void* b_vals, size_t nb,
              void* out_keys, void* out_vals,
              void* temp, size_t temp_bytes, void* stream);

/* --------------------------------------------------------------------------
 * k-way merge.  The device counterpart of __gnu_parallel::multiway_merge of
 * sorted runs at reference src/gowanlock_sort.cuh:91-92 (MergeChunks and
 * MergeMergeGroups), for runs that sit in device memory.  runs and run_lens
 * are HOST arrays of k <= MGS_MULTIWAY_MAX_RUNS device pointers and lengths;
 * every run sorted in enc order.  Writes the stable merge (equal keys in run
 * order, then index order) of all runs to out, which must not overlap them.
 * Runs that are not sorted give an unspecified order and may set
 * MGS_ERR_UNSORTED in the status word (mgs_temp_status).  Keys only.
 * Scratch: mgs_multiway_merge_temp_bytes(kt, k, sum of run_lens).  The call
 * copies the run table from host memory, so it is not for graph capture.
 * ------------------------------------------------------------------------ */
#define MGS_MULTIWAY_MAX_RUNS 256
size_t mgs_multiway_merge_temp_bytes(mgs_key kt, size_t k, size_t total_keys);

int mgs_multiway_merge(mgs_key kt, const void* const* runs, const size_t* run_lens, size_t k,
                       void* out, void* temp, size_t temp_bytes, void* stream);

/* The same merge in `portions` consecutive stretches of the output, for
 * callers that move the output on while the rest merges (the heterogeneous
 * sort's copy back to the host).  Waits on the host once, after the
 * splitters, for the stretch boundaries: portion_end[p] is the output index
 * where stretch p ends, and portion_done[p] (an event of the caller,
 * hipEvent_t, or portion_done NULL) is recorded on `stream` once stretch p is
 * written. */
int mgs_multiway_merge_portions(mgs_key kt, const void* const* runs, const size_t* run_lens, size_t k,
                                void* out, void* temp, size_t temp_bytes, void* stream,
                                size_t portions, size_t* portion_end, void* const* portion_done);

/* --------------------------------------------------------------------------
 * Tanasic pivot.  Replaces SelectPivot/FindPivot at reference
 * src/tanasic_sort.cuh:15-87.  left_parts/right_parts are HOST arrays of
 * num_parts DEVICE pointers (num_parts <= 8), each partition part_size keys
 * sorted in enc order; L and R are their concatenations.  Writes
 *   p = max k in [0, M] with k == 0 or enc(L[M-k]) > enc(R[k-1]),
 *   M = part_size * num_parts,
 * as a uint64 to *result, which may be device memory or mapped pinned host
 * memory (as the reference's HostAllocator slot, src/host_allocator.cuh:11-12).
 * ------------------------------------------------------------------------ */
int mgs_select_pivot(mgs_key kt,
                     const void* const* left_parts, const void* const* right_parts,
                     size_t part_size, size_t num_parts,
                     uint64_t* result, void* stream);

/* --------------------------------------------------------------------------
 * Range swap.  Replaces the pairwise exchange of SwapPartitions at reference
 * src/tanasic_sort.cuh:103-143 for two partitions on one device: exchanges
 * a[0..n) with b[0..n) in place (element order kept).  The ranges must not
 * overlap.  a_vals/b_vals are ignored when vt == MGS_VAL_NONE.
 * ------------------------------------------------------------------------ */
int mgs_swap_ranges(mgs_key kt, mgs_val vt,
                    void* a_keys, void* a_vals, void* b_keys, void* b_vals,
                    size_t n, void* stream);

/* --------------------------------------------------------------------------
 * Filler.  Replaces thrust::fill(..., numeric_limits<T>::max()) at reference
 * src/tanasic_sort.cuh:256-259 and src/gowanlock_sort.cuh:180-183,329-332.
 * Writes the key whose enc() is all ones (for floats a +NaN, which sorts
 * after +inf, where the reference's FLT_MAX would not), so fillers always
 * sort last.
 * ------------------------------------------------------------------------ */
int mgs_fill_max(mgs_key kt, void* keys, size_t n, void* stream);

/* --------------------------------------------------------------------------
 * Segmented sort (no reference counterpart: the batching layer).  Sorts
 * nseg independent segments stably by enc(); segment s is
 * [offsets[s], offsets[s+1]) of keys/vals.  d_offsets is a DEVICE array of
 * nseg+1 uint64 offsets, non-decreasing, offsets[0] may be > 0.
 * Offsets must not decrease (a segment starting before an earlier offset, or
 * reaching past offsets[0] + max_total_keys, is left unsorted).
 * keys_in may equal keys_out.  `tuning` may be NULL.  One call enqueues a
 * fixed launch sequence independent of the segment lengths, so it can be
 * captured once in a hipGraph and replayed.  max_total_keys bounds
 * offsets[nseg] - offsets[0] and sizes the scratch.
 * ------------------------------------------------------------------------ */
size_t mgs_segmented_sort_temp_bytes(mgs_key kt, mgs_val vt,
                                     size_t max_total_keys, size_t nseg);

/* The same for calls that pass `tuning` (its max_segment_keys shrinks the
 * look-back descriptors, and its large_variant sets the number of large
 * tiles); NULL gives mgs_segmented_sort_temp_bytes. */
size_t mgs_segmented_sort_temp_bytes_tuned(mgs_key kt, mgs_val vt,
                                           size_t max_total_keys, size_t nseg,
                                           const mgs_tuning* tuning);

/* Tile geometries of the large segments' passes (no reference counterpart:
 * the search space of the batching layer's autotuner).  Variants are
 * 0 .. mgs_segmented_variant_count() - 1 for every (kt, vt); variant v of
 * (kt, vt) runs workgroups of *block threads with *items_per_thread keys
 * each (tiles of their product).  Returns MGS_ERR_INVALID for a variant out
 * of range. */
int mgs_segmented_variant_count(void);
int mgs_segmented_variant(mgs_key kt, mgs_val vt, int variant,
                          int32_t* block, int32_t* items_per_thread);

int mgs_segmented_sort(mgs_key kt, mgs_val vt,
                       const void* keys_in, void* keys_out,
                       const void* vals_in, void* vals_out,
                       const uint64_t* d_offsets, size_t nseg,
                       size_t max_total_keys, const mgs_tuning* tuning,
                       void* temp, size_t temp_bytes, void* stream);

/* Ranges form of the segmented sort (no reference counterpart: the serving
 * layer sorts requests that sit anywhere in its device pool).  Segment s is
 * keys[d_ranges[2s], d_ranges[2s+1]), sorted in place with its payloads
 * vals[v, v + len) where v = d_val_begin[s] (d_val_begin NULL: v =
 * d_ranges[2s]); segments must not overlap and need not be adjacent or
 * ordered.  d_ranges (2*nseg uint64) and d_val_begin (nseg uint64) are
 * DEVICE arrays of element indices; the lengths may add up to at
 * most max_total_keys (segments past that bound, or with end < begin, are left
 * unsorted).  Scratch: mgs_segmented_sort_temp_bytes(kt, vt, max_total_keys,
 * nseg).  Same fixed launch sequence as mgs_segmented_sort. */
int mgs_segmented_sort_ranges(mgs_key kt, mgs_val vt,
                              void* keys, void* vals,
                              const uint64_t* d_ranges, const uint64_t* d_val_begin,
                              size_t nseg,
                              size_t max_total_keys, const mgs_tuning* tuning,
                              void* temp, size_t temp_bytes, void* stream);

/* Device-side status of the last mgs_sort / mgs_sort_timed /
 * mgs_segmented_sort / mgs_segmented_sort_ranges call that used `temp` (no
 * reference counterpart: the reference ignores device errors,
 * src/cuda_error.cuh:5-10).  Those calls' kernels record what they detect
 * while running in the first word of temp: a look-back wait that reached its
 * bound, a look-back chain whose tiles were not all taken, or segments left
 * unsorted.  Waits for `stream`, then returns MGS_OK, MGS_ERR_DEVICE,
 * MGS_ERR_SKIPPED or MGS_ERR_HIP.  Calls that needed no scratch
 * (temp_bytes == 0) always succeed on the device: pass temp = NULL. */
int mgs_temp_status(const void* temp, size_t temp_bytes, void* stream);

/* Status text; never NULL. */
const char* mgs_status_str(int status);

/* The hipError_t of the last MGS_ERR_HIP on this thread (0 if none). */
int mgs_last_hip_error(void);

/* Library version string. */
const char* mgs_version(void);

#ifdef __cplusplus
}
#endif

#endif /* MGSORT_H_ */
