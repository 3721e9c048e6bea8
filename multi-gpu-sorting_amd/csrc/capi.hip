// extern "C" entry points of libmgsort (declared in include/mgsort.h).
#include "internal.hpp"
#include "radix_sort.hpp"

namespace mgs {

namespace {
thread_local int g_last_hip_error = 0;
}

int record_hip(hipError_t e) {
  if (e == hipSuccess) return MGS_OK;
  g_last_hip_error = (int)e;
  return MGS_ERR_HIP;
}

int last_hip_error() { return g_last_hip_error; }

}  // namespace mgs

using mgs::key_bytes;
using mgs::val_bytes;

namespace {
inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }
constexpr size_t kMaxKeys = MGS_MAX_KEYS;  // kernels index with 32 bits, tiles of <= 65536 keys
static_assert(kMaxKeys + 65536 <= (size_t(1) << 32), "the last tile must end below 2^32");
}  // namespace

extern "C" {

size_t mgs_sort_temp_bytes(mgs_key kt, mgs_val vt, size_t n) { return mgs::sort_temp_bytes(kt, vt, n); }

int mgs_sort(mgs_key kt, mgs_val vt, const void* keys_in, void* keys_out, const void* vals_in, void* vals_out,
             size_t n, void* temp, size_t temp_bytes, void* stream) {
  if (key_bytes(kt) < 0 || val_bytes(vt) < 0) return MGS_ERR_BAD_TYPE;
  if (n == 0) return MGS_OK;
  if (n > kMaxKeys) return MGS_ERR_TOO_LARGE;
  if (keys_in == nullptr || keys_out == nullptr) return MGS_ERR_INVALID;
  if (vt != MGS_VAL_NONE && (vals_in == nullptr || vals_out == nullptr)) return MGS_ERR_INVALID;
  return mgs::sort_dispatch(kt, vt, keys_in, keys_out, vals_in, vals_out, n, temp, temp_bytes, as_stream(stream));
}

int mgs_sort_noalloc(mgs_key kt, mgs_val vt, const void* keys_in, void* keys_out, const void* vals_in, void* vals_out,
                     size_t n, void* stream) {
  if (key_bytes(kt) < 0 || val_bytes(vt) < 0) return MGS_ERR_BAD_TYPE;
  if (n == 0) return MGS_OK;
  if (n > kMaxKeys) return MGS_ERR_TOO_LARGE;
  if (keys_in == nullptr || keys_out == nullptr) return MGS_ERR_INVALID;
  if (vt != MGS_VAL_NONE && (vals_in == nullptr || vals_out == nullptr)) return MGS_ERR_INVALID;
  return mgs::sort_noalloc(kt, vt, keys_in, keys_out, vals_in, vals_out, n, as_stream(stream));
}

int mgs_sort_timed(mgs_key kt, mgs_val vt, const void* keys_in, void* keys_out, const void* vals_in, void* vals_out,
                   size_t n, void* temp, size_t temp_bytes, void* stream, void* const* events, int num_events,
                   int* num_marks) {
  if (key_bytes(kt) < 0 || val_bytes(vt) < 0) return MGS_ERR_BAD_TYPE;
  if (n == 0) {
    if (num_marks) *num_marks = 0;
    return MGS_OK;
  }
  if (n > kMaxKeys) return MGS_ERR_TOO_LARGE;
  if (keys_in == nullptr || keys_out == nullptr || (events == nullptr && num_events > 0)) return MGS_ERR_INVALID;
  if (vt != MGS_VAL_NONE && (vals_in == nullptr || vals_out == nullptr)) return MGS_ERR_INVALID;
  return mgs::sort_dispatch(kt, vt, keys_in, keys_out, vals_in, vals_out, n, temp, temp_bytes, as_stream(stream),
                            reinterpret_cast<hipEvent_t*>(const_cast<void**>(events)), num_events, num_marks);
}

const char* mgs_sort_timed_stages(void) { return mgs::last_stage_names(); }

int mgs_sort_path(mgs_key kt, mgs_val vt, size_t n, const void* temp, size_t temp_bytes, void* stream) {
  return mgs::sort_path(kt, vt, n, temp, temp_bytes, as_stream(stream));
}

size_t mgs_merge_temp_bytes(mgs_key kt, mgs_val vt, size_t na, size_t nb) {
  return mgs::merge_temp_bytes(kt, vt, na, nb);
}

int mgs_merge(mgs_key kt, mgs_val vt, const void* a_keys, const void* a_vals, size_t na, const void* b_keys,
              const void* b_vals, size_t nb, void* out_keys, void* out_vals, void* temp, size_t temp_bytes,
              void* stream) {
  if (key_bytes(kt) < 0 || val_bytes(vt) < 0) return MGS_ERR_BAD_TYPE;
  if (na + nb == 0) return MGS_OK;
  if ((na > 0 && a_keys == nullptr) || (nb > 0 && b_keys == nullptr) || out_keys == nullptr) return MGS_ERR_INVALID;
  if (vt != MGS_VAL_NONE &&
      ((na > 0 && a_vals == nullptr) || (nb > 0 && b_vals == nullptr) || out_vals == nullptr))
    return MGS_ERR_INVALID;
  return mgs::merge_dispatch(kt, vt, a_keys, a_vals, na, b_keys, b_vals, nb, out_keys, out_vals, temp, temp_bytes,
                             as_stream(stream));
}

int mgs_select_pivot(mgs_key kt, const void* const* left_parts, const void* const* right_parts, size_t part_size,
                     size_t num_parts, uint64_t* result, void* stream) {
  if (key_bytes(kt) < 0) return MGS_ERR_BAD_TYPE;
  return mgs::pivot_dispatch(kt, left_parts, right_parts, part_size, num_parts, result, as_stream(stream));
}

int mgs_swap_ranges(mgs_key kt, mgs_val vt, void* a_keys, void* a_vals, void* b_keys, void* b_vals, size_t n,
                    void* stream) {
  return mgs::swap_dispatch(kt, vt, a_keys, a_vals, b_keys, b_vals, n, as_stream(stream));
}

size_t mgs_multiway_merge_temp_bytes(mgs_key kt, size_t k, size_t total_keys) {
  return mgs::multiway_temp_bytes(kt, k, total_keys);
}

static int multiway_checks(mgs_key kt, const void* const* runs, const size_t* run_lens, size_t k, const void* out,
                           size_t* total) {
  if (key_bytes(kt) < 0) return MGS_ERR_BAD_TYPE;
  if (k > MGS_MULTIWAY_MAX_RUNS || (k > 0 && (runs == nullptr || run_lens == nullptr))) return MGS_ERR_INVALID;
  size_t n = 0;
  for (size_t j = 0; j < k; ++j) {
    if (run_lens[j] > 0 && runs[j] == nullptr) return MGS_ERR_INVALID;
    if (run_lens[j] > kMaxKeys || n + run_lens[j] < n) return MGS_ERR_TOO_LARGE;
    n += run_lens[j];
  }
  if (n > 0 && out == nullptr) return MGS_ERR_INVALID;
  *total = n;
  return MGS_OK;
}

int mgs_multiway_merge(mgs_key kt, const void* const* runs, const size_t* run_lens, size_t k, void* out, void* temp,
                       size_t temp_bytes, void* stream) {
  size_t n = 0;
  const int rc = multiway_checks(kt, runs, run_lens, k, out, &n);
  if (rc != MGS_OK) return rc;
  if (n == 0) return MGS_OK;
  return mgs::multiway_dispatch(kt, runs, run_lens, k, out, temp, temp_bytes, as_stream(stream), 0, nullptr, nullptr);
}

int mgs_multiway_merge_portions(mgs_key kt, const void* const* runs, const size_t* run_lens, size_t k, void* out,
                                void* temp, size_t temp_bytes, void* stream, size_t portions, size_t* portion_end,
                                void* const* portion_done) {
  size_t n = 0;
  const int rc = multiway_checks(kt, runs, run_lens, k, out, &n);
  if (rc != MGS_OK) return rc;
  if (portions == 0 || portion_end == nullptr) return MGS_ERR_INVALID;
  return mgs::multiway_dispatch(kt, runs, run_lens, k, out, temp, temp_bytes, as_stream(stream), portions, portion_end,
                                reinterpret_cast<hipEvent_t*>(const_cast<void**>(portion_done)));
}

int mgs_fill_max(mgs_key kt, void* keys, size_t n, void* stream) {
  if (key_bytes(kt) < 0) return MGS_ERR_BAD_TYPE;
  return mgs::fill_max_dispatch(kt, keys, n, as_stream(stream));
}

size_t mgs_segmented_sort_temp_bytes(mgs_key kt, mgs_val vt, size_t max_total_keys, size_t nseg) {
  return mgs::segmented_temp_bytes(kt, vt, max_total_keys, nseg, nullptr);
}

size_t mgs_segmented_sort_temp_bytes_tuned(mgs_key kt, mgs_val vt, size_t max_total_keys, size_t nseg,
                                           const mgs_tuning* tuning) {
  return mgs::segmented_temp_bytes(kt, vt, max_total_keys, nseg, tuning);
}

int mgs_segmented_variant_count(void) { return mgs::segmented_variant_count(); }

int mgs_segmented_variant(mgs_key kt, mgs_val vt, int variant, int32_t* block, int32_t* items_per_thread) {
  return mgs::segmented_variant(kt, vt, variant, block, items_per_thread);
}

int mgs_segmented_sort(mgs_key kt, mgs_val vt, const void* keys_in, void* keys_out, const void* vals_in,
                       void* vals_out, const uint64_t* d_offsets, size_t nseg, size_t max_total_keys,
                       const mgs_tuning* tuning, void* temp, size_t temp_bytes, void* stream) {
  if (key_bytes(kt) < 0 || val_bytes(vt) < 0) return MGS_ERR_BAD_TYPE;
  if (nseg == 0 || max_total_keys == 0) return MGS_OK;
  if (max_total_keys > kMaxKeys) return MGS_ERR_TOO_LARGE;
  if (keys_in == nullptr || keys_out == nullptr || d_offsets == nullptr) return MGS_ERR_INVALID;
  if (vt != MGS_VAL_NONE && (vals_in == nullptr || vals_out == nullptr)) return MGS_ERR_INVALID;
  return mgs::segmented_dispatch(kt, vt, keys_in, keys_out, vals_in, vals_out, d_offsets, false, nullptr, nseg, max_total_keys,
                                 tuning, temp, temp_bytes, as_stream(stream));
}

int mgs_segmented_sort_ranges(mgs_key kt, mgs_val vt, void* keys, void* vals, const uint64_t* d_ranges,
                              const uint64_t* d_val_begin, size_t nseg, size_t max_total_keys,
                              const mgs_tuning* tuning, void* temp, size_t temp_bytes, void* stream) {
  if (key_bytes(kt) < 0 || val_bytes(vt) < 0) return MGS_ERR_BAD_TYPE;
  if (nseg == 0 || max_total_keys == 0) return MGS_OK;
  if (max_total_keys > kMaxKeys) return MGS_ERR_TOO_LARGE;
  if (keys == nullptr || d_ranges == nullptr) return MGS_ERR_INVALID;
  if (vt != MGS_VAL_NONE && vals == nullptr) return MGS_ERR_INVALID;
  return mgs::segmented_dispatch(kt, vt, keys, keys, vals, vals, d_ranges, true, d_val_begin, nseg, max_total_keys, tuning, temp,
                                 temp_bytes, as_stream(stream));
}

int mgs_temp_status(const void* temp, size_t temp_bytes, void* stream) {
  if (temp == nullptr || temp_bytes < sizeof(uint32_t)) return MGS_OK;
  uint32_t word = 0;
  hipError_t e = hipMemcpyAsync(&word, temp, sizeof(word), hipMemcpyDeviceToHost, as_stream(stream));
  if (e == hipSuccess) e = hipStreamSynchronize(as_stream(stream));
  if (e != hipSuccess) return mgs::record_hip(e);
  if (word & (mgs::kStatusLookback | mgs::kStatusChainUnserved)) return MGS_ERR_DEVICE;
  if (word & mgs::kStatusSkipped) return MGS_ERR_SKIPPED;
  if (word & mgs::kStatusUnsorted) return MGS_ERR_UNSORTED;
  return MGS_OK;
}

const char* mgs_status_str(int status) {
  switch (status) {
    case MGS_OK: return "ok";
    case MGS_ERR_BAD_TYPE: return "unknown key or value type";
    case MGS_ERR_TEMP_TOO_SMALL: return "temporary buffer too small";
    case MGS_ERR_HIP: return "HIP runtime error";
    case MGS_ERR_INVALID: return "invalid argument";
    case MGS_ERR_TOO_LARGE: return "too many keys";
    case MGS_ERR_DEVICE: return "a kernel detected an invalid look-back state; output not valid";
    case MGS_ERR_SKIPPED: return "segments left unsorted (malformed or out-of-bound segments)";
    case MGS_ERR_UNSORTED: return "merge input runs not sorted; output not valid";
  }
  return "unknown status";
}

int mgs_last_hip_error(void) { return mgs::last_hip_error(); }

const char* mgs_version(void) { return "mgsort-mi355x 0.1 (gfx950)"; }

}  // extern "C"
