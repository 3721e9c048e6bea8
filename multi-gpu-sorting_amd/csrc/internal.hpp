// Host-side declarations shared by the translation units of libmgsort.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/mgsort.h"

namespace mgs {

// Records `e` as this thread's last HIP error and returns MGS_ERR_HIP, or
// MGS_OK when e == hipSuccess.
int record_hip(hipError_t e);
int last_hip_error();

// Stage marks of a launch sequence (mgs_sort_timed): records events[i], when
// given, before stage i, and collects the stage names.
struct Marks {
  hipEvent_t* ev;
  int cap, used;
  std::string* names;
  void operator()(hipStream_t s, const char* name = nullptr) {
    if (ev != nullptr && used < cap) (void)hipEventRecord(ev[used], s);
    if (names != nullptr && name != nullptr) {
      if (!names->empty()) names->push_back(',');
      names->append(name);
    }
    ++used;
  }
};

// Checks the launch(es) just enqueued.
inline int check_launch() { return record_hip(hipGetLastError()); }

inline int key_bytes(mgs_key kt) {
  switch (kt) {
    case MGS_KEY_U32: case MGS_KEY_I32: case MGS_KEY_F32: return 4;
    case MGS_KEY_U64: case MGS_KEY_I64: case MGS_KEY_F64: return 8;
  }
  return -1;
}
inline int val_bytes(mgs_val vt) {
  switch (vt) {
    case MGS_VAL_NONE: return 0;
    case MGS_VAL_U32: return 4;
    case MGS_VAL_U64: return 8;
  }
  return -1;
}

size_t sort_temp_bytes(mgs_key kt, mgs_val vt, size_t n);
int sort_dispatch(mgs_key kt, mgs_val vt, const void* kin, void* kout, const void* vin, void* vout, size_t n,
                  void* temp, size_t temp_bytes, hipStream_t s, hipEvent_t* events = nullptr, int num_events = 0,
                  int* num_marks = nullptr);

// mgs_sort_noalloc (noalloc.hip).
int sort_noalloc(mgs_key kt, mgs_val vt, const void* kin, void* kout, const void* vin, void* vout, size_t n,
                 hipStream_t s);

// mgs_sort_path.
int sort_path(mgs_key kt, mgs_val vt, size_t n, const void* temp, size_t temp_bytes, hipStream_t s);

// Comma-separated stage names of this thread's last mgs_sort_timed call.
const char* last_stage_names();

size_t merge_temp_bytes(mgs_key kt, mgs_val vt, size_t na, size_t nb);
int merge_dispatch(mgs_key kt, mgs_val vt, const void* ak, const void* av, size_t na, const void* bk, const void* bv,
                   size_t nb, void* ok, void* ov, void* temp, size_t temp_bytes, hipStream_t s);
size_t multiway_temp_bytes(mgs_key kt, size_t k, size_t total);
int multiway_dispatch(mgs_key kt, const void* const* runs, const size_t* lens, size_t k, void* out, void* temp,
                      size_t temp_bytes, hipStream_t s, size_t portions, size_t* portion_end,
                      hipEvent_t* portion_done);
int pivot_dispatch(mgs_key kt, const void* const* left, const void* const* right, size_t part_size, size_t num_parts,
                   uint64_t* result, hipStream_t s);
int swap_dispatch(mgs_key kt, mgs_val vt, void* ak, void* av, void* bk, void* bv, size_t n, hipStream_t s);
int fill_max_dispatch(mgs_key kt, void* keys, size_t n, hipStream_t s);

size_t segmented_temp_bytes(mgs_key kt, mgs_val vt, size_t max_total, size_t nseg, const mgs_tuning* tuning);
int segmented_variant(mgs_key kt, mgs_val vt, int var, int32_t* block, int32_t* ipt);
int segmented_variant_count();
// ranges == false: d_bounds holds nseg+1 offsets; true: nseg (begin, end)
// pairs, with payloads at d_val_begin[s] (NULL: at the keys' begin).
int segmented_dispatch(mgs_key kt, mgs_val vt, const void* kin, void* kout, const void* vin, void* vout,
                       const uint64_t* d_bounds, bool ranges, const uint64_t* d_val_begin, size_t nseg,
                       size_t max_total,
                       const mgs_tuning* tuning, void* temp, size_t temp_bytes, hipStream_t s);

}  // namespace mgs
