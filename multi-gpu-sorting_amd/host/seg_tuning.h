// The segmented sort's tuning table (written by mgsort.tuning, one entry per
// key type, payload type and size bin) for C++ callers such as the serving
// layer.  No counterpart in the reference.
#pragma once

#include <string>
#include <vector>

#include "mgsort.h"

namespace mgsort {

struct SegTuningEntry {
  mgs_key kt;
  mgs_val vt;
  int size_log2;
  mgs_tuning tuning;
};

class SegTuningTable {
 public:
  // Reads a table file; throws mgsort::Error on a malformed file.  An empty
  // path gives an empty table.
  static SegTuningTable Load(const std::string& path);
  // The entry of the nearest size bin for (kt, vt) (the smaller bin on a
  // tie), or nullptr when the table has none for the pair.
  const mgs_tuning* Lookup(mgs_key kt, mgs_val vt, size_t total_keys) const;
  size_t size() const { return entries_.size(); }

 private:
  std::vector<SegTuningEntry> entries_;
};

}  // namespace mgsort
