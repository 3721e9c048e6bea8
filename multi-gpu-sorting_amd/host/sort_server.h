// Serving layer: a request loop that sorts arrays of mixed size and key type
// on one GPU with continuous batching.
//
//   admission   FIFO; a request is admitted when its data fits in the device
//               pool (the HBM budget, one allocation made at start-up)
//   batching    admitted requests of one (key, payload) type that arrive
//               while earlier batches are on the device form the next batch:
//               one mgs_segmented_sort_ranges launch sequence over all of
//               them, in place in their pool slots
//   overlap     up to max_inflight batches are on the device at once: host to
//               device copies of batch b+1, the sort of batch b and the copies
//               back of batch b-1 run on their own streams ordered by events;
//               consecutive batches sort on sort_streams streams in turn, so
//               a batch that leaves CUs idle (a mid-size solo sort) overlaps
//               the next one
//   tuning      segmented launches take the tuning table's entry for their
//               type and size (seg_tuning_path; empty: library defaults)
//   large       in batched mode a request of at least solo_min_keys keys is
//               sorted alone with mgs_sort: the plain sort already runs at
//               its per-type rate there, and a batch holding it would make
//               the small requests wait for it
//   plain mode  the same loop with one request per launch sequence (mgs_sort)
//
// Requests either bring host arrays (host_data: copied in on admission and
// back on completion) or already sit in a pool slot the client obtained with
// AllocSlot (sorted in place there).  Latency = Submit() to the completion
// event of the request's batch.  No counterpart in the reference.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "device_pool.h"
#include "mgsort.h"
#include "seg_tuning.h"

namespace mgsort {

struct SortRequest {
  mgs_key kt = MGS_KEY_U32;
  mgs_val vt = MGS_VAL_NONE;
  size_t n = 0;
  // host_data mode: host arrays (pinned memory gives asynchronous copies)
  const void* keys_in = nullptr;
  const void* vals_in = nullptr;
  void* keys_out = nullptr;
  void* vals_out = nullptr;
  // device mode: a slot from SortServer::AllocSlot holding the data
  void* slot = nullptr;
  // filled in by the server (seconds, steady clock)
  double t_submit = 0, t_launch = 0, t_done = 0;
  int status = MGS_OK;
  std::atomic<bool> done{false};
};

struct ServerConfig {
  int device = 0;
  size_t hbm_budget = size_t(64) << 30;
  bool batched = true;
  bool host_data = true;
  int max_inflight = 3;
  size_t max_batch_keys = size_t(1) << 30;
  size_t max_batch_requests = 4096;
  size_t solo_min_keys = size_t(1) << 21;  // 0: batch every request
  int sort_streams = 2;
  std::string seg_tuning_path;
};

struct ServerStats {
  uint64_t completed = 0, failed = 0, batches = 0, keys = 0;
  double first_submit = 0, last_done = 0;
};

class SortServer {
 public:
  explicit SortServer(const ServerConfig& cfg);
  ~SortServer();
  SortServer(const SortServer&) = delete;
  SortServer& operator=(const SortServer&) = delete;

  // Device mode: a pool slot for n keys (and payloads, stored after the keys),
  // or nullptr when the budget has no room.  SlotVals gives the payload part.
  void* AllocSlot(mgs_key kt, mgs_val vt, size_t n);
  void FreeSlot(void* slot);
  static size_t SlotBytes(mgs_key kt, mgs_val vt, size_t n);
  static void* SlotVals(void* slot, mgs_key kt, size_t n);

  // Queues r (which must outlive its completion); returns at once.
  void Submit(SortRequest* r);
  // Returns when every submitted request has completed.
  void Drain();
  ServerStats Stats();
  static double Now();

 private:
  struct Batch;
  void SchedulerLoop();
  void CompletionLoop();
  std::unique_ptr<Batch> FormBatch(std::unique_lock<std::mutex>& lk);
  void Launch(Batch& b);
  void Finish(Batch& b);

  ServerConfig cfg_;
  std::unique_ptr<DevicePool> pool_;
  SegTuningTable tuning_;
  hipStream_t s_in_ = nullptr, s_out_ = nullptr;
  std::vector<hipStream_t> s_sort_;
  uint64_t launches_ = 0;

  std::mutex mu_;
  std::condition_variable cv_sched_, cv_done_, cv_drain_;
  std::deque<SortRequest*> queue_;
  std::deque<std::unique_ptr<Batch>> inflight_;
  uint64_t submitted_ = 0, finished_ = 0;
  bool stop_ = false;
  ServerStats stats_;

  // pinned staging of segment bounds, one ring entry per in-flight batch
  std::vector<uint64_t*> h_bounds_;
  std::vector<bool> bounds_busy_;

  std::thread sched_, done_;
};

}  // namespace mgsort
