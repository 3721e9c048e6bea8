#include "sort_server.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>

#include "hip_check.h"

namespace mgsort {

namespace {

size_t KeyBytes(mgs_key kt) { return (kt == MGS_KEY_U32 || kt == MGS_KEY_I32 || kt == MGS_KEY_F32) ? 4 : 8; }
size_t ValBytes(mgs_val vt) { return vt == MGS_VAL_NONE ? 0 : (vt == MGS_VAL_U32 ? 4 : 8); }
size_t AlignUp(size_t x) { return (x + DevicePool::kAlign - 1) / DevicePool::kAlign * DevicePool::kAlign; }

}  // namespace

struct SortServer::Batch {
  mgs_key kt = MGS_KEY_U32;
  mgs_val vt = MGS_VAL_NONE;
  std::vector<SortRequest*> reqs;
  std::vector<char*> slots;     // per request: where its keys sit in the pool
  std::vector<bool> own_slot;   // slot allocated by the server (host_data)
  char* scratch = nullptr;      // sort temp + device copy of the bounds
  size_t temp_bytes = 0;
  uint64_t* d_bounds = nullptr;
  int ring = -1;
  size_t keys = 0;
  bool solo = false;  // one request, sorted with mgs_sort
  bool tuned = false;  // segmented launch with `tuning` (else the defaults)
  mgs_tuning tuning{};
  hipEvent_t ev_in = nullptr, ev_sort = nullptr, ev_out = nullptr;
};

double SortServer::Now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

size_t SortServer::SlotBytes(mgs_key kt, mgs_val vt, size_t n) {
  return AlignUp(n * KeyBytes(kt)) + n * ValBytes(vt);
}

void* SortServer::SlotVals(void* slot, mgs_key kt, size_t n) {
  return static_cast<char*>(slot) + AlignUp(n * KeyBytes(kt));
}

SortServer::SortServer(const ServerConfig& cfg) : cfg_(cfg) {
  if (cfg_.max_inflight < 1) cfg_.max_inflight = 1;
  if (!cfg_.batched) cfg_.max_batch_requests = 1;
  if (cfg_.sort_streams < 1) cfg_.sort_streams = 1;
  tuning_ = SegTuningTable::Load(cfg_.seg_tuning_path);
  HIP_CHECK(hipSetDevice(cfg_.device));
  pool_ = std::make_unique<DevicePool>(cfg_.hbm_budget);
  HIP_CHECK(hipStreamCreateWithFlags(&s_in_, hipStreamNonBlocking));
  s_sort_.assign(cfg_.sort_streams, nullptr);
  for (auto& st : s_sort_) HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  HIP_CHECK(hipStreamCreateWithFlags(&s_out_, hipStreamNonBlocking));
  h_bounds_.resize(cfg_.max_inflight, nullptr);
  bounds_busy_.assign(cfg_.max_inflight, false);
  for (auto& h : h_bounds_)
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h), 3 * cfg_.max_batch_requests * sizeof(uint64_t),
                            hipHostMallocDefault));
  sched_ = std::thread([this] { SchedulerLoop(); });
  done_ = std::thread([this] { CompletionLoop(); });
}

SortServer::~SortServer() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_sched_.notify_all();
  cv_done_.notify_all();
  sched_.join();
  done_.join();
  for (auto* h : h_bounds_) (void)hipHostFree(h);
  (void)hipStreamDestroy(s_in_);
  for (auto st : s_sort_) (void)hipStreamDestroy(st);
  (void)hipStreamDestroy(s_out_);
}

void* SortServer::AllocSlot(mgs_key kt, mgs_val vt, size_t n) { return pool_->Allocate(SlotBytes(kt, vt, n)); }
void SortServer::FreeSlot(void* slot) { pool_->Free(slot); }

void SortServer::Submit(SortRequest* r) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    r->done.store(false);
    r->status = MGS_OK;
    r->t_submit = Now();
    if (stats_.first_submit == 0) stats_.first_submit = r->t_submit;
    queue_.push_back(r);
    ++submitted_;
  }
  cv_sched_.notify_one();
}

void SortServer::Drain() {
  std::unique_lock<std::mutex> lk(mu_);
  cv_drain_.wait(lk, [&] { return finished_ == submitted_; });
}

ServerStats SortServer::Stats() {
  std::lock_guard<std::mutex> lk(mu_);
  return stats_;
}

// Called with mu_ held.  Takes the queue's front request and every later
// queued request of the same type that fits the batch limits and the pool.
// Returns nullptr when the pool cannot hold even the front request now.
std::unique_ptr<SortServer::Batch> SortServer::FormBatch(std::unique_lock<std::mutex>&) {
  auto b = std::make_unique<Batch>();
  SortRequest* front = queue_.front();
  b->kt = front->kt;
  b->vt = front->vt;
  auto large = [&](const SortRequest* r) { return cfg_.solo_min_keys > 0 && r->n >= cfg_.solo_min_keys; };
  b->solo = !cfg_.batched || large(front);
  std::vector<size_t> taken;  // queue positions
  for (size_t i = 0; i < queue_.size() && b->reqs.size() < cfg_.max_batch_requests; ++i) {
    SortRequest* r = queue_[i];
    if (r->kt != b->kt || r->vt != b->vt) continue;
    if (!b->reqs.empty() && b->solo) break;
    if (!b->reqs.empty() && large(r)) continue;  // sorted alone when it reaches the front
    if (!b->reqs.empty() && b->keys + r->n > cfg_.max_batch_keys) break;
    char* slot = static_cast<char*>(r->slot);
    bool own = false;
    if (cfg_.host_data) {
      slot = pool_->Allocate(SlotBytes(r->kt, r->vt, r->n));
      if (slot == nullptr) break;  // FIFO within the type: stop at the first that does not fit
      own = true;
    }
    b->reqs.push_back(r);
    b->slots.push_back(slot);
    b->own_slot.push_back(own);
    b->keys += r->n;
    taken.push_back(i);
  }
  auto release_from = [&](size_t k) {  // give back requests k.. of the batch
    for (size_t j = k; j < b->reqs.size(); ++j)
      if (b->own_slot[j]) pool_->Free(b->slots[j]);
    b->keys = 0;
    for (size_t j = 0; j < k; ++j) b->keys += b->reqs[j]->n;
    b->reqs.resize(k);
    b->slots.resize(k);
    b->own_slot.resize(k);
    taken.resize(k);
  };
  while (!b->reqs.empty()) {
    const size_t nseg = b->reqs.size();
    const mgs_tuning* tn = b->solo ? nullptr : tuning_.Lookup(b->kt, b->vt, b->keys);
    b->tuned = tn != nullptr;
    if (tn) b->tuning = *tn;
    b->temp_bytes = b->solo ? mgs_sort_temp_bytes(b->kt, b->vt, b->keys)
                            : mgs_segmented_sort_temp_bytes_tuned(b->kt, b->vt, b->keys, nseg, tn);
    b->scratch = pool_->Allocate(AlignUp(b->temp_bytes) + 3 * nseg * sizeof(uint64_t));
    if (b->scratch != nullptr) break;
    release_from(nseg / 2);  // smaller batch, smaller scratch
  }
  if (b->reqs.empty()) return nullptr;
  b->d_bounds = reinterpret_cast<uint64_t*>(b->scratch + AlignUp(b->temp_bytes));
  for (int i = 0; i < cfg_.max_inflight; ++i)
    if (!bounds_busy_[i]) {
      b->ring = i;
      bounds_busy_[i] = true;
      break;
    }
  // drop the taken requests from the queue (positions ascending)
  std::deque<SortRequest*> rest;
  size_t t = 0;
  for (size_t i = 0; i < queue_.size(); ++i) {
    if (t < taken.size() && taken[t] == i) {
      ++t;
      continue;
    }
    rest.push_back(queue_[i]);
  }
  queue_.swap(rest);
  return b;
}

void SortServer::Launch(Batch& b) {
  const size_t K = KeyBytes(b.kt), V = ValBytes(b.vt), nseg = b.reqs.size();
  HIP_CHECK(hipEventCreateWithFlags(&b.ev_in, hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&b.ev_sort, hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&b.ev_out, hipEventDisableTiming));
  const double now = Now();
  for (size_t i = 0; i < nseg; ++i) {
    SortRequest* r = b.reqs[i];
    r->t_launch = now;
    if (cfg_.host_data) {
      HIP_CHECK(hipMemcpyAsync(b.slots[i], r->keys_in, r->n * K, hipMemcpyHostToDevice, s_in_));
      if (V) HIP_CHECK(hipMemcpyAsync(SlotVals(b.slots[i], b.kt, r->n), r->vals_in, r->n * V,
                                      hipMemcpyHostToDevice, s_in_));
    }
  }
  hipStream_t s_sort = s_sort_[launches_++ % s_sort_.size()];
  HIP_CHECK(hipEventRecord(b.ev_in, s_in_));
  HIP_CHECK(hipStreamWaitEvent(s_sort, b.ev_in, 0));
  char* base = pool_->base();
  int st = MGS_OK;
  if (!b.solo) {
    uint64_t* h = h_bounds_[b.ring];
    for (size_t i = 0; i < nseg; ++i) {
      const uint64_t kb = (uint64_t)(b.slots[i] - base) / K;
      h[2 * i] = kb;
      h[2 * i + 1] = kb + b.reqs[i]->n;
      h[2 * nseg + i] = V ? (uint64_t)(static_cast<char*>(SlotVals(b.slots[i], b.kt, b.reqs[i]->n)) - base) / V : 0;
    }
    HIP_CHECK(hipMemcpyAsync(b.d_bounds, h, 3 * nseg * sizeof(uint64_t), hipMemcpyHostToDevice, s_sort));
    st = mgs_segmented_sort_ranges(b.kt, b.vt, base, V ? base : nullptr, b.d_bounds, V ? b.d_bounds + 2 * nseg : nullptr,
                                   nseg, b.keys, b.tuned ? &b.tuning : nullptr, b.scratch, b.temp_bytes, s_sort);
  } else {
    SortRequest* r = b.reqs[0];
    void* vals = V ? SlotVals(b.slots[0], b.kt, r->n) : nullptr;
    st = mgs_sort(b.kt, b.vt, b.slots[0], b.slots[0], vals, vals, r->n, b.scratch, b.temp_bytes, s_sort);
  }
  for (auto* r : b.reqs) r->status = st;
  HIP_CHECK(hipEventRecord(b.ev_sort, s_sort));
  HIP_CHECK(hipStreamWaitEvent(s_out_, b.ev_sort, 0));
  if (cfg_.host_data) {
    for (size_t i = 0; i < nseg; ++i) {
      SortRequest* r = b.reqs[i];
      HIP_CHECK(hipMemcpyAsync(r->keys_out, b.slots[i], r->n * K, hipMemcpyDeviceToHost, s_out_));
      if (V) HIP_CHECK(hipMemcpyAsync(r->vals_out, SlotVals(b.slots[i], b.kt, r->n), r->n * V,
                                      hipMemcpyDeviceToHost, s_out_));
    }
  }
  HIP_CHECK(hipEventRecord(b.ev_out, s_out_));
}

void SortServer::SchedulerLoop() {
  uint64_t seen_finished = 0;
  for (;;) {
    std::unique_lock<std::mutex> lk(mu_);
    cv_sched_.wait(lk, [&] {
      return stop_ || (!queue_.empty() && (int)inflight_.size() < cfg_.max_inflight);
    });
    if (stop_) return;
    std::unique_ptr<Batch> b = FormBatch(lk);
    if (!b) {
      if (inflight_.empty()) {
        // nothing on the device will ever free memory: the front request
        // cannot fit the budget at all
        SortRequest* r = queue_.front();
        queue_.pop_front();
        r->status = MGS_ERR_TOO_LARGE;
        r->t_done = Now();
        r->done.store(true);
        ++stats_.failed;
        ++finished_;
        cv_drain_.notify_all();
        continue;
      }
      seen_finished = finished_;
      cv_sched_.wait(lk, [&] { return stop_ || finished_ != seen_finished; });
      continue;
    }
    lk.unlock();
    try {
      Launch(*b);
    } catch (const std::exception&) {
      for (auto* r : b->reqs) r->status = MGS_ERR_HIP;
    }
    lk.lock();
    inflight_.push_back(std::move(b));
    ++stats_.batches;
    lk.unlock();
    cv_done_.notify_one();
  }
}

void SortServer::Finish(Batch& b) {
  if (b.ev_out) (void)hipEventSynchronize(b.ev_out);
  const double now = Now();
  std::lock_guard<std::mutex> lk(mu_);
  for (size_t i = 0; i < b.reqs.size(); ++i) {
    SortRequest* r = b.reqs[i];
    r->t_done = now;
    if (r->status == MGS_OK) {
      ++stats_.completed;
      stats_.keys += r->n;
    } else {
      ++stats_.failed;
    }
    if (b.own_slot[i]) pool_->Free(b.slots[i]);
    r->done.store(true);
  }
  stats_.last_done = now;
  pool_->Free(b.scratch);
  if (b.ring >= 0) bounds_busy_[b.ring] = false;
  for (hipEvent_t e : {b.ev_in, b.ev_sort, b.ev_out})
    if (e) (void)hipEventDestroy(e);
  finished_ += b.reqs.size();
}

void SortServer::CompletionLoop() {
  for (;;) {
    Batch* b = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_done_.wait(lk, [&] { return !inflight_.empty() || stop_; });
      if (inflight_.empty()) return;  // stop_ with nothing left
      b = inflight_.front().get();
    }
    Finish(*b);
    {
      std::lock_guard<std::mutex> lk(mu_);
      inflight_.pop_front();
    }
    cv_sched_.notify_all();
    cv_drain_.notify_all();
  }
}

}  // namespace mgsort
