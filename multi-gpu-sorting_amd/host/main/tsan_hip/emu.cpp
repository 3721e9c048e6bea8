// Host emulation of the HIP stream API and of the libmgsort entry points the
// serving layer calls (see hip/hip_runtime.h beside this file).  The sorts
// run on the stream's worker thread as stable sorts by the library's encoded
// key order, so every output can be compared with a host sort.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "key_traits.h"
#include "mgsort.h"

struct EmuState {
  std::mutex m;
  std::condition_variable cv;
  uint64_t recorded = 0, done = 0;
};

struct EmuEvent {
  std::shared_ptr<EmuState> st = std::make_shared<EmuState>();
};

struct EmuStream {
  std::mutex m;
  std::condition_variable cv;
  std::deque<std::function<void()>> q;
  uint64_t enq = 0, ran = 0;
  bool stop = false;
  std::thread worker;
  EmuStream() {
    worker = std::thread([this] {
      for (;;) {
        std::function<void()> op;
        {
          std::unique_lock<std::mutex> lk(m);
          cv.wait(lk, [&] { return stop || !q.empty(); });
          if (q.empty()) return;
          op = std::move(q.front());
          q.pop_front();
        }
        op();
        {
          std::lock_guard<std::mutex> lk(m);
          ++ran;
        }
        cv.notify_all();
      }
    });
  }
  void Push(std::function<void()> op) {
    {
      std::lock_guard<std::mutex> lk(m);
      q.push_back(std::move(op));
      ++enq;
    }
    cv.notify_all();
  }
  void Sync() {
    std::unique_lock<std::mutex> lk(m);
    const uint64_t target = enq;
    cv.wait(lk, [&] { return ran >= target; });
  }
  ~EmuStream() {
    {
      std::lock_guard<std::mutex> lk(m);
      stop = true;
    }
    cv.notify_all();
    worker.join();
  }
};

hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "emulated HIP error"; }
hipError_t hipMalloc(void** p, size_t bytes) {
  *p = std::aligned_alloc(256, (bytes + 255) / 256 * 256);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) { return hipMalloc(p, bytes); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
hipError_t hipHostGetDevicePointer(void** dev, void* host, unsigned) {
  *dev = host;
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  *s = new EmuStream();
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  if (s) {
    s->Sync();
    delete s;
  }
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
  s->Sync();
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) {
  *e = new EmuEvent();
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;  // operations still queued hold the state
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  std::shared_ptr<EmuState> st = e->st;
  uint64_t gen;
  {
    std::lock_guard<std::mutex> lk(st->m);
    gen = ++st->recorded;
  }
  s->Push([st, gen] {
    {
      std::lock_guard<std::mutex> lk(st->m);
      st->done = std::max(st->done, gen);
    }
    st->cv.notify_all();
  });
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
  std::shared_ptr<EmuState> st = e->st;
  std::unique_lock<std::mutex> lk(st->m);
  const uint64_t gen = st->recorded;
  st->cv.wait(lk, [&] { return st->done >= gen; });
  return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) {
  std::shared_ptr<EmuState> st = e->st;
  uint64_t gen;
  {
    std::lock_guard<std::mutex> lk(st->m);
    gen = st->recorded;
  }
  s->Push([st, gen] {
    std::unique_lock<std::mutex> lk(st->m);
    st->cv.wait(lk, [&] { return st->done >= gen; });
  });
  return hipSuccess;
}
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind) {
  std::memcpy(dst, src, bytes);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t s) {
  s->Push([dst, src, bytes] { std::memcpy(dst, src, bytes); });
  return hipSuccess;
}

// ---------------------------------------------------------------------------
// libmgsort entry points: stable sorts on the stream's worker
// ---------------------------------------------------------------------------
namespace {

template <typename K, typename V>
void StableSort(K* k, V* v, size_t n) {
  std::vector<size_t> idx(n);
  std::iota(idx.begin(), idx.end(), size_t(0));
  std::stable_sort(idx.begin(), idx.end(),
                   [&](size_t a, size_t b) { return mgsort::Encode(k[a]) < mgsort::Encode(k[b]); });
  std::vector<K> kk(n);
  for (size_t i = 0; i < n; ++i) kk[i] = k[idx[i]];
  std::memcpy(k, kk.data(), n * sizeof(K));
  if (v) {
    std::vector<V> vv(n);
    for (size_t i = 0; i < n; ++i) vv[i] = v[idx[i]];
    std::memcpy(v, vv.data(), n * sizeof(V));
  }
}

template <typename K>
void SortByVal(mgs_val vt, void* k, void* v, size_t n) {
  if (vt == MGS_VAL_U32)
    StableSort(static_cast<K*>(k), static_cast<uint32_t*>(v), n);
  else if (vt == MGS_VAL_U64)
    StableSort(static_cast<K*>(k), static_cast<uint64_t*>(v), n);
  else
    StableSort(static_cast<K*>(k), static_cast<uint32_t*>(nullptr), n);
}

void SortAny(mgs_key kt, mgs_val vt, void* k, void* v, size_t n) {
  switch (kt) {
    case MGS_KEY_U32: SortByVal<uint32_t>(vt, k, v, n); break;
    case MGS_KEY_U64: SortByVal<uint64_t>(vt, k, v, n); break;
    case MGS_KEY_I32: SortByVal<int32_t>(vt, k, v, n); break;
    case MGS_KEY_I64: SortByVal<int64_t>(vt, k, v, n); break;
    case MGS_KEY_F32: SortByVal<float>(vt, k, v, n); break;
    case MGS_KEY_F64: SortByVal<double>(vt, k, v, n); break;
  }
}

size_t KeyB(mgs_key kt) { return (kt == MGS_KEY_U32 || kt == MGS_KEY_I32 || kt == MGS_KEY_F32) ? 4 : 8; }
size_t ValB(mgs_val vt) { return vt == MGS_VAL_NONE ? 0 : (vt == MGS_VAL_U32 ? 4 : 8); }

}  // namespace

extern "C" {

size_t mgs_sort_temp_bytes(mgs_key, mgs_val, size_t) { return 256; }

int mgs_sort(mgs_key kt, mgs_val vt, const void* kin, void* kout, const void* vin, void* vout, size_t n, void*,
             size_t, void* stream) {
  static_cast<hipStream_t>(stream)->Push([=] {
    if (kin != kout) std::memcpy(kout, kin, n * KeyB(kt));
    if (vt != MGS_VAL_NONE && vin != vout) std::memcpy(vout, vin, n * ValB(vt));
    SortAny(kt, vt, kout, vout, n);
  });
  return MGS_OK;
}

size_t mgs_segmented_sort_temp_bytes_tuned(mgs_key, mgs_val, size_t, size_t, const mgs_tuning*) { return 256; }

int mgs_segmented_variant_count(void) { return 3; }

int mgs_segmented_sort_ranges(mgs_key kt, mgs_val vt, void* keys, void* vals, const uint64_t* d_ranges,
                              const uint64_t* d_val_begin, size_t nseg, size_t, const mgs_tuning*, void*, size_t,
                              void* stream) {
  static_cast<hipStream_t>(stream)->Push([=] {
    for (size_t s = 0; s < nseg; ++s) {
      const uint64_t b = d_ranges[2 * s], e = d_ranges[2 * s + 1];
      const uint64_t v = d_val_begin ? d_val_begin[s] : b;
      SortAny(kt, vt, static_cast<char*>(keys) + b * KeyB(kt),
              vt == MGS_VAL_NONE ? nullptr : static_cast<char*>(vals) + v * ValB(vt), e - b);
    }
  });
  return MGS_OK;
}

}  // extern "C"
