// mgs_sort_noalloc: mgs_sort with its scratch in a static device array, for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops).  Same kernels and launch sequence as mgs_sort.
//
// The array (kNoAllocScratchBytes) exists once per GPU the library runs on
// and stays allocated until the process ends.  Its address is looked up once
// per device with hipGetSymbolAddress, with that device current; host code
// uses only that address, never the array's name (which in host code is a
// stand-in, not a device address).  Every call on a GPU shares that GPU's
// array: calls that overlap on one GPU (concurrent streams, or host threads)
// race on it.  Nothing needs to start at zero: the sort's first kernels clear
// the histograms, counters, status word and look-back descriptors they use.
#include <mutex>

#include "internal.hpp"

namespace mgs {

namespace {

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (looked up on first use).
int scratch_ptr(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return record_hip(e);
  if (dev < 0 || dev >= kMaxDevices) return MGS_ERR_INVALID;
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return record_hip(e);
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return MGS_OK;
}

}  // namespace

size_t noalloc_scratch_bytes() { return kNoAllocScratchBytes; }

int sort_noalloc(mgs_key kt, mgs_val vt, const void* kin, void* kout, const void* vin, void* vout, size_t n,
                 hipStream_t s) {
  const size_t need = sort_temp_bytes(kt, vt, n);
  if (need > kNoAllocScratchBytes) return MGS_ERR_TOO_LARGE;
  void* p = nullptr;
  if (need > 0) {
    const int rc = scratch_ptr(&p);
    if (rc != MGS_OK) return rc;
  }
  return sort_dispatch(kt, vt, kin, kout, vin, vout, n, p, need > 0 ? kNoAllocScratchBytes : 0, s);
}

}  // namespace mgs
