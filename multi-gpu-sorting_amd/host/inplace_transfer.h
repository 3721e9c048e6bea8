// The in-place exchange through one device buffer (InplaceMemcpy,
// src/inplace_memcpy.cuh:7-69): the first out_bytes of `dev` go to host_out
// while the first in_bytes of host_in take their place, block by block.
// Unlike the reference, which waits on the host after every block pair, the
// order is kept by events: block k leaves on `d2h`, an event is recorded
// behind it, and block k's refill on `h2d` waits for that event, so block
// k+1 leaves while block k refills.  Blocks of host_in past out_bytes wait
// for nothing (the caller orders them if it must).  Enqueues only.
// block_out must hold at least ceil(max(in_bytes, out_bytes) / block) events.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <memory>
#include <vector>

#include "device_buffer.h"
#include "hip_check.h"

namespace mgsort {

inline void EnqueueInplaceTransfer(char* dev, const char* host_in, size_t in_bytes, char* host_out, size_t out_bytes,
                                   size_t block, hipStream_t h2d, hipStream_t d2h,
                                   const std::vector<std::unique_ptr<Event>>& block_out) {
  if (block == 0) throw Error("in-place transfer: block size 0");
  const size_t total = std::max(out_bytes, in_bytes);
  if ((total + block - 1) / block > block_out.size()) throw Error("in-place transfer: too few block events");
  size_t k = 0;
  for (size_t at = 0; at < total; at += block, ++k) {
    const size_t ob = at < out_bytes ? std::min(block, out_bytes - at) : 0;
    const size_t ib = at < in_bytes ? std::min(block, in_bytes - at) : 0;
    if (ob) {
      HIP_CHECK(hipMemcpyAsync(host_out + at, dev + at, ob, hipMemcpyDeviceToHost, d2h));
      block_out[k]->Record(d2h);
    }
    if (ib) {
      // the block's old contents must be out first
      if (ob) HIP_CHECK(hipStreamWaitEvent(h2d, *block_out[k], 0));
      HIP_CHECK(hipMemcpyAsync(dev + at, host_in + at, ib, hipMemcpyHostToDevice, h2d));
    }
  }
}

}  // namespace mgsort
