// Device-to-host transfers through two pinned staging buffers: the copy of
// chunk k + 1 (DMA over PCIe at the pinned rate) runs while the host consumes
// chunk k (writes it to a file, or spreads it over pageable memory with
// several threads). Used by the GEP download and the PETSc writer.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>

namespace fem2d {

class PinnedStaging {
 public:
  static constexpr size_t kChunk = size_t(32) << 20;

  PinnedStaging() = default;
  PinnedStaging(const PinnedStaging&) = delete;
  PinnedStaging& operator=(const PinnedStaging&) = delete;
  ~PinnedStaging() {
    if (stream_) (void)hipStreamSynchronize(stream_);  // no copy is in flight into a freed buffer
    for (int k = 0; k < 2; ++k) {
      if (done_[k]) (void)hipEventDestroy(done_[k]);
      if (buf_[k]) (void)hipHostFree(buf_[k]);
    }
  }

  hipError_t init(hipStream_t st) {
    stream_ = st;
    for (int k = 0; k < 2; ++k) {
      hipError_t e = hipHostMalloc(&buf_[k], kChunk, hipHostMallocDefault);
      if (e != hipSuccess) return e;
      e = hipEventCreateWithFlags(&done_[k], hipEventDisableTiming);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }

  // Streams device bytes [src, src + bytes): consume(host_chunk, offset, len)
  // is called for each chunk in order and returns 0 or a status code, which
  // ends the transfer. A HIP failure is returned in *err (and as -1).
  template <class F>
  int run(const void* src, size_t bytes, F&& consume, hipError_t* err) {
    const unsigned char* s = static_cast<const unsigned char*>(src);
    const size_t n_chunks = (bytes + kChunk - 1) / kChunk;
    auto issue = [&](size_t c) -> hipError_t {
      const size_t off = c * kChunk, len = std::min(kChunk, bytes - off);
      hipError_t e = hipMemcpyAsync(buf_[c % 2], s + off, len, hipMemcpyDeviceToHost, stream_);
      return e == hipSuccess ? hipEventRecord(done_[c % 2], stream_) : e;
    };
    *err = hipSuccess;
    if (n_chunks && (*err = issue(0)) != hipSuccess) return -1;
    for (size_t c = 0; c < n_chunks; ++c) {
      *err = hipEventSynchronize(done_[c % 2]);
      if (*err == hipSuccess && c + 1 < n_chunks) *err = issue(c + 1);
      if (*err != hipSuccess) return -1;
      const size_t off = c * kChunk;
      if (int rc = consume(static_cast<const unsigned char*>(buf_[c % 2]), off, std::min(kChunk, bytes - off)))
        return rc;
    }
    return 0;
  }

 private:
  void* buf_[2] = {nullptr, nullptr};
  hipEvent_t done_[2] = {nullptr, nullptr};
  hipStream_t stream_ = nullptr;
};

}  // namespace fem2d
