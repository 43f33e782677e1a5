// Device buffers for the benchmarks.  Replaces the reference's
// PrefetchedBuffer (include/memory_wrappers.h:39-92): the allocation and the
// host-to-device copy happen in initialize(), i.e. inside the untimed setup(),
// and results come back to the host only through toHost() in verify().
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <vector>

#include "hip_device.h"

template <class T>
class DeviceBuffer {
public:
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), count_(o.count_) {
    o.ptr_ = nullptr;
    o.count_ = 0;
  }
  ~DeviceBuffer() { release(); }

  // Allocates `count` elements (contents undefined).
  void initialize(HipDevice&, std::size_t count) {
    release();
    count_ = count;
    if(count) hipCheck(hipMalloc(reinterpret_cast<void**>(&ptr_), count * sizeof(T)), "hipMalloc");
  }
  // Allocates and copies `count` elements from the host (stream-ordered).
  void initialize(HipDevice& d, const T* host, std::size_t count) {
    initialize(d, count);
    if(count)
      hipCheck(hipMemcpyAsync(ptr_, host, count * sizeof(T), hipMemcpyHostToDevice, d.stream()), "hipMemcpyAsync");
  }
  // Device-to-device copy of another buffer of the same size.
  void copyFrom(HipDevice& d, const DeviceBuffer& src) {
    if(src.count_ != count_) initialize(d, src.count_);
    if(count_)
      hipCheck(hipMemcpyAsync(ptr_, src.ptr_, count_ * sizeof(T), hipMemcpyDeviceToDevice, d.stream()), "hipMemcpyAsync");
  }
  std::vector<T> toHost(HipDevice& d) const {
    std::vector<T> h(count_);
    if(count_) {
      hipCheck(hipMemcpyAsync(h.data(), ptr_, count_ * sizeof(T), hipMemcpyDeviceToHost, d.stream()), "hipMemcpyAsync");
      d.synchronize();
    }
    return h;
  }

  T* get() const { return ptr_; }
  std::size_t size() const { return count_; }

private:
  void release() {
    if(ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    count_ = 0;
  }
  T* ptr_ = nullptr;
  std::size_t count_ = 0;
};
