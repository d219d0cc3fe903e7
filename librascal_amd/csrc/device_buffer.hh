// Growable device buffer owned by a context (hipMalloc / hipFree).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "common.hh"

namespace rgpu {

// Set while an allocation-free entry point runs on this thread: a buffer
// that would have to grow throws instead of allocating.
inline thread_local bool g_no_alloc = false;

struct NoAllocScope {
  NoAllocScope() { g_no_alloc = true; }
  ~NoAllocScope() { g_no_alloc = false; }
  NoAllocScope(const NoAllocScope&) = delete;
  NoAllocScope& operator=(const NoAllocScope&) = delete;
};

[[noreturn]] inline void refuse_allocation() {
  throw RuntimeError(
      "this call allocates no device memory, and a buffer of the context is too small for the "
      "current batch: run rgpu_gap_predict once for this batch first");
}

template <class T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  ~DeviceBuffer() { release(); }

  // Ensure capacity for n elements (contents are not preserved on growth).
  T* reserve(size_t n) {
    if (n > cap_) {
      if (g_no_alloc) refuse_allocation();
      release();
      size_t bytes = (n == 0 ? 1 : n) * sizeof(T);
      RGPU_HIP(hipMalloc(reinterpret_cast<void**>(&ptr_), bytes));
      cap_ = n;
    }
    size_ = n;
    return ptr_;
  }
  void release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    cap_ = size_ = 0;
  }
  T* get() const { return ptr_; }
  size_t size() const { return size_; }
  size_t bytes() const { return size_ * sizeof(T); }
  size_t capacity() const { return cap_; }

 private:
  T* ptr_ = nullptr;
  size_t cap_ = 0, size_ = 0;
};

// Growable page-locked host buffer (hipHostMalloc / hipHostFree) for
// staging small transfers asynchronously.
template <class T>
class PinnedBuffer {
 public:
  PinnedBuffer() = default;
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;
  ~PinnedBuffer() { release(); }

  T* reserve(size_t n) {
    if (n > cap_) {
      if (g_no_alloc) refuse_allocation();
      release();
      RGPU_HIP(hipHostMalloc(reinterpret_cast<void**>(&ptr_), (n == 0 ? 1 : n) * sizeof(T),
                             hipHostMallocDefault));
      cap_ = n;
    }
    return ptr_;
  }
  void release() {
    if (ptr_) (void)hipHostFree(ptr_);
    ptr_ = nullptr;
    dev_ = nullptr;
    cap_ = 0;
  }
  T* get() const { return ptr_; }
  // The device address of the buffer: page-locked memory is mapped into the
  // device's address space, so kernels read and write it directly (no copy
  // engine, no wait between engines).
  T* device() {
    if (ptr_ && !dev_) RGPU_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&dev_), ptr_, 0));
    return dev_;
  }

 private:
  T* ptr_ = nullptr;
  T* dev_ = nullptr;
  size_t cap_ = 0;
};

}  // namespace rgpu
