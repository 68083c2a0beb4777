// Owned device allocation (hipMalloc / hipFree), movable, not copyable.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstddef>
#include <utility>

#include "common.hpp"

namespace stl {

// Bytes currently held by DevArray allocations of this process.
std::atomic<size_t>& device_bytes_live();
std::atomic<size_t>& device_bytes_peak();

template <typename T>
class DevArray {
 public:
  DevArray() = default;
  explicit DevArray(size_t n) { resize(n); }
  ~DevArray() { release(); }
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
  DevArray(DevArray&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevArray& operator=(DevArray&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_; n_ = o.n_;
      o.p_ = nullptr; o.n_ = 0;
    }
    return *this;
  }

  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) {
      STL_HIP(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)));
      n_ = n;
      device_bytes_live() += bytes();
      size_t live = device_bytes_live().load();
      if (live > device_bytes_peak().load()) device_bytes_peak() = live;
    }
  }

  void release() {
    if (p_) {
      (void)hipFree(p_);
      device_bytes_live() -= bytes();
    }
    p_ = nullptr;
    n_ = 0;
  }

  void zero(hipStream_t s) {
    if (n_) STL_HIP(hipMemsetAsync(p_, 0, bytes(), s));
  }

  T* data() { return p_; }
  const T* data() const { return p_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// Page-locked host allocation (hipHostMalloc): staging of the statistics and
// cotangents, so that their copies are plain DMA transfers.
template <typename T>
class PinnedArray {
 public:
  PinnedArray() = default;
  ~PinnedArray() { release(); }
  PinnedArray(const PinnedArray&) = delete;
  PinnedArray& operator=(const PinnedArray&) = delete;

  void resize(size_t n) {
    if (n == n_) return;
    release();
    if (n) {
      STL_HIP(hipHostMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T), hipHostMallocDefault));
      n_ = n;
    }
  }
  void release() {
    if (p_) (void)hipHostFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* data() { return p_; }
  const T* data() const { return p_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

}  // namespace stl
