// RAII device buffer over the C-ABI device layer (the reference's src/Buffer.h:16-102).
#pragma once

#include "owl_hip.h"

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

struct GpuError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void owlCheck(int rc, const char* what) {
  if (rc) { throw GpuError(std::string(what) + ": " + owl_last_error()); }
}

template<typename T>
class Buffer {
  T* ptr_ = nullptr;
  size_t size_ = 0;

public:
  Buffer() = default;
  explicit Buffer(size_t n) : size_{n} {
    void* p = nullptr;
    owlCheck(owl_hip_alloc(&p, n * sizeof(T)), "alloc");
    ptr_ = static_cast<T*>(p);
  }
  Buffer(owl_queue q, const std::vector<T>& host) : Buffer(host.size()) { write(q, host); }
  ~Buffer() { owl_hip_free(ptr_); }
  Buffer(const Buffer&) = delete;
  Buffer& operator=(const Buffer&) = delete;
  Buffer(Buffer&& o) noexcept : ptr_{o.ptr_}, size_{o.size_} {
    o.ptr_ = nullptr;
    o.size_ = 0;
  }
  Buffer& operator=(Buffer&& o) noexcept {
    if (this != &o) {
      owl_hip_free(ptr_);
      ptr_ = o.ptr_;
      size_ = o.size_;
      o.ptr_ = nullptr;
      o.size_ = 0;
    }
    return *this;
  }

  T* get() const { return ptr_; }
  size_t size() const { return size_; }

  void write(owl_queue q, const std::vector<T>& host) {
    if (host.size() > size_) { throw std::invalid_argument("buffer write too large"); }
    owlCheck(owl_hip_write(q, ptr_, host.data(), host.size() * sizeof(T), 1), "write");
  }
  std::vector<T> read(owl_queue q, size_t n) const {
    std::vector<T> out(n);
    owlCheck(owl_hip_read(q, out.data(), ptr_, n * sizeof(T), 1), "read");
    return out;
  }
  std::vector<T> read(owl_queue q) const { return read(q, size_); }
  void zero(owl_queue q) {
    if ((size_ * sizeof(T)) % 4) { throw std::invalid_argument("zero: size not a multiple of 4"); }
    owlCheck(owl_hip_fill32(q, ptr_, 0, size_ * sizeof(T) / 4), "fill");
  }
};
