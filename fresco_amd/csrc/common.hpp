// Shared host-side utilities of libfresco_hip: status/error plumbing, device
// buffers with failure injection, and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace fresco_hip {

// Exceptions thrown inside the library and mapped to status codes at the C ABI.
struct InvalidArgument : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct RuntimeFailure : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct HipFailure : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void set_last_error(const std::string& msg);
bool consume_fault();  // true if an injected allocation failure is pending

#define FRESCO_HIP_CHECK(expr)                                                      \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess) {                                                         \
      throw ::fresco_hip::HipFailure(std::string(#expr) + " failed: " +             \
                                     hipGetErrorString(_e) + " at " + __FILE__ +    \
                                     ":" + std::to_string(__LINE__));               \
    }                                                                               \
  } while (0)

// Device allocation: owning (hipMalloc'd by allocate, freed on destruction)
// or borrowed (borrow: a view of memory someone else owns, never freed).
template <typename T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) { allocate(n); }
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), n_(o.n_), owned_(o.owned_) {
    o.ptr_ = nullptr;
    o.n_ = 0;
  }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      release();
      ptr_ = o.ptr_;
      n_ = o.n_;
      owned_ = o.owned_;
      o.ptr_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  void allocate(size_t n) {
    release();
    if (n == 0) return;
    if (consume_fault()) throw HipFailure("injected device allocation failure");
    void* p = nullptr;
    FRESCO_HIP_CHECK(hipMalloc(&p, n * sizeof(T)));
    ptr_ = static_cast<T*>(p);
    n_ = n;
    owned_ = true;
  }
  // n elements at p (device memory the caller keeps alive); no allocation
  void borrow(T* p, size_t n) {
    release();
    ptr_ = n ? p : nullptr;
    n_ = n;
    owned_ = false;
  }
  void release() {
    if (ptr_ && owned_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    n_ = 0;
    owned_ = false;
  }
  void upload(const T* host, size_t n, hipStream_t s = nullptr) {
    if (n > n_) throw RuntimeFailure("upload larger than buffer");
    FRESCO_HIP_CHECK(hipMemcpyAsync(ptr_, host, n * sizeof(T), hipMemcpyHostToDevice, s));
    FRESCO_HIP_CHECK(hipStreamSynchronize(s));
  }
  T* get() const { return ptr_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

 private:
  T* ptr_ = nullptr;
  size_t n_ = 0;
  bool owned_ = false;
};

// Carves 256-byte-aligned pieces out of one caller-owned device region.
class Carve {
 public:
  Carve(void* base, size_t bytes) : p_(static_cast<char*>(base)), left_(bytes) {}
  static size_t rounded(size_t bytes) { return (bytes + 255) & ~(size_t)255; }
  template <typename T>
  T* take(size_t n) {
    const size_t b = rounded(n * sizeof(T));
    if (b > left_) throw InvalidArgument("workspace too small");
    T* r = reinterpret_cast<T*>(p_);
    p_ += b;
    left_ -= b;
    return r;
  }
  template <typename T>
  void take(DeviceBuffer<T>& buf, size_t n) {
    buf.borrow(take<T>(n), n);
  }

 private:
  char* p_;
  size_t left_;
};

inline unsigned blocks_for(size_t n, unsigned threads) {
  return static_cast<unsigned>((n + threads - 1) / threads);
}

}  // namespace fresco_hip
