// Device memory, streams and error plumbing shared by the host-side modules.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace swfwi {

// Status codes of the C ABI (mirrored in include/swfwi_hip.h).
enum Status : int {
  kOk = 0,
  kErrArg = 1,      // invalid argument / geometry (the reference exit(1)s here)
  kErrState = 2,    // call out of order (e.g. no velocity bound)
  kErrHip = 3,      // HIP runtime failure
  kErrNumeric = 4,  // numerical failure (e.g. eigen-solve did not converge)
  kErrRange = 5,    // size beyond a supported maximum
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Error(code, msg); }

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(kErrHip, std::string(what) + ": " + hipGetErrorString(e));
}
#define SWFWI_HIP(call) ::swfwi::hip_check((call), #call)

// Owning device buffer (hipMalloc'd), move-only, grow-only on resize.
template <typename T>
class DBuf {
 public:
  DBuf() = default;
  explicit DBuf(size_t n) { resize(n); }
  ~DBuf() { release(); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p_(o.p_), n_(o.n_), cap_(o.cap_) { o.p_ = nullptr; o.n_ = o.cap_ = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; cap_ = o.cap_; o.p_ = nullptr; o.n_ = o.cap_ = 0; }
    return *this;
  }
  void resize(size_t n) {
    if (n > cap_) {
      release();
      void* q = nullptr;
      SWFWI_HIP(hipMalloc(&q, (n ? n : 1) * sizeof(T)));
      p_ = static_cast<T*>(q);
      cap_ = n;
    }
    n_ = n;
  }
  void upload(const T* h, size_t n, hipStream_t s) {
    resize(n);
    if (n) SWFWI_HIP(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& h, hipStream_t s) { upload(h.data(), h.size(), s); }
  void download(T* h, size_t n, hipStream_t s) const {
    if (n) SWFWI_HIP(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = cap_ = 0;
  }

 private:
  T* p_ = nullptr;
  size_t n_ = 0, cap_ = 0;
};

// A HIP stream owned by one handle; default priority on purpose.
class Stream {
 public:
  Stream() { SWFWI_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking)); }
  ~Stream() { if (s_) (void)hipStreamDestroy(s_); }
  Stream(const Stream&) = delete;
  Stream& operator=(const Stream&) = delete;
  hipStream_t get() const { return s_; }
  void sync() const { SWFWI_HIP(hipStreamSynchronize(s_)); }

 private:
  hipStream_t s_ = nullptr;
};

}  // namespace swfwi
