// Shared host-side helpers: error type, HIP status checks, device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <exception>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace gpuga {

// Every failure inside the library becomes a gpuga::Error; the C-ABI layer
// turns it into a non-zero status plus gpuga_last_error() text, and the CLI
// prints it the way the reference's print_error does (src/error.cu:21-25).
struct Error : std::runtime_error {
  explicit Error(const std::string& s) : std::runtime_error(s) {}
};

// f(i) for i in [0, n) on up to 16 host threads, in contiguous blocks; f
// must write only outputs owned by i, so results do not depend on the
// thread count.  An exception from any f is rethrown after all have joined.
template <class F>
void parallel_for(int n, F&& f)
{
  const int hw = (int)std::thread::hardware_concurrency();
  const int T = std::max(1, std::min({hw > 0 ? hw : 1, 16, n / 16}));
  if (T <= 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  std::vector<std::exception_ptr> err(T);
  for (int t = 0; t < T; ++t)
    th.emplace_back([&, t] {
      try {
        for (int i = (int)((long long)n * t / T); i < (int)((long long)n * (t + 1) / T); ++i)
          f(i);
      } catch (...) {
        err[t] = std::current_exception();
      }
    });
  for (auto& x : th) x.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
}

inline std::string strfmt(const char* fmt, ...)
{
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  return buf;
}

[[noreturn]] inline void fail(const std::string& s) { throw Error(s); }

#define GPUGA_HIP(call)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      ::gpuga::fail(::gpuga::strfmt("HIP error %d (%s) at %s:%d: %s", (int)e_,             \
                                    hipGetErrorString(e_), __FILE__, __LINE__, #call));    \
  } while (0)

// Owning device allocation (hipMalloc / hipFree), move-only.
template <typename T>
class DeviceBuffer {
public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) { resize(n); }
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
  {
    if (this != &o) {
      release();
      p_ = o.p_; n_ = o.n_;
      o.p_ = nullptr; o.n_ = 0;
    }
    return *this;
  }
  void resize(size_t n)
  {
    if (n == n_) return;
    release();
    if (n) GPUGA_HIP(hipMalloc((void**)&p_, n * sizeof(T)));
    n_ = n;
  }
  void upload(const T* h, size_t n, hipStream_t s = nullptr)
  {
    resize(n);
    if (n) GPUGA_HIP(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s = nullptr) { upload(v.data(), v.size(), s); }
  void download(T* h, size_t n, hipStream_t s = nullptr) const
  {
    if (n) GPUGA_HIP(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  void zero(hipStream_t s = nullptr)
  {
    if (n_) GPUGA_HIP(hipMemsetAsync(p_, 0, n_ * sizeof(T), s));
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  size_t bytes() const { return n_ * sizeof(T); }

private:
  void release()
  {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* p_ = nullptr;
  size_t n_ = 0;
};

// Page-locked host allocation for asynchronous copies (hipHostMalloc).
template <typename T>
class PinnedBuffer {
public:
  PinnedBuffer() = default;
  ~PinnedBuffer() { release(); }
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;
  void resize(size_t n)
  {
    if (n <= cap_) { n_ = n; return; }
    release();
    GPUGA_HIP(hipHostMalloc((void**)&p_, n * sizeof(T), hipHostMallocDefault));
    cap_ = n_ = n;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  T& operator[](size_t i) { return p_[i]; }

private:
  void release()
  {
    if (p_) (void)hipHostFree(p_);
    p_ = nullptr;
    n_ = cap_ = 0;
  }
  T* p_ = nullptr;
  size_t n_ = 0, cap_ = 0;
};

} // namespace gpuga
