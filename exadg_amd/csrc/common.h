// Shared host-side helpers: error type, HIP checks, a small thread pool loop.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/exadg_amd.h"

namespace ExaDG
{
// Error carrying an exadg_status code; thrown inside the library, converted to a status
// code at the C ABI (the role of AssertThrow/ExcMessage in the reference).
struct Error : std::runtime_error
{
  int code;
  Error(int c, std::string const & msg) : std::runtime_error(msg), code(c) {}
};

#define EXADG_THROW(code, msg) throw ::ExaDG::Error((code), (msg))

#define HIP_CHECK(expr)                                                                        \
  do                                                                                           \
  {                                                                                            \
    hipError_t const e_ = (expr);                                                              \
    if(e_ != hipSuccess)                                                                       \
      EXADG_THROW(EXADG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));          \
  } while(0)

// parallel loop over [0, n) on the host (geometry setup); f(begin, end)
template<typename F>
void parallel_range(int64_t n, F const & f)
{
  unsigned const nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  if(n < 64 || nt == 1)
  {
    f(int64_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  int64_t const            chunk = (n + nt - 1) / nt;
  for(unsigned t = 0; t < nt; ++t)
  {
    int64_t const b = t * chunk, e = std::min(n, b + chunk);
    if(b < e)
      th.emplace_back([&f, b, e] { f(b, e); });
  }
  for(auto & t : th)
    t.join();
}

// Owning device buffer (hipMalloc'd), move-only.
template<typename T>
class DeviceBuffer
{
public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) { resize(n); }
  ~DeviceBuffer() { release(); }
  DeviceBuffer(DeviceBuffer const &) = delete;
  DeviceBuffer & operator=(DeviceBuffer const &) = delete;
  DeviceBuffer(DeviceBuffer && o) noexcept : ptr_(o.ptr_), n_(o.n_)
  {
    o.ptr_ = nullptr;
    o.n_   = 0;
  }
  DeviceBuffer & operator=(DeviceBuffer && o) noexcept
  {
    if(this != &o)
    {
      release();
      ptr_   = o.ptr_;
      n_     = o.n_;
      o.ptr_ = nullptr;
      o.n_   = 0;
    }
    return *this;
  }
  void resize(size_t n)
  {
    release();
    if(n > 0)
      HIP_CHECK(hipMalloc(&ptr_, n * sizeof(T)));
    n_ = n;
  }
  void upload(std::vector<T> const & h)
  {
    resize(h.size());
    if(!h.empty())
      HIP_CHECK(hipMemcpy(ptr_, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  }
  T *       data() { return ptr_; }
  T const * data() const { return ptr_; }
  size_t    size() const { return n_; }

private:
  void release()
  {
    if(ptr_)
      (void)hipFree(ptr_);
    ptr_ = nullptr;
    n_   = 0;
  }
  T *    ptr_ = nullptr;
  size_t n_   = 0;
};

}  // namespace ExaDG
