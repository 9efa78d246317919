// Host-side helpers shared by the C ABI sources: HIP error checks that
// return MF_ERR_HIP with a message, and an owning device buffer.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "mf_internal.hpp"

#define MF_HIP(call)                                                   \
  do {                                                                 \
    hipError_t e_ = (call);                                            \
    if (e_ != hipSuccess) {                                            \
      ::mf::set_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
      return MF_ERR_HIP;                                               \
    }                                                                  \
  } while (0)

namespace mf {

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    release();
    n = count;
    if (count == 0) return hipSuccess;
    return hipMalloc(reinterpret_cast<void **>(&p), sizeof(T) * count);
  }
  hipError_t upload(const std::vector<T> &v, hipStream_t s) {
    hipError_t e = alloc(v.size());
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(),
                          hipMemcpyHostToDevice, s);
  }
  hipError_t ensure(size_t count) {
    if (count <= n && p) return hipSuccess;
    return alloc(count);
  }
};

}  // namespace mf
