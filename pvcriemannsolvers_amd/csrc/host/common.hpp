// Shared host-side helpers of the C ABI: error capture, HIP checks and
// owning device arrays.
#ifndef MINICFD_HOST_COMMON_HPP_
#define MINICFD_HOST_COMMON_HPP_

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/minicfd.h"

namespace mcfd {

// Last error of the calling thread, returned by mcfd_last_error().
std::string &LastError();

struct Failure {
  int code;
  std::string what;
};

inline void Check(hipError_t e, const char *where) {
  if (e != hipSuccess) throw Failure{ MCFD_EHIP, std::string(where) + ": " + hipGetErrorString(e) };
}

// Runs f and maps exceptions to MCFD_* codes; nothing crosses the C ABI.
template <class F>
int Guard(F &&f) {
  try {
    f();
    return MCFD_OK;
  } catch (Failure const &e) {
    LastError() = e.what;
    return e.code;
  } catch (std::exception const &e) {
    LastError() = e.what();
    return MCFD_EINVAL;
  }
}

inline int Fail(int code, const char *what) {
  LastError() = what;
  return code;
}

template <class T>
struct DeviceArray {
  T *ptr = nullptr;
  size_t n = 0;
  DeviceArray() = default;
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  void Allocate(size_t count) {
    Release();
    if (count == 0) return;
    Check(hipMalloc(&ptr, count * sizeof(T)), "hipMalloc");
    n = count;
  }
  void Upload(const std::vector<T> &v, hipStream_t s) {
    Allocate(v.size());
    if (n) Check(hipMemcpyAsync(ptr, v.data(), n * sizeof(T), hipMemcpyHostToDevice, s), "upload");
  }
  void Release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  ~DeviceArray() { Release(); }
};

}  // namespace mcfd

#endif  // MINICFD_HOST_COMMON_HPP_
