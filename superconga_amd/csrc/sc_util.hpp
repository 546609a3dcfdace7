// sc_util.hpp -- device-memory helpers shared by the engine's translation
// units.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace scamd {

#define SC_HIP(expr)                                                            \
  do {                                                                          \
    const hipError_t err_ = (expr);                                             \
    if (err_ != hipSuccess)                                                     \
      throw std::runtime_error(std::string("HIP error in ") + #expr + ": " +    \
                               hipGetErrorString(err_));                        \
  } while (0)


// Allocates and clears.  The clear runs on the null stream, which the
// engine's non-blocking stream does not wait for, so it is completed here
// before any stream work can touch the buffer.
template <class T> inline T *devAlloc(size_t n) {
  T *p = nullptr;
  if (n == 0) n = 1;
  SC_HIP(hipMalloc(reinterpret_cast<void **>(&p), n * sizeof(T)));
  SC_HIP(hipMemset(p, 0, n * sizeof(T)));
  SC_HIP(hipDeviceSynchronize());
  return p;
}
template <class T> inline T *devUpload(const std::vector<T> &v) {
  T *p = devAlloc<T>(v.size());
  if (!v.empty())
    SC_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}
template <class T> inline void devFree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}
// Records the message sc_last_error() returns on this thread and returns
// code (defined with the C ABI in sc_capi.hip).
int reportError(int code, const std::string &message);

inline unsigned blocksFor(long long n, int bs = 256) {
  return static_cast<unsigned>((n + bs - 1) / bs);
}


} // namespace scamd
