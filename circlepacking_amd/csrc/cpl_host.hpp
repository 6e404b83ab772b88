// Host-side helpers shared by the library's translation units: the error
// string behind cpl_last_error(), HIP error propagation and a growable device
// buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "cpl.h"

namespace cpl_host {

inline thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                          \
  do {                                                                                         \
    hipError_t _e = (expr);                                                                    \
    if (_e != hipSuccess)                                                                      \
      return ::cpl_host::fail(CPL_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t need) {
    if (need <= bytes) return CPL_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, need);
    if (e != hipSuccess) return fail(CPL_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    bytes = need;
    return CPL_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// Library context accessors (defined in cpl_hip.hip).
int ensure_init();
hipStream_t stream();

}  // namespace cpl_host
