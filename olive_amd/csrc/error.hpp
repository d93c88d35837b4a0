// Status/error plumbing shared by every C-ABI entry point.
//
// The reference reports errors with CUDA_CHECK -> stderr + assert(false)
// (common.h:105-113) and asserts on contract violations (oliver.h:59-60).
// Across a C ABI nothing may abort or throw, so each entry point returns an
// OLV_* status and leaves a thread-local message for olv_last_error().
#pragma once
#include <cstdarg>
#include <cstdio>

namespace olv {

char *error_buffer();

// Record a formatted message and return `code`.
inline int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(error_buffer(), 1024, fmt, ap);
  va_end(ap);
  return code;
}

int host_threads();

}  // namespace olv

// Return OLV_EHIP from the enclosing function when a HIP call fails.
#define OLV_HIP(call)                                                        \
  do {                                                                       \
    hipError_t olv_e_ = (call);                                              \
    if (olv_e_ != hipSuccess)                                                \
      return ::olv::fail(OLV_EHIP, "%s:%d: %s failed: %s", __FILE__,         \
                         __LINE__, #call, hipGetErrorString(olv_e_));        \
  } while (0)

// Propagate a non-OK status from a helper.
#define OLV_TRY(expr)                   \
  do {                                  \
    int olv_s_ = (expr);                \
    if (olv_s_ != OLV_OK) return olv_s_; \
  } while (0)
