// Shared host-side helpers of libqpsdm: error reporting and HIP checks.
#ifndef QPSDM_COMMON_H_
#define QPSDM_COMMON_H_

#include <stdarg.h>
#include <stdio.h>

#include "qpsdm.h"

namespace qpsdm {

// Records a formatted message as the thread's last error; returns `code`.
int set_error(int code, const char *fmt, ...);

}  // namespace qpsdm

#define QPSDM_REQUIRE(cond, ...)                                 \
  do {                                                           \
    if (!(cond)) return qpsdm::set_error(QPSDM_EINVAL, __VA_ARGS__); \
  } while (0)

#define QPSDM_HIP(call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return qpsdm::set_error(QPSDM_EHIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

#endif
