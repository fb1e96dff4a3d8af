// Shared helpers of the op2hip C-ABI library: error recording and HIP call
// checking.  No C++ exception crosses the ABI: every entry point returns an
// int status and leaves a message for op2hip_last_error().
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

namespace op2hip {

void set_error(const char* fmt, ...);

}  // namespace op2hip

enum {
  OP2HIP_OK = 0,
  OP2HIP_ERR_HIP = -1,
  OP2HIP_ERR_COMPILE = -2,
  OP2HIP_ERR_ARG = -3,
  OP2HIP_ERR_ALLOC = -4,
  OP2HIP_ERR_INTERNAL = -5,
};

#define OP2HIP_CHECK(expr)                                                    \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      op2hip::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                        __FILE__, __LINE__);                                  \
      return OP2HIP_ERR_HIP;                                                  \
    }                                                                         \
  } while (0)

#define OP2HIP_REQUIRE(cond, ...)      \
  do {                                 \
    if (!(cond)) {                     \
      op2hip::set_error(__VA_ARGS__);  \
      return OP2HIP_ERR_ARG;           \
    }                                  \
  } while (0)
