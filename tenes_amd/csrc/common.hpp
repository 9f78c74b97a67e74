// Shared helpers for the MI355X (gfx950) CTMRG library: error handling and
// small device utilities.  Everything here is host+device C++17 compiled by
// hipcc for gfx950 only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace tn {

// Every failure inside the library becomes a tn::Error; the C ABI converts it
// to a negative status code and a per-context error string.
struct Error : public std::runtime_error {
  int code;
  explicit Error(const std::string &msg, int code = -1)
      : std::runtime_error(msg), code(code) {}
};

#define TN_HIP_CHECK(expr)                                                    \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      throw ::tn::Error(std::string("HIP error '") + hipGetErrorString(_e) +  \
                            "' at " + __FILE__ + ":" +                        \
                            std::to_string(__LINE__) + " in " #expr,          \
                        -2);                                                  \
    }                                                                         \
  } while (0)

#define TN_REQUIRE(cond, msg)                                                 \
  do {                                                                        \
    if (!(cond)) throw ::tn::Error(std::string(msg), -3);                     \
  } while (0)

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

}  // namespace tn
