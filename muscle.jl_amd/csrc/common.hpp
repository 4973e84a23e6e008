// Internal definitions shared by the host layer and the kernel launchers.
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "muscle_hip.h"

namespace mh {

// Error carried up to the C ABI, which turns it into a status code and a
// thread-local message.
struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

[[noreturn]] inline void fail(int status, const std::string& msg) { throw Error(status, msg); }

inline bool is_complex(int dt) { return dt == MUSCLE_HIP_C32 || dt == MUSCLE_HIP_C64; }
inline bool is_double(int dt) { return dt == MUSCLE_HIP_F64 || dt == MUSCLE_HIP_C64; }
inline bool valid_dtype(int dt) { return dt >= MUSCLE_HIP_F32 && dt <= MUSCLE_HIP_C64; }
inline int make_dtype(bool dbl, bool cplx) {
  return cplx ? (dbl ? MUSCLE_HIP_C64 : MUSCLE_HIP_C32) : (dbl ? MUSCLE_HIP_F64 : MUSCLE_HIP_F32);
}
// Julia's promote_type over {Float32, Float64, ComplexF32, ComplexF64}.
inline int promote_dtype(int a, int b) {
  return make_dtype(is_double(a) || is_double(b), is_complex(a) || is_complex(b));
}
inline size_t dtype_size(int dt) {
  switch (dt) {
    case MUSCLE_HIP_F32: return 4;
    case MUSCLE_HIP_F64: return 8;
    case MUSCLE_HIP_C32: return 8;
    case MUSCLE_HIP_C64: return 16;
  }
  return 0;
}
inline const char* dtype_name(int dt) {
  switch (dt) {
    case MUSCLE_HIP_F32: return "Float32";
    case MUSCLE_HIP_F64: return "Float64";
    case MUSCLE_HIP_C32: return "ComplexF32";
    case MUSCLE_HIP_C64: return "ComplexF64";
  }
  return "?";
}

}  // namespace mh

#define MH_HIP_CHECK(expr)                                                              \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      ::mh::fail(MUSCLE_HIP_HIP_ERROR, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)
