// Shared plumbing of the HIP library: status codes, the thread-local error
// message, argument checks and HIP error mapping.  Compiled into every .hip
// translation unit of libgaffa_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <stdexcept>
#include <string>

#include "gaffa_hip.h"

namespace gaffa_hip {

void set_error(const std::string& message);

struct status_error : std::runtime_error {
  int code;
  status_error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(bool condition, const char* message) {
  if (!condition) throw status_error(GAFFA_HIP_INVALID_ARGUMENT, message);
}

inline void check_overflow(bool condition, const char* message) {
  if (!condition) throw status_error(GAFFA_HIP_OVERFLOW, message);
}

inline void hip_check(hipError_t err, const char* what) {
  if (err != hipSuccess) {
    throw status_error(GAFFA_HIP_RUNTIME,
                       std::string(what) + ": " + hipGetErrorString(err));
  }
}

// Launch errors surface through hipGetLastError right after the launch.
inline void launch_check(const char* kernel) {
  hip_check(hipGetLastError(), kernel);
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return GAFFA_HIP_OK;
  } catch (const status_error& e) {
    set_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return GAFFA_HIP_RUNTIME;
  } catch (const std::exception& e) {
    set_error(e.what());
    return GAFFA_HIP_RUNTIME;
  }
}

inline hipStream_t as_stream(gaffa_hip_stream s) {
  return reinterpret_cast<hipStream_t>(s);
}

constexpr uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

}  // namespace gaffa_hip
