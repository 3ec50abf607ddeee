// Shared definitions for the MI355X thermal-solve library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#define THD __host__ __device__ inline

namespace thesis {

// Error raised inside the library; converted to a status code at the C-ABI.
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
  kOk = 0,
  kBadInput = 1,      // malformed or out-of-scope input
  kBadArgument = 2,   // invalid argument to an entry point
  kDevice = 3,        // HIP runtime failure
  kCapacity = 4,      // a caller-provided buffer is too small
  kNoDevice = 5,      // no gfx950 device visible
};

void hip_fail(hipError_t e, const char* what, const char* file, int line);

#define THESIS_HIP(call)                                              \
  do {                                                                \
    hipError_t e_ = (call);                                           \
    if (e_ != hipSuccess) ::thesis::hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)

}  // namespace thesis
