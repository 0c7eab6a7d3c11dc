// Shared helpers of the compatibility headers: status -> exception mapping.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "mi355sssp.h"

namespace mi355_compat {

// Throws the exception type the reference family uses for each failure
// (std::invalid_argument for arguments, std::out_of_range for the source of
// delta-stepping, std::overflow_error for sizes, std::runtime_error for HIP /
// capacity failures; DS/src/delta_stepping_hip_CSR.cpp:142-185).
inline void throw_on_error(int status) {
  if (status == MI355_OK) return;
  const std::string msg = mi355_last_error();
  switch (status) {
    case MI355_EINVAL: throw std::invalid_argument(msg);
    case MI355_ERANGE: throw std::out_of_range(msg);
    case MI355_EOVERFLOW: throw std::overflow_error(msg);
    default: throw std::runtime_error(msg);
  }
}

// hipError_t for the reference's minplus_* functions, which never throw.
inline hipError_t to_hip_error(int status) {
  if (status == MI355_OK) return hipSuccess;
  if (status == MI355_EHIP && mi355_last_hip_error() != 0) return (hipError_t)mi355_last_hip_error();
  if (status == MI355_ENOMEM) return hipErrorOutOfMemory;
  return hipErrorInvalidValue;
}

}  // namespace mi355_compat
