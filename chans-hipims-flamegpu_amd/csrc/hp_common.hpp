// Error reporting shared by every C-ABI entry point: a negative status code
// plus a per-thread message readable through hp_last_error().
#pragma once

#include <hip/hip_runtime.h>
#include <string>

#include "../../include/hipims_abi.h"

namespace hp {

std::string& last_error_slot();

inline int fail(int code, const std::string& msg) {
  last_error_slot() = msg;
  return code;
}

// Report a rejected launch; clears HIP's sticky launch error.
inline int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(HP_ELAUNCH, std::string(what) + ": " + hipGetErrorString(e));
  return HP_OK;
}

}  // namespace hp
