// Error reporting shared by every C-ABI entry point (a negative status code
// plus a per-thread message readable through hp_last_error()) and the tile
// order shared by the tiled kernels.
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

// What a calibration step records once its time control has run: the depths
// at n plane offsets into `row` (grid dtype) and, when sched is set, (t, dt)
// of the next step into sched[k_next].
struct StepRecord {
  const void* h;
  const int64_t* off;
  int n;
  void* row;
  double* sched;
  int k_next;
  int f32;
};

// Workgroups of a launch are handed to the 8 XCDs round-robin (block b runs
// on XCD b % 8), and each XCD has its own L2.  Tile numbers are reassigned so
// that every XCD walks one contiguous band of tiles: x- and y-neighbours then
// share their halo lines in one L2 instead of fetching them from HBM twice.
constexpr int kXcds = 8;

// A domain descriptor read in place from constant memory (a kernel's own
// arguments, or a descriptor array no launch writes): its fields are fetched
// by scalar loads where they are used.  launder() hides the pointer's value
// from the compiler (an empty asm that must be taken to change it), so the
// fields are fetched again after it instead of being held in scalar
// registers across long kernels, which spills them.  The pointer must be
// uniform across the wave.
typedef __attribute__((address_space(4))) const hp_grid_desc* KDesc;
template <class DP> __device__ inline DP launder(DP p) {
  asm volatile("" : "+s"(p));
  return p;
}
// The descriptor of a kernel whose first argument is its hp_grid_desc.
__device__ inline KDesc kernarg_desc() { return (KDesc)__builtin_amdgcn_kernarg_segment_ptr(); }
__device__ inline int xcd_band(int b, int n) {
  const int q = n / kXcds, rem = n - q * kXcds;
  const int x = b % kXcds, local = b / kXcds;
  return x * q + min(x, rem) + local;
}

}  // namespace hp
