// Rain source terms and vector plane access shared by the step and its adjoint.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hipims_abi.h"

namespace hp {

// Vector access to kVec consecutive plane elements (16 B in float).
constexpr int kVec = 4;
template <class T> struct alignas(sizeof(T) * kVec) VecT {
  T x[kVec];
};
struct alignas(kVec) VecU8 {
  uint8_t x[kVec];
};
struct alignas(2 * kVec) VecI16 {
  int16_t x[kVec];
};
struct alignas(4 * kVec) VecI32 {
  int32_t x[kVec];
};
struct alignas(4 * kVec) VecF32 {
  float x[kVec];
};
struct alignas(8 * kVec) VecD {
  double x[kVec];
};
template <class V, class T> __device__ inline V ldv(const T* p) { return *reinterpret_cast<const V*>(p); }
template <class V, class T> __device__ inline void stv(T* p, const V& v) { *reinterpret_cast<V*>(p) = v; }

// Station rain at time t_now (SWE_CUDA.py:472-495): linear in time between
// the table rows that bracket t_now, 0 from the table's last row on, scaled
// by rain_scale.  The bracketing rows are found once per block; the passes
// keep the first stations' rates in LDS and evaluate any further station's
// rate per cell from the two rows.
struct RainRows {
  const double* lo;  // station columns of the rows bracketing t_now
  const double* hi;
  double per;        // interpolation weight
  int n_st;
  bool zero;         // every station rate is 0 (after the table, or one row)
};

// The descriptor through any pointer type (generic, or the kernel-argument
// address space).
template <class DP> __device__ inline RainRows rain_rows_at(DP a, double t_now) {
  RainRows r{nullptr, nullptr, 0.0, 0, true};
  if (a->rain_mode != HP_RAIN_STATION || a->rain_table == nullptr || a->rain_rows <= 0) return r;
  const int rows = a->rain_rows, cols = a->rain_cols;
  const double* tab = a->rain_table;
  r.n_st = cols - 1;
  const bool after = !(t_now < tab[(rows - 1) * cols]);
  int i = 0;
  while (i + 1 < rows - 1 && !(t_now < tab[(i + 1) * cols])) i++;
  r.zero = after || rows <= 1;
  if (!r.zero) {
    r.lo = tab + i * cols + 1;
    r.hi = tab + (i + 1) * cols + 1;
    r.per = (t_now - tab[i * cols]) / (tab[(i + 1) * cols] - tab[i * cols]);
  }
  return r;
}

__device__ inline RainRows rain_rows(const hp_grid_desc& a, double t_now) { return rain_rows_at(&a, t_now); }

// Rain rate of a cell whose station index is k.
template <class T> __device__ inline T station_rate(const RainRows& rr, int k, double scale) {
  if (rr.zero || k < 0 || k >= rr.n_st) return T(0);
  const double lo = rr.lo[k], hi = rr.hi[k];
  return T((lo + rr.per * (hi - lo)) * scale);
}

}  // namespace hp
