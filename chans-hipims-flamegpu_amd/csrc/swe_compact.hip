// The reference's torch-extension op surface on its own compact layout:
// arrays of the N valid cells in row-major raster order and the int32[5*N]
// neighbour table (row 0 mask code, rows 1..4 the E, N, W, S neighbour or -1),
// built by Godunov.init__fluidField_tensor (SWE_CUDA.py:115-227).
//
// These entry points exist so that code calling the reference's ops directly
// (fluxMask.update, ...addFlux, addStation_Precipitation,
// addFriction_eulerUpdate, updateTimestep, euler_update.update) can switch to
// this library unchanged.  They share the edge physics of swe_physics.hpp
// with the structured path; each thread owns one listed cell, as in the
// reference, so all writes are to the thread's own cell.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hipims_abi.h"
#include "hp_common.hpp"
#include "swe_physics.hpp"

namespace hp {

constexpr int kCThreads = 256;

inline unsigned nblocks(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

template <class T>
__device__ inline int c_series_row(const T* tab, int rows, int w, T t) {
  int r = 0;
  while (r < rows - 1 && t > tab[r * w]) r++;
  return r;
}
template <class T>
__device__ inline T c_series_at(const T* tab, int rows, int w, int row, int col) {
  if (tab == nullptr || col < 0 || col >= w || row < 0 || row >= rows) return T(0);
  return tab[row * w + col];
}

template <class T, int ORDER>
__global__ __launch_bounds__(kCThreads) void k_flux_compact(
    int64_t M, int64_t N, const int32_t* __restrict__ wet, T* __restrict__ dh, T* __restrict__ dqx,
    T* __restrict__ dqy, const T* __restrict__ h, const T* __restrict__ wl, const T* __restrict__ z,
    const T* __restrict__ qx, const T* __restrict__ qy, const int32_t* __restrict__ index, const T* gd, int gd_rows,
    int gd_w, const T* gwl, int gwl_rows, int gwl_w, const T* gq, int gq_rows, int gq_w, const T* dxp, const T* tp,
    const T* dtp) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= M) return;
  const T dx = dxp[0], t = tp[0], dt = dtp[0];
  const T dt_dx = dt / dx;
  const int32_t i = wet[w];
  const int32_t code = index[i];
  const T zref = Prec<T>::local_datum ? z[i] : T(0);
  T ah = dh[i], aqx = dqx[i], aqy = dqy[i];
  for (int d = 0; d < 4; ++d) {
    const int32_t rid = index[(int64_t)(d + 1) * N + i];
    const int32_t lid = index[(int64_t)(((d + 2) & 3) + 1) * N + i];
    FrameCell<T> c;
    c.h = h[i]; c.qn = rot_n(qx[i], qy[i], d); c.qt = rot_t(qx[i], qy[i], d); c.z = z[i] - zref;
    const T c_wl = wl[i] - zref;
    EdgeOut<T> o;
    if (rid == -1) {
      int type, id;
      decode_code(code, type, id);
      const T bd = c_series_at(gd, gd_rows, gd_w, c_series_row(gd, gd_rows, gd_w, t), id + 1);
      const T bw = c_series_at(gwl, gwl_rows, gwl_w, c_series_row(gwl, gwl_rows, gwl_w, t), id + 1) - zref;
      const int rq = c_series_row(gq, gq_rows, gq_w, t);
      const T bqx = c_series_at(gq, gq_rows, gq_w, rq, 1 + 2 * id);
      const T bqy = c_series_at(gq, gq_rows, gq_w, rq, 2 + 2 * id);
      const FrameCell<T> g = bc_ghost<T>(type, c, d, bd, bw, bqx, bqy);
      const T etaC = ORDER == 1 ? c.h + c.z : c_wl;
      const T etaR = ORDER == 1 ? g.h + g.z : ghost_wl<T>(type, c_wl, g, bw);
      o = solve_edge<T, ORDER>(c, c, g, g, true, etaC, etaR, dx, dt, dt_dx);
    } else {
      const int32_t rrid = index[(int64_t)(d + 1) * N + rid];
      FrameCell<T> r, l, rr;
      r.h = h[rid]; r.qn = rot_n(qx[rid], qy[rid], d); r.qt = rot_t(qx[rid], qy[rid], d); r.z = z[rid] - zref;
      if (lid == -1) {
        l = c;
      } else {
        l.h = h[lid]; l.qn = rot_n(qx[lid], qy[lid], d); l.qt = rot_t(qx[lid], qy[lid], d); l.z = z[lid] - zref;
      }
      if (rrid == -1) {
        rr = r;
      } else {
        rr.h = h[rrid]; rr.qn = rot_n(qx[rrid], qy[rrid], d); rr.qt = rot_t(qx[rrid], qy[rrid], d);
        rr.z = z[rrid] - zref;
      }
      const T etaC = ORDER == 1 ? c.h + c.z : c_wl;
      const T etaR = ORDER == 1 ? r.h + r.z : wl[rid] - zref;
      o = solve_edge<T, ORDER>(c, l, r, rr, false, etaC, etaR, dx, dt, dt_dx);
    }
    if (!o.active) continue;
    T gx, gy;
    switch (d) {
      case 0: gx = o.fn; gy = o.ft; break;
      case 1: gx = -o.ft; gy = o.fn; break;
      case 2: gx = -o.fn; gy = -o.ft; break;
      default: gx = o.ft; gy = -o.fn; break;
    }
    ah -= o.fh;
    aqx -= gx;
    aqy -= gy;
    if (d == 0 || d == 2) aqx -= o.bed_c * T(nrm_x(d));
    else aqy -= o.bed_c * T(nrm_y(d));
  }
  dh[i] = ah;
  dqx[i] = aqx;
  dqy[i] = aqy;
}

template <class T>
__global__ void k_flux_mask(int64_t N, uint8_t* __restrict__ mask, const T* __restrict__ h,
                            const int32_t* __restrict__ index, const T* tp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (tp[0] == T(0)) {
    mask[i] = 1;
    return;
  }
  bool m = gt_hs(h[i]);
  for (int k = 1; k < 5; ++k) {
    const int32_t n = index[(int64_t)k * N + i];
    if (n > 0) m = m || gt_hs(h[n]);  // id 0 never counts (fluxMaskGenerator_Kernel.cu:36)
  }
  mask[i] = m ? 1 : 0;
}

template <class T, class R>
__global__ void k_station_rain(int64_t N, T* __restrict__ dh, const int16_t* __restrict__ st,
                               const R* __restrict__ rain, const T* dtp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  dh[i] += T(rain[st[i]]) * dtp[0];
}

template <class T>
__global__ void k_friction_compact(int64_t K, const int32_t* __restrict__ wet, T* __restrict__ dh,
                                   T* __restrict__ dqx, T* __restrict__ dqy, T* __restrict__ dz,
                                   const uint8_t* __restrict__ lu, T* __restrict__ h, T* __restrict__ wl,
                                   T* __restrict__ qx, T* __restrict__ qy, T* __restrict__ z,
                                   const T* __restrict__ man, const T* dtp) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= K) return;
  const int32_t i = wet[j];
  const T dt = dtp[0];
  T hi = h[i];
  const bool dry = new_depth_dry(hi, dh[i]);
  T qxi, qyi;
  if (dry) {
    qxi = T(0);
    qyi = T(0);
  } else {
    qxi = qx[i] + dqx[i];
    qyi = qy[i] + dqy[i];
    implicit_friction<T>(hi, man[lu[i]], dt, qxi, qyi);
  }
  h[i] = hi;
  qx[i] = qxi;
  qy[i] = qyi;
  const T zi = z[i] + dz[i];
  z[i] = zi;
  wl[i] = zi + hi;
  dh[i] = T(0);
  dqx[i] = T(0);
  dqy[i] = T(0);
  dz[i] = T(0);
}

template <class T>
__global__ void k_time_control_compact(int64_t K, const int32_t* __restrict__ wet, T* __restrict__ accel,
                                       T* __restrict__ h_max, const T* __restrict__ h, const T* __restrict__ qx,
                                       const T* __restrict__ qy, const T* dxp, const T* cflp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K) return;
  const int32_t j = wet[i];
  const T hj = h[j];
  h_max[j] = fmax(hj, h_max[j]);
  accel[i] = cfl_candidate<T>(hj, qx[j], qy[j], cflp[0] * dxp[0]);
}

template <class T>
__global__ void k_euler_compact(int64_t K, const int32_t* __restrict__ wet, const T* __restrict__ dh,
                                const T* __restrict__ dqx, const T* __restrict__ dqy, const T* __restrict__ dz,
                                T* __restrict__ h, T* __restrict__ wl, T* __restrict__ z, T* __restrict__ qx,
                                T* __restrict__ qy) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= K) return;
  const int32_t i = wet[j];
  const T hi = h[i] + dh[i];
  const T zi = z[i] + dz[i];
  h[i] = hi;
  qx[i] += dqx[i];
  qy[i] += dqy[i];
  z[i] = zi;
  wl[i] = hi + zi;
}

// infiltrationCUDA_Kernel.cu:20-46, one thread per listed cell.
template <class T>
__global__ void k_infiltration_compact(int64_t M, const int32_t* __restrict__ wet, T* __restrict__ dh,
                                       const uint8_t* __restrict__ landuse, const T* __restrict__ h,
                                       const T* __restrict__ ks, const T* __restrict__ phi,
                                       const T* __restrict__ dtheta, T* __restrict__ F, const T* __restrict__ dtp) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  const int32_t i = wet[j];
  const int c = landuse[i];
  const T dF = green_ampt<T>(h[i], F[i], ks[c], phi[c], dtheta[c], dtp[0]);
  F[i] += dF;
  dh[i] -= dF;
}

static int check_dtype(int dtype) {
  if (dtype != HP_F32 && dtype != HP_F64) return fail(HP_EINVAL, "dtype must be HP_F32 or HP_F64");
  return HP_OK;
}

template <class T, int ORDER>
static void launch_flux_compact(int64_t M, int64_t N, const int32_t* wet, void* dh, void* dqx, void* dqy,
                                const void* h, const void* wl, const void* z, const void* qx, const void* qy,
                                const int32_t* index, const void* gd, int gd_rows, int gd_w, const void* gwl,
                                int gwl_rows, int gwl_w, const void* gq, int gq_rows, int gq_w, const void* dx,
                                const void* t, const void* dt, hipStream_t st) {
  hipLaunchKernelGGL((k_flux_compact<T, ORDER>), dim3(nblocks(M, kCThreads)), dim3(kCThreads), 0, st, M, N, wet,
                     (T*)dh, (T*)dqx, (T*)dqy, (const T*)h, (const T*)wl, (const T*)z, (const T*)qx, (const T*)qy,
                     index, (const T*)gd, gd_rows, gd_w, (const T*)gwl, gwl_rows, gwl_w, (const T*)gq, gq_rows, gq_w,
                     (const T*)dx, (const T*)t, (const T*)dt);
}

}  // namespace hp

using namespace hp;

extern "C" {

int hp_flux_mask(int dtype, int64_t N, uint8_t* mask, const void* h, const int32_t* index, const void* t,
                 void* stream) {
  if (int rc = check_dtype(dtype)) return rc;
  if (N < 0) return fail(HP_EINVAL, "N must be >= 0");
  if (N == 0) return HP_OK;
  if (!mask || !h || !index || !t) return fail(HP_EINVAL, "null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == HP_F32)
    hipLaunchKernelGGL(k_flux_mask<float>, dim3(nblocks(N, 512)), dim3(512), 0, st, N, mask, (const float*)h, index,
                       (const float*)t);
  else
    hipLaunchKernelGGL(k_flux_mask<double>, dim3(nblocks(N, 512)), dim3(512), 0, st, N, mask, (const double*)h, index,
                       (const double*)t);
  return check_launch("k_flux_mask");
}

int hp_flux_compact(int dtype, int order, int64_t M, int64_t N, const int32_t* wet, void* dh, void* dqx, void* dqy,
                    const void* h, const void* wl, const void* z, const void* qx, const void* qy,
                    const int32_t* index, const void* normal, const void* given_depth, int32_t gd_rows,
                    int32_t gd_w, const void* given_wl, int32_t gwl_rows, int32_t gwl_w, const void* given_q,
                    int32_t gq_rows, int32_t gq_w, const void* dx, const void* t, const void* dt, void* stream) {
  (void)normal;  // the E, N, W, S unit normals are built in (SWE_CUDA.py:185-188)
  if (int rc = check_dtype(dtype)) return rc;
  if (order != 1 && order != 2) return fail(HP_EINVAL, "order must be 1 or 2");
  if (M < 0 || N < 0 || M > N) return fail(HP_EINVAL, "need 0 <= M <= N");
  if (M == 0) return HP_OK;  // empty wet list: nothing to do (fluxCal_jh_modified_surface_Kernel.cu:647-650)
  if (!wet || !dh || !dqx || !dqy || !h || !wl || !z || !qx || !qy || !index || !dx || !t || !dt)
    return fail(HP_EINVAL, "null pointer");
  if (!given_depth || !given_wl || !given_q || gd_rows < 1 || gd_w < 1 || gwl_rows < 1 || gwl_w < 1 || gq_rows < 1 ||
      gq_w < 1)
    return fail(HP_EINVAL, "boundary series tables need >= 1 row and column");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == HP_F32) {
    if (order == 2)
      launch_flux_compact<float, 2>(M, N, wet, dh, dqx, dqy, h, wl, z, qx, qy, index, given_depth, gd_rows, gd_w,
                                    given_wl, gwl_rows, gwl_w, given_q, gq_rows, gq_w, dx, t, dt, st);
    else
      launch_flux_compact<float, 1>(M, N, wet, dh, dqx, dqy, h, wl, z, qx, qy, index, given_depth, gd_rows, gd_w,
                                    given_wl, gwl_rows, gwl_w, given_q, gq_rows, gq_w, dx, t, dt, st);
  } else {
    if (order == 2)
      launch_flux_compact<double, 2>(M, N, wet, dh, dqx, dqy, h, wl, z, qx, qy, index, given_depth, gd_rows, gd_w,
                                     given_wl, gwl_rows, gwl_w, given_q, gq_rows, gq_w, dx, t, dt, st);
    else
      launch_flux_compact<double, 1>(M, N, wet, dh, dqx, dqy, h, wl, z, qx, qy, index, given_depth, gd_rows, gd_w,
                                     given_wl, gwl_rows, gwl_w, given_q, gq_rows, gq_w, dx, t, dt, st);
  }
  return check_launch("k_flux_compact");
}

int hp_station_rain(int dtype, int64_t N, void* dh, const int16_t* station, const void* rain, int rain_is_f64,
                    const void* dt, void* stream) {
  if (int rc = check_dtype(dtype)) return rc;
  if (N < 0) return fail(HP_EINVAL, "N must be >= 0");
  if (N == 0) return HP_OK;
  if (!dh || !station || !rain || !dt) return fail(HP_EINVAL, "null pointer");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = nblocks(N, 1024);
  if (dtype == HP_F32) {
    if (rain_is_f64)
      hipLaunchKernelGGL((k_station_rain<float, double>), dim3(nb), dim3(1024), 0, st, N, (float*)dh, station,
                         (const double*)rain, (const float*)dt);
    else
      hipLaunchKernelGGL((k_station_rain<float, float>), dim3(nb), dim3(1024), 0, st, N, (float*)dh, station,
                         (const float*)rain, (const float*)dt);
  } else {
    if (rain_is_f64)
      hipLaunchKernelGGL((k_station_rain<double, double>), dim3(nb), dim3(1024), 0, st, N, (double*)dh, station,
                         (const double*)rain, (const double*)dt);
    else
      hipLaunchKernelGGL((k_station_rain<double, float>), dim3(nb), dim3(1024), 0, st, N, (double*)dh, station,
                         (const float*)rain, (const double*)dt);
  }
  return check_launch("k_station_rain");
}

int hp_friction_update_compact(int dtype, int64_t K, const int32_t* wet, void* dh, void* dqx, void* dqy, void* dz,
                               const uint8_t* landuse, void* h, void* wl, void* qx, void* qy, void* z,
                               const void* manning, const void* dt, void* stream) {
  if (int rc = check_dtype(dtype)) return rc;
  if (K < 0) return fail(HP_EINVAL, "K must be >= 0");
  if (K == 0) return HP_OK;
  if (!wet || !dh || !dqx || !dqy || !dz || !landuse || !h || !wl || !qx || !qy || !z || !manning || !dt)
    return fail(HP_EINVAL, "null pointer");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = nblocks(K, 1024);
  if (dtype == HP_F32)
    hipLaunchKernelGGL(k_friction_compact<float>, dim3(nb), dim3(1024), 0, st, K, wet, (float*)dh, (float*)dqx,
                       (float*)dqy, (float*)dz, landuse, (float*)h, (float*)wl, (float*)qx, (float*)qy, (float*)z,
                       (const float*)manning, (const float*)dt);
  else
    hipLaunchKernelGGL(k_friction_compact<double>, dim3(nb), dim3(1024), 0, st, K, wet, (double*)dh, (double*)dqx,
                       (double*)dqy, (double*)dz, landuse, (double*)h, (double*)wl, (double*)qx, (double*)qy,
                       (double*)z, (const double*)manning, (const double*)dt);
  return check_launch("k_friction_compact");
}

int hp_time_control_compact(int dtype, int64_t K, const int32_t* wet, void* accel, void* h_max, const void* h,
                            const void* qx, const void* qy, const void* dx, const void* cfl, void* stream) {
  if (int rc = check_dtype(dtype)) return rc;
  if (K < 0) return fail(HP_EINVAL, "K must be >= 0");
  if (K == 0) return HP_OK;
  if (!wet || !accel || !h_max || !h || !qx || !qy || !dx || !cfl) return fail(HP_EINVAL, "null pointer");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = nblocks(K, 1024);
  if (dtype == HP_F32)
    hipLaunchKernelGGL(k_time_control_compact<float>, dim3(nb), dim3(1024), 0, st, K, wet, (float*)accel,
                       (float*)h_max, (const float*)h, (const float*)qx, (const float*)qy, (const float*)dx,
                       (const float*)cfl);
  else
    hipLaunchKernelGGL(k_time_control_compact<double>, dim3(nb), dim3(1024), 0, st, K, wet, (double*)accel,
                       (double*)h_max, (const double*)h, (const double*)qx, (const double*)qy, (const double*)dx,
                       (const double*)cfl);
  return check_launch("k_time_control_compact");
}

int hp_euler_update_compact(int dtype, int64_t K, const int32_t* wet, const void* dh, const void* dqx,
                            const void* dqy, const void* dz, void* h, void* wl, void* z, void* qx, void* qy,
                            void* stream) {
  if (int rc = check_dtype(dtype)) return rc;
  if (K < 0) return fail(HP_EINVAL, "K must be >= 0");
  if (K == 0) return HP_OK;
  if (!wet || !dh || !dqx || !dqy || !dz || !h || !wl || !z || !qx || !qy) return fail(HP_EINVAL, "null pointer");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = nblocks(K, 1024);
  if (dtype == HP_F32)
    hipLaunchKernelGGL(k_euler_compact<float>, dim3(nb), dim3(1024), 0, st, K, wet, (const float*)dh,
                       (const float*)dqx, (const float*)dqy, (const float*)dz, (float*)h, (float*)wl, (float*)z,
                       (float*)qx, (float*)qy);
  else
    hipLaunchKernelGGL(k_euler_compact<double>, dim3(nb), dim3(1024), 0, st, K, wet, (const double*)dh,
                       (const double*)dqx, (const double*)dqy, (const double*)dz, (double*)h, (double*)wl, (double*)z,
                       (double*)qx, (double*)qy);
  return check_launch("k_euler_compact");
}

int hp_infiltration_compact(int dtype, int64_t M, const int32_t* wet, void* dh, const uint8_t* landuse,
                            const void* h, const void* ks, const void* phi, const void* dtheta, void* F,
                            const void* dt, void* stream) {
  if (int rc = check_dtype(dtype)) return rc;
  if (M < 0) return fail(HP_EINVAL, "M must be >= 0");
  if (M == 0) return HP_OK;
  if (!wet || !dh || !landuse || !h || !ks || !phi || !dtheta || !F || !dt) return fail(HP_EINVAL, "null pointer");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = nblocks(M, 1024);
  if (dtype == HP_F32)
    hipLaunchKernelGGL(k_infiltration_compact<float>, dim3(nb), dim3(1024), 0, st, M, wet, (float*)dh, landuse,
                       (const float*)h, (const float*)ks, (const float*)phi, (const float*)dtheta, (float*)F,
                       (const float*)dt);
  else
    hipLaunchKernelGGL(k_infiltration_compact<double>, dim3(nb), dim3(1024), 0, st, M, wet, (double*)dh, landuse,
                       (const double*)h, (const double*)ks, (const double*)phi, (const double*)dtheta, (double*)F,
                       (const double*)dt);
  return check_launch("k_infiltration_compact");
}

}  // extern "C"
