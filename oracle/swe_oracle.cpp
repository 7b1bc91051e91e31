// CPU oracle for the HiPIMS 2-D shallow-water step: TEST INFRASTRUCTURE ONLY.
//
// This file restates, on the CPU and in the reference's own data layout (the
// compact valid-cell list plus the int32[5*N] neighbour table built at
// CHANS-HiPIMS/hipims/pythonHipims/SWE_CUDA.py:115-227), every op on the
// reference's hot path.  It is the checker for the HIP product path and the
// CPU baseline timed beside it.  Nothing in the shipped library links it; only
// tests, the smoke check and the bench's cpu_baseline leg load it.
//
// Citations (paths relative to the reference root, prefix CHANS-HiPIMS/hipims/):
//   flux mask ........ cuda/fluxMaskGenerator_Kernel.cu:19-42
//   order-1 flux ..... cuda/fluxCal_jh_modified_surface_Kernel.cu:43-615
//   order-2 flux ..... cuda/flucCal_2ndOrder_jh_improved_Kernel.cu:43-506
//   station rain ..... cuda/stationPrecipitation_Kernel.cu:18-29
//   friction+update .. cuda/friction_implicit_andUpdate_jh_Kernel.cu:22-146
//   CFL accelerator .. cuda/timeControl_Kernel.cu:164-198
//   euler update ..... cuda/euler_update_Kernel.cu:18-36
//   infiltration ..... cuda/infiltrationCUDA_Kernel.cu:20-46
//   driver step ...... pythonHipims/SWE_CUDA.py:339-424, 436-446, 472-502, 780-831
//
// Every per-cell loop writes only the cell it owns, as the reference kernels
// do, so the OpenMP loops are race-free.  Floating-point contraction is off
// (see the Makefile) so each expression rounds as written.
//
// Pinning (which check fixes which function): see tests/test_oracle_*.py.
// The FP32 instantiation computes in pure float; the reference's FP32 kernels
// mix in double through `auto` and double literals (e.g.
// fluxCal_jh_modified_surface_Kernel.cu:181-185).  That mix is NOT restated:
// FP32 oracle results are parity unpinned; they serve as CPU timing and as
// the accuracy bar of float arithmetic run through the reference's algorithm.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <chrono>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <class T> struct C {
  static constexpr T g = T(9.81);
  static constexpr T hs = T(1.0e-6);
};

// minmod limiter: fluxCal_jh_modified_surface_Kernel.cu:321-328,
// flucCal_2ndOrder_jh_improved_Kernel.cu:74-81.
template <class T> inline T minmod(T x, T y) {
  if (x * y <= T(0)) return T(0);
  return std::fabs(x) < std::fabs(y) ? x : y;
}

// field layout used by the reference kernels: 0 h, 1 wl, 2 qx, 3 qy, 4 z.
enum { H = 0, WL = 1, QX = 2, QY = 3, Z = 4 };

// Order-1 ("jiaheng paper") hydrostatic reconstruction,
// fluxCal_jh_modified_surface_Kernel.cu:96-139.  chz / rhz are the cell
// (h, z) pairs; L / R are the edge states, modified in place.
template <class T> inline void hydro_o1(const T* chz, const T* rhz, T* L, T* R) {
  L[WL] = L[H] + L[Z];
  L[WL] = std::fmin(L[WL], std::fmax(chz[0] + chz[1], rhz[0] + rhz[1]));
  L[WL] = std::fmax(L[WL], std::fmin(chz[0] + chz[1], rhz[0] + rhz[1]));
  R[WL] = R[H] + R[Z];
  R[WL] = std::fmin(R[WL], std::fmax(chz[0] + chz[1], rhz[0] + rhz[1]));
  R[WL] = std::fmax(R[WL], std::fmin(chz[0] + chz[1], rhz[0] + rhz[1]));
  if (((chz[0] + chz[1]) - (rhz[0] + rhz[1])) * (L[WL] - R[WL]) < T(0)) {
    T tmp = (L[WL] + R[WL]) / T(2);
    L[WL] = tmp;
    R[WL] = tmp;
  }
  L[Z] = L[WL] - L[H];
  R[Z] = R[WL] - R[H];
  T zbm = std::fmax(L[Z], R[Z]);
  L[H] = std::fmax(L[WL] - zbm, T(0));
  R[H] = std::fmax(R[WL] - zbm, T(0));
  L[Z] = L[WL] - L[H];
  R[Z] = R[WL] - R[H];
}

// Order-2 hydrostatic reconstruction, flucCal_2ndOrder_jh_improved_Kernel.cu:43-72.
// cl / cr are the full cell states (their wl is the cell water level).
template <class T> inline void hydro_o2(const T* cl, const T* cr, T* L, T* R) {
  L[WL] = L[H] + L[Z];
  L[WL] = std::fmin(L[WL], std::fmax(cl[WL], cr[WL]));
  L[WL] = std::fmax(L[WL], std::fmin(cl[WL], cr[WL]));
  R[WL] = R[H] + R[Z];
  R[WL] = std::fmin(R[WL], std::fmax(cl[WL], cr[WL]));
  R[WL] = std::fmax(R[WL], std::fmin(cl[WL], cr[WL]));
  if (cl[WL] > cr[WL]) {
    L[WL] = std::fmax(L[WL], R[WL]);
  } else {
    R[WL] = std::fmax(L[WL], R[WL]);
  }
  L[Z] = L[WL] - L[H];
  R[Z] = R[WL] - R[H];
  T zbm = std::fmax(L[Z], R[Z]);
  L[H] = std::fmax(L[WL] - zbm, T(0));
  R[H] = std::fmax(R[WL] - zbm, T(0));
  L[Z] = L[WL] - L[H];
  R[Z] = R[WL] - R[H];
}

// HLLC flux with the reconstruction applied first; identical in both orders
// (fluxCal_jh_modified_surface_Kernel.cu:173-260,
//  flucCal_2ndOrder_jh_improved_Kernel.cu:109-195).  Accumulates into
// fh/fqx/fqy exactly as the reference does.
template <class T, int ORDER>
inline void hllc(const T* a, const T* b, T* L, T* R, const T* n, const T* tg,
                 T& fh, T& fqx, T& fqy, T dx, T dt) {
  const T hs = C<T>::hs, g = C<T>::g;
  T ux_l = L[H] < hs ? T(0) : L[QX] / L[H];
  T uy_l = L[H] < hs ? T(0) : L[QY] / L[H];
  T ux_r = R[H] < hs ? T(0) : R[QX] / R[H];
  T uy_r = R[H] < hs ? T(0) : R[QY] / R[H];
  if (ORDER == 1) hydro_o1(a, b, L, R); else hydro_o2(a, b, L, R);
  if (L[H] < hs && R[H] < hs) return;
  T u_L = ux_l * n[0] + uy_l * n[1];
  T u_R = ux_r * n[0] + uy_r * n[1];
  T v_L = -ux_l * n[1] + uy_l * n[0];
  T v_R = -ux_r * n[1] + uy_r * n[0];
  T qx_L = u_L * L[H], qy_L = v_L * L[H];
  T qx_R = u_R * R[H], qy_R = v_R * R[H];
  T a_L = std::sqrt(g * L[H]);
  T a_R = std::sqrt(g * R[H]);
  T base = (a_L + a_R) / T(2) + (u_L - u_R) / T(4);
  T h_star = (base * base) / g;
  T u_star = (u_L + u_R) / T(2) + a_L - a_R;
  T a_star = std::sqrt(g * h_star);
  T s_L = L[H] <= hs ? u_R - T(2) * a_R : std::fmin(u_L - a_L, u_star - a_star);
  T s_R = R[H] <= hs ? u_L + T(2) * a_L : std::fmax(u_R + a_R, u_star + a_star);
  T s_M = (s_L * R[H] * (u_R - s_R) - s_R * L[H] * (u_L - s_L)) /
          (R[H] * (u_R - s_R) - L[H] * (u_L - s_L));
  T h_flux_L = qx_L;
  T qx_flux_L = u_L * qx_L + T(0.5) * g * L[H] * L[H];
  T qy_flux_L = u_L * qy_L;
  T h_flux_R = qx_R;
  T qx_flux_R = u_R * qx_R + T(0.5) * g * R[H] * R[H];
  T qy_flux_R = u_R * qy_R;
  T h_flux_M = (s_R * h_flux_L - s_L * h_flux_R + s_L * s_R * (R[H] - L[H])) / (s_R - s_L);
  T qx_flux_M = (s_R * qx_flux_L - s_L * qx_flux_R + s_L * s_R * (qx_R - qx_L)) / (s_R - s_L);
  T hf, qxf, qyf;
  if (T(0) <= s_L) {
    hf = h_flux_L; qxf = qx_flux_L; qyf = qy_flux_L;
  } else if (s_L < T(0) && T(0) <= s_M) {
    hf = h_flux_M; qxf = qx_flux_M; qyf = h_flux_M * v_L;
  } else if (s_M < T(0) && T(0) <= s_R) {
    hf = h_flux_M; qxf = qx_flux_M; qyf = h_flux_M * v_R;
  } else {
    hf = h_flux_R; qxf = qx_flux_R; qyf = qy_flux_R;
  }
  fh -= (hf / dx) * dt;
  fqx -= ((qxf * n[0] + qyf * tg[0]) / dx) * dt;
  fqy -= ((qxf * n[1] + qyf * tg[1]) / dx) * dt;
}

// Boundary ghost state from the cell state c and the decoded type,
// fluxCal_jh_modified_surface_Kernel.cu:262-435 (type 9 "fall" exists only in
// order 1: flucCal_2ndOrder_jh_improved_Kernel.cu:300-332 has no case 9 and
// leaves the ghost undefined; this restatement uses the order-1 definition
// there too).  A type outside 3..9 leaves the ghost undefined in the
// reference; here it is the cell state itself (a transmissive edge).
template <class T>
inline void bc_ghost(int type, const T* c, T* b, const T* n, T gd, T gwl, T gqx, T gqy) {
  for (int k = 0; k < 5; ++k) b[k] = c[k];
  switch (type) {
    case 3:  // "rigid": reflects only on E/W edges (the N/S expression is the identity)
      b[QX] = (n[1] * n[1] - n[0] * n[0]) * c[QX];
      b[QY] = -(n[0] * n[0] - n[1] * n[1]) * c[QY];
      break;
    case 4:
      b[QX] = (n[1] * n[1] - n[0] * n[0]) * c[QX];
      b[QY] = (n[0] * n[0] - n[1] * n[1]) * c[QY];
      break;
    case 5:
      break;
    case 6:
      b[H] = gd; b[Z] = c[Z]; b[WL] = b[Z] + b[H];
      break;
    case 7:
      b[H] = c[H]; b[Z] = c[Z]; b[WL] = b[Z] + b[H]; b[QX] = gqx; b[QY] = gqy;
      break;
    case 8:
      b[Z] = c[Z]; b[WL] = gwl > b[Z] ? gwl : b[Z]; b[H] = b[WL] - b[Z];
      break;
    case 9:
      b[H] = T(0); b[QX] = T(0); b[QY] = T(0);
      break;
    default:
      break;
  }
}

// Mask-code decode, fluxCal_jh_modified_surface_Kernel.cu:528-535: leading
// digit = type, the remaining digits = id (50 -> open #0, 612 -> H-given #12).
inline void decode_code(int32_t code, int& type, int& id) {
  int t = code, n = 0;
  while (t >= 10) { t = (t - (t % 10)) / 10; n++; }
  int p = 1;
  for (int k = 0; k < n; ++k) p *= 10;
  type = t;
  id = code - t * p;
}

// Series row lookup, fluxCal_jh_modified_surface_Kernel.cu:459-470: the first
// row whose time is >= t, else the last row.
template <class T> inline int series_row(const T* tab, int rows, int w, T t) {
  int r = 0;
  while (t > tab[r * w] && r < rows - 1) r++;
  return r;
}

struct Tables {
  const void* gd; int gd_rows, gd_w;
  const void* gwl; int gwl_rows, gwl_w;
  const void* gq; int gq_rows, gq_w;
};

template <class T>
inline T table_at(const void* tab, int rows, int w, int row, int col) {
  // Out-of-range columns are read out of bounds by the reference; here 0.
  if (col < 0 || col >= w || row < 0 || row >= rows) return T(0);
  return static_cast<const T*>(tab)[row * w + col];
}

//
// Slope source (hb, qxb, qyb; test construct, not in the reference): when
// given, the order-2 reconstruction takes its limited slopes and its
// "cell below hs carries no momentum" rule from that state instead of from
// (h, qx, qy): face = own cell value + slope of the source state.  Run with
// the source = an unperturbed trajectory, it is the forward model whose
// derivative the frozen-slope adjoint computes, so central differences of it
// check that adjoint.
template <class T, int ORDER>
void flux_compact(int64_t M, int64_t N, const int32_t* wet, T* dh, T* dqx, T* dqy,
                  const T* h, const T* wl, const T* z, const T* qx, const T* qy,
                  const int32_t* index, const T* normal, const Tables& tb,
                  const T* dxp, const T* tp, const T* dtp,
                  const T* hb = nullptr, const T* qxb = nullptr, const T* qyb = nullptr) {
  const T hs = C<T>::hs, g = C<T>::g;
  const T dx = dxp[0], t = tp[0], dt = dtp[0];
  const int ld = series_row<T>(static_cast<const T*>(tb.gd), tb.gd_rows, tb.gd_w, t);
  const int lw = series_row<T>(static_cast<const T*>(tb.gwl), tb.gwl_rows, tb.gwl_w, t);
  const int lq = series_row<T>(static_cast<const T*>(tb.gq), tb.gq_rows, tb.gq_w, t);
#pragma omp parallel for schedule(static)
  for (int64_t w = 0; w < M; ++w) {
    const int32_t i = wet[w];
    const int32_t c_type = index[i];
    for (int j = 0; j < 4; ++j) {
      T fc[5] = {h[i], wl[i], qx[i], qy[i], z[i]};
      T fr[5];
      const T nrm[2] = {normal[2 * j], normal[2 * j + 1]};
      const T tg[2] = {-nrm[1], nrm[0]};
      int rdir = j + 1;
      int ldir = (rdir + 2) > 4 ? rdir - 2 : rdir + 2;
      const int32_t l_id = index[(int64_t)ldir * N + i];
      const int32_t r_id = index[(int64_t)rdir * N + i];
      bool boundary = (r_id == -1);
      T ghost_vals[4] = {0, 0, 0, 0};
      int btype = 0;
      if (boundary) {
        int bid;
        decode_code(c_type, btype, bid);
        ghost_vals[0] = table_at<T>(tb.gd, tb.gd_rows, tb.gd_w, ld, bid + 1);
        ghost_vals[1] = table_at<T>(tb.gwl, tb.gwl_rows, tb.gwl_w, lw, bid + 1);
        ghost_vals[2] = table_at<T>(tb.gq, tb.gq_rows, tb.gq_w, lq, 1 + 2 * bid);
        ghost_vals[3] = table_at<T>(tb.gq, tb.gq_rows, tb.gq_w, lq, 2 + 2 * bid);
        bc_ghost<T>(btype, fc, fr, nrm, ghost_vals[0], ghost_vals[1], ghost_vals[2], ghost_vals[3]);
      }
      if (ORDER == 1) {
        // fluxCal_jh_modified_surface_Kernel.cu:501-611
        T chz[2] = {h[i], z[i]};
        T rhz[2];
        T zl = (l_id == -1) ? z[i] : z[l_id];
        if (boundary) {
          rhz[0] = fr[H]; rhz[1] = fr[Z];
        } else {
          int32_t rr_id = index[(int64_t)rdir * N + r_id];
          T zrr = (rr_id == -1) ? z[r_id] : z[rr_id];
          fr[H] = h[r_id]; fr[WL] = wl[r_id]; fr[QX] = qx[r_id]; fr[QY] = qy[r_id]; fr[Z] = z[r_id];
          rhz[0] = fr[H]; rhz[1] = fr[Z];
          T rb = minmod<T>(fc[Z] - zl, fr[Z] - fc[Z]);
          if (fr[H] > hs) fc[Z] += T(0.5) * rb;
          rb = minmod<T>(fr[Z] - zrr, chz[1] - fr[Z]);
          if (fc[H] > hs) fr[Z] += T(0.5) * rb;
        }
        if (fc[H] < hs && fr[H] < hs) continue;
        hllc<T, 1>(chz, rhz, fc, fr, nrm, tg, dh[i], dqx[i], dqy[i], dx, dt);
        dqx[i] -= ((T(0.5) * g * (fc[H] + chz[0]) * (fc[Z] - chz[1]) * nrm[0]) / dx) * dt;
        dqy[i] -= ((T(0.5) * g * (fc[H] + chz[0]) * (fc[Z] - chz[1]) * nrm[1]) / dx) * dt;
      } else {
        // flucCal_2ndOrder_jh_improved_Kernel.cu:399-502
        T fl[5], el[5], er[5];
        if (l_id == -1) {
          for (int k = 0; k < 5; ++k) fl[k] = fc[k];
        } else {
          fl[H] = h[l_id]; fl[WL] = wl[l_id]; fl[QX] = qx[l_id]; fl[QY] = qy[l_id]; fl[Z] = z[l_id];
        }
        if (boundary) {
          for (int k = 0; k < 5; ++k) { er[k] = fr[k]; el[k] = fc[k]; }
        } else {
          int32_t rr_id = index[(int64_t)rdir * N + r_id];
          T frr[5];
          fr[H] = h[r_id]; fr[WL] = wl[r_id]; fr[QX] = qx[r_id]; fr[QY] = qy[r_id]; fr[Z] = z[r_id];
          if (rr_id == -1) {
            for (int k = 0; k < 5; ++k) frr[k] = fr[k];
          } else {
            frr[H] = h[rr_id]; frr[WL] = wl[rr_id]; frr[QX] = qx[rr_id]; frr[QY] = qy[rr_id]; frr[Z] = z[rr_id];
          }
          // singleSlope_MUSCL_reconstruction, :83-107
          auto muscl = [&](const T* c, const T* l, const T* r, T* e) {
            e[H] = c[H] + T(0.5) * minmod<T>(c[H] - l[H], r[H] - c[H]);
            e[WL] = c[WL] + T(0.5) * minmod<T>(c[WL] - l[WL], r[WL] - c[WL]);
            e[Z] = c[Z] + T(0.5) * minmod<T>(c[Z] - l[Z], r[Z] - c[Z]);
            if (c[H] < hs) {
              e[QX] = T(0); e[QY] = T(0);
            } else {
              e[QX] = c[QX] + T(0.5) * minmod<T>(c[QX] - l[QX], r[QX] - c[QX]);
              e[QY] = c[QY] + T(0.5) * minmod<T>(c[QY] - l[QY], r[QY] - c[QY]);
            }
          };
          if (hb) {
            // the same reconstruction with the slopes of the source state
            auto frozen = [&](int32_t ci, int32_t li, int32_t ri, const T* c, T* e) {
              auto sl = [&](const T* v) { return T(0.5) * minmod<T>(v[ci] - v[li], v[ri] - v[ci]); };
              e[H] = c[H] + sl(hb);
              e[WL] = c[WL];
              e[Z] = c[Z] + T(0.5) * minmod<T>(z[ci] - z[li], z[ri] - z[ci]);
              if (hb[ci] < hs) {
                e[QX] = T(0); e[QY] = T(0);
              } else {
                e[QX] = c[QX] + sl(qxb);
                e[QY] = c[QY] + sl(qyb);
              }
            };
            const int32_t li = (l_id == -1) ? i : l_id, rri = (rr_id == -1) ? r_id : rr_id;
            frozen(i, li, r_id, fc, el);
            frozen(r_id, rri, i, fr, er);
          } else {
            muscl(fc, fl, fr, el);
            muscl(fr, frr, fc, er);
          }
        }
        if (fc[H] < hs && fr[H] < hs) continue;
        hllc<T, 2>(fc, fr, el, er, nrm, tg, dh[i], dqx[i], dqy[i], dx, dt);
        dqx[i] -= ((T(0.5) * g * (fc[H] + el[H]) * (el[Z] - fc[Z]) * nrm[0]) / dx) * dt;
        dqy[i] -= ((T(0.5) * g * (fc[H] + el[H]) * (el[Z] - fc[Z]) * nrm[1]) / dx) * dt;
      }
    }
  }
}

// fluxMaskGenerator_Kernel.cu:19-42.  Neighbour id 0 is never counted (the
// reference tests `n_index > 0`).
template <class T>
void flux_mask(int64_t N, uint8_t* mask, const T* h, const int32_t* index, const T* tp) {
  const T hs = C<T>::hs;
  const T t = tp[0];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    if (t == T(0)) { mask[i] = 1; continue; }
    bool m = h[i] > hs;
    for (int nei = 1; nei < 5; ++nei) {
      int32_t k = index[(int64_t)nei * N + i];
      if (k > 0) m = m || (h[k] > hs);
    }
    mask[i] = m ? 1 : 0;
  }
}

// stationPrecipitation_Kernel.cu:18-29.  The rain table is double in the
// reference driver (SWE_CUDA.py:483-489); rain_is_f64 says which it is.
template <class T>
void station_rain(int64_t N, T* dh, const int16_t* station, const void* rain, int rain_is_f64, const T* dtp) {
  const T dt = dtp[0];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < N; ++i) {
    T r = rain_is_f64 ? T(static_cast<const double*>(rain)[station[i]])
                      : T(static_cast<const float*>(rain)[station[i]]);
    dh[i] += r * dt;
  }
}

// friction_implicit_andUpdate_jh_Kernel.cu:22-146.
template <class T>
void friction_update(int64_t K, const int32_t* wet, T* dh, T* dqx, T* dqy, T* dz,
                     const uint8_t* landuse, T* h, T* wl, T* qx, T* qy, T* z,
                     const T* manning, const T* dtp) {
  const T hs = C<T>::hs, g = C<T>::g;
  const T dt = dtp[0];
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < K; ++j) {
    const int32_t i = wet[j];
    h[i] += dh[i];
    if (h[i] < hs) {
      qx[i] = T(0);
      qy[i] = T(0);
    } else {
      const T n = manning[landuse[i]];
      T C_f = g * n * n * T(std::pow((double)h[i], -1.0 / 3.0));
      qx[i] += dqx[i];
      qy[i] += dqy[i];
      T q_norm = std::sqrt(qx[i] * qx[i] + qy[i] * qy[i]);
      if (n > T(0)) {
        if (std::fabs(qx[i]) > T(1.0e-10)) {
          T tmp = (T(1) - std::sqrt(T(1) + (T(4) * dt * C_f) / (h[i] * h[i]) * q_norm)) /
                  (T(-2) * dt * C_f / (h[i] * h[i]) * (q_norm / qx[i]));
          if (std::fabs(tmp) > T(5) * h[i] * std::sqrt(h[i] * g)) {
            qx[i] = T(2) * tmp - qx[i];
            if (qx[i] * tmp <= T(0)) qx[i] = T(0);
          } else {
            qx[i] = tmp;
          }
        }
        if (std::fabs(qy[i]) > T(1.0e-10)) {
          T tmp = (T(1) - std::sqrt(T(1) + (T(4) * dt * C_f) / (h[i] * h[i]) * q_norm)) /
                  (T(-2) * dt * C_f / (h[i] * h[i]) * (q_norm / qy[i]));
          if (std::fabs(tmp) > T(5) * h[i] * std::sqrt(h[i] * g)) {
            qy[i] = T(2) * tmp - qy[i];
            if (qy[i] * tmp <= T(0)) qy[i] = T(0);
          } else {
            qy[i] = tmp;
          }
        }
      }
    }
    z[i] += dz[i];
    wl[i] = z[i] + h[i];
    dh[i] = T(0);
    dqx[i] = T(0);
    dqy[i] = T(0);
    dz[i] = T(0);
  }
}

// timeControl_Kernel.cu:164-198.  The reference calls single-precision sqrtf
// on g*h even in the double instantiation (:188); restated with the
// correctly rounded sqrtf.
template <class T>
void time_control(int64_t K, const int32_t* wet, T* accel, T* h_max, const T* h,
                  const T* qx, const T* qy, const T* dxp, const T* cflp) {
  const T hs = C<T>::hs, g = C<T>::g;
  const T dx = dxp[0], cfl = cflp[0];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < K; ++i) {
    const int32_t j = wet[i];
    h_max[j] = std::fmax(h[j], h_max[j]);
    accel[i] = h[j] > hs ? cfl * dx / (T(sqrtf((float)(g * h[j]))) +
                                       std::fmax(std::fabs(qx[j]), std::fabs(qy[j])) / h[j])
                         : T(60.0);
  }
}

// euler_update_Kernel.cu:18-36.
template <class T>
void euler_update(int64_t K, const int32_t* wet, const T* dh, const T* dqx, const T* dqy,
                  const T* dz, T* h, T* wl, T* z, T* qx, T* qy) {
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < K; ++j) {
    const int32_t i = wet[j];
    h[i] += dh[i];
    qx[i] += dqx[i];
    qy[i] += dqy[i];
    z[i] += dz[i];
    wl[i] = h[i] + z[i];
  }
}

// Green-Ampt infiltration over the wet list, infiltrationCUDA_Kernel.cu:20-46:
// F1 = (A + sqrt(A^2 + 4 dt Ks (phi + h) dtheta)) / 2 with A = F0 + dt Ks,
// dF = min(h, F1 - F0); F += dF; dh -= dF.  Parameters are per landuse class.
template <class T>
void infiltration(int64_t M, const int32_t* wet, T* dh, const uint8_t* landuse, const T* h, const T* ks,
                  const T* phi, const T* dth, T* F, const T* dtp) {
  const T dt = *dtp;
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < M; ++j) {
    const int32_t i = wet[j];
    const int c = landuse[i];
    const T K = ks[c], head = phi[c] + h[i], d = dth[c], F0 = F[i];
    const T a = F0 + dt * K;
    const T F1 = T(0.5) * (a + std::sqrt(a * a + T(4) * dt * K * head * d));
    const T dF = std::fmin(h[i], F1 - F0);
    F[i] += dF;
    dh[i] -= dF;
  }
}

}  // namespace

// The normal table used when a caller passes none: SWE_CUDA.py:185-188.
static const double kNormal64[8] = {1, 0, 0, 1, -1, 0, 0, -1};
static const float kNormal32[8] = {1, 0, 0, 1, -1, 0, 0, -1};

template <class T> const T* default_normal();
template <> const double* default_normal<double>() { return kNormal64; }
template <> const float* default_normal<float>() { return kNormal32; }

namespace {

// ---------------------------------------------------------------------------
// Whole-step driver: the server's call sequence (CHANS-HiPIMS_server.py:246-250)
// addFlux (SWE_CUDA.py:397-423 / 370-396), station rain (:497-502),
// time_friction_euler_update_cuda (:780-831).  Used for the CPU baseline and
// for multi-step oracle runs.
// ---------------------------------------------------------------------------
template <class T> struct Sim {
  int64_t N;
  const int32_t* index;  // [5N]
  T *h, *wl, *qx, *qy, *z, *h_max;
  T *dh, *dqx, *dqy, *dz;
  const uint8_t* landuse;
  const T* manning;
  const int16_t* station;      // may be null (no rain)
  const T *ks, *phi, *dth;     // Green-Ampt class tables; null: no infiltration
  const T *hb = nullptr, *qxb = nullptr, *qyb = nullptr;  // order-2 slope source; null: own state
  T* F;                        // cumulative infiltrated depth
  const double* rain_table;    // [rows, 1+S] in m/s (SWE_CUDA.py:472-495)
  int rain_rows, rain_cols;
  int order;
  Tables tb;
  T dx, cfl;
  double export_T;
  int export_n;
  // device-scalar equivalents
  T t, dt;
  int rain_index;  // _rainfall_station_time_index
  std::vector<int32_t> list;
  std::vector<uint8_t> fmask;
  std::vector<T> accel;
  std::vector<double> rain_now;
  int n_exports;
};

// SWE_CUDA.py:472-495, the host-side linear interpolation of the station table.
template <class T> void rain_interp(Sim<T>& s) {
  const double* tab = s.rain_table;
  const int cols = s.rain_cols, rows = s.rain_rows;
  const double t = (double)s.t;
  int S = cols - 1;
  if ((int)s.rain_now.size() != S) s.rain_now.assign(S, 0.0);
  for (;;) {
    if (t < tab[(rows - 1) * cols]) {
      int i = s.rain_index;
      if (t < tab[(i + 1) * cols]) {
        double per = (t - tab[i * cols]) / (tab[(i + 1) * cols] - tab[i * cols]);
        for (int k = 0; k < S; ++k)
          s.rain_now[k] = tab[i * cols + 1 + k] + per * (tab[(i + 1) * cols + 1 + k] - tab[i * cols + 1 + k]);
        return;
      }
      s.rain_index++;
    } else {
      for (int k = 0; k < S; ++k) s.rain_now[k] -= s.rain_now[k];
      return;
    }
  }
}

template <class T> void sim_step(Sim<T>& s) {
  const int64_t N = s.N;
  // addFlux: fluxMask + nonzero + flux kernel
  s.fmask.resize(N);
  flux_mask<T>(N, s.fmask.data(), s.h, s.index, &s.t);
  s.list.clear();
  for (int64_t i = 0; i < N; ++i)
    if (s.fmask[i]) s.list.push_back((int32_t)i);
  if (!s.list.empty()) {
    if (s.order == 2)
      flux_compact<T, 2>((int64_t)s.list.size(), N, s.list.data(), s.dh, s.dqx, s.dqy, s.h, s.wl, s.z, s.qx,
                         s.qy, s.index, default_normal<T>(), s.tb, &s.dx, &s.t, &s.dt, s.hb, s.qxb, s.qyb);
    else
      flux_compact<T, 1>((int64_t)s.list.size(), N, s.list.data(), s.dh, s.dqx, s.dqy, s.h, s.wl, s.z, s.qx,
                         s.qy, s.index, default_normal<T>(), s.tb, &s.dx, &s.t, &s.dt);
  }
  // station rain
  if (s.station && s.rain_table) {
    rain_interp<T>(s);
    station_rain<T>(N, s.dh, s.station, s.rain_now.data(), 1, &s.dt);
  }
  // addInfiltrationSource over the flux wet list (SWE_CUDA.py:439-443,
  // called after the rain in CatchFlood_main.py:164-166)
  if (s.ks && !s.list.empty())
    infiltration<T>((int64_t)s.list.size(), s.list.data(), s.dh, s.landuse, s.h, s.ks, s.phi, s.dth, s.F, &s.dt);
  // time_friction_euler_update_cuda
  const T dt_old = s.dt;
  s.list.clear();
  for (int64_t i = 0; i < N; ++i)
    if (std::fabs(s.dh[i]) > T(0) || s.h[i] >= T(0)) s.list.push_back((int32_t)i);
  const int64_t K = (int64_t)s.list.size();
  friction_update<T>(K, s.list.data(), s.dh, s.dqx, s.dqy, s.dz, s.landuse, s.h, s.wl, s.qx, s.qy, s.z,
                     s.manning, &s.dt);
  s.accel.assign(K, T(60.0));
  time_control<T>(K, s.list.data(), s.accel.data(), s.h_max, s.h, s.qx, s.qy, &s.dx, &s.cfl);
  T dt_new = s.dt;
  if (K > 0) {
    T m = s.accel[0];
    for (int64_t i = 1; i < K; ++i) m = std::fmin(m, s.accel[i]);
    dt_new = m;
  }
  const T upper = dt_old * T(10);
  dt_new = (upper < dt_new) ? upper : dt_new;  // python min(dt, time_upper)
  if ((double)(dt_new + s.t) >= (double)(s.export_n + 1) * s.export_T) {
    dt_new = T((double)(s.export_n + 1) * s.export_T) - s.t;
    s.export_n++;
    s.n_exports++;
  }
  s.dt = dt_new;
  s.t += s.dt;
}

template <class T, int ORDER>
void flux_entry(int64_t M, int64_t N, const int32_t* wet, void* dh, void* dqx, void* dqy, const void* h,
                const void* wl, const void* z, const void* qx, const void* qy, const int32_t* index,
                const void* normal, const void* gd, int gd_rows, int gd_w, const void* gwl, int gwl_rows,
                int gwl_w, const void* gq, int gq_rows, int gq_w, const void* dx, const void* t, const void* dt) {
  Tables tb{gd, gd_rows, gd_w, gwl, gwl_rows, gwl_w, gq, gq_rows, gq_w};
  flux_compact<T, ORDER>(M, N, wet, (T*)dh, (T*)dqx, (T*)dqy, (const T*)h, (const T*)wl, (const T*)z,
                         (const T*)qx, (const T*)qy, index, normal ? (const T*)normal : default_normal<T>(), tb,
                         (const T*)dx, (const T*)t, (const T*)dt);
}
}  // namespace

extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

void orc_flux_mask(int is_f64, int64_t N, uint8_t* mask, const void* h, const int32_t* index, const void* t) {
  if (is_f64) flux_mask<double>(N, mask, (const double*)h, index, (const double*)t);
  else flux_mask<float>(N, mask, (const float*)h, index, (const float*)t);
}

void orc_flux(int is_f64, int order, int64_t M, int64_t N, const int32_t* wet, void* dh, void* dqx, void* dqy,
              const void* h, const void* wl, const void* z, const void* qx, const void* qy, const int32_t* index,
              const void* normal, const void* gd, int gd_rows, int gd_w, const void* gwl, int gwl_rows, int gwl_w,
              const void* gq, int gq_rows, int gq_w, const void* dx, const void* t, const void* dt) {
  if (is_f64) {
    if (order == 2)
      flux_entry<double, 2>(M, N, wet, dh, dqx, dqy, h, wl, z, qx, qy, index, normal, gd, gd_rows, gd_w, gwl,
                            gwl_rows, gwl_w, gq, gq_rows, gq_w, dx, t, dt);
    else
      flux_entry<double, 1>(M, N, wet, dh, dqx, dqy, h, wl, z, qx, qy, index, normal, gd, gd_rows, gd_w, gwl,
                            gwl_rows, gwl_w, gq, gq_rows, gq_w, dx, t, dt);
  } else {
    if (order == 2)
      flux_entry<float, 2>(M, N, wet, dh, dqx, dqy, h, wl, z, qx, qy, index, normal, gd, gd_rows, gd_w, gwl,
                           gwl_rows, gwl_w, gq, gq_rows, gq_w, dx, t, dt);
    else
      flux_entry<float, 1>(M, N, wet, dh, dqx, dqy, h, wl, z, qx, qy, index, normal, gd, gd_rows, gd_w, gwl,
                           gwl_rows, gwl_w, gq, gq_rows, gq_w, dx, t, dt);
  }
}

void orc_station_rain(int is_f64, int64_t N, void* dh, const int16_t* station, const void* rain, int rain_is_f64,
                      const void* dt) {
  if (is_f64) station_rain<double>(N, (double*)dh, station, rain, rain_is_f64, (const double*)dt);
  else station_rain<float>(N, (float*)dh, station, rain, rain_is_f64, (const float*)dt);
}

void orc_friction_update(int is_f64, int64_t K, const int32_t* wet, void* dh, void* dqx, void* dqy, void* dz,
                         const uint8_t* landuse, void* h, void* wl, void* qx, void* qy, void* z,
                         const void* manning, const void* dt) {
  if (is_f64)
    friction_update<double>(K, wet, (double*)dh, (double*)dqx, (double*)dqy, (double*)dz, landuse, (double*)h,
                            (double*)wl, (double*)qx, (double*)qy, (double*)z, (const double*)manning,
                            (const double*)dt);
  else
    friction_update<float>(K, wet, (float*)dh, (float*)dqx, (float*)dqy, (float*)dz, landuse, (float*)h,
                           (float*)wl, (float*)qx, (float*)qy, (float*)z, (const float*)manning,
                           (const float*)dt);
}

void orc_time_control(int is_f64, int64_t K, const int32_t* wet, void* accel, void* h_max, const void* h,
                      const void* qx, const void* qy, const void* dx, const void* cfl) {
  if (is_f64)
    time_control<double>(K, wet, (double*)accel, (double*)h_max, (const double*)h, (const double*)qx,
                         (const double*)qy, (const double*)dx, (const double*)cfl);
  else
    time_control<float>(K, wet, (float*)accel, (float*)h_max, (const float*)h, (const float*)qx,
                        (const float*)qy, (const float*)dx, (const float*)cfl);
}

void orc_euler_update(int is_f64, int64_t K, const int32_t* wet, const void* dh, const void* dqx, const void* dqy,
                      const void* dz, void* h, void* wl, void* z, void* qx, void* qy) {
  if (is_f64)
    euler_update<double>(K, wet, (const double*)dh, (const double*)dqx, (const double*)dqy, (const double*)dz,
                         (double*)h, (double*)wl, (double*)z, (double*)qx, (double*)qy);
  else
    euler_update<float>(K, wet, (const float*)dh, (const float*)dqx, (const float*)dqy, (const float*)dz,
                        (float*)h, (float*)wl, (float*)z, (float*)qx, (float*)qy);
}

void orc_infiltration(int is_f64, int64_t M, const int32_t* wet, void* dh, const uint8_t* landuse, const void* h,
                      const void* ks, const void* phi, const void* dth, void* F, const void* dt) {
  if (is_f64)
    infiltration<double>(M, wet, (double*)dh, landuse, (const double*)h, (const double*)ks, (const double*)phi,
                         (const double*)dth, (double*)F, (const double*)dt);
  else
    infiltration<float>(M, wet, (float*)dh, landuse, (const float*)h, (const float*)ks, (const float*)phi,
                        (const float*)dth, (float*)F, (const float*)dt);
}

// Run `nsteps` whole steps of the server sequence on caller-owned compact
// arrays.  ctrl (in/out, double[4]) = {t, dt, export_n, export_T}.
// rain_table may be null (no rain).  Returns the number of export clamps hit,
// and writes the wall time of the loop in seconds to *seconds when non-null.
int orc_run_steps(int is_f64, int order, int64_t N, int nsteps, const int32_t* index, void* h, void* wl, void* qx,
                  void* qy, void* z, void* h_max, void* dh, void* dqx, void* dqy, void* dz,
                  const uint8_t* landuse, const void* manning, const int16_t* station, const double* rain_table,
                  int rain_rows, int rain_cols, int* rain_index, const void* gd, int gd_rows, int gd_w,
                  const void* gwl, int gwl_rows, int gwl_w, const void* gq, int gq_rows, int gq_w, double dx,
                  double cfl, const void* ks, const void* phi, const void* dth, void* F, double* ctrl,
                  double* seconds) {
  auto run = [&](auto tag) -> int {
    using T = decltype(tag);
    Sim<T> s;
    s.N = N; s.index = index;
    s.h = (T*)h; s.wl = (T*)wl; s.qx = (T*)qx; s.qy = (T*)qy; s.z = (T*)z; s.h_max = (T*)h_max;
    s.dh = (T*)dh; s.dqx = (T*)dqx; s.dqy = (T*)dqy; s.dz = (T*)dz;
    s.landuse = landuse; s.manning = (const T*)manning; s.station = station;
    s.ks = (const T*)ks; s.phi = (const T*)phi; s.dth = (const T*)dth; s.F = (T*)F;
    s.rain_table = rain_table; s.rain_rows = rain_rows; s.rain_cols = rain_cols;
    s.order = order;
    s.tb = Tables{gd, gd_rows, gd_w, gwl, gwl_rows, gwl_w, gq, gq_rows, gq_w};
    s.dx = T(dx); s.cfl = T(cfl);
    s.t = T(ctrl[0]); s.dt = T(ctrl[1]); s.export_n = (int)ctrl[2]; s.export_T = ctrl[3];
    s.rain_index = rain_index ? *rain_index : 0;
    s.n_exports = 0;
    auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < nsteps; ++k) sim_step<T>(s);
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    ctrl[0] = (double)s.t; ctrl[1] = (double)s.dt; ctrl[2] = (double)s.export_n;
    if (rain_index) *rain_index = s.rain_index;
    return s.n_exports;
  };
  return is_f64 ? run(double(0)) : run(float(0));
}

// Frozen-slope loss (test construct for the adjoint checks, not in the
// reference; FP64, order 2): advances a source trajectory (state 0) and a
// perturbed one (state 1, its own Manning table and rain table) together on
// a given schedule sched[k] = (t_k, dt_k); the perturbed flux reconstructs
// with the source state's slopes (flux_compact's slope source).  Returns
// sum_k sum_g (h1[gauge_g] after step k - obs[k * G + g])^2.
double orc_frozen_slope_loss(int64_t N, int nsteps, const int32_t* index, double* const* s0, double* const* s1,
                             const double* z, const uint8_t* landuse, const double* man0, const double* man1,
                             const int16_t* station, const double* rain0, const double* rain1, int rain_rows,
                             int rain_cols, double dx, double cfl, const double* sched, const int32_t* gauges,
                             int G, const double* obs) {
  static const double zero2[2] = {0, 0}, zero3[3] = {0, 0, 0};
  std::vector<double> work(2 * 5 * (size_t)N, 0.0), zc(2 * (size_t)N);
  Sim<double> sim[2];
  for (int v = 0; v < 2; ++v) {
    Sim<double>& s = sim[v];
    double* const* a = v ? s1 : s0;  // h, wl, qx, qy, h_max
    double* w = work.data() + (size_t)v * 5 * N;
    std::copy(z, z + N, zc.begin() + (size_t)v * N);
    s.N = N; s.index = index;
    s.h = a[0]; s.wl = a[1]; s.qx = a[2]; s.qy = a[3]; s.h_max = a[4]; s.z = zc.data() + (size_t)v * N;
    s.dh = w; s.dqx = w + N; s.dqy = w + 2 * N; s.dz = w + 3 * N; s.F = w + 4 * N;
    s.landuse = landuse; s.manning = v ? man1 : man0; s.station = station;
    s.ks = s.phi = s.dth = nullptr;
    s.rain_table = v ? rain1 : rain0; s.rain_rows = rain_rows; s.rain_cols = rain_cols;
    s.order = 2;
    s.tb = Tables{zero2, 1, 2, zero2, 1, 2, zero3, 1, 3};
    s.dx = dx; s.cfl = cfl;
    s.export_n = 0; s.export_T = 1.0e30;
    s.rain_index = 0;
    s.n_exports = 0;
  }
  sim[1].hb = sim[0].h; sim[1].qxb = sim[0].qx; sim[1].qyb = sim[0].qy;
  double loss = 0.0;
  for (int k = 0; k < nsteps; ++k) {
    for (int v = 1; v >= 0; --v) {  // the perturbed step reads the source state before it advances
      sim[v].t = sched[2 * k];
      sim[v].dt = sched[2 * k + 1];
      sim_step<double>(sim[v]);
    }
    for (int g = 0; g < G; ++g) {
      const double d = sim[1].h[gauges[g]] - obs[(size_t)k * G + g];
      loss += d * d;
    }
  }
  return loss;
}

}  // extern "C"

