// Device-side physics of the HiPIMS shallow-water step, written for gfx950.
//
// Everything here works in an *edge frame*: an edge between cell c and its
// neighbour r in direction d (E, N, W, S; N is raster row - 1) is solved with
// the normal momentum qn = q . n and the tangential momentum qt = q . t,
// t = (-n_y, n_x).  One call produces the edge's mass flux and both sides'
// bed-slope terms, so a stencil kernel can solve every edge once and hand the
// negated flux to the other cell.
//
// Behaviour follows the reference kernels:
//   minmod ............ CHANS-HiPIMS/hipims/cuda/fluxCal_jh_modified_surface_Kernel.cu:321-328
//   order-1 bed recon . fluxCal_jh_modified_surface_Kernel.cu:579-592
//   order-1 hydro recon fluxCal_jh_modified_surface_Kernel.cu:96-139
//   order-2 MUSCL ..... flucCal_2ndOrder_jh_improved_Kernel.cu:83-107
//   order-2 hydro recon flucCal_2ndOrder_jh_improved_Kernel.cu:43-72
//   HLLC .............. fluxCal_jh_modified_surface_Kernel.cu:173-260
//   bed-slope source .. fluxCal_jh_modified_surface_Kernel.cu:603-610,
//                       flucCal_2ndOrder_jh_improved_Kernel.cu:494-501
//   boundary ghosts ... fluxCal_jh_modified_surface_Kernel.cu:262-435
//   implicit friction . friction_implicit_andUpdate_jh_Kernel.cu:44-139
//   CFL accelerator ... timeControl_Kernel.cu:186-189
//
// Precision policy (template parameter T):
//   double: expressions are evaluated in the reference's order with IEEE
//           division and square root, except four forms that agree with the
//           reference's to a few ulp at a fraction of the instructions: the
//           friction's h^(-1/3) (a float seed and two Newton steps in place of
//           pow), the star celerity |base| (in place of sqrt(g * base^2 / g)),
//           the edge terms scaled by dt / dx formed once (in place of
//           (x / dx) * dt per term) and one reciprocal per face for both
//           velocities and per edge for both star-region fluxes; results
//           agree with the reference update to rounding.
//   float:  the reconstruction works relative to the cell's own bed (a local
//           datum) so thin films are not lost in the rounding of a 50 m
//           water level; the implicit-friction quotient uses its
//           cancellation-free form 2q / (1 + sqrt(1 + x)); divisions and
//           square roots use the hardware reciprocal / square-root (<= 1 ulp);
//           the HLLC star celerity sqrt(g * base^2 / g) is taken as |base|.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hp {

constexpr double kG = 9.81;
constexpr double kHs = 1.0e-6;

// Wet/dry threshold tests of a depth against h_small = 1e-6 m, decided as
// FP64 comparisons in both precisions.  For a float x no FP64 arithmetic is
// needed: fl(1e-6) = 9.99999997e-7 lies just below 1e-6 and no float lies
// between the two, so x < 1e-6 (and x <= 1e-6) exactly when x <= fl(1e-6),
// and x > 1e-6 (and x >= 1e-6) exactly when x > fl(1e-6).
__host__ __device__ inline bool lt_hs(double x) { return x < kHs; }
__host__ __device__ inline bool le_hs(double x) { return x <= kHs; }
__host__ __device__ inline bool gt_hs(double x) { return x > kHs; }
__host__ __device__ inline bool lt_hs(float x) { return x <= float(kHs); }
__host__ __device__ inline bool le_hs(float x) { return x <= float(kHs); }
__host__ __device__ inline bool gt_hs(float x) { return x > float(kHs); }

// h += dh, and whether the new depth is below h_small, decided on the exact
// sum in FP64 (the sum of two floats is exact in double, and rounding it to
// float gives the float sum).
__host__ __device__ inline bool new_depth_dry(double& h, double dh) {
  h += dh;
  return h < kHs;
}
__host__ __device__ inline bool new_depth_dry(float& h, float dh) {
  const double hn = (double)h + (double)dh;
  h = (float)hn;
  return hn < kHs;
}

template <class T> __host__ __device__ inline T minmod(T x, T y) {
  const T m = fabs(x) < fabs(y) ? x : y;
  return x * y <= T(0) ? T(0) : m;
}

template <class T> struct Prec;
template <> struct Prec<double> {
  static constexpr bool local_datum = false;
  __host__ __device__ static inline double limit(double x, double y) { return minmod<double>(x, y); }
  __host__ __device__ static inline double clamp(double x, double lo, double hi) { return fmax(fmin(x, hi), lo); }
  __host__ __device__ static inline double max2(double x, double y) { return fmax(x, y); }
  __host__ __device__ static inline double min2(double x, double y) { return fmin(x, y); }
  __host__ __device__ static inline double div(double a, double b) { return a / b; }
  __host__ __device__ static inline double sqrt_(double x) { return __builtin_sqrt(x); }
  // h^(-1/3) for a wet depth (h > 1e-6): a single-precision seed (relative
  // error ~1e-7) refined by two Newton steps of y <- y (4 - h y^3) / 3, the
  // root of y^-3 = h; the error squares each step (1e-7 -> 1e-13 -> rounding,
  // a few ulp), at about a tenth of the instructions of the double pow.
  __host__ __device__ static inline double rcbrt_friction(double h) {
    double y = (double)exp2f(log2f((float)h) * (-1.0f / 3.0f));
    y = y * (4.0 - h * y * y * y) * (1.0 / 3.0);
    y = y * (4.0 - h * y * y * y) * (1.0 / 3.0);
    return y;
  }
  // x * (dt / dx), the quotient formed once by the caller (the reference
  // forms (x / dx) * dt per term: five divisions per edge; the two differ by
  // rounding only)
  __host__ __device__ static inline double scale(double x, double, double, double dt_dx) { return x * dt_dx; }
  // star-region celerity sqrt(g h*) with h* = base^2 / g (the reference's
  // form, a division and a square root per edge) is |base|; the two differ
  // by rounding only (at most 2 ulp)
  __host__ __device__ static inline double star_celerity(double base, double) { return fabs(base); }
  // d|base| / d base = sign(base), 0 at base = 0
  __host__ __device__ static inline double star_celerity_adj(double base, double, double, double a_astar) {
    return base > 0.0 ? a_astar : (base < 0.0 ? -a_astar : 0.0);
  }
};
template <> struct Prec<float> {
  static constexpr bool local_datum = true;
  // minmod(x, y) is the median of x, y and 0: 0 when the signs differ, else
  // the one nearer zero; one v_med3_f32 instead of a product, two compares
  // and two selects (it differs from the product form only when x * y
  // underflows to zero)
  __device__ static inline float limit(float x, float y) { return __builtin_amdgcn_fmed3f(x, y, 0.0f); }
  // clamp to [lo, hi] (lo <= hi) as one median-of-three
  __device__ static inline float clamp(float x, float lo, float hi) { return __builtin_amdgcn_fmed3f(x, lo, hi); }
  // max / min as a median with an infinity: no quieting of the (loaded,
  // possibly signalling) inputs that fmaxf / fminf need in IEEE mode
  __device__ static inline float max2(float x, float y) { return __builtin_amdgcn_fmed3f(x, y, INFINITY); }
  __device__ static inline float min2(float x, float y) { return __builtin_amdgcn_fmed3f(x, y, -INFINITY); }
  __device__ static inline float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  __device__ static inline float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
  // h^(-1/3) = exp2(-log2(h) / 3) with the hardware log2 / exp2 (h >= 1e-6)
  __device__ static inline float rcbrt_friction(float h) {
    return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(h) * (-1.0f / 3.0f));
  }
  // x * (dt / dx) with the quotient computed once per block
  __device__ static inline float scale(float x, float, float, float dt_dx) { return x * dt_dx; }
  // sqrt(g * base^2 / g) is |base|; no square root needed
  __device__ static inline float star_celerity(float base, float) { return fabsf(base); }
  // d|base| / d base = sign(base), 0 at base = 0
  __device__ static inline float star_celerity_adj(float base, float, float, float a_astar) {
    return base > 0.0f ? a_astar : (base < 0.0f ? -a_astar : 0.0f);
  }
};

// A cell seen from an edge frame.
template <class T> struct FrameCell {
  T h, qn, qt, z;
};

// Normal of direction d (E, N, W, S) as small integers.
__host__ __device__ inline int nrm_x(int d) { return d == 0 ? 1 : (d == 2 ? -1 : 0); }
__host__ __device__ inline int nrm_y(int d) { return d == 1 ? 1 : (d == 3 ? -1 : 0); }

// q . n and q . t with t = (-n_y, n_x); exact for unit axis normals.
template <class T> __host__ __device__ inline T rot_n(T qx, T qy, int d) {
  switch (d) {
    case 0: return qx;
    case 1: return qy;
    case 2: return -qx;
    default: return -qy;
  }
}
template <class T> __host__ __device__ inline T rot_t(T qx, T qy, int d) {
  switch (d) {
    case 0: return qy;
    case 1: return -qx;
    case 2: return -qy;
    default: return qx;
  }
}

// Boundary ghost in the edge frame, built from the cell state c (frame
// values) and the decoded boundary type.  The reference builds the ghost in
// the global frame (boundary_WallNonSlip etc.); the frame form below is the
// same map rotated:
//   3 "rigid":  E/W edges reverse qx and qy; on N/S edges the reference's
//               expressions reduce to the identity (the ghost copies c).
//   4 slip:     reverse the normal component only.
//   5 open:     copy.
//   6 H given:  h = given depth, q copied.
//   7 Q given:  q = given (global qx, qy), h copied.
//   8 WL given: h = max(given_wl, z) - z, q copied.
//   9 fall:     h = 0, q = 0.
// Any other type: copy (the reference leaves the ghost undefined there).
// gwl is passed relative to the same datum as c.z.
template <class T>
__host__ __device__ inline FrameCell<T> bc_ghost(int type, const FrameCell<T>& c, int d, T gd, T gwl, T gqx, T gqy) {
  FrameCell<T> b = c;
  switch (type) {
    case 3:
      if (d == 0 || d == 2) { b.qn = -c.qn; b.qt = -c.qt; }
      break;
    case 4:
      b.qn = -c.qn;
      break;
    case 6:
      b.h = gd;
      break;
    case 7:
      b.qn = rot_n(gqx, gqy, d);
      b.qt = rot_t(gqx, gqy, d);
      break;
    case 8: {
      T wlb = gwl > c.z ? gwl : c.z;
      b.h = wlb - c.z;
      break;
    }
    case 9:
      b.h = T(0); b.qn = T(0); b.qt = T(0);
      break;
    default:
      break;
  }
  return b;
}

// Water level of a boundary ghost as the reference stores it in field_b[1]
// (fluxCal_jh_modified_surface_Kernel.cu:262-396): only H/Q/WL-given ghosts
// get a new one; the others keep the cell's wl.
template <class T>
__host__ __device__ inline T ghost_wl(int type, T c_wl, const FrameCell<T>& g, T gwl_rel) {
  switch (type) {
    case 6:
    case 7: return g.z + g.h;
    case 8: return gwl_rel > g.z ? gwl_rel : g.z;
    default: return c_wl;
  }
}

// Result of one edge solve, already scaled by dt/dx as the reference
// accumulates it:  e = (raw / dx) * dt.
template <class T> struct EdgeOut {
  T fh;      // mass flux leaving c through the edge
  T fn, ft;  // normal / tangential momentum flux (edge frame)
  T bed_c;   // bed-slope term of c along the normal (sign of n included below)
  T bed_r;   // bed-slope term of r along the same normal
  bool active;
};

// HLLC with the hydrostatic reconstruction applied first.
// eL / eR: edge states before reconstruction (h, qn, qt, z) - velocities come
// from these.  etaC / etaR: the cells' water levels used by the clamp.
// On return eL.h/eL.z and eR.h/eR.z hold the reconstructed depth and bed.
template <class T, int ORDER>
__host__ __device__ inline void hllc_edge(FrameCell<T>& eL, FrameCell<T>& eR, T etaC, T etaR, T& fh, T& fn, T& ft) {
  // Written as straight-line code with selects: a wave whose edges fall in
  // different wave-speed regions evaluates one path instead of all three
  // under divergent branches.  The selected values are the same expressions
  // the branches would produce.
  using P = Prec<T>;
  const T g = T(kG);
  // one reciprocal per face serves both velocities (in FP64 two IEEE
  // divisions become one; in FP32 the hardware reciprocal was shared anyway)
  const T rhL = P::div(T(1), eL.h), rhR = P::div(T(1), eR.h);
  const T uL0 = lt_hs(eL.h) ? T(0) : eL.qn * rhL;
  const T vL0 = lt_hs(eL.h) ? T(0) : eL.qt * rhL;
  const T uR0 = lt_hs(eR.h) ? T(0) : eR.qn * rhR;
  const T vR0 = lt_hs(eR.h) ? T(0) : eR.qt * rhR;

  // hydrostatic reconstruction
  T wL = eL.h + eL.z;
  T wR = eR.h + eR.z;
  const T emax = P::max2(etaC, etaR), emin = P::min2(etaC, etaR);
  wL = P::clamp(wL, emin, emax);
  wR = P::clamp(wR, emin, emax);
  if (ORDER == 1) {
    const T m = (wL + wR) / T(2);
    const bool avg = (etaC - etaR) * (wL - wR) < T(0);
    wL = avg ? m : wL;
    wR = avg ? m : wR;
  } else {
    const T w = fmax(wL, wR);
    const bool left_high = etaC > etaR;
    wL = left_high ? w : wL;
    wR = left_high ? wR : w;
  }
  const T zbm = fmax(wL - eL.h, wR - eR.h);
  eL.h = fmax(wL - zbm, T(0));
  eR.h = fmax(wR - zbm, T(0));
  eL.z = wL - eL.h;
  eR.z = wR - eR.h;

  const T hL = eL.h, hR = eR.h;
  const T qnL = uL0 * hL, qtL = vL0 * hL;
  const T qnR = uR0 * hR, qtR = vR0 * hR;
  const T aL = P::sqrt_(g * hL);
  const T aR = P::sqrt_(g * hR);
  const T base = (aL + aR) / T(2) + (uL0 - uR0) / T(4);
  const T u_star = (uL0 + uR0) / T(2) + aL - aR;
  const T a_star = P::star_celerity(base, g);
  const T sL = le_hs(hL) ? uR0 - T(2) * aR : fmin(uL0 - aL, u_star - a_star);
  const T sR = le_hs(hR) ? uL0 + T(2) * aL : fmax(uR0 + aR, u_star + a_star);
  const T sM = P::div(sL * hR * (uR0 - sR) - sR * hL * (uL0 - sL), hR * (uR0 - sR) - hL * (uL0 - sL));
  const T qnfL = uL0 * qnL + T(0.5) * g * hL * hL;
  const T qnfR = uR0 * qnR + T(0.5) * g * hR * hR;
  const T rS = P::div(T(1), sR - sL);
  const T fh_s = (sR * qnL - sL * qnR + sL * sR * (hR - hL)) * rS;
  const T fn_s = (sR * qnfL - sL * qnfR + sL * sR * (qnR - qnL)) * rS;
  const bool star_l = sL < T(0) && T(0) <= sM;
  const T ft_s = fh_s * (star_l ? vL0 : vR0);
  const bool left = T(0) <= sL;
  const bool star = star_l || (sM < T(0) && T(0) <= sR);
  const bool dry = lt_hs(hL) && lt_hs(hR);
  fh = dry ? T(0) : (left ? qnL : (star ? fh_s : qnR));
  fn = dry ? T(0) : (left ? qnfL : (star ? fn_s : qnfR));
  ft = dry ? T(0) : (left ? uL0 * qtL : (star ? ft_s : uR0 * qtR));
}

// Solve one edge from cell c towards r (direction d).
//   l:  c's neighbour opposite to r (c itself when that neighbour is nodata)
//   rr: r's neighbour beyond r (r itself when nodata)
//   r_is_ghost: r is a boundary ghost (c's neighbour in direction d is nodata);
//               then l and rr are unused and no reconstruction is applied.
// etaC / etaR: the two cells' water levels.  Order 1 uses h + z of the
// unreconstructed cells (the reference rebuilds them from (h, z)); order 2
// uses the stored cell water level wl (a ghost's wl for a boundary edge).
// Frame values carry z relative to c's datum (c.z == 0 in float).
// The returned terms are scaled by dt/dx: (raw / dx) * dt in double, raw *
// dt_dx in float (dt_dx = dt / dx, computed once by the caller).
template <class T, int ORDER>
__host__ __device__ inline EdgeOut<T> solve_edge(const FrameCell<T>& c, const FrameCell<T>& l, const FrameCell<T>& r,
                                        const FrameCell<T>& rr, bool r_is_ghost, T etaC, T etaR, T dx, T dt,
                                        T dt_dx) {
  const T g = T(kG);
  EdgeOut<T> o;
  o.fh = T(0); o.fn = T(0); o.ft = T(0); o.bed_c = T(0); o.bed_r = T(0);
  o.active = !(lt_hs(c.h) && lt_hs(r.h));
  if (!o.active) return o;

  FrameCell<T> eL, eR;
  if (ORDER == 1) {
    eL = c;
    eR = r;
    if (!r_is_ghost) {
      const T dz = r.z - c.z;
      const T zl = eL.z + T(0.5) * Prec<T>::limit(c.z - l.z, dz);
      const T zr = eR.z + T(0.5) * Prec<T>::limit(r.z - rr.z, -dz);
      eL.z = gt_hs(r.h) ? zl : eL.z;
      eR.z = gt_hs(c.h) ? zr : eR.z;
    }
  } else {
    if (r_is_ghost) {
      eL = c;
      eR = r;
    } else {
      // r - c serves both sides: c - r is its exact negation, and the
      // limiter gives zero whenever either argument is zero
      const T dh = r.h - c.h, dz = r.z - c.z, dqn = r.qn - c.qn, dqt = r.qt - c.qt;
      eL.h = c.h + T(0.5) * Prec<T>::limit(c.h - l.h, dh);
      eL.z = c.z + T(0.5) * Prec<T>::limit(c.z - l.z, dz);
      eR.h = r.h + T(0.5) * Prec<T>::limit(r.h - rr.h, -dh);
      eR.z = r.z + T(0.5) * Prec<T>::limit(r.z - rr.z, -dz);
      const T qnL = c.qn + T(0.5) * Prec<T>::limit(c.qn - l.qn, dqn);
      const T qtL = c.qt + T(0.5) * Prec<T>::limit(c.qt - l.qt, dqt);
      const T qnR = r.qn + T(0.5) * Prec<T>::limit(r.qn - rr.qn, -dqn);
      const T qtR = r.qt + T(0.5) * Prec<T>::limit(r.qt - rr.qt, -dqt);
      eL.qn = lt_hs(c.h) ? T(0) : qnL;
      eL.qt = lt_hs(c.h) ? T(0) : qtL;
      eR.qn = lt_hs(r.h) ? T(0) : qnR;
      eR.qt = lt_hs(r.h) ? T(0) : qtR;
    }
  }
  T fh, fn, ft;
  hllc_edge<T, ORDER>(eL, eR, etaC, etaR, fh, fn, ft);
  using P = Prec<T>;
  o.fh = P::scale(fh, dx, dt, dt_dx);
  o.fn = P::scale(fn, dx, dt, dt_dx);
  o.ft = P::scale(ft, dx, dt, dt_dx);
  o.bed_c = P::scale(T(0.5) * g * (eL.h + c.h) * (eL.z - c.z), dx, dt, dt_dx);
  o.bed_r = P::scale(T(0.5) * g * (eR.h + r.h) * (eR.z - r.z), dx, dt, dt_dx);
  return o;
}

// Implicit Manning friction for one wet cell (h >= hs), updating qx, qy in
// place after they have received their flux increments.
template <class T>
__host__ __device__ inline void implicit_friction(T h, T n, T dt, T& qx, T& qy) {
  using P = Prec<T>;
  const T g = T(kG);
  const T C_f = g * n * n * Prec<T>::rcbrt_friction(h);
  const T q_norm = P::sqrt_(qx * qx + qy * qy);
  if (!(n > T(0))) return;
  const T lim = T(5) * h * P::sqrt_(h * g);
  if (P::local_datum) {
    // cancellation-free: (1 - sqrt(1 + x)) / (-(x/2) / q_k) == 2 q_k / (1 + sqrt(1 + x))
    const T x = P::div(T(4) * dt * C_f, h * h) * q_norm;
    const T k = P::div(T(2), T(1) + P::sqrt_(T(1) + x));
    // per component, as selects: q' = k q, or past the Froude limit the
    // reflected 2 q' - q, zeroed on a sign change; |q| <= 1e-10 unchanged
    const T tx = k * qx, ty = k * qy;
    const T rx = T(2) * tx - qx, ry = T(2) * ty - qy;
    const T ox = fabs(tx) > lim ? (rx * tx <= T(0) ? T(0) : rx) : tx;
    const T oy = fabs(ty) > lim ? (ry * ty <= T(0) ? T(0) : ry) : ty;
    qx = fabs(qx) > T(1.0e-10) ? ox : qx;
    qy = fabs(qy) > T(1.0e-10) ? oy : qy;
  } else {
    if (fabs(qx) > T(1.0e-10)) {
      const T tmp = (T(1) - P::sqrt_(T(1) + (T(4) * dt * C_f) / (h * h) * q_norm)) /
                    (T(-2) * dt * C_f / (h * h) * (q_norm / qx));
      if (fabs(tmp) > lim) {
        qx = T(2) * tmp - qx;
        if (qx * tmp <= T(0)) qx = T(0);
      } else {
        qx = tmp;
      }
    }
    if (fabs(qy) > T(1.0e-10)) {
      const T tmp = (T(1) - P::sqrt_(T(1) + (T(4) * dt * C_f) / (h * h) * q_norm)) /
                    (T(-2) * dt * C_f / (h * h) * (q_norm / qy));
      if (fabs(tmp) > lim) {
        qy = T(2) * tmp - qy;
        if (qy * tmp <= T(0)) qy = T(0);
      } else {
        qy = tmp;
      }
    }
  }
}

// Green-Ampt increment of one listed cell (infiltrationCUDA_Kernel.cu:30-44):
// F1 = (A + sqrt(A^2 + 4 dt Ks (phi + h) dtheta)) / 2 with A = F0 + dt Ks, the
// positive root of F^2 - A F - dt Ks (phi + h) dtheta = 0; the increment is
// capped by the depth.
// In float, F1 - F0 loses the increment to cancellation once F0 >> dt Ks;
// the increment is then taken from the rationalised form
// dF = 2 (F0 dt Ks + c) / (s + F0 - dt Ks), s = sqrt(A^2 + 4c), which has no
// cancellation for F0 >= dt Ks (in double the reference's form is kept).
template <class T>
__host__ __device__ inline T green_ampt(T h, T F0, T ks, T phi, T dtheta, T dt) {
  const T dks = dt * ks;
  const T a = F0 + dks;
  const T c4 = T(4) * dt * ks * (phi + h) * dtheta;
  const T s = Prec<T>::sqrt_(a * a + c4);
  T dF;
  if (Prec<T>::local_datum && F0 >= dks) {
    const T den = s + (F0 - dks);
    dF = den > T(0) ? Prec<T>::div(T(2) * F0 * dks + T(0.5) * c4, den) : T(0);
  } else {
    dF = T(0.5) * (a + s) - F0;
  }
  return fmin(h, dF);
}

// CFL time-step candidate of one updated cell (timeControl_Kernel.cu:186-189):
// the square root is single precision in both instantiations, as in the
// reference - correctly rounded for FP64 (parity with the reference's FP64
// time step), the hardware square root for FP32 (one instruction; the
// reference's build uses --use_fast_math, cuda/setup.py:78, whose sqrtf is not
// correctly rounded either).
template <class T>
__device__ inline T cfl_candidate(T h, T qx, T qy, T cfl_dx) {
  if (!gt_hs(h)) return T(60.0);
  const float gh = (float)(T(kG) * h);
  const float a = sizeof(T) == 4 ? __builtin_amdgcn_sqrtf(gh) : __builtin_sqrtf(gh);
  return Prec<T>::div(cfl_dx, T(a) + Prec<T>::div(fmax(fabs(qx), fabs(qy)), h));
}

// Mask-code decode (fluxCal_jh_modified_surface_Kernel.cu:528-535): leading
// decimal digit = boundary type, remaining digits = series id.
__host__ __device__ inline void decode_code(int32_t code, int& type, int& id) {
  int t = code, n = 0;
  while (t >= 10) {
    t = (t - (t % 10)) / 10;
    n++;
  }
  int p = 1;
  for (int k = 0; k < n; ++k) p *= 10;
  type = t;
  id = code - t * p;
}

}  // namespace hp
