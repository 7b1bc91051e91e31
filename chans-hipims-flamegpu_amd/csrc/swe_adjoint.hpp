// Reverse-mode derivatives of the step's physics (no counterpart in the
// reference).  Each function recomputes the forward expression of
// swe_physics.hpp with its branch decisions (minmod selections, wet/dry
// tests, clamps, HLLC region, friction limiter) and propagates cotangents
// back through the branch that was taken, so the result is the exact
// derivative of the forward map wherever that map is differentiable.
// Bed elevation is static and the time step is frozen, so neither gets a
// cotangent.
#pragma once

#include "swe_physics.hpp"

namespace hp {

template <class T> struct FrameAdj {
  T h, qn, qt;
};

template <class T> __host__ __device__ inline FrameAdj<T> zero_adj() {
  FrameAdj<T> a;
  a.h = T(0); a.qn = T(0); a.qt = T(0);
  return a;
}

// Which argument minmod(x, y) returned: 0 none (opposite signs), 1 x, 2 y.
template <class T> __host__ __device__ inline int minmod_pick(T x, T y) {
  if (x * y <= T(0)) return 0;
  return fabs(x) < fabs(y) ? 1 : 2;
}

// Cotangent of e = c + 0.5 * minmod(c - l, r - c) pushed to (c, l, r).
template <class T> __host__ __device__ inline void muscl_adj(T c, T l, T r, T a_e, T& a_c, T& a_l, T& a_r) {
  a_c += a_e;
  const int p = minmod_pick<T>(c - l, r - c);
  const T m = T(0.5) * a_e;
  if (p == 1) { a_c += m; a_l -= m; }
  else if (p == 2) { a_r += m; a_c -= m; }
}

// VJP of solve_edge (swe_physics.hpp) with cotangents (v_fh, v_fn, v_ft,
// v_bc, v_br) of its five scaled outputs.  Accumulates into the cotangents of
// c, l, r, rr (h, qn, qt) and of the two water levels.  Returns whether the
// edge was active.
template <class T, int ORDER>
__host__ __device__ inline bool solve_edge_vjp(const FrameCell<T>& c, const FrameCell<T>& l, const FrameCell<T>& r,
                                               const FrameCell<T>& rr, bool r_is_ghost, T etaC, T etaR, T dx, T dt,
                                               T v_fh, T v_fn, T v_ft, T v_bc, T v_br, FrameAdj<T>& ac,
                                               FrameAdj<T>& al, FrameAdj<T>& ar, FrameAdj<T>& arr, T& a_etaC,
                                               T& a_etaR, bool frozen_slopes = false) {
  using P = Prec<T>;
  const T g = T(kG);
  if (lt_hs(c.h) && lt_hs(r.h)) return false;

  // ---- forward: edge states ----
  FrameCell<T> eL, eR;
  bool zeroL = false, zeroR = false;
  if (ORDER == 1 || r_is_ghost) {
    eL = c;
    eR = r;
    if (ORDER == 1 && !r_is_ghost) {
      T rb = Prec<T>::limit(c.z - l.z, r.z - c.z);
      if (gt_hs(r.h)) eL.z += T(0.5) * rb;
      rb = Prec<T>::limit(r.z - rr.z, c.z - r.z);
      if (gt_hs(c.h)) eR.z += T(0.5) * rb;
    }
  } else {
    eL.h = c.h + T(0.5) * Prec<T>::limit(c.h - l.h, r.h - c.h);
    eL.z = c.z + T(0.5) * Prec<T>::limit(c.z - l.z, r.z - c.z);
    eR.h = r.h + T(0.5) * Prec<T>::limit(r.h - rr.h, c.h - r.h);
    eR.z = r.z + T(0.5) * Prec<T>::limit(r.z - rr.z, c.z - r.z);
    zeroL = lt_hs(c.h);
    zeroR = lt_hs(r.h);
    if (zeroL) { eL.qn = T(0); eL.qt = T(0); }
    else {
      eL.qn = c.qn + T(0.5) * Prec<T>::limit(c.qn - l.qn, r.qn - c.qn);
      eL.qt = c.qt + T(0.5) * Prec<T>::limit(c.qt - l.qt, r.qt - c.qt);
    }
    if (zeroR) { eR.qn = T(0); eR.qt = T(0); }
    else {
      eR.qn = r.qn + T(0.5) * Prec<T>::limit(r.qn - rr.qn, c.qn - r.qn);
      eR.qt = r.qt + T(0.5) * Prec<T>::limit(r.qt - rr.qt, c.qt - r.qt);
    }
  }
  // ---- forward: velocities and hydrostatic reconstruction ----
  const bool dryL0 = lt_hs(eL.h), dryR0 = lt_hs(eR.h);
  const T rhL = P::div(T(1), eL.h), rhR = P::div(T(1), eR.h);  // as hllc_edge forms them
  const T uL0 = dryL0 ? T(0) : eL.qn * rhL, vL0 = dryL0 ? T(0) : eL.qt * rhL;
  const T uR0 = dryR0 ? T(0) : eR.qn * rhR, vR0 = dryR0 ? T(0) : eR.qt * rhR;
  const T wL0 = eL.h + eL.z, wR0 = eR.h + eR.z;
  const bool maxC = etaC >= etaR;  // fmax(etaC, etaR) took etaC
  const bool minC = etaC <= etaR;  // fmin(etaC, etaR) took etaC
  const T emax = maxC ? etaC : etaR, emin = minC ? etaC : etaR;
  const bool L1a = wL0 <= emax, R1a = wR0 <= emax;
  const T wL1 = L1a ? wL0 : emax, wR1 = R1a ? wR0 : emax;
  const bool L2a = wL1 >= emin, R2a = wR1 >= emin;
  const T wL2 = L2a ? wL1 : emin, wR2 = R2a ? wR1 : emin;
  T wL3 = wL2, wR3 = wR2;
  bool avg = false, liftL = false, takeL = false;
  if (ORDER == 1) {
    avg = (etaC - etaR) * (wL2 - wR2) < T(0);
    if (avg) { wL3 = wR3 = (wL2 + wR2) / T(2); }
  } else {
    liftL = etaC > etaR;
    takeL = wL2 >= wR2;
    const T m = takeL ? wL2 : wR2;
    if (liftL) wL3 = m; else wR3 = m;
  }
  const T zL0 = wL3 - eL.h, zR0 = wR3 - eR.h;
  const bool zbL = zL0 >= zR0;
  const T zbm = zbL ? zL0 : zR0;
  const bool posL = wL3 - zbm > T(0), posR = wR3 - zbm > T(0);
  const T hL = posL ? wL3 - zbm : T(0), hR = posR ? wR3 - zbm : T(0);
  const T zL = wL3 - hL, zR = wR3 - hR;
  const bool flux_on = !(lt_hs(hL) && lt_hs(hR));

  const T s = dt / dx;
  const T Gb = T(0.5) * g * s;
  // cotangents of intermediates
  T a_hL = T(0), a_hR = T(0), a_zL = T(0), a_zR = T(0);
  T a_uL0 = T(0), a_vL0 = T(0), a_uR0 = T(0), a_vR0 = T(0);

  // ---- reverse: bed terms ----
  a_hL += v_bc * Gb * (zL - c.z);
  a_zL += v_bc * Gb * (hL + c.h);
  ac.h += v_bc * Gb * (zL - c.z);
  a_hR += v_br * Gb * (zR - r.z);
  a_zR += v_br * Gb * (hR + r.h);
  ar.h += v_br * Gb * (zR - r.z);

  // ---- reverse: HLLC ----
  if (flux_on) {
    const T qnL = uL0 * hL, qtL = vL0 * hL, qnR = uR0 * hR, qtR = vR0 * hR;
    const T aL = Prec<T>::sqrt_(g * hL), aR = Prec<T>::sqrt_(g * hR);
    const T base = (aL + aR) / T(2) + (uL0 - uR0) / T(4);
    const T u_star = (uL0 + uR0) / T(2) + aL - aR;
    const T a_star = P::star_celerity(base, g);
    const bool sLdry = le_hs(hL), sRdry = le_hs(hR);
    const T x1 = uL0 - aL, x2 = u_star - a_star;
    const T y1 = uR0 + aR, y2 = u_star + a_star;
    const bool sLx1 = x1 <= x2, sRy1 = y1 >= y2;
    const T sL = sLdry ? uR0 - T(2) * aR : (sLx1 ? x1 : x2);
    const T sR = sRdry ? uL0 + T(2) * aL : (sRy1 ? y1 : y2);
    const T sM = P::div(sL * hR * (uR0 - sR) - sR * hL * (uL0 - sL), hR * (uR0 - sR) - hL * (uL0 - sL));
    const T b_fh = v_fh * s, b_fn = v_fn * s, b_ft = v_ft * s;
    T a_qnL = T(0), a_qtL = T(0), a_qnR = T(0), a_qtR = T(0), a_sL = T(0), a_sR = T(0);
    if (T(0) <= sL) {
      a_qnL += b_fh + b_fn * uL0;
      a_uL0 += b_fn * qnL + b_ft * qtL;
      a_hL += b_fn * g * hL;
      a_qtL += b_ft * uL0;
    } else if ((sL < T(0) && T(0) <= sM) || (sM < T(0) && T(0) <= sR)) {
      const bool useL = (sL < T(0) && T(0) <= sM);
      const T vstar = useL ? vL0 : vR0;
      const T D = sR - sL;
      const T qnfL = uL0 * qnL + T(0.5) * g * hL * hL;
      const T qnfR = uR0 * qnR + T(0.5) * g * hR * hR;
      const T Nh = sR * qnL - sL * qnR + sL * sR * (hR - hL);
      const T Nn = sR * qnfL - sL * qnfR + sL * sR * (qnR - qnL);
      const T rD = P::div(T(1), D);
      const T fh = Nh * rD, fn = Nn * rD;
      const T b_fh_tot = b_fh + b_ft * vstar;
      if (useL) a_vL0 += b_ft * fh; else a_vR0 += b_ft * fh;
      T a_D = P::div(-(b_fh_tot * fh + b_fn * fn), D);
      const T a_Nh = P::div(b_fh_tot, D), a_Nn = P::div(b_fn, D);
      a_sR += a_Nh * (qnL + sL * (hR - hL));
      a_sL += a_Nh * (-qnR + sR * (hR - hL));
      a_qnL += a_Nh * sR;
      a_qnR -= a_Nh * sL;
      a_hR += a_Nh * sL * sR;
      a_hL -= a_Nh * sL * sR;
      T a_qnfL = a_Nn * sR, a_qnfR = -a_Nn * sL;
      a_sR += a_Nn * (qnfL + sL * (qnR - qnL));
      a_sL += a_Nn * (-qnfR + sR * (qnR - qnL));
      a_qnR += a_Nn * sL * sR;
      a_qnL -= a_Nn * sL * sR;
      a_qnL += a_qnfL * uL0; a_uL0 += a_qnfL * qnL; a_hL += a_qnfL * g * hL;
      a_qnR += a_qnfR * uR0; a_uR0 += a_qnfR * qnR; a_hR += a_qnfR * g * hR;
      a_sR += a_D;
      a_sL -= a_D;
    } else {
      a_qnR += b_fh + b_fn * uR0;
      a_uR0 += b_fn * qnR + b_ft * qtR;
      a_hR += b_fn * g * hR;
      a_qtR += b_ft * uR0;
    }
    // wave speeds
    T a_aL = T(0), a_aR = T(0), a_ustar = T(0), a_astar = T(0);
    if (sLdry) { a_uR0 += a_sL; a_aR -= T(2) * a_sL; }
    else if (sLx1) { a_uL0 += a_sL; a_aL -= a_sL; }
    else { a_ustar += a_sL; a_astar -= a_sL; }
    if (sRdry) { a_uL0 += a_sR; a_aL += T(2) * a_sR; }
    else if (sRy1) { a_uR0 += a_sR; a_aR += a_sR; }
    else { a_ustar += a_sR; a_astar += a_sR; }
    const T a_base = P::star_celerity_adj(base, g, a_star, a_astar);
    a_uL0 += a_ustar / T(2); a_uR0 += a_ustar / T(2); a_aL += a_ustar; a_aR -= a_ustar;
    a_aL += a_base / T(2); a_aR += a_base / T(2); a_uL0 += a_base / T(4); a_uR0 -= a_base / T(4);
    if (aL > T(0)) a_hL += P::div(a_aL * g, T(2) * aL);
    if (aR > T(0)) a_hR += P::div(a_aR * g, T(2) * aR);
    a_uL0 += a_qnL * hL; a_hL += a_qnL * uL0;
    a_vL0 += a_qtL * hL; a_hL += a_qtL * vL0;
    a_uR0 += a_qnR * hR; a_hR += a_qnR * uR0;
    a_vR0 += a_qtR * hR; a_hR += a_qtR * vR0;
  }

  // ---- reverse: hydrostatic reconstruction ----
  T a_wL3 = a_zL, a_wR3 = a_zR;
  a_hL -= a_zL;
  a_hR -= a_zR;
  T a_zbm = T(0);
  if (posL) { a_wL3 += a_hL; a_zbm -= a_hL; }
  if (posR) { a_wR3 += a_hR; a_zbm -= a_hR; }
  T a_zL0 = zbL ? a_zbm : T(0), a_zR0 = zbL ? T(0) : a_zbm;
  a_wL3 += a_zL0;
  a_wR3 += a_zR0;
  T a_eLh = -a_zL0, a_eRh = -a_zR0;
  T a_wL2 = T(0), a_wR2 = T(0);
  if (ORDER == 1) {
    if (avg) {
      const T m = (a_wL3 + a_wR3) / T(2);
      a_wL2 += m; a_wR2 += m;
    } else {
      a_wL2 += a_wL3; a_wR2 += a_wR3;
    }
  } else {
    const T a_m = liftL ? a_wL3 : a_wR3;
    if (liftL) a_wR2 += a_wR3; else a_wL2 += a_wL3;
    if (takeL) a_wL2 += a_m; else a_wR2 += a_m;
  }
  T a_emax = T(0), a_emin = T(0);
  T a_wL1 = L2a ? a_wL2 : T(0), a_wR1 = R2a ? a_wR2 : T(0);
  if (!L2a) a_emin += a_wL2;
  if (!R2a) a_emin += a_wR2;
  T a_wL0 = L1a ? a_wL1 : T(0), a_wR0 = R1a ? a_wR1 : T(0);
  if (!L1a) a_emax += a_wL1;
  if (!R1a) a_emax += a_wR1;
  if (maxC) a_etaC += a_emax; else a_etaR += a_emax;
  if (minC) a_etaC += a_emin; else a_etaR += a_emin;
  a_eLh += a_wL0;
  a_eRh += a_wR0;
  T a_eLqn = T(0), a_eLqt = T(0), a_eRqn = T(0), a_eRqt = T(0);
  if (!dryL0) {
    a_eLqn += P::div(a_uL0, eL.h);
    a_eLqt += P::div(a_vL0, eL.h);
    a_eLh -= P::div(a_uL0 * uL0 + a_vL0 * vL0, eL.h);
  }
  if (!dryR0) {
    a_eRqn += P::div(a_uR0, eR.h);
    a_eRqt += P::div(a_vR0, eR.h);
    a_eRh -= P::div(a_uR0 * uR0 + a_vR0 * vR0, eR.h);
  }

  // ---- reverse: edge states ----
  if (ORDER == 1 || r_is_ghost || frozen_slopes) {
    // (with frozen slopes the limited slope terms are constants)
    if (ORDER == 2 && !r_is_ghost) {
      if (zeroL) a_eLqn = a_eLqt = T(0);
      if (zeroR) a_eRqn = a_eRqt = T(0);
    }
    ac.h += a_eLh; ac.qn += a_eLqn; ac.qt += a_eLqt;
    ar.h += a_eRh; ar.qn += a_eRqn; ar.qt += a_eRqt;
  } else {
    muscl_adj<T>(c.h, l.h, r.h, a_eLh, ac.h, al.h, ar.h);
    muscl_adj<T>(r.h, rr.h, c.h, a_eRh, ar.h, arr.h, ac.h);
    if (!zeroL) {
      muscl_adj<T>(c.qn, l.qn, r.qn, a_eLqn, ac.qn, al.qn, ar.qn);
      muscl_adj<T>(c.qt, l.qt, r.qt, a_eLqt, ac.qt, al.qt, ar.qt);
    }
    if (!zeroR) {
      muscl_adj<T>(r.qn, rr.qn, c.qn, a_eRqn, ar.qn, arr.qn, ac.qn);
      muscl_adj<T>(r.qt, rr.qt, c.qt, a_eRqt, ar.qt, arr.qt, ac.qt);
    }
  }
  return true;
}

// Cotangent of a boundary ghost (bc_ghost) and of its water level pushed to
// the cell it mirrors.  Given values are constants.
template <class T, int ORDER>
__host__ __device__ inline void ghost_adj(int type, int d, const FrameAdj<T>& ag, T a_etaG, FrameAdj<T>& ac) {
  T gh = T(0), gn = T(0), gt = T(0);  // d ghost / d cell for h, qn, qt
  switch (type) {
    case 3:
      gh = T(1);
      gn = (d == 0 || d == 2) ? T(-1) : T(1);
      gt = gn;
      break;
    case 4: gh = T(1); gn = T(-1); gt = T(1); break;
    case 6: gh = T(0); gn = T(1); gt = T(1); break;
    case 7: gh = T(1); gn = T(0); gt = T(0); break;
    case 8: gh = T(0); gn = T(1); gt = T(1); break;
    case 9: gh = T(0); gn = T(0); gt = T(0); break;
    default: gh = T(1); gn = T(1); gt = T(1); break;
  }
  ac.h += gh * ag.h;
  ac.qn += gn * ag.qn;
  ac.qt += gt * ag.qt;
  // water level of the ghost
  if (ORDER == 1) {
    ac.h += gh * a_etaG;  // h_g + z_g
  } else {
    switch (type) {
      case 6: break;                         // z + given depth
      case 7: ac.h += a_etaG; break;         // z + h_c
      case 8: break;                         // given level
      default: ac.h += a_etaG; break;        // the cell's own wl
    }
  }
}

// VJP of implicit_friction for one wet cell, in its cancellation-free form
// q' = 2 q / (1 + sqrt(1 + x)), x = 4 dt C_f |q| / h^2, C_f = g n^2 h^(-1/3),
// with the Froude limiter branch.  (qx, qy) are the momenta the friction
// acts on; (a_ox, a_oy) the cotangents of its outputs.  Adds the cotangents
// of qx, qy, h and n.
template <class T>
__host__ __device__ inline void friction_vjp(T h, T n, T dt, T qx, T qy, T a_ox, T a_oy, T& a_qx, T& a_qy, T& a_h,
                                             T& a_n) {
  const T g = T(kG);
  if (!(n > T(0))) {
    a_qx += a_ox;
    a_qy += a_oy;
    return;
  }
  using P = Prec<T>;
  const T h13 = P::rcbrt_friction(h);
  const T C_f = g * n * n * h13;
  const T qn2 = qx * qx + qy * qy;
  const T q_norm = P::sqrt_(qn2);
  const T lim = T(5) * h * P::sqrt_(h * g);
  const T x = P::div(T(4) * dt * C_f, h * h) * q_norm;
  const T S = P::sqrt_(T(1) + x);
  const T k = P::div(T(2), T(1) + S);
  T a_k = T(0);
  const T qc[2] = {qx, qy};
  const T ac_[2] = {a_ox, a_oy};
  T aq[2] = {T(0), T(0)};
  for (int m = 0; m < 2; ++m) {
    const T q = qc[m];
    if (!(fabs(q) > T(1.0e-10))) {
      aq[m] += ac_[m];
      continue;
    }
    const T tmp = k * q;
    T a_tmp;
    if (fabs(tmp) > lim) {
      const T out = T(2) * tmp - q;
      if (out * tmp <= T(0)) continue;  // clipped to 0
      a_tmp = T(2) * ac_[m];
      aq[m] -= ac_[m];
    } else {
      a_tmp = ac_[m];
    }
    a_k += a_tmp * q;
    aq[m] += a_tmp * k;
  }
  const T dk_dx = P::div(-T(1), (T(1) + S) * (T(1) + S) * S);
  const T a_x = a_k * dk_dx;
  const T coef = P::div(T(4) * dt, h * h);
  const T a_Cf = a_x * coef * q_norm;
  T a_qnorm = a_x * coef * C_f;
  a_h += a_x * P::div(T(-2) * x, h);
  if (q_norm > T(0)) {
    aq[0] += P::div(a_qnorm * qx, q_norm);
    aq[1] += P::div(a_qnorm * qy, q_norm);
  }
  a_n += P::div(a_Cf * T(2) * C_f, n);
  a_h += P::div(a_Cf * (T(-1.0 / 3.0)) * C_f, h);
  a_qx += aq[0];
  a_qy += aq[1];
}

}  // namespace hp
