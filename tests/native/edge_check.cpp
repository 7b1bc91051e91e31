// Host-side FP64 harness for the CPU tests: exposes the device physics of
// swe_physics.hpp / swe_adjoint.hpp (compiled here for the host) so the
// hand-written reverse derivatives can be checked against central finite
// differences without a GPU.
#include <stdint.h>

#include "../../chans-hipims-flamegpu_amd/csrc/swe_adjoint.hpp"

using namespace hp;

namespace {
// cells: 4 cells (l, c, r, rr) x (h, qn, qt, z)
FrameCell<double> cell(const double* p) {
  FrameCell<double> f;
  f.h = p[0]; f.qn = p[1]; f.qt = p[2]; f.z = p[3];
  return f;
}
}  // namespace

extern "C" {

// h^(-1/3) as the FP64 friction forms it (seed + two Newton steps).
double ec_rcbrt(double h) { return Prec<double>::rcbrt_friction(h); }

// Interior edge: water levels are the cells' h + z.  out: fh, fn, ft, bed_c, bed_r.
int ec_edge(int order, const double* cells, double dx, double dt, double* out) {
  const FrameCell<double> l = cell(cells), c = cell(cells + 4), r = cell(cells + 8), rr = cell(cells + 12);
  EdgeOut<double> o = order == 1 ? solve_edge<double, 1>(c, l, r, rr, false, c.h + c.z, r.h + r.z, dx, dt, dt / dx)
                                  : solve_edge<double, 2>(c, l, r, rr, false, c.h + c.z, r.h + r.z, dx, dt, dt / dx);
  out[0] = o.fh; out[1] = o.fn; out[2] = o.ft; out[3] = o.bed_c; out[4] = o.bed_r;
  return o.active ? 1 : 0;
}

// grad[12] = d(v . out)/d(h, qn, qt of l, c, r, rr)
int ec_edge_vjp(int order, const double* cells, double dx, double dt, const double* v, double* grad) {
  const FrameCell<double> l = cell(cells), c = cell(cells + 4), r = cell(cells + 8), rr = cell(cells + 12);
  FrameAdj<double> a[4] = {zero_adj<double>(), zero_adj<double>(), zero_adj<double>(), zero_adj<double>()};
  double aec = 0.0, aer = 0.0;
  bool act = order == 1 ? solve_edge_vjp<double, 1>(c, l, r, rr, false, c.h + c.z, r.h + r.z, dx, dt, v[0], v[1],
                                                    v[2], v[3], v[4], a[1], a[0], a[2], a[3], aec, aer)
                        : solve_edge_vjp<double, 2>(c, l, r, rr, false, c.h + c.z, r.h + r.z, dx, dt, v[0], v[1],
                                                    v[2], v[3], v[4], a[1], a[0], a[2], a[3], aec, aer);
  a[1].h += aec;
  a[2].h += aer;
  for (int k = 0; k < 4; ++k) {
    grad[3 * k] = a[k].h; grad[3 * k + 1] = a[k].qn; grad[3 * k + 2] = a[k].qt;
  }
  return act ? 1 : 0;
}

// Interior order-2 edge whose faces take their limited slopes (and the
// dry-cell momentum rule) from `base` (4 cells like `cells`) instead of from
// the cells themselves: the model the frozen-slope reverse differentiates.
int ec_edge_frozen(const double* cells, const double* base, double dx, double dt, double* out) {
  const FrameCell<double> c = cell(cells + 4), r = cell(cells + 8);
  const FrameCell<double> bl = cell(base), bc = cell(base + 4), br = cell(base + 8), brr = cell(base + 12);
  using P = Prec<double>;
  const double g = kG;
  EdgeOut<double> o;
  o.fh = o.fn = o.ft = o.bed_c = o.bed_r = 0.0;
  o.active = !(lt_hs(c.h) && lt_hs(r.h));
  if (!o.active) {
    out[0] = out[1] = out[2] = out[3] = out[4] = 0.0;
    return 0;
  }
  FrameCell<double> eL, eR;
  eL.h = c.h + 0.5 * P::limit(bc.h - bl.h, br.h - bc.h);
  eL.z = c.z + 0.5 * P::limit(bc.z - bl.z, br.z - bc.z);
  eR.h = r.h + 0.5 * P::limit(br.h - brr.h, bc.h - br.h);
  eR.z = r.z + 0.5 * P::limit(br.z - brr.z, bc.z - br.z);
  eL.qn = lt_hs(bc.h) ? 0.0 : c.qn + 0.5 * P::limit(bc.qn - bl.qn, br.qn - bc.qn);
  eL.qt = lt_hs(bc.h) ? 0.0 : c.qt + 0.5 * P::limit(bc.qt - bl.qt, br.qt - bc.qt);
  eR.qn = lt_hs(br.h) ? 0.0 : r.qn + 0.5 * P::limit(br.qn - brr.qn, bc.qn - br.qn);
  eR.qt = lt_hs(br.h) ? 0.0 : r.qt + 0.5 * P::limit(br.qt - brr.qt, bc.qt - br.qt);
  double fh, fn, ft;
  hllc_edge<double, 2>(eL, eR, c.h + c.z, r.h + r.z, fh, fn, ft);
  out[0] = P::scale(fh, dx, dt, dt / dx);
  out[1] = P::scale(fn, dx, dt, dt / dx);
  out[2] = P::scale(ft, dx, dt, dt / dx);
  out[3] = P::scale(0.5 * g * (eL.h + c.h) * (eL.z - c.z), dx, dt, dt / dx);
  out[4] = P::scale(0.5 * g * (eR.h + r.h) * (eR.z - r.z), dx, dt, dt / dx);
  return 1;
}

// The frozen-slope VJP at `cells` (grad[12] as ec_edge_vjp; l and rr get none).
int ec_edge_vjp_frozen(const double* cells, double dx, double dt, const double* v, double* grad) {
  const FrameCell<double> l = cell(cells), c = cell(cells + 4), r = cell(cells + 8), rr = cell(cells + 12);
  FrameAdj<double> a[4] = {zero_adj<double>(), zero_adj<double>(), zero_adj<double>(), zero_adj<double>()};
  double aec = 0.0, aer = 0.0;
  const bool act = solve_edge_vjp<double, 2>(c, l, r, rr, false, c.h + c.z, r.h + r.z, dx, dt, v[0], v[1], v[2],
                                             v[3], v[4], a[1], a[0], a[2], a[3], aec, aer, true);
  a[1].h += aec;
  a[2].h += aer;
  for (int k = 0; k < 4; ++k) {
    grad[3 * k] = a[k].h; grad[3 * k + 1] = a[k].qn; grad[3 * k + 2] = a[k].qt;
  }
  return act ? 1 : 0;
}

// Boundary edge of cell c (h, qn, qt, z) in direction d with ghost `type`.
int ec_ghost_edge(int order, int type, int d, const double* c4, const double* given, double dx, double dt,
                  double* out) {
  const FrameCell<double> c = cell(c4);
  const FrameCell<double> g = bc_ghost<double>(type, c, d, given[0], given[1], given[2], given[3]);
  const double etaG = order == 1 ? g.h + g.z : ghost_wl<double>(type, c.h + c.z, g, given[1]);
  EdgeOut<double> o = order == 1 ? solve_edge<double, 1>(c, c, g, g, true, c.h + c.z, etaG, dx, dt, dt / dx)
                                  : solve_edge<double, 2>(c, c, g, g, true, c.h + c.z, etaG, dx, dt, dt / dx);
  out[0] = o.fh; out[1] = o.fn; out[2] = o.ft; out[3] = o.bed_c; out[4] = o.bed_r;
  return o.active ? 1 : 0;
}

int ec_ghost_edge_vjp(int order, int type, int d, const double* c4, const double* given, double dx, double dt,
                      const double* v, double* grad3) {
  const FrameCell<double> c = cell(c4);
  const FrameCell<double> g = bc_ghost<double>(type, c, d, given[0], given[1], given[2], given[3]);
  const double etaG = order == 1 ? g.h + g.z : ghost_wl<double>(type, c.h + c.z, g, given[1]);
  FrameAdj<double> ac = zero_adj<double>(), ag = zero_adj<double>(), unused = zero_adj<double>();
  FrameAdj<double> unused2 = zero_adj<double>();
  double aec = 0.0, aeg = 0.0;
  bool act;
  if (order == 1) {
    act = solve_edge_vjp<double, 1>(c, c, g, g, true, c.h + c.z, etaG, dx, dt, v[0], v[1], v[2], v[3], v[4], ac,
                                    unused, ag, unused2, aec, aeg);
    ghost_adj<double, 1>(type, d, ag, aeg, ac);
  } else {
    act = solve_edge_vjp<double, 2>(c, c, g, g, true, c.h + c.z, etaG, dx, dt, v[0], v[1], v[2], v[3], v[4], ac,
                                    unused, ag, unused2, aec, aeg);
    ghost_adj<double, 2>(type, d, ag, aeg, ac);
  }
  ac.h += aec;
  grad3[0] = ac.h; grad3[1] = ac.qn; grad3[2] = ac.qt;
  return act ? 1 : 0;
}

void ec_friction(double h, double n, double dt, double qx, double qy, double* out2) {
  implicit_friction<double>(h, n, dt, qx, qy);
  out2[0] = qx;
  out2[1] = qy;
}

// grad4 = d(a . out)/d(qx, qy, h, n)
void ec_friction_vjp(double h, double n, double dt, double qx, double qy, const double* a2, double* grad4) {
  double aqx = 0, aqy = 0, ah = 0, an = 0;
  friction_vjp<double>(h, n, dt, qx, qy, a2[0], a2[1], aqx, aqy, ah, an);
  grad4[0] = aqx; grad4[1] = aqy; grad4[2] = ah; grad4[3] = an;
}

}  // extern "C"
