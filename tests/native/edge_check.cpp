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
