// FP32 step as a column walk (gfx950): the flux stencil of both step forms
// (flux pass, single-pass step) for single-precision domains.
//
// A workgroup is one wave64.  Its lanes load 64 consecutive columns
// [x0 - 2, x0 + 62) and walk down rows y0 - 2 .. y0 + R + 1 of the planes,
// finishing rows y0 .. y0 + R - 1 of the 60 columns [x0, x0 + 60) (lanes
// 2 .. 61; the two lanes on either side are the x halo).  Nothing is staged
// in LDS:
//   * a cell's x neighbours come from the lanes beside it through DPP
//     wavefront shifts (one VALU move per value), its y neighbours from the
//     rows the lane already holds (a rolling window of four rows: the row
//     being finished, the next two, and the one being prefetched);
//   * every limited slope (MUSCL: h, qx, qy, z along x and along y; order 1:
//     z) is formed once per cell and serves the cell's two faces on that
//     axis (the reference's per-edge form evaluates it twice; minmod is
//     symmetric and odd, so the faces are the same numbers);
//   * each lane solves the west edge of its cell (the west neighbour's east
//     face arrives by a lane shift) and the south edge between the row being
//     finished and the next (both faces in its registers); the west
//     neighbour's share of the x edge goes back to it by the opposite shift,
//     the next row's share of the y edge stays in the lane for the next row.
// The per-edge arithmetic is the flux tile's (solve in the edge frame on the
// west / north cell's datum, HLLC after the hydrostatic reconstruction, bed
// terms on both sides), and a cell sums its E, N, W, S shares in the
// reference's order.
//
// Geometry: a unit is (strip bx, row block by); 60 columns x R rows.  Each
// interior x edge is solved once, except the one between two strips (solved
// by both); each interior y edge once, except the one between two row
// blocks.

namespace hp {

constexpr int kWalkLanes = 64;
constexpr int kWalkOwn = 60;

// Lane i receives lane i - 1's value (lane 0 reads 0); all 64 lanes must be
// active.  Lanes 0 and 63 are halo lanes whose results are never used, so
// the out-of-range value does not matter, and without an `old` operand the
// shift is one instruction.
__device__ inline int from_west(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138 /* wave_shr:1 */, 0xf, 0xf, true); }
__device__ inline float from_west(float v) { return __builtin_bit_cast(float, from_west(__builtin_bit_cast(int, v))); }
// Lane i receives lane i + 1's value (lane 63 reads 0).
__device__ inline int from_east(int v) { return __builtin_amdgcn_mov_dpp(v, 0x130 /* wave_shl:1 */, 0xf, 0xf, true); }
__device__ inline float from_east(float v) { return __builtin_bit_cast(float, from_east(__builtin_bit_cast(int, v))); }

struct WRow {
  float h, qx, qy, z;
  int ty;  // type slot, 0 for nodata or outside the domain
};

// Limited slopes of one cell along one axis (order 1 uses z only).
struct WSlope {
  float h, qx, qy, z;
};

// One edge's two shares (FP32 record words, as FluxTile's kRec = 4 record
// split by side): mass, momentum across the edge axis, momentum along it
// with the side's bed term folded in.
struct WPart {
  float m, across, along;
};

// The walk reads the domain descriptor through a KDesc that is laundered at
// every row, so descriptor fields are fetched by scalar loads where they are
// used instead of all being held in scalar registers across the walk (which
// spills them).  The single-domain kernel points at its own kernel
// arguments, the ensemble kernel at the descriptor array in its workspace.

// Plane access by 32-bit byte offset from the plane's base (scalar base +
// vector offset addressing); e is the plane element.  The planes are device
// memory: the accesses name the global address space, so they compile to
// global loads and stores whatever memory the descriptor holding the base
// came from.
#define HP_GLOBAL __attribute__((address_space(1)))
template <class V> __device__ inline V pl_ld(const void* base, uint32_t e) {
  return *(const HP_GLOBAL V*)((const HP_GLOBAL char*)base + e * (uint32_t)sizeof(V));
}
template <class V> __device__ inline void pl_st(void* base, uint32_t e, V v) {
  *(HP_GLOBAL V*)((HP_GLOBAL char*)base + e * (uint32_t)sizeof(V)) = v;
}

struct WalkEnv {
  float dt_dx;
  float dx, dt;
  float cfl_dx;
  bool t_zero;  // t == 0: every valid cell is listed
  int n_rates;  // station rates in LDS (the table's stations at this t)
};

template <class DP> __device__ inline BcValues<float> walk_bc_values(DP a, int slot, float t) {
  BcValues<float> b;
  int id;
  decode_code(a->slot_code[slot], b.type, id);
  const int rd = a->given_depth ? series_row<float>((const float*)a->given_depth, a->gd_rows, a->gd_w, t) : 0;
  const int rw = a->given_wl ? series_row<float>((const float*)a->given_wl, a->gwl_rows, a->gwl_w, t) : 0;
  const int rq = a->given_q ? series_row<float>((const float*)a->given_q, a->gq_rows, a->gq_w, t) : 0;
  b.gd = series_at<float>((const float*)a->given_depth, a->gd_rows, a->gd_w, rd, id + 1);
  b.gwl = series_at<float>((const float*)a->given_wl, a->gwl_rows, a->gwl_w, rw, id + 1);
  b.gqx = series_at<float>((const float*)a->given_q, a->gq_rows, a->gq_w, rq, 1 + 2 * id);
  b.gqy = series_at<float>((const float*)a->given_q, a->gq_rows, a->gq_w, rq, 2 + 2 * id);
  return b;
}

template <class DP> __device__ inline WRow walk_load(DP a, int gx, int gy, bool col_in) {
  WRow w{0.0f, 0.0f, 0.0f, 0.0f, 0};
  if (col_in && gy >= 0 && gy < a->ny) {
    const uint32_t e = (uint32_t)gy * (uint32_t)a->pitch + (uint32_t)gx;
    w.ty = pl_ld<uint8_t>(a->type_slot, e);
    w.h = pl_ld<float>(a->h, e);
    w.qx = pl_ld<float>(a->qx, e);
    w.qy = pl_ld<float>(a->qy, e);
    w.z = pl_ld<float>(a->z, e);
  }
  return w;
}

// minmod(c - l, r - c) with a nodata neighbour replaced by the cell itself.
template <int ORDER>
__device__ inline WSlope walk_slope(const WRow& c, const WRow& l, const WRow& r) {
  using P = Prec<float>;
  WSlope s;
  const bool lv = l.ty != 0, rv = r.ty != 0;
  s.z = P::limit(c.z - (lv ? l.z : c.z), (rv ? r.z : c.z) - c.z);
  if (ORDER == 2) {
    s.h = P::limit(c.h - (lv ? l.h : c.h), (rv ? r.h : c.h) - c.h);
    s.qx = P::limit(c.qx - (lv ? l.qx : c.qx), (rv ? r.qx : c.qx) - c.qx);
    s.qy = P::limit(c.qy - (lv ? l.qy : c.qy), (rv ? r.qy : c.qy) - c.qy);
  } else {
    s.h = s.qx = s.qy = 0.0f;
  }
  return s;
}

// Interior edge between valid cells A and B (B east of A for D = 0, south
// for D = 3).  eL / eR: A's and B's faces toward the edge in the edge frame,
// with z holding the half bed slope (A's face bed is z_A + eL.z, B's
// z_B - eR.z; order 1 applies each only where the other side is wet).
// Returns A's and B's shares.
template <int ORDER, int D>
__device__ inline void walk_edge(const WalkEnv& env, const WRow& A, const WRow& B, FrameCell<float> eL,
                                 FrameCell<float> eR, WPart& pa, WPart& pb) {
  const float g = float(kG);
  pa = WPart{0.0f, 0.0f, 0.0f};
  pb = pa;
  if (lt_hs(A.h) && lt_hs(B.h)) return;
  const float zB = B.z - A.z;  // B's bed on A's datum (A's is 0)
  if (ORDER == 1) {
    // bed faces only where the other side is wet (eL.z / eR.z hold the
    // half slopes on entry)
    eL.z = gt_hs(B.h) ? eL.z : 0.0f;
    eR.z = gt_hs(A.h) ? zB - eR.z : zB;
  } else {
    eR.z = zB - eR.z;
  }
  float fh, fn, ft;
  const float etaC = A.h, etaR = B.h + zB;
  hllc_edge<float, ORDER>(eL, eR, etaC, etaR, fh, fn, ft);
  // raw fluxes: the cell scales its summed increments by dt / dx once
  const float ofh = fh, ofn = fn, oft = ft;
  const float bed_c = 0.5f * g * (eL.h + A.h) * eL.z;
  const float bed_r = 0.5f * g * (eR.h + B.h) * (eR.z - zB);
  constexpr float nA = D == 0 ? 1.0f : -1.0f;
  const float gx = D == 0 ? ofn : oft, gy = D == 0 ? oft : -ofn;
  const float axis = D == 0 ? gx : gy;
  const float across = D == 0 ? gy : gx;
  pa = WPart{ofh, across, axis + bed_c * nA};
  pb = WPart{ofh, across, axis - bed_r * (-nA)};
}

// Boundary edge: exactly one of A, B is valid; solved from the valid side
// with its outward normal (FluxTile's boundary branch).
template <int ORDER, int D, class DP>
__device__ inline void walk_boundary_edge(const WalkEnv& env, DP a, const WRow& A, const WRow& B, WPart& pa,
                                          WPart& pb) {
  pa = WPart{0.0f, 0.0f, 0.0f};
  pb = pa;
  const bool from_a = A.ty != 0;
  const WRow& C = from_a ? A : B;
  const int d = from_a ? D : (D + 2) & 3;
  FrameCell<float> c;
  c.h = C.h;
  c.qn = rot_n(C.qx, C.qy, d);
  c.qt = rot_t(C.qx, C.qy, d);
  c.z = 0.0f;  // the valid cell's own datum
  const BcValues<float> bv = walk_bc_values(a, C.ty, float(a->ctrl->t));
  const float gwl = bv.gwl - C.z;
  const FrameCell<float> gh = bc_ghost<float>(bv.type, c, d, bv.gd, gwl, bv.gqx, bv.gqy);
  const float etaG = ORDER == 1 ? gh.h + gh.z : ghost_wl<float>(bv.type, c.h + c.z, gh, gwl);
  // unscaled (dt_dx = 1): the cell scales its sum
  const EdgeOut<float> o = solve_edge<float, ORDER>(c, c, gh, gh, true, c.h + c.z, etaG, env.dx, env.dt, 1.0f);
  if (!o.active) return;
  float gx, gy;
  to_global(o.fn, o.ft, d, gx, gy);
  constexpr bool axis_x = D == 0;
  const float bed = o.bed_c * float(axis_x ? nrm_x(d) : nrm_y(d));
  // the record from A's view (store_record): mass, momentum across the
  // axis, along the axis + A's bed / - B's bed
  const float fh = from_a ? o.fh : -o.fh;
  const float rx = from_a ? gx : -gx, ry = from_a ? gy : -gy;
  const float axis = axis_x ? rx : ry, across = axis_x ? ry : rx;
  if (from_a) pa = WPart{fh, across, axis + bed};
  else pb = WPart{fh, across, axis - bed};
}

// Edge between A and B of any validity; `need`: some owned cell reads the
// edge (the halo lanes' boundary edges are skipped: their shares are unused).
template <int ORDER, int D, class DP>
__device__ inline void walk_any_edge(const WalkEnv& env, DP a, const WRow& A, const WRow& B, const FrameCell<float>& eL,
                                     const FrameCell<float>& eR, bool need, WPart& pa, WPart& pb) {
  if (A.ty != 0 && B.ty != 0) {
    walk_edge<ORDER, D>(env, A, B, eL, eR, pa, pb);
  } else if (need && (A.ty != 0 || B.ty != 0)) {
    walk_boundary_edge<ORDER, D>(env, a, A, B, pa, pb);
  } else {
    pa = WPart{0.0f, 0.0f, 0.0f};
    pb = pa;
  }
}

// The faces of an edge frame from a cell, its slopes and the side (+1: the
// face toward increasing index along the axis, -1 the other).  D = 0: qn =
// qx, qt = qy; D = 3 (south): qn = -qy, qt = qx.  z carries the half slope.
template <int ORDER, int D>
__device__ inline FrameCell<float> walk_face(const WRow& c, const WSlope& s, float side) {
  FrameCell<float> f;
  if (ORDER == 2) {
    f.h = fmaf(0.5f * side, s.h, c.h);
    const float fqx = fmaf(0.5f * side, s.qx, c.qx), fqy = fmaf(0.5f * side, s.qy, c.qy);
    const bool dry = lt_hs(c.h);
    f.qn = dry ? 0.0f : (D == 0 ? fqx : -fqy);
    f.qt = dry ? 0.0f : (D == 0 ? fqy : fqx);
  } else {
    f.h = c.h;
    f.qn = D == 0 ? c.qx : -c.qy;
    f.qt = D == 0 ? c.qy : c.qx;
  }
  f.z = 0.5f * s.z;  // the half slope; walk_edge turns it into the face bed
  return f;
}

// Per-cell inputs of the update for the lane's cell of row gy.
template <class DP> __device__ inline CellIn<float> walk_cell_in(DP a, int gx, int gy, bool own) {
  CellIn<float> in{0.0f, 0, 0.0f, 0.0f, 0.0f, 0, false};
  if (own && gy >= 0 && gy < a->ny) {
    const uint32_t e = (uint32_t)gy * (uint32_t)a->pitch + (uint32_t)gx;
    const int cls = pl_ld<uint8_t>(a->landuse, e);
    in.on = cls != 255;
    in.cls = cls;
    if (in.on) in.n = ((const float*)a->manning)[cls < a->n_classes ? cls : a->n_classes - 1];
    if (a->rain_mode == HP_RAIN_STATION) in.st = a->station ? (int)pl_ld<int16_t>(a->station, e) : 0;
    else if (a->rain_mode == HP_RAIN_FIELD) in.rain = pl_ld<float>(a->rain_field, e) * float(a->rain_scale);
    in.hmax = pl_ld<float>(a->h_max, e);
    if (a->infil_ks) in.F = pl_ld<float>(a->infil_F, e);
  }
  return in;
}

// Walk state carried from one row to the next.
struct WalkCarry {
  WSlope sy;          // y slopes of the row being finished
  WPart north;        // its share of the edge with the row above
  bool wet_up;        // the cell above is valid, wet and not compact cell 0 (listing rule)
  float mn;           // CFL minimum so far
  double* sums;       // this lane's FP64 volume / flux / rain sums (LDS; with mass sums only)
};

// Finish row gy: `cur` is row gy, `nxt` gy + 1, `nw` gy + 2 (all loaded);
// `pre` receives row gy + 3.
template <int ORDER, bool FUSED, class DP>
__device__ inline void walk_row(const WalkEnv& env, DP a, const WRow& cur, const WRow& nxt, const WRow& nw, WRow& pre,
                                WalkCarry& k, int gx, int gy, bool own, bool need_x, bool col_in, const float* s_rate,
                                bool save_flux, bool time_control) {
  a = launder(a);
  pre = walk_load(a, gx, gy + 3, col_in);
  // this row's update inputs (their latency hides behind the edge work)
  CellIn<float> in_row{};
  if (FUSED) in_row = walk_cell_in(a, gx, gy, own);

  // y slopes of the next row; the edge between this row and the next
  const WSlope sy_n = walk_slope<ORDER>(nxt, cur, nw);
  WPart ps, pn_next;
  walk_any_edge<ORDER, 3>(env, a, cur, nxt, walk_face<ORDER, 3>(cur, k.sy, 1.0f), walk_face<ORDER, 3>(nxt, sy_n, -1.0f),
                          own, ps, pn_next);

  // x neighbours by lane shifts, x slopes, the west edge of every lane
  WRow w, e;
  w.h = from_west(cur.h); w.qx = from_west(cur.qx); w.qy = from_west(cur.qy); w.z = from_west(cur.z);
  w.ty = from_west(cur.ty);
  e.h = from_east(cur.h); e.qx = from_east(cur.qx); e.qy = from_east(cur.qy); e.z = from_east(cur.z);
  e.ty = from_east(cur.ty);
  const WSlope sx = walk_slope<ORDER>(cur, w, e);
  const FrameCell<float> east_face = walk_face<ORDER, 0>(cur, sx, 1.0f);
  FrameCell<float> eL;  // the west neighbour's east face
  eL.h = from_west(east_face.h);
  eL.qn = from_west(east_face.qn);
  eL.qt = from_west(east_face.qt);
  eL.z = from_west(east_face.z);
  WPart pw_a, pw;  // the west edge's shares: the west neighbour's, this cell's
  walk_any_edge<ORDER, 0>(env, a, w, cur, eL, walk_face<ORDER, 0>(cur, sx, -1.0f), need_x, pw_a, pw);
  WPart pe;  // this cell's share of its east edge, solved by the east lane
  pe.m = from_east(pw_a.m);
  pe.across = from_east(pw_a.across);
  pe.along = from_east(pw_a.along);

  // listing rule (fluxMask): t == 0, wet, or a wet valid neighbour E, N, W,
  // S that is not compact cell 0 (fluxMaskGenerator_Kernel.cu:36)
  bool listed = false;
  if (cur.ty != 0) {
    listed = env.t_zero | gt_hs(cur.h);
    if (!listed) {
      const int c0x = a->cell0_x, c0y = a->cell0_y;
      const bool w_e = (e.ty != 0) & gt_hs(e.h) & !(gx + 1 == c0x && gy == c0y);
      const bool w_n = k.wet_up;
      const bool w_w = (w.ty != 0) & gt_hs(w.h) & !(gx - 1 == c0x && gy == c0y);
      const bool w_s = (nxt.ty != 0) & gt_hs(nxt.h) & !(gx == c0x && gy + 1 == c0y);
      listed = w_e | w_n | w_w | w_s;
    }
  }
  float ah = 0.0f, aqx = 0.0f, aqy = 0.0f;
  if (listed) {  // the reference's order: E, N, W, S
    ah -= pe.m; aqx -= pe.along; aqy -= pe.across;
    ah += k.north.m; aqx += k.north.across; aqy += k.north.along;
    ah += pw.m; aqx += pw.along; aqy += pw.across;
    ah -= ps.m; aqx -= ps.across; aqy -= ps.along;
    ah *= env.dt_dx;
    aqx *= env.dt_dx;
    aqy *= env.dt_dx;
  }

  if (own && gy < a->ny) {
    const uint32_t gi = (uint32_t)gy * (uint32_t)a->pitch + (uint32_t)gx;
    if (!FUSED || save_flux) {
      pl_st<float>(a->dh, gi, ah);
      pl_st<float>(a->dqx, gi, aqx);
      pl_st<float>(a->dqy, gi, aqy);
    }
    if (FUSED) {
      CellIn<float> in = in_row;
      if (a->rain_mode == HP_RAIN_STATION) in.rain = (in.st >= 0 && in.st < env.n_rates) ? s_rate[in.st] : 0.0f;
      float hn = cur.h, qxn = cur.qx, qyn = cur.qy, hm = in.hmax;
      float dF = 0.0f;
      if (a->infil_ks && listed) {  // infiltration over the flux list (SWE_CUDA.py:439-443)
        dF = green_ampt<float>(hn, in.F, ((const float*)a->infil_ks)[in.cls], ((const float*)a->infil_phi)[in.cls],
                               ((const float*)a->infil_dtheta)[in.cls], env.dt);
        pl_st<float>(a->infil_F, gi, in.F + dF);
      }
      float cand, vh, fh, rh;
      update_state<float>(in, hn, qxn, qyn, ah, aqx, aqy, dF, env.dt, env.cfl_dx, hm, cand, vh, fh, rh, time_control);
      if (time_control) {
        k.mn = fmin(k.mn, cand);
        if (a->mass_sums) {
          k.sums[0] += (double)vh;
          k.sums[kWalkLanes] += (double)fh;
          k.sums[2 * kWalkLanes] += (double)rh;
        }
      }
      pl_st<float>(a->h_out, gi, hn);
      pl_st<float>(a->qx_out, gi, qxn);
      pl_st<float>(a->qy_out, gi, qyn);
      if (hm != in.hmax) pl_st<float>(a->h_max, gi, hm);
    }
  }
  k.sy = sy_n;
  k.north = pn_next;
  k.wet_up = (cur.ty != 0) & gt_hs(cur.h) & !(gx == a->cell0_x && gy == a->cell0_y);
}

// One unit: strip bx (columns [60 bx, 60 bx + 60)), rows [R by, R by + R).
// FUSED false: write the flux accumulators (flux pass).  FUSED true: finish
// the step into h_out / qx_out / qy_out with one partial record per unit
// (none without time control); save_flux also writes the flux; force_copy
// makes a stopped domain copy its state through even when marked in step.
// (inlined into each kernel: the launder needs the descriptor pointer known
// uniform, which a call shared by two kernels would not let the compiler see)
template <int ORDER, int R, bool FUSED, class DP>
__device__ __forceinline__ void walk_unit(DP a, int bx, int by, const Partials& part, int64_t pidx, bool save_flux,
                                 bool time_control, bool force_copy) {
  static_assert(R >= 2 && R % 2 == 0, "rows per unit: even (the four-row window's first two rows end a unit)");
  __shared__ float s_rate[FUSED ? kMaxStations : 1];
  __shared__ double s_sums[FUSED ? 3 * kWalkLanes : 1];
  const int lane = (int)threadIdx.x;
  const int gx = bx * kWalkOwn - 2 + lane;
  const bool col_in = gx >= 0 && gx < a->nx;
  const bool own = lane >= 2 && lane < 2 + kWalkOwn && gx < a->nx;
  // lanes whose west edge an owned cell reads: the owned ones and the lane
  // east of them (the last owned cell's east edge)
  const bool need_x = lane >= 2 && lane <= 2 + kWalkOwn && gx <= a->nx;
  const int y0 = by * R;
  if (stopped(a->ctrl)) {
    if (FUSED) {
      if (lane == 0) part.put(pidx, INFINITY, 0.0, 0.0, 0.0, a->mass_sums != 0);
      if (a->ctrl->stopped_synced != 0.0 && !force_copy) return;
      if (own) {
        for (int r = 0; r < R; ++r) {
          const int gy = y0 + r;
          if (gy >= a->ny) break;
          const uint32_t gi = (uint32_t)gy * (uint32_t)a->pitch + (uint32_t)gx;
          pl_st<float>(a->h_out, gi, pl_ld<float>(a->h, gi));
          pl_st<float>(a->qx_out, gi, pl_ld<float>(a->qx, gi));
          pl_st<float>(a->qy_out, gi, pl_ld<float>(a->qy, gi));
        }
      }
    }
    return;
  }
  WalkEnv env;
  env.dt = float(a->ctrl->dt);
  env.dx = float(a->dx);
  env.dt_dx = env.dt / env.dx;
  env.cfl_dx = float(a->cfl) * float(a->dx);
  env.t_zero = float(a->ctrl->t) == 0.0f;
  env.n_rates = 0;
  if (FUSED && a->rain_mode == HP_RAIN_STATION) {
    // the table's station rates at this step's t, in LDS (as many as the
    // table has stations, at most 1024, which the host enforces: a cell
    // naming another station, or any cell once the table has ended, gets
    // no rain, as station_rate gives)
    const RainRows rr = rain_rows_at(a, a->ctrl->t);
    env.n_rates = rr.zero ? 0 : min(rr.n_st, kMaxStations);
    for (int st = lane; st < env.n_rates; st += kWalkLanes) s_rate[st] = station_rate<float>(rr, st, a->rain_scale);
    __syncthreads();
  }

  // prologue: rows y0 - 2 .. y0 + 1, the edge between y0 - 1 and y0
  WRow A = walk_load(a, gx, y0 - 2, col_in), B = walk_load(a, gx, y0 - 1, col_in);
  WRow C = walk_load(a, gx, y0, col_in), D = walk_load(a, gx, y0 + 1, col_in);
  WalkCarry k;
  {
    const WSlope sy_b = walk_slope<ORDER>(B, A, C);
    k.sy = walk_slope<ORDER>(C, B, D);
    WPart unused;
    walk_any_edge<ORDER, 3>(env, a, B, C, walk_face<ORDER, 3>(B, sy_b, 1.0f), walk_face<ORDER, 3>(C, k.sy, -1.0f), own,
                            unused, k.north);
  }
  k.wet_up = (B.ty != 0) & gt_hs(B.h) & !(gx == a->cell0_x && y0 - 1 == a->cell0_y);
  k.mn = INFINITY;
  // the FP64 sums live in LDS, one column per lane, so they hold no
  // registers across the walk
  k.sums = s_sums + lane;
  if (FUSED && a->mass_sums) k.sums[0] = k.sums[kWalkLanes] = k.sums[2 * kWalkLanes] = 0.0;
  A = walk_load(a, gx, y0 + 2, col_in);
  for (int r = 0; r < R; r += 4) {
    walk_row<ORDER, FUSED>(env, a, C, D, A, B, k, gx, y0 + r, own, need_x, col_in, s_rate, save_flux,
                           time_control);
    walk_row<ORDER, FUSED>(env, a, D, A, B, C, k, gx, y0 + r + 1, own, need_x, col_in, s_rate, save_flux,
                           time_control);
    if (r + 2 == R) break;  // R = 4k + 2: the unit ends halfway through the window
    walk_row<ORDER, FUSED>(env, a, A, B, C, D, k, gx, y0 + r + 2, own, need_x, col_in, s_rate, save_flux,
                           time_control);
    walk_row<ORDER, FUSED>(env, a, B, C, D, A, k, gx, y0 + r + 3, own, need_x, col_in, s_rate, save_flux,
                           time_control);
  }
  if (FUSED && time_control) {
    const bool sums = a->mass_sums != 0;
    const float mn = wave_min(k.mn);
    double vol = 0.0, fs = 0.0, rs = 0.0;
    if (sums) {
      vol = k.sums[0];
      fs = k.sums[kWalkLanes];
      rs = k.sums[2 * kWalkLanes];
    }
    if (sums) {
      vol = wave_sum(vol);
      fs = wave_sum(fs);
      rs = wave_sum(rs);
    }
    if (lane == 0) part.put(pidx, (double)mn, vol, fs, rs, sums);
  }
}

// The descriptor is the kernel's first argument: read it in place.
// flags: 1 also save the step's flux in dh, dqx, dqy; 2 no time-control
// outputs (a frozen schedule); 4 a stopped domain copies its state through
// even when marked in step.
template <int ORDER, int R, bool FUSED>
__global__ __launch_bounds__(kWalkLanes) void k_walk(hp_grid_desc a, int nstrips, int flags) {
  const KDesc pa = kernarg_desc();
  const int unit = xcd_band(blockIdx.x, gridDim.x);
  walk_unit<ORDER, R, FUSED>(pa, unit % nstrips, unit / nstrips, Partials(a.work, gridDim.x), unit, (flags & 1) != 0,
                             (flags & 2) == 0, (flags & 4) != 0);
}

// Every unit of every domain of an ensemble; map[i] = (domain, unit), domain
// by domain, so a domain's partial records are one contiguous range.
template <int ORDER, int R, bool FUSED>
__global__ __launch_bounds__(kWalkLanes) void k_walk_multi(const hp_grid_desc* __restrict__ doms,
                                                           const int2* __restrict__ map, void* parts, int n_total) {
  const int slot = xcd_band(blockIdx.x, gridDim.x);
  // (the map entry is the same for the whole wave)
  const int dom = __builtin_amdgcn_readfirstlane(map[slot].x), unit = __builtin_amdgcn_readfirstlane(map[slot].y);
  // the descriptor array is read-only for the launch: read it as constant
  // memory, by scalar loads, as the single-domain kernel reads its argument
  const KDesc a = (KDesc)(doms + dom);
  const int nstrips = (a->nx + kWalkOwn - 1) / kWalkOwn;
  walk_unit<ORDER, R, FUSED>(a, unit % nstrips, unit / nstrips, Partials(parts, n_total), slot, false, true, false);
}

}  // namespace hp
