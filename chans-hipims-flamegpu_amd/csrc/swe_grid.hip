// Structured-grid HiPIMS step for gfx950.
//
// Layout: every field is a plane of ny rows x pitch elements (pitch is a
// multiple of 64, so each row starts 256-byte aligned in float).  Raster
// (row r, col c) of the reference's input maps to plane element r*pitch + c,
// so the raster's own nodata ring doubles as the stencil apron.
//
// Per step:
//   flux    tile of 64 x TY cells + 2-cell cross halo staged in LDS; every
//           edge touching the tile is solved once (swe_physics.hpp), by the
//           thread that owns its east/south cell, and both cells take their
//           share (the owner's in registers, the other through LDS); listed
//           cells (the reference's fluxMask rule) write dh/dqx/dqy.
//   single-pass step: the flux tile finishes its own cells (the update
//           below) into a second set of state planes.
//   update  a contiguous chunk of plane elements per block: rain, implicit
//           friction + update, h_max, the CFL minimum (wave64 shuffle + LDS),
//           FP64 volume / flux / rain sums; one partial record per block.
//   finalize one block per domain folds its partials and advances dt, t and
//           the export clamp in the domain's control block (SWE_CUDA.py:818-831).
// No host synchronisation anywhere in the step.
//
// Each pass is a device function over (domain, tile or chunk).  Single-domain
// launches take the domain descriptor as a kernel argument; ensemble launches
// take a device array of descriptors plus a (domain, local index) map per
// block, so one launch advances every domain of an ensemble by its own dt.
// A domain whose control block has t_stop > 0 and t >= t_stop is skipped
// (the single-pass step copies its state through to the other set).

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <cstring>
#include <mutex>
#include <utility>
#include <string>
#include <vector>

#include "../../include/hipims_abi.h"
#include "hp_common.hpp"
#include "swe_physics.hpp"
#include "swe_source.hpp"

namespace hp {

constexpr int kTX = 64;
constexpr int kThreads = 256;
constexpr int kChunk = 2048;  // plane elements per update block
constexpr int kMaxStations = 1024;
constexpr int kRateCache = 64;  // station rates kept in LDS per block; further stations are evaluated per cell

__device__ inline bool stopped(const hp_ctrl* c) { return c->t_stop > 0.0 && c->t >= c->t_stop; }

// --- boundary series ------------------------------------------------------
template <class T>
__device__ inline int series_row(const T* tab, int rows, int w, T t) {
  int r = 0;
  while (r < rows - 1 && t > tab[r * w]) r++;
  return r;
}

template <class T>
__device__ inline T series_at(const T* tab, int rows, int w, int row, int col) {
  if (tab == nullptr || col < 0 || col >= w || row < 0 || row >= rows) return T(0);
  return tab[row * w + col];
}

struct SeriesRows {
  int d, w, q;
};

template <class T> struct BcValues {
  int type;
  T gd, gwl, gqx, gqy;
};

template <class T>
__device__ inline BcValues<T> bc_values(const hp_grid_desc& a, int slot, const SeriesRows& rows) {
  BcValues<T> b;
  int id;
  decode_code(a.slot_code[slot], b.type, id);
  b.gd = series_at<T>((const T*)a.given_depth, a.gd_rows, a.gd_w, rows.d, id + 1);
  b.gwl = series_at<T>((const T*)a.given_wl, a.gwl_rows, a.gwl_w, rows.w, id + 1);
  b.gqx = series_at<T>((const T*)a.given_q, a.gq_rows, a.gq_w, rows.q, 1 + 2 * id);
  b.gqy = series_at<T>((const T*)a.given_q, a.gq_rows, a.gq_w, rows.q, 2 + 2 * id);
  return b;
}

// Global (qx, qy) of an edge-frame momentum flux (fn, ft) for direction d.
template <class T> __device__ inline void to_global(T fn, T ft, int d, T& gx, T& gy) {
  switch (d) {
    case 0: gx = fn; gy = ft; break;
    case 1: gx = -ft; gy = fn; break;
    case 2: gx = -fn; gy = -ft; break;
    default: gx = ft; gy = -fn; break;
  }
}

// --- reductions ------------------------------------------------------------
// Time control of one domain from its step's folded CFL minimum and FP64
// sums (one thread): dt = min of the candidates (kept when no cell was
// listed, SWE_CUDA.py:818-822), capped at 10x the previous dt, clamped onto
// the next export time, then t advanced by it (SWE_CUDA.py:824-831).
// c: the control block as it was read at the start of the pass (nothing
// else writes it meanwhile).
__device__ inline void advance_ctrl(hp_ctrl* ctrl, const hp_ctrl& c, double mn, double vol, double fs, double rs,
                                    double dx, bool sums) {
  const double area = dx * dx;
  const double t = c.t;
  const double dt_old = c.dt;
  double dt_new = isinf(mn) ? dt_old : mn;
  const double upper = dt_old * 10.0;
  dt_new = (upper < dt_new) ? upper : dt_new;
  const double next_export = (c.export_n + 1.0) * c.export_T;
  if (dt_new + t >= next_export) {
    dt_new = next_export - t;
    ctrl->export_n = c.export_n + 1.0;
    ctrl->export_pending = c.export_pending + 1.0;
  }
  ctrl->dt = dt_new;
  ctrl->t = t + dt_new;
  if (sums) {
    ctrl->volume = vol * area;
    ctrl->flux_volume = c.flux_volume + fs * area;
    ctrl->rain_volume = c.rain_volume + rs * area;
  }
  ctrl->steps = c.steps + 1.0;
  ctrl->stopped_synced = 0.0;
}

// Per-block results of a pass that ends a step, in the workspace: the CFL
// minimum of every block in one array (all the finalize pass reads when the
// mass sums are off), then, when they are on, one FP64 record per block.
struct Sums {
  double vol, flux, rain, pad;
};

__host__ __device__ constexpr int64_t align256(int64_t x) { return (x + 255) & ~int64_t(255); }
__host__ __device__ constexpr int64_t partials_bytes(int64_t n) { return align256(n * 8) + n * (int64_t)sizeof(Sums); }

struct Partials {
  double* mins;
  Sums* sums;
  __host__ __device__ Partials(void* work, int64_t n)
      : mins((double*)work), sums((Sums*)((char*)work + align256(n * 8))) {}
  __device__ void put(int64_t i, double mn, double vol, double flux, double rain, bool with_sums) const {
    mins[i] = mn;
    if (with_sums) sums[i] = Sums{vol, flux, rain, 0.0};
  }
};

template <class T> __device__ inline T wave_min(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// LDS tile of one flux block.
template <class T, int TY> struct FluxTile {
  static constexpr int LW = kTX + 4, LH = TY + 4;
  T h[LH][LW], qx[LH][LW], qy[LH][LW], z[LH][LW];
  uint8_t ty[LH][LW];
  // Edge record (solve_tile_edge), A = west / north cell, B = east / south
  // cell, fluxes from A's view (subtract at A, add at B):
  //   double (kRec 5): 0 mass, 1 qx, 2 qy, 3 bed term of A, 4 bed term of B
  //     (the reference's accumulation order: flux terms, then the bed term);
  //   float (kRec 4): 0 mass, 1 momentum across the edge axis, 2 momentum
  //     along the axis + A's bed term, 3 the same - B's bed term.
  static constexpr int kRec = sizeof(T) == 4 ? 4 : 5;
  // A record splits into the part its A cell reads and the part its B cell
  // reads (kSide words each: mass, the two momentum terms and that side's
  // bed term, folded as above).  The thread that owns a cell solves the
  // cell's west x-edge and north y-edge, where the cell is B: it keeps the B
  // part in registers and leaves the A part here, under the A cell - the
  // west neighbour's east-edge part (a[0]) and the north neighbour's
  // south-edge part (a[1]).
  static constexpr int kSide = kRec - 1;
  T a[2][kSide][TY][kTX];
};

// The part of record e that cell A (side 0) or cell B (side 1) reads.
template <class T, int kRec>
__device__ inline T side_word(const T* e, int side, int k) {
  if (kRec == 5) return k < 3 ? e[k] : e[3 + side];
  return k < 2 ? e[k] : e[2 + side];
}

template <class T, int TY>
__device__ inline FrameCell<T> frame_cell(const FluxTile<T, TY>& s, int r, int c, int d, T zref) {
  FrameCell<T> f;
  const T qx = s.qx[r][c], qy = s.qy[r][c];
  f.h = s.h[r][c];
  f.qn = rot_n(qx, qy, d);
  f.qt = rot_t(qx, qy, d);
  f.z = s.z[r][c] - zref;
  return f;
}

// Frame view for a compile-time direction D (0 E for x-edges, 3 S for
// y-edges): the rotation is a fixed renaming, no selects.
template <int D, class T, int TY>
__device__ inline FrameCell<T> frame_cell_d(const FluxTile<T, TY>& s, int r, int c, T zref) {
  FrameCell<T> f;
  const T qx = s.qx[r][c], qy = s.qy[r][c];
  f.h = s.h[r][c];
  f.qn = D == 0 ? qx : -qy;
  f.qt = D == 0 ? qy : qx;
  f.z = s.z[r][c] - zref;
  return f;
}

// Write one edge record (layouts in FluxTile).  gx, gy: global momentum flux
// from A's view; bedA / bedB: the two sides' bed terms along the edge axis.
template <class T, int D, int kRec>
__device__ inline void store_record(T* out, T fh, T gx, T gy, T bedA, T bedB) {
  out[0] = fh;
  if (kRec == 5) {
    out[1] = gx;
    out[2] = gy;
    out[3] = bedA;
    out[4] = bedB;
  } else {
    const T axis = D == 0 ? gx : gy;
    out[1] = D == 0 ? gy : gx;
    out[2] = axis + bedA;
    out[3] = axis - bedB;
  }
}

// Solve the edge between LDS cells A and B, where B lies in direction D from
// A (D = 0: x-edge, B east of A; D = 3: y-edge, B south of A).  A is
// (ra, ca); L is A's neighbour opposite to B, RR is B's neighbour beyond B.
// Writes the edge record out[0 .. kRec-1].
template <class T, int ORDER, int TY, int D>
__device__ inline void solve_tile_edge(const FluxTile<T, TY>& s, const hp_grid_desc& a, const SeriesRows& rows,
                                       int ra, int ca, T dx, T dt, T dt_dx, T* out) {
  constexpr int di = D == 0 ? 0 : 1, dj = D == 0 ? 1 : 0;  // B - A in (row, col)
  const int rb = ra + di, cb = ca + dj;
  const int rl = ra - di, cl = ca - dj;
  const int rr2 = rb + di, cr2 = cb + dj;
  constexpr int kRec = FluxTile<T, TY>::kRec;
#pragma unroll
  for (int k = 0; k < kRec; ++k) out[k] = T(0);
  const int tA = s.ty[ra][ca], tB = s.ty[rb][cb];
  if (tA == 0 && tB == 0) return;
  constexpr T nA = T(D == 0 ? 1 : -1);  // A's outward normal component along the edge axis
  if (tA != 0 && tB != 0) {
    const T zref = Prec<T>::local_datum ? s.z[ra][ca] : T(0);
    FrameCell<T> c = frame_cell_d<D>(s, ra, ca, zref);
    if (Prec<T>::local_datum) c.z = T(0);  // its own datum: z - z is exactly 0, known here at compile time
    const FrameCell<T> r = frame_cell_d<D>(s, rb, cb, zref);
    const FrameCell<T> lc = frame_cell_d<D>(s, rl, cl, zref);
    const FrameCell<T> rc = frame_cell_d<D>(s, rr2, cr2, zref);
    const FrameCell<T> l = s.ty[rl][cl] ? lc : c;
    const FrameCell<T> rr = s.ty[rr2][cr2] ? rc : r;
    const EdgeOut<T> o = solve_edge<T, ORDER>(c, l, r, rr, false, c.h + c.z, r.h + r.z, dx, dt, dt_dx);
    if (!o.active) return;
    store_record<T, D, kRec>(out, o.fh, D == 0 ? o.fn : o.ft, D == 0 ? o.ft : -o.fn, o.bed_c * nA, o.bed_r * (-nA));
  } else {
    // boundary edge, solved from the valid side with its outward normal
    const bool from_a = tA != 0;
    const int rc = from_a ? ra : rb, cc = from_a ? ca : cb, d = from_a ? D : (D + 2) & 3;
    const bool axis_x = D == 0;
    const T zref = Prec<T>::local_datum ? s.z[rc][cc] : T(0);
    const FrameCell<T> c = frame_cell(s, rc, cc, d, zref);
    const BcValues<T> bv = bc_values<T>(a, from_a ? tA : tB, rows);
    const FrameCell<T> g = bc_ghost<T>(bv.type, c, d, bv.gd, bv.gwl - zref, bv.gqx, bv.gqy);
    const T etaG = ORDER == 1 ? g.h + g.z : ghost_wl<T>(bv.type, c.h + c.z, g, bv.gwl - zref);
    const EdgeOut<T> o = solve_edge<T, ORDER>(c, c, g, g, true, c.h + c.z, etaG, dx, dt, dt_dx);
    if (!o.active) return;
    T gx, gy;
    to_global(o.fn, o.ft, d, gx, gy);
    const T bed = o.bed_c * T(axis_x ? nrm_x(d) : nrm_y(d));
    if (from_a) store_record<T, D, kRec>(out, o.fh, gx, gy, bed, T(0));
    else store_record<T, D, kRec>(out, -o.fh, -gx, -gy, T(0), bed);
  }
}

// Fold a block's CFL minimum and FP64 sums into its partial slot.
template <class T>
__device__ inline void block_partial(T mn, double vol, double fsum, double rsum, const Partials& part, int64_t idx,
                                     bool sums) {
  __shared__ double s_red[4][kThreads / 64];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  mn = wave_min(mn);
  if (sums) {
    vol = wave_sum(vol);
    fsum = wave_sum(fsum);
    rsum = wave_sum(rsum);
  }
  if (lane == 0) {
    s_red[0][w] = (double)mn;
    s_red[1][w] = vol;
    s_red[2][w] = fsum;
    s_red[3][w] = rsum;
  }
  __syncthreads();
  if (tid == 0) {
    double m = s_red[0][0], v = s_red[1][0], f = s_red[2][0], r = s_red[3][0];
    for (int k = 1; k < kThreads / 64; ++k) {
      m = fmin(m, s_red[0][k]);
      v += s_red[1][k];
      f += s_red[2][k];
      r += s_red[3][k];
    }
    part.put(idx, m, v, f, r, sums);
  }
}

// Per-cell inputs of the update that do not depend on the fluxes; loaded
// before the flux work so their latency hides behind it.
template <class T> struct CellIn {
  T rain;   // rain rate (m/s) of a field; of a station once resolved
  int st;   // station index (station rain)
  T hmax;   // h_max before the step
  T n;      // Manning n of the cell's class
  T F;      // cumulative infiltrated depth (when infiltration is on)
  int cls;  // landuse class
  bool on;  // landuse class != 255 (the reference's update covers the cell)
};

// Rain + implicit friction + update of one cell (HP_UPDATE_STEP semantics,
// SWE_CUDA.py:780-831).  Returns the new state in (h, qx, qy) and the new
// h_max, plus the cell's contributions to the block's reductions: its CFL
// candidate (INFINITY if none), its depth for the volume sum (0 if the cell
// is not covered), and its flux and rain depth increments (0 if not
// updated).  Every output is assigned on every path, so the caller folds
// them in unconditionally and the accumulators stay in registers.
template <class T>
__device__ inline void update_state(const CellIn<T>& in, T& h, T& qx, T& qy, T dhf, T dqx, T dqy, T dF, T dt,
                                    T cfl_dx, T& hmax, T& cand, T& vol_h, T& flux_h, T& rain_h,
                                    bool time_control = true) {
  cand = T(INFINITY);
  vol_h = T(0);
  flux_h = T(0);
  rain_h = T(0);
  if (!in.on) return;
  const T rdt = in.rain * dt;
  const T dht = (dhf + rdt) - dF;  // flux, + rain, - infiltration (the reference's order)
  if (!(fabs(dht) > T(0) || h >= T(0))) {
    vol_h = h;
    return;
  }
  if (new_depth_dry(h, dht)) {
    qx = T(0);
    qy = T(0);
  } else {
    qx += dqx;
    qy += dqy;
    implicit_friction<T>(h, in.n, dt, qx, qy);
  }
  hmax = fmax(h, hmax);
  if (time_control) cand = cfl_candidate<T>(h, qx, qy, cfl_dx);
  vol_h = h;
  flux_h = dhf;
  rain_h = rdt;
}

// FUSED = false: write the flux accumulators dh/dqx/dqy (two-pass step).
// FUSED = true: finish the step in the same block - rain, implicit friction
// and update of the tile's own cells into h_out/qx_out/qy_out (the input
// planes stay untouched for the neighbouring tiles' halos), h_max, and one
// partial record (CFL minimum, FP64 sums) per tile.
// time_control false (single-pass step of a frozen schedule, whose time
// control never runs): no CFL candidates, sums or partial record.
template <class T, int ORDER, int TY, bool FUSED>
__device__ inline void flux_tile(const hp_grid_desc& a, int bx, int by, const Partials& part, int64_t pidx,
                                 bool save_flux = false, bool time_control = true, bool force_copy = false) {
  __shared__ FluxTile<T, TY> s;
  __shared__ T s_rate[FUSED ? kRateCache : 1];
  using Tile = FluxTile<T, TY>;
  constexpr int LH = Tile::LH;
  constexpr int CPT = TY * kTX / kThreads;  // own cells per thread
  static_assert(TY * kTX % kThreads == 0, "tile cells per thread");
  if (stopped(a.ctrl)) {
    if (FUSED) {
      if (threadIdx.x == 0) part.put(pidx, INFINITY, 0.0, 0.0, 0.0, a.mass_sums != 0);
      // the state passes through to the output planes as well, so that a
      // caller alternating the two sets of planes keeps the final state;
      // once the time control has marked it done, both sets hold it (unless
      // the caller asks for the copy: the first step of a call, after which
      // the host may have written the state)
      if (a.ctrl->stopped_synced != 0.0 && !force_copy) return;
      const int gx = bx * kTX + (int)(threadIdx.x & (kTX - 1));
#pragma unroll
      for (int k = 0; k < CPT; ++k) {
        const int gy = by * TY + (int)(threadIdx.x / kTX) + (kThreads / kTX) * k;
        if (gx < a.nx && gy < a.ny) {
          const int64_t gi = (int64_t)gy * a.pitch + gx;
          ((T*)a.h_out)[gi] = ((const T*)a.h)[gi];
          ((T*)a.qx_out)[gi] = ((const T*)a.qx)[gi];
          ((T*)a.qy_out)[gi] = ((const T*)a.qy)[gi];
        }
      }
    }
    return;
  }
  const int tid = threadIdx.x;
  const int lane = tid & (kTX - 1), wv = tid / kTX;
  const int x0 = bx * kTX, y0 = by * TY;
  const int nx = a.nx, ny = a.ny;
  const int64_t pitch = a.pitch;
  const T* __restrict__ h = (const T*)a.h;
  const T* __restrict__ qx = (const T*)a.qx;
  const T* __restrict__ qy = (const T*)a.qy;
  const T* __restrict__ z = (const T*)a.z;
  const uint8_t* __restrict__ ty = a.type_slot;

  // the thread's own cells: column x0 + lane, rows y0 + wv + 4k
  CellIn<T> cin[FUSED ? CPT : 1];
  RainRows rr{};
  if (FUSED) {
    rr = rain_rows(a, a.ctrl->t);
    if (a.rain_mode == HP_RAIN_STATION && tid < kRateCache) s_rate[tid] = station_rate<T>(rr, tid, a.rain_scale);
    const int gx = x0 + lane;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int gy = y0 + wv + 4 * k;
      cin[k] = CellIn<T>{T(0), 0, T(0), T(0), T(0), 0, false};
      if (gx < nx && gy < ny) {
        const int64_t gi = (int64_t)gy * pitch + gx;
        const int cls = a.landuse[gi];
        cin[k].on = cls != 255;
        cin[k].cls = cls;
        // the class's Manning value from the (cached) table: no LDS copy, so
        // that six blocks fit a CU; a class beyond the table (undefined in
        // the reference) reads its last entry, never past it
        if (cin[k].on) cin[k].n = ((const T*)a.manning)[cls < a.n_classes ? cls : a.n_classes - 1];
        if (a.rain_mode == HP_RAIN_STATION) cin[k].st = a.station ? (int)a.station[gi] : 0;
        else if (a.rain_mode == HP_RAIN_FIELD) cin[k].rain = T(a.rain_field[gi]) * T(a.rain_scale);
        cin[k].hmax = ((const T*)a.h_max)[gi];
        if (a.infil_ks) cin[k].F = ((const T*)a.infil_F)[gi];
      }
    }
  }

  // Stage the tile and a 2-cell halo: the 64 interior columns row by row
  // (one wave per row, coalesced), then the 4 halo columns.
  {
    const int gx = x0 + lane;
    for (int r = wv; r < LH; r += kThreads / kTX) {
      const int gy = y0 - 2 + r;
      const bool in = gx < nx && gy >= 0 && gy < ny;
      T vh = T(0), vqx = T(0), vqy = T(0), vz = T(0);
      uint8_t vt = 0;
      if (in) {
        const int64_t gi = (int64_t)gy * pitch + gx;
        vt = ty[gi]; vh = h[gi]; vqx = qx[gi]; vqy = qy[gi]; vz = z[gi];
      }
      s.ty[r][lane + 2] = vt; s.h[r][lane + 2] = vh; s.qx[r][lane + 2] = vqx; s.qy[r][lane + 2] = vqy;
      s.z[r][lane + 2] = vz;
    }
    static_assert(4 * LH <= kThreads, "halo staging");
    if (tid < 4 * LH) {
      const int r = tid >> 2, k = tid & 3;
      const int c = k < 2 ? k : kTX + k;  // local columns 0, 1, 66, 67
      const int gy = y0 - 2 + r, gxh = x0 - 2 + c;
      const bool in = gxh >= 0 && gxh < nx && gy >= 0 && gy < ny;
      T vh = T(0), vqx = T(0), vqy = T(0), vz = T(0);
      uint8_t vt = 0;
      if (in) {
        const int64_t gi = (int64_t)gy * pitch + gxh;
        vt = ty[gi]; vh = h[gi]; vqx = qx[gi]; vqy = qy[gi]; vz = z[gi];
      }
      s.ty[r][c] = vt; s.h[r][c] = vh; s.qx[r][c] = vqx; s.qy[r][c] = vqy; s.z[r][c] = vz;
    }
  }

  const T t = T(a.ctrl->t);
  const T dt = T(a.ctrl->dt);
  const T dx = T(a.dx);
  const T dt_dx = dt / dx;
  SeriesRows rows;
  rows.d = a.given_depth ? series_row<T>((const T*)a.given_depth, a.gd_rows, a.gd_w, t) : 0;
  rows.w = a.given_wl ? series_row<T>((const T*)a.given_wl, a.gwl_rows, a.gwl_w, t) : 0;
  rows.q = a.given_q ? series_row<T>((const T*)a.given_q, a.gq_rows, a.gq_w, t) : 0;
  __syncthreads();

  // Edges touching the tile: x-edges (row ly, between local columns lx-1 = A
  // and lx = B, lx = 0..64) and y-edges (column lx, between local rows ly-1 =
  // A and ly = B, ly = 0..TY).  The owner of cell (ly, lx) solves its west
  // and north edges, in the same passes for every wave (one direction per
  // pass); the last x-edge column runs in wave 0 and the last y-edge row in
  // wave 1.  B parts stay in registers, A parts go to LDS.
  constexpr int kRec = Tile::kRec, kSide = Tile::kSide;
  T bw[CPT][kSide], bn[CPT][kSide];  // the own cells' parts of their west and north edges
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int ly = wv + 4 * k, lx = lane;
    T rec[kRec];
    solve_tile_edge<T, ORDER, TY, 0>(s, a, rows, ly + 2, lx + 1, dx, dt, dt_dx, rec);
#pragma unroll
    for (int w = 0; w < kSide; ++w) {
      bw[k][w] = side_word<T, kRec>(rec, 1, w);
      if (lx > 0) s.a[0][w][ly][lx - 1] = side_word<T, kRec>(rec, 0, w);
    }
  }
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int ly = wv + 4 * k, lx = lane;
    T rec[kRec];
    solve_tile_edge<T, ORDER, TY, 3>(s, a, rows, ly + 1, lx + 2, dx, dt, dt_dx, rec);
#pragma unroll
    for (int w = 0; w < kSide; ++w) {
      bn[k][w] = side_word<T, kRec>(rec, 1, w);
      if (ly > 0) s.a[1][w][ly - 1][lx] = side_word<T, kRec>(rec, 0, w);
    }
  }
  static_assert(TY <= kTX && kThreads >= 2 * kTX, "edge tail layout");
  if (tid < TY) {
    T rec[kRec];
    solve_tile_edge<T, ORDER, TY, 0>(s, a, rows, tid + 2, kTX + 1, dx, dt, dt_dx, rec);
#pragma unroll
    for (int w = 0; w < kSide; ++w) s.a[0][w][tid][kTX - 1] = side_word<T, kRec>(rec, 0, w);
  } else if (tid >= kTX && tid < 2 * kTX) {
    const int lx = tid - kTX;
    T rec[kRec];
    solve_tile_edge<T, ORDER, TY, 3>(s, a, rows, TY + 1, lx + 2, dx, dt, dt_dx, rec);
#pragma unroll
    for (int w = 0; w < kSide; ++w) s.a[1][w][TY - 1][lx] = side_word<T, kRec>(rec, 0, w);
  }
  __syncthreads();

  const bool t_zero = (t == T(0));
  T* __restrict__ dh = (T*)a.dh;
  T* __restrict__ dqx = (T*)a.dqx;
  T* __restrict__ dqy = (T*)a.dqy;
  const T cfl_dx = T(a.cfl) * T(a.dx);
  const bool sums = a.mass_sums != 0;
  T mn = T(INFINITY);
  double vol = 0.0, fsum = 0.0, rsum = 0.0;
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int ly = wv + 4 * k, lx = lane;
    const int gy = y0 + ly, gx = x0 + lx;
    if (gy >= ny || gx >= nx) continue;
    const int r = ly + 2, c = lx + 2;
    T ah = T(0), aqx = T(0), aqy = T(0);
    bool listed = false;
    if (s.ty[r][c] != 0) {
      // fluxMask rule: t == 0, wet, or a wet valid neighbour E, N, W, S that
      // is not compact cell 0 (fluxMaskGenerator_Kernel.cu:36); the
      // neighbours are looked at only for a dry cell (the four LDS reads of
      // that case issue together)
      listed = t_zero | gt_hs(s.h[r][c]);
      if (!listed) {
        const bool w_e = (s.ty[r][c + 1] != 0) & gt_hs(s.h[r][c + 1]) & !(gx + 1 == a.cell0_x && gy == a.cell0_y);
        const bool w_n = (s.ty[r - 1][c] != 0) & gt_hs(s.h[r - 1][c]) & !(gx == a.cell0_x && gy - 1 == a.cell0_y);
        const bool w_w = (s.ty[r][c - 1] != 0) & gt_hs(s.h[r][c - 1]) & !(gx - 1 == a.cell0_x && gy == a.cell0_y);
        const bool w_s = (s.ty[r + 1][c] != 0) & gt_hs(s.h[r + 1][c]) & !(gx == a.cell0_x && gy + 1 == a.cell0_y);
        listed = w_e | w_n | w_w | w_s;
      }
      if (listed) {
        T eE[kSide], eS[kSide];  // this cell is A of its east and south edges
#pragma unroll
        for (int w = 0; w < kSide; ++w) {
          eE[w] = s.a[0][w][ly][lx];
          eS[w] = s.a[1][w][ly][lx];
        }
        const T* eN = bn[k];  // B of its north edge
        const T* eW = bw[k];  // B of its west edge
        // the reference's order: E, N, W, S; per edge HLLC terms then bed term
        if (kRec == 5) {
          ah -= eE[0]; aqx -= eE[1]; aqx -= eE[3]; aqy -= eE[2];
          ah -= -eN[0]; aqx -= -eN[1]; aqy -= -eN[2]; aqy -= eN[3];
          ah -= -eW[0]; aqx -= -eW[1]; aqx -= eW[3]; aqy -= -eW[2];
          ah -= eS[0]; aqx -= eS[1]; aqy -= eS[2]; aqy -= eS[3];
        } else {
          ah -= eE[0]; aqx -= eE[2]; aqy -= eE[1];
          ah += eN[0]; aqx += eN[1]; aqy += eN[2];
          ah += eW[0]; aqx += eW[2]; aqy += eW[1];
          ah -= eS[0]; aqx -= eS[1]; aqy -= eS[2];
        }
      }
    }
    const int64_t gi = (int64_t)gy * pitch + gx;
    if (!FUSED) {
      dh[gi] = ah;
      dqx[gi] = aqx;
      dqy[gi] = aqy;
    } else {
      if (save_flux) {  // the step's flux as well (a calibration run's saved steps)
        dh[gi] = ah;
        dqx[gi] = aqx;
        dqy[gi] = aqy;
      }
      CellIn<T> in = cin[k];
      if (a.rain_mode == HP_RAIN_STATION)
        in.rain = (in.st >= 0 && in.st < kRateCache) ? s_rate[in.st] : station_rate<T>(rr, in.st, a.rain_scale);
      T hn = s.h[r][c], qxn = s.qx[r][c], qyn = s.qy[r][c], hm = in.hmax;
      // infiltration over the flux list (SWE_CUDA.py:439-443)
      T dF = T(0);
      if (a.infil_ks && listed) {
        dF = green_ampt<T>(hn, in.F, ((const T*)a.infil_ks)[in.cls], ((const T*)a.infil_phi)[in.cls],
                           ((const T*)a.infil_dtheta)[in.cls], dt);
        ((T*)a.infil_F)[gi] = in.F + dF;
      }
      T cand, vh, fh, rh;
      update_state<T>(in, hn, qxn, qyn, ah, aqx, aqy, dF, dt, cfl_dx, hm, cand, vh, fh, rh, time_control);
      if (time_control) {
        mn = fmin(mn, cand);
        if (sums) {
          vol += (double)vh;
          fsum += (double)fh;
          rsum += (double)rh;
        }
      }
      ((T*)a.h_out)[gi] = hn;
      ((T*)a.qx_out)[gi] = qxn;
      ((T*)a.qy_out)[gi] = qyn;
      if (hm != in.hmax) ((T*)a.h_max)[gi] = hm;
    }
  }
  if (FUSED && time_control) block_partial<T>(mn, vol, fsum, rsum, part, pidx, sums);
}

template <class T, int ORDER, int TY>
__global__ __launch_bounds__(kThreads) void k_flux(hp_grid_desc a, int ntx) {
  const int tile = xcd_band(blockIdx.x, gridDim.x);
  flux_tile<T, ORDER, TY, false>(a, tile % ntx, tile / ntx, Partials(nullptr, 0), 0);
}

template <class T, int ORDER, int TY>
__global__ __launch_bounds__(kThreads) void k_flux_multi(const hp_grid_desc* __restrict__ doms,
                                                         const int2* __restrict__ map) {
  const int2 m = map[xcd_band(blockIdx.x, gridDim.x)];
  const hp_grid_desc& a = doms[m.x];
  const int ntx = (a.nx + kTX - 1) / kTX;
  flux_tile<T, ORDER, TY, false>(a, m.y % ntx, m.y / ntx, Partials(nullptr, 0), 0);
}

// Single-pass step of every tile of every domain of an ensemble; map[i] =
// (domain, tile), listed domain by domain, so the partial records of a domain
// are one contiguous range.
template <class T, int ORDER, int TY>
__global__ __launch_bounds__(kThreads) void k_step_fused_multi(const hp_grid_desc* __restrict__ doms,
                                                               const int2* __restrict__ map, void* parts,
                                                               int n_total) {
  const int slot = xcd_band(blockIdx.x, gridDim.x);
  const int2 m = map[slot];
  const hp_grid_desc& a = doms[m.x];
  const int ntx = (a.nx + kTX - 1) / kTX;
  flux_tile<T, ORDER, TY, true>(a, m.y % ntx, m.y / ntx, Partials(parts, n_total), slot);
}

// Single-pass step: reads (h, qx, qy), writes (h_out, qx_out, qy_out).
template <class T, int ORDER, int TY>
// flags: 1 also save the step's flux in dh, dqx, dqy; 2 no time-control
// outputs (a frozen schedule); 4 a stopped domain copies its state to the
// output planes even when the time control marked both sets in step.
__global__ __launch_bounds__(kThreads) void k_step_fused(hp_grid_desc a, int ntx, int flags) {
  // the partial record goes to the tile's own slot (tile order, as in an
  // ensemble), not the block's
  const int tile = xcd_band(blockIdx.x, gridDim.x);
  flux_tile<T, ORDER, TY, true>(a, tile % ntx, tile / ntx, Partials(a.work, gridDim.x), tile, (flags & 1) != 0,
                                (flags & 2) == 0, (flags & 4) != 0);
}

}  // namespace hp

#include "swe_walk.hpp"

namespace hp {

// Infiltration pass over plane elements: a valid cell listed by the flux
// rule of the current state (t == 0, wet, or a wet valid neighbour E, N, W,
// S that is not compact cell 0: the wet list the reference's
// addInfiltrationSource receives from addFlux) takes the Green-Ampt
// increment; dF is written for every element (0 elsewhere) so the update
// pass can subtract it after the rain.
template <class T>
__global__ __launch_bounds__(kThreads) void k_infiltration(hp_grid_desc a) {
  if (stopped(a.ctrl)) return;
  const int64_t pitch = a.pitch, total = pitch * (int64_t)a.ny;
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i >= total) return;
  const T* __restrict__ h = (const T*)a.h;
  const uint8_t* __restrict__ ty = a.type_slot;
  const int row = (int)(i / pitch), col = (int)(i - row * pitch);
  T dF = T(0);
  if (col < a.nx && ty[i] != 0) {
    bool listed = T(a.ctrl->t) == T(0) || gt_hs(h[i]);
    const int nr[4] = {row, row - 1, row, row + 1};
    const int nc[4] = {col + 1, col, col - 1, col};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (nr[k] < 0 || nr[k] >= a.ny || nc[k] < 0 || nc[k] >= a.nx) continue;
      const int64_t j = (int64_t)nr[k] * pitch + nc[k];
      const bool is0 = nc[k] == a.cell0_x && nr[k] == a.cell0_y;
      listed = listed || (ty[j] != 0 && !is0 && gt_hs(h[j]));
    }
    if (listed) {
      const int cls = a.landuse[i];
      T* F = (T*)a.infil_F;
      dF = green_ampt<T>(h[i], F[i], ((const T*)a.infil_ks)[cls], ((const T*)a.infil_phi)[cls],
                         ((const T*)a.infil_dtheta)[cls], T(a.ctrl->dt));
      F[i] += dF;
    }
  }
  ((T*)a.infil_dF)[i] = dF;
}

// MODE selects which of the reference driver's update methods the pass is:
//   HP_UPDATE_STEP      time_friction_euler_update_cuda (SWE_CUDA.py:780-831):
//                       cells with |dh| > 0 or h >= 0; friction; CFL + dt/t
//   HP_UPDATE_FRICTION  friction_euler_update_cuda (:766-778): same list, no CFL
//   HP_UPDATE_EULER     euler_update (:712-724): cells with |dh| > 0 or
//                       h > 1e-6; q += dq without friction; no CFL
//   HP_UPDATE_CFL       time_update_cuda (:726-764): cells with h > 1e-6; no
//                       state change and no rain; h_max, CFL + dt/t
// Processes plane elements [chunk*kChunk, (chunk+1)*kChunk) and writes one
// partial record.
template <class T, int MODE>
__device__ inline void update_chunk(const hp_grid_desc& a, int64_t chunk, const Partials& part, int64_t pidx) {
  __shared__ T s_man[256];
  __shared__ T s_rate[kRateCache];
  const int tid = threadIdx.x;
  if (stopped(a.ctrl)) {
    if (tid == 0) part.put(pidx, INFINITY, 0.0, 0.0, 0.0, a.mass_sums != 0);
    return;
  }
  const T dt = T(a.ctrl->dt);
  const T cfl_dx = T(a.cfl) * T(a.dx);
  for (int k = tid; k < a.n_classes; k += kThreads) s_man[k] = ((const T*)a.manning)[k];
  const RainRows rr = rain_rows(a, a.ctrl->t);
  const int mode = a.rain_mode;
  if (mode == HP_RAIN_STATION && tid < kRateCache) s_rate[tid] = station_rate<T>(rr, tid, a.rain_scale);
  __syncthreads();

  const T* __restrict__ dhp = (const T*)a.dh;
  const T* __restrict__ dqxp = (const T*)a.dqx;
  const T* __restrict__ dqyp = (const T*)a.dqy;
  T* __restrict__ hp_ = (T*)a.h;
  T* __restrict__ qxp = (T*)a.qx;
  T* __restrict__ qyp = (T*)a.qy;
  T* __restrict__ hmp = (T*)a.h_max;
  const uint8_t* __restrict__ lu = a.landuse;
  const int16_t* __restrict__ st = a.station;
  const float* __restrict__ rf = a.rain_field;
  const T* __restrict__ dfp = (const T*)a.infil_dF;
  const T rscale = T(a.rain_scale);

  const bool sums = a.mass_sums != 0;
  T mn = T(INFINITY);
  double vol = 0.0, fsum = 0.0, rsum = 0.0;
  const int64_t total = a.pitch * (int64_t)a.ny;  // a multiple of 64
  const int64_t base = chunk * kChunk;
  for (int it = 0; it < kChunk / (kThreads * kVec); ++it) {
    const int64_t i0 = base + ((int64_t)it * kThreads + tid) * kVec;
    if (i0 >= total) break;
    const VecU8 vl = ldv<VecU8>(lu + i0);
    VecT<T> vh = ldv<VecT<T>>(hp_ + i0);
    VecT<T> vqx = ldv<VecT<T>>(qxp + i0);
    VecT<T> vqy = ldv<VecT<T>>(qyp + i0);
    VecT<T> vhm = ldv<VecT<T>>(hmp + i0);
    VecT<T> vdh{}, vdqx{}, vdqy{}, vr{}, vdf{};
    if (MODE != HP_UPDATE_CFL) {
      if (dfp) vdf = ldv<VecT<T>>(dfp + i0);
      vdh = ldv<VecT<T>>(dhp + i0);
      vdqx = ldv<VecT<T>>(dqxp + i0);
      vdqy = ldv<VecT<T>>(dqyp + i0);
      if (mode == HP_RAIN_STATION) {
        VecI16 vs{};
        if (st) vs = ldv<VecI16>(st + i0);
#pragma unroll
        for (int j = 0; j < kVec; ++j) {
          const int k = vs.x[j];
          vr.x[j] = (k >= 0 && k < kRateCache) ? s_rate[k] : station_rate<T>(rr, k, a.rain_scale);
        }
      } else if (mode == HP_RAIN_FIELD) {
        const VecF32 vf = ldv<VecF32>(rf + i0);
#pragma unroll
        for (int j = 0; j < kVec; ++j) vr.x[j] = T(vf.x[j]) * rscale;
      }
    }
#pragma unroll
    for (int j = 0; j < kVec; ++j) {
      const int cls = vl.x[j];
      if (cls == 255) continue;
      T h = vh.x[j];
      if (MODE == HP_UPDATE_CFL) {
        if (gt_hs(h)) {
          vhm.x[j] = fmax(h, vhm.x[j]);
          mn = fmin(mn, cfl_candidate<T>(h, vqx.x[j], vqy.x[j], cfl_dx));
        }
        if (sums) vol += (double)h;
        continue;
      }
      const T dhf = vdh.x[j];
      const T rdt = vr.x[j] * dt;
      const T dht = (dhf + rdt) - vdf.x[j];  // flux, + rain, - infiltration (the reference's order)
      const bool listed = (MODE == HP_UPDATE_EULER) ? (fabs(dht) > T(0) || gt_hs(h)) : (fabs(dht) > T(0) || h >= T(0));
      if (!listed) {
        if (sums) vol += (double)h;
        continue;
      }
      const bool dry = new_depth_dry(h, dht);
      T qx, qy;
      if (MODE == HP_UPDATE_EULER) {
        qx = vqx.x[j] + vdqx.x[j];
        qy = vqy.x[j] + vdqy.x[j];
      } else if (dry) {
        qx = T(0);
        qy = T(0);
      } else {
        qx = vqx.x[j] + vdqx.x[j];
        qy = vqy.x[j] + vdqy.x[j];
        implicit_friction<T>(h, s_man[cls], dt, qx, qy);
      }
      vh.x[j] = h;
      vqx.x[j] = qx;
      vqy.x[j] = qy;
      if (MODE == HP_UPDATE_STEP) {
        vhm.x[j] = fmax(h, vhm.x[j]);
        mn = fmin(mn, cfl_candidate<T>(h, qx, qy, cfl_dx));
      }
      if (sums) {
        vol += (double)h;
        fsum += (double)dhf;
        rsum += (double)rdt;
      }
    }
    if (MODE != HP_UPDATE_CFL) {
      stv(hp_ + i0, vh);
      stv(qxp + i0, vqx);
      stv(qyp + i0, vqy);
    }
    if (MODE == HP_UPDATE_STEP || MODE == HP_UPDATE_CFL) stv(hmp + i0, vhm);
  }
  block_partial<T>(mn, vol, fsum, rsum, part, pidx, sums);
}

template <class T, int MODE>
__global__ __launch_bounds__(kThreads) void k_update(hp_grid_desc a) {
  update_chunk<T, MODE>(a, blockIdx.x, Partials(a.work, gridDim.x), blockIdx.x);
}

template <class T, int MODE>
__global__ __launch_bounds__(kThreads) void k_update_multi(const hp_grid_desc* __restrict__ doms,
                                                           const int2* __restrict__ map, void* parts) {
  const int2 m = map[blockIdx.x];
  update_chunk<T, MODE>(doms[m.x], m.y, Partials(parts, gridDim.x), blockIdx.x);
}

// Fold a domain's partials in a fixed order and advance its time control.
// The minima are read 24 per thread before any is used, so one block covers
// 24 576 partials (a 2048^2 step's 17 920 four-row units) in one memory round
// trip; the control block is read alongside them rather than ahead of them.
// single_pass: the step just run was the single-pass one.  On a stopped
// domain the time control records whether that step left the second set of
// planes holding the state (the single-pass kernel copies it there on the
// first step past the stop, and no more once this is set).
__device__ inline void finalize_domain(hp_ctrl* ctrl, const double* __restrict__ mins, const Sums* __restrict__ part,
                                       int n, double dx, bool sums, bool single_pass) {
  constexpr int kU = 24;
  __shared__ double s_red[4][16];
  const int tid = threadIdx.x;
  const hp_ctrl c = *ctrl;
  double mn = INFINITY, vol = 0.0, fs = 0.0, rs = 0.0;
  for (int base = tid; base < n; base += blockDim.x * kU) {
    double v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int k = base + u * (int)blockDim.x;
      v[u] = k < n ? mins[k] : INFINITY;
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) mn = fmin(mn, v[u]);
  }
  if (stopped(&c)) {
    if (threadIdx.x == 0) ctrl->stopped_synced = single_pass ? 1.0 : 0.0;
    return;
  }
  if (sums) {
    for (int k = tid; k < n; k += blockDim.x) {
      const Sums p = part[k];
      vol += p.vol;
      fs += p.flux;
      rs += p.rain;
    }
  }
  mn = wave_min(mn);
  if (sums) {
    vol = wave_sum(vol);
    fs = wave_sum(fs);
    rs = wave_sum(rs);
  }
  const int lane = tid & 63, w = tid >> 6;
  const int nw = (blockDim.x + 63) >> 6;
  if (lane == 0) {
    s_red[0][w] = mn;
    s_red[1][w] = vol;
    s_red[2][w] = fs;
    s_red[3][w] = rs;
  }
  __syncthreads();
  if (tid == 0) {
    mn = s_red[0][0];
    vol = s_red[1][0];
    fs = s_red[2][0];
    rs = s_red[3][0];
    for (int k = 1; k < nw; ++k) {
      mn = fmin(mn, s_red[0][k]);
      vol += s_red[1][k];
      fs += s_red[2][k];
      rs += s_red[3][k];
    }
    advance_ctrl(ctrl, c, mn, vol, fs, rs, dx, sums);
  }
}

__global__ __launch_bounds__(1024) void k_finalize(hp_ctrl* ctrl, void* work, int n, double dx, int sums,
                                                   int single_pass) {
  const Partials P(work, n);
  finalize_domain(ctrl, P.mins, P.sums, n, dx, sums != 0, single_pass != 0);
}

// k_finalize followed by a calibration step's record (thread 0 wrote the
// control block, so it also reads the next step's t and dt).  The gauge
// depths (one per thread; more are read after the fold) are loaded before
// the fold so that their latency hides behind it.
__global__ __launch_bounds__(1024) void k_finalize_record(hp_ctrl* ctrl, void* work, int n, double dx, int sums,
                                                          StepRecord rec) {
  const Partials P(work, n);
  const int i0 = threadIdx.x;
  double v0 = 0.0;
  if (i0 < rec.n) v0 = rec.f32 ? (double)((const float*)rec.h)[rec.off[i0]] : ((const double*)rec.h)[rec.off[i0]];
  finalize_domain(ctrl, P.mins, P.sums, n, dx, sums != 0, true);
  if (i0 < rec.n) {
    if (rec.f32) ((float*)rec.row)[i0] = (float)v0;
    else ((double*)rec.row)[i0] = v0;
  }
  for (int i = i0 + blockDim.x; i < rec.n; i += blockDim.x) {
    if (rec.f32) ((float*)rec.row)[i] = ((const float*)rec.h)[rec.off[i]];
    else ((double*)rec.row)[i] = ((const double*)rec.h)[rec.off[i]];
  }
  if (rec.sched && threadIdx.x == 0) {
    rec.sched[2 * (int64_t)rec.k_next] = ctrl->t;
    rec.sched[2 * (int64_t)rec.k_next + 1] = ctrl->dt;
  }
}

// One block per domain, with k_finalize's block size, so that a domain's
// partials are folded in the same order as in a standalone run.
__global__ __launch_bounds__(1024) void k_finalize_multi(const hp_grid_desc* __restrict__ doms,
                                                        const int2* __restrict__ ranges, void* parts,
                                                        int n_total, int single_pass) {
  const int2 r = ranges[blockIdx.x];
  const hp_grid_desc& a = doms[blockIdx.x];
  const Partials P(parts, n_total);
  finalize_domain(a.ctrl, P.mins + r.x, P.sums + r.x, r.y, a.dx, a.mass_sums != 0, single_pass != 0);
}

// --- host side -------------------------------------------------------------
static int64_t update_blocks(const hp_grid_desc* g) {
  const int64_t total = g->pitch * (int64_t)g->ny;
  return (total + kChunk - 1) / kChunk;
}

// Rows per flux work unit: FP32 column-walk units of 60 columns x 2, 4, 6,
// 8, 10, 12 or 16 rows (default 8); FP64 LDS tiles of 64 x 4 or 64 x 8 cells
// (default 4).  Every launch of a domain's units takes its height from here.
static int flux_rows(const hp_grid_desc* g) {
  const int r = g->tile_rows;
  if (g->dtype == HP_F32) return (r == 2 || r == 4 || r == 6 || r == 10 || r == 12 || r == 16) ? r : 8;
  return r == 8 ? 8 : 4;
}

static int flux_cols(const hp_grid_desc* g) { return g->dtype == HP_F32 ? kWalkOwn : kTX; }

// Flux work units of a domain: one partial record each in the single-pass step.
static int64_t flux_tiles(const hp_grid_desc* g) {
  const int ty = flux_rows(g), tx = flux_cols(g);
  return (int64_t)((g->nx + tx - 1) / tx) * (int64_t)((g->ny + ty - 1) / ty);
}

static int check_desc(const hp_grid_desc* g) {
  if (!g) return fail(HP_EINVAL, "null descriptor");
  if (g->dtype != HP_F32 && g->dtype != HP_F64) return fail(HP_EINVAL, "dtype must be HP_F32 or HP_F64");
  if (g->order != 1 && g->order != 2) return fail(HP_EINVAL, "order must be 1 or 2");
  if (g->nx <= 0 || g->ny <= 0) return fail(HP_EINVAL, "nx and ny must be positive");
  if (g->pitch < g->nx) return fail(HP_EINVAL, "pitch must be >= nx");
  if (g->pitch % kVec != 0) return fail(HP_EINVAL, "pitch must be a multiple of 4");
  if (!(g->dx > 0.0)) return fail(HP_EINVAL, "dx must be positive");
  if (!g->h || !g->qx || !g->qy || !g->z || !g->h_max || !g->dh || !g->dqx || !g->dqy)
    return fail(HP_EINVAL, "a state plane pointer is null");
  if (!g->type_slot || !g->landuse || !g->slot_code || !g->manning || !g->ctrl)
    return fail(HP_EINVAL, "type_slot, landuse, slot_code, manning and ctrl are required");
  if (g->n_classes <= 0 || g->n_classes > 255) return fail(HP_EINVAL, "n_classes must be in 1..255");
  {
    // the update pass moves kVec consecutive elements per access
    const uintptr_t va = (uintptr_t)kVec * (g->dtype == HP_F32 ? 4 : 8);
    const void* planes[] = {g->h, g->qx, g->qy, g->h_max, g->dh, g->dqx, g->dqy};
    for (const void* p : planes)
      if ((uintptr_t)p % va) return fail(HP_EINVAL, "state planes must be aligned to 4 elements");
    if ((uintptr_t)g->landuse % kVec || (uintptr_t)g->station % (2 * kVec) || (uintptr_t)g->rain_field % (4 * kVec))
      return fail(HP_EINVAL, "landuse, station and rain_field planes must be aligned to 4 elements");
  }
  if (g->rain_mode == HP_RAIN_STATION) {
    if (!g->rain_table || g->rain_rows < 1 || g->rain_cols < 2)
      return fail(HP_EINVAL, "station rain needs a table of >= 1 row and >= 2 columns");
    if (g->rain_cols - 1 > kMaxStations) return fail(HP_EINVAL, "at most 1024 rain stations");
  } else if (g->rain_mode == HP_RAIN_FIELD) {
    if (!g->rain_field) return fail(HP_EINVAL, "field rain needs rain_field");
  } else if (g->rain_mode != HP_RAIN_NONE) {
    return fail(HP_EINVAL, "unknown rain_mode");
  }
  if ((g->given_depth && (g->gd_rows < 1 || g->gd_w < 1)) || (g->given_wl && (g->gwl_rows < 1 || g->gwl_w < 1)) ||
      (g->given_q && (g->gq_rows < 1 || g->gq_w < 1)))
    return fail(HP_EINVAL, "boundary series tables need >= 1 row and column");
  if (g->infil_ks && (!g->infil_phi || !g->infil_dtheta || !g->infil_F))
    return fail(HP_EINVAL, "infiltration needs infil_phi, infil_dtheta and infil_F");
  if ((uintptr_t)g->infil_dF % ((uintptr_t)kVec * (g->dtype == HP_F32 ? 4 : 8)))
    return fail(HP_EINVAL, "infil_dF must be aligned to 4 elements");
  if (flux_tiles(g) > 0x7fffffffLL || update_blocks(g) > 0x7fffffffLL) return fail(HP_EINVAL, "domain too large");
  if (g->dtype == HP_F32 && g->pitch * (int64_t)g->ny > 0xffffffffLL)
    return fail(HP_EINVAL, "float planes are limited to 2^32 elements");
  return HP_OK;
}

template <class T, int ORDER, int TY>
static int launch_flux_single(const hp_grid_desc* g, hipStream_t st) {
  const int ntx = (g->nx + kTX - 1) / kTX, nty = (g->ny + TY - 1) / TY;
  hipLaunchKernelGGL((k_flux<T, ORDER, TY>), dim3(ntx * nty), dim3(kThreads), 0, st, *g, ntx);
  return check_launch("k_flux");
}

// FP32: column-walk units (k_walk), 64-lane workgroups.
template <int ORDER, int R, bool FUSED>
static int launch_walk_r(const hp_grid_desc* g, hipStream_t st, int flags) {
  const int ns = (g->nx + kWalkOwn - 1) / kWalkOwn, nb = (g->ny + R - 1) / R;
  hipLaunchKernelGGL((k_walk<ORDER, R, FUSED>), dim3(ns * nb), dim3(kWalkLanes), 0, st, *g, ns, flags);
  return check_launch(FUSED ? "k_walk (single-pass step)" : "k_walk (flux pass)");
}

template <int ORDER, bool FUSED>
static int launch_walk_o(const hp_grid_desc* g, hipStream_t st, int flags) {
  switch (flux_rows(g)) {
    case 2: return launch_walk_r<ORDER, 2, FUSED>(g, st, flags);
    case 4: return launch_walk_r<ORDER, 4, FUSED>(g, st, flags);
    case 6: return launch_walk_r<ORDER, 6, FUSED>(g, st, flags);
    case 10: return launch_walk_r<ORDER, 10, FUSED>(g, st, flags);
    case 12: return launch_walk_r<ORDER, 12, FUSED>(g, st, flags);
    case 16: return launch_walk_r<ORDER, 16, FUSED>(g, st, flags);
    default: return launch_walk_r<ORDER, 8, FUSED>(g, st, flags);
  }
}

template <bool FUSED>
static int launch_walk(const hp_grid_desc* g, hipStream_t st, int flags) {
  return g->order == 2 ? launch_walk_o<2, FUSED>(g, st, flags) : launch_walk_o<1, FUSED>(g, st, flags);
}

template <int ORDER>
static int launch_flux_f64(const hp_grid_desc* g, hipStream_t st) {
  return flux_rows(g) == 8 ? launch_flux_single<double, ORDER, 8>(g, st) : launch_flux_single<double, ORDER, 4>(g, st);
}

static int launch_flux(const hp_grid_desc* g, hipStream_t st) {
  if (g->dtype == HP_F32) return launch_walk<false>(g, st, 0);
  return g->order == 2 ? launch_flux_f64<2>(g, st) : launch_flux_f64<1>(g, st);
}

template <class T, int ORDER, int TY>
static int launch_fused_t(const hp_grid_desc* g, hipStream_t st, int flags) {
  const int ntx = (g->nx + kTX - 1) / kTX, nty = (g->ny + TY - 1) / TY;
  hipLaunchKernelGGL((k_step_fused<T, ORDER, TY>), dim3(ntx * nty), dim3(kThreads), 0, st, *g, ntx, flags);
  return check_launch("k_step_fused");
}

template <int ORDER>
static int launch_fused_f64(const hp_grid_desc* g, hipStream_t st, int flags) {
  return flux_rows(g) == 8 ? launch_fused_t<double, ORDER, 8>(g, st, flags) : launch_fused_t<double, ORDER, 4>(g, st, flags);
}

static int launch_step_fused(const hp_grid_desc* g, hipStream_t st, int flags = 0) {
  if (g->dtype == HP_F32) return launch_walk<true>(g, st, flags);
  return g->order == 2 ? launch_fused_f64<2>(g, st, flags) : launch_fused_f64<1>(g, st, flags);
}

static int launch_finalize_fused(const hp_grid_desc* g, hipStream_t st, const StepRecord* rec = nullptr) {
  if (rec) {
    hipLaunchKernelGGL(k_finalize_record, dim3(1), dim3(1024), 0, st, g->ctrl, g->work, (int)flux_tiles(g), g->dx,
                       g->mass_sums, *rec);
    return check_launch("k_finalize_record");
  }
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, st, g->ctrl, g->work, (int)flux_tiles(g), g->dx,
                     g->mass_sums, 1);
  return check_launch("k_finalize");
}

static int launch_infiltration(const hp_grid_desc* g, hipStream_t st) {
  const int64_t total = g->pitch * (int64_t)g->ny;
  const unsigned nb = (unsigned)((total + kThreads - 1) / kThreads);
  if (g->dtype == HP_F32) hipLaunchKernelGGL(k_infiltration<float>, dim3(nb), dim3(kThreads), 0, st, *g);
  else hipLaunchKernelGGL(k_infiltration<double>, dim3(nb), dim3(kThreads), 0, st, *g);
  return check_launch("k_infiltration");
}

template <int MODE>
static void launch_update_mode(const hp_grid_desc* g, int64_t nb, hipStream_t st) {
  if (g->dtype == HP_F32) hipLaunchKernelGGL((k_update<float, MODE>), dim3((unsigned)nb), dim3(kThreads), 0, st, *g);
  else hipLaunchKernelGGL((k_update<double, MODE>), dim3((unsigned)nb), dim3(kThreads), 0, st, *g);
}

static int launch_update(const hp_grid_desc* g, int mode, hipStream_t st) {
  const int64_t nb = update_blocks(g);
  if (!g->work || g->work_bytes < partials_bytes(nb))
    return fail(HP_ESPACE, "workspace smaller than hp_grid_workspace_bytes()");
  switch (mode) {
    case HP_UPDATE_STEP: launch_update_mode<HP_UPDATE_STEP>(g, nb, st); break;
    case HP_UPDATE_FRICTION: launch_update_mode<HP_UPDATE_FRICTION>(g, nb, st); break;
    case HP_UPDATE_EULER: launch_update_mode<HP_UPDATE_EULER>(g, nb, st); break;
    case HP_UPDATE_CFL: launch_update_mode<HP_UPDATE_CFL>(g, nb, st); break;
    default: return fail(HP_EINVAL, "unknown update mode");
  }
  int rc = check_launch("k_update");
  if (rc) return rc;
  if (mode == HP_UPDATE_STEP || mode == HP_UPDATE_CFL) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, st, g->ctrl, g->work, (int)nb, g->dx,
                       g->mass_sums, 0);
    return check_launch("k_finalize");
  }
  return HP_OK;
}

// ---- ensemble ---------------------------------------------------------------
// Workspace of an ensemble: the descriptors twice (the second copy with the
// two sets of state planes exchanged, for the odd steps of the single-pass
// step), the (domain, tile) and (domain, chunk) maps, the per-domain ranges of
// partial records for the two-pass and the single-pass step, the records.
struct EnsembleLayout {
  int64_t off_doms, off_fmap, off_umap, off_fin, off_fin1, off_parts, total;
  int64_t n_ftiles, n_ublocks;
};


static EnsembleLayout ensemble_layout(int32_t n, const hp_grid_desc* d) {
  EnsembleLayout L{};
  for (int k = 0; k < n; ++k) {
    L.n_ftiles += flux_tiles(&d[k]);
    L.n_ublocks += update_blocks(&d[k]);
  }
  L.off_doms = 0;
  L.off_fmap = align256(2 * (int64_t)n * (int64_t)sizeof(hp_grid_desc));
  L.off_umap = L.off_fmap + align256(L.n_ftiles * 8);
  L.off_fin = L.off_umap + align256(L.n_ublocks * 8);
  L.off_fin1 = L.off_fin + align256((int64_t)n * 8);
  L.off_parts = L.off_fin1 + align256((int64_t)n * 8);
  L.total = L.off_parts + align256(partials_bytes(L.n_ublocks > L.n_ftiles ? L.n_ublocks : L.n_ftiles));
  return L;
}

template <class T, int ORDER, int TY>
static void launch_flux_multi(const hp_ensemble* e, hipStream_t st) {
  hipLaunchKernelGGL((k_flux_multi<T, ORDER, TY>), dim3((unsigned)e->n_flux_tiles), dim3(kThreads), 0, st,
                     (const hp_grid_desc*)e->doms, (const int2*)e->flux_map);
}

template <class T>
static void launch_update_multi(const hp_ensemble* e, hipStream_t st) {
  hipLaunchKernelGGL((k_update_multi<T, HP_UPDATE_STEP>), dim3((unsigned)e->n_update_blocks), dim3(kThreads), 0, st,
                     (const hp_grid_desc*)e->doms, (const int2*)e->update_map, e->partials);
}

template <class T, int ORDER, int TY>
static void launch_fused_multi(const hp_ensemble* e, int parity, hipStream_t st) {
  hipLaunchKernelGGL((k_step_fused_multi<T, ORDER, TY>), dim3((unsigned)e->n_flux_tiles), dim3(kThreads), 0, st,
                     (const hp_grid_desc*)e->doms + (parity ? e->n_domains : 0), (const int2*)e->flux_map,
                     e->partials, (int)e->n_flux_tiles);
}

template <int ORDER, int R, bool FUSED>
static void launch_walk_multi(const hp_ensemble* e, int parity, hipStream_t st) {
  hipLaunchKernelGGL((k_walk_multi<ORDER, R, FUSED>), dim3((unsigned)e->n_flux_tiles), dim3(kWalkLanes), 0, st,
                     (const hp_grid_desc*)e->doms + (parity ? e->n_domains : 0), (const int2*)e->flux_map,
                     e->partials, (int)e->n_flux_tiles);
}

template <int ORDER, bool FUSED>
static void launch_walk_multi_o(const hp_ensemble* e, int parity, hipStream_t st) {
  switch (e->tile_rows) {
    case 2: return launch_walk_multi<ORDER, 2, FUSED>(e, parity, st);
    case 4: return launch_walk_multi<ORDER, 4, FUSED>(e, parity, st);
    case 6: return launch_walk_multi<ORDER, 6, FUSED>(e, parity, st);
    case 10: return launch_walk_multi<ORDER, 10, FUSED>(e, parity, st);
    case 12: return launch_walk_multi<ORDER, 12, FUSED>(e, parity, st);
    case 16: return launch_walk_multi<ORDER, 16, FUSED>(e, parity, st);
    default: return launch_walk_multi<ORDER, 8, FUSED>(e, parity, st);
  }
}

template <bool FUSED>
static void launch_walk_multi_any(const hp_ensemble* e, int parity, hipStream_t st) {
  if (e->order == 2) launch_walk_multi_o<2, FUSED>(e, parity, st);
  else launch_walk_multi_o<1, FUSED>(e, parity, st);
}

// One single-pass step of every domain: even steps read the domains' own
// planes and write their second set, odd steps the reverse.
static int ensemble_fused_step_once(const hp_ensemble* e, int parity, hipStream_t st) {
  if (e->dtype == HP_F32) {
    launch_walk_multi_any<true>(e, parity, st);
  } else {
    const bool t8 = e->tile_rows == 8;
    if (e->order == 2) (t8 ? launch_fused_multi<double, 2, 8> : launch_fused_multi<double, 2, 4>)(e, parity, st);
    else (t8 ? launch_fused_multi<double, 1, 8> : launch_fused_multi<double, 1, 4>)(e, parity, st);
  }
  int rc = check_launch("k_step_fused_multi");
  if (rc) return rc;
  hipLaunchKernelGGL(k_finalize_multi, dim3((unsigned)e->n_domains), dim3(1024), 0, st,
                     (const hp_grid_desc*)e->doms, (const int2*)e->fused_finalize_map, e->partials,
                     (int)e->n_flux_tiles, 1);
  return check_launch("k_finalize_multi");
}

static int ensemble_step_once(const hp_ensemble* e, hipStream_t st) {
  const bool f32 = e->dtype == HP_F32;
  if (f32) {
    launch_walk_multi_any<false>(e, 0, st);
  } else {
    const bool t8 = e->tile_rows == 8;
    if (e->order == 2) (t8 ? launch_flux_multi<double, 2, 8> : launch_flux_multi<double, 2, 4>)(e, st);
    else (t8 ? launch_flux_multi<double, 1, 8> : launch_flux_multi<double, 1, 4>)(e, st);
  }
  int rc = check_launch("k_flux_multi");
  if (rc) return rc;
  if (f32) launch_update_multi<float>(e, st);
  else launch_update_multi<double>(e, st);
  rc = check_launch("k_update_multi");
  if (rc) return rc;
  hipLaunchKernelGGL(k_finalize_multi, dim3((unsigned)e->n_domains), dim3(1024), 0, st,
                     (const hp_grid_desc*)e->doms, (const int2*)e->finalize_map, e->partials,
                     (int)e->n_update_blocks, 0);
  return check_launch("k_finalize_multi");
}

template <class T>
__global__ void k_gather(const T* __restrict__ plane, const int64_t* __restrict__ off, int64_t n, T* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = plane[off[i]];
}
template <class T>
__global__ void k_scatter(T* __restrict__ plane, const int64_t* __restrict__ off, int64_t n, const T* __restrict__ in) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) plane[off[i]] = in[i];
}

// Single-pass step (h, qx, qy -> h_out, qx_out, qy_out; with save_flux the
// step's flux also into dh, dqx, dqy) and, with finalize, its time control,
// followed by rec when given; without time_control the step computes no
// CFL candidates, sums or partial records (no time control may follow).
// Shared with the calibration runs.
int grid_fused_launch(const hp_grid_desc* g, bool save_flux, bool finalize, hipStream_t st, const StepRecord* rec,
                      bool time_control) {
  int rc = check_desc(g);
  if (rc) return rc;
  if (!g->h_out || !g->qx_out || !g->qy_out) return fail(HP_EINVAL, "the fused step needs h_out, qx_out, qy_out");
  if (save_flux && (!g->dh || !g->dqx || !g->dqy)) return fail(HP_EINVAL, "saving the flux needs dh, dqx, dqy");
  if (!g->work || g->work_bytes < partials_bytes(flux_tiles(g)))
    return fail(HP_ESPACE, "workspace smaller than hp_grid_workspace_bytes()");
  if (finalize && !time_control) return fail(HP_EINVAL, "a time control needs the step's CFL candidates");
  rc = launch_step_fused(g, st, (save_flux ? 1 : 0) | (time_control ? 0 : 2));
  if (rc || !finalize) return rc;
  return launch_finalize_fused(g, st, rec);
}

// Flux pass launch shared with the adjoint (which recomputes it).
int grid_flux_launch(const hp_grid_desc* g, hipStream_t st) {
  int rc = check_desc(g);
  if (rc) return rc;
  return launch_flux(g, st);
}

// ---- allocation-free single-pass step ---------------------------------------
// The single-pass step's only scratch memory is its per-tile partial records
// (CFL minimum, FP64 sums), normally in the caller's workspace.  For callers
// that cannot allocate, a file-scope device array holds them instead: every
// GPU the library runs on keeps its own copy until the process ends.  Every
// tile writes its record in every step before the time control reads it, so
// nothing in it must start at zero; calls that overlap on one GPU share the
// array and race.
constexpr int64_t kScratchBytes = int64_t(8) << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];
constexpr int kMaxScratchDevices = 64;
static void* g_scratch[kMaxScratchDevices];
static std::mutex g_scratch_mu;

// Device address of the current device's scratch array, looked up once per
// device (hipGetSymbolAddress with that device current).
static int scratch_ptr(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(HP_ELAUNCH, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxScratchDevices) return fail(HP_EINVAL, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(HP_ELAUNCH, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return HP_OK;
}

// Workspace of a domain: the partial records of the larger of the update
// pass and the flux / single-pass step.
static int64_t workspace_need(const hp_grid_desc* g) {
  const int64_t nt = flux_tiles(g), nu = update_blocks(g);
  return partials_bytes(nu > nt ? nu : nt);
}

// The checks every single-pass entry point makes.
static int check_fused(const hp_grid_desc* g) {
  int rc = check_desc(g);
  if (rc) return rc;
  if (!g->h_out || !g->qx_out || !g->qy_out) return fail(HP_EINVAL, "the fused step needs h_out, qx_out, qy_out");
  if (!g->work || g->work_bytes < workspace_need(g))
    return fail(HP_ESPACE, "workspace smaller than hp_grid_workspace_bytes()");
  return HP_OK;
}

// One single-pass step and its time control.  first: the first step of a
// call (a stopped domain copies its state through even when marked in step).
static int single_pass_step(const hp_grid_desc* d, hipStream_t st, bool first) {
  int rc = launch_step_fused(d, st, first ? 4 : 0);
  if (rc) return rc;
  return launch_finalize_fused(d, st);
}

// After an odd number of single-pass steps the state sits in the second set
// of planes: bring it back to the caller's (h, qx, qy).
static int copy_state_back(const hp_grid_desc* g, hipStream_t st) {
  const size_t bytes = (size_t)(g->pitch * (int64_t)g->ny) * (g->dtype == HP_F32 ? 4 : 8);
  hipError_t e = hipMemcpyAsync(g->h, g->h_out, bytes, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(g->qx, g->qx_out, bytes, hipMemcpyDeviceToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(g->qy, g->qy_out, bytes, hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return fail(HP_ELAUNCH, std::string("state copy: ") + hipGetErrorString(e));
  return HP_OK;
}

}  // namespace hp

using namespace hp;

extern "C" {

int64_t hp_grid_workspace_bytes(const hp_grid_desc* g) {
  if (!g || g->pitch <= 0 || g->ny <= 0 || g->nx <= 0) return 0;
  return workspace_need(g);
}

int hp_grid_step_fused(const hp_grid_desc* g, int32_t nsteps, void* stream) {
  int rc = check_fused(g);
  if (rc) return rc;
  if (nsteps < 0) return fail(HP_EINVAL, "nsteps must be >= 0");
  hipStream_t st = (hipStream_t)stream;
  hp_grid_desc d = *g;
  for (int32_t k = 0; k < nsteps; ++k) {
    // the first step of a call copies a stopped domain's state through even
    // if an earlier call left both sets in step: the host may have written
    // the state in between
    rc = single_pass_step(&d, st, k == 0);
    if (rc) return rc;
    std::swap(d.h, d.h_out);
    std::swap(d.qx, d.qx_out);
    std::swap(d.qy, d.qy_out);
  }
  if (nsteps & 1) return copy_state_back(g, st);
  return HP_OK;
}

int64_t hp_grid_noalloc_max_bytes(void) { return kScratchBytes; }

int hp_grid_step_fused_noalloc(const hp_grid_desc* g, int32_t nsteps, void* stream) {
  int rc = check_desc(g);
  if (rc) return rc;
  const int64_t need = workspace_need(g);
  if (need > kScratchBytes)
    return fail(HP_ESPACE, "domain needs " + std::to_string(need) + " bytes of per-tile records; the allocation-free "
                           "step holds at most " + std::to_string(kScratchBytes) + " (use hp_grid_step_fused)");
  void* p = nullptr;
  rc = scratch_ptr(&p);
  if (rc) return rc;
  hp_grid_desc d = *g;
  d.work = p;
  d.work_bytes = kScratchBytes;
  return hp_grid_step_fused(&d, nsteps, stream);
}

int hp_grid_step_fused_timed(const hp_grid_desc* g, int32_t nsteps, int32_t sample_every, float* kernel_ms,
                             float* step_ms, void* stream) {
  int rc = check_fused(g);
  if (rc) return rc;
  if (nsteps <= 0 || sample_every <= 0) return fail(HP_EINVAL, "nsteps and sample_every must be positive");
  if (!kernel_ms || !step_ms) return fail(HP_EINVAL, "kernel_ms and step_ms are required");
  hipStream_t st = (hipStream_t)stream;
  // ev[0] / ev[1]: around all nsteps steps; ev[2 + 2j] / ev[3 + 2j]: around
  // the single-pass kernel of the j-th sampled step
  const int32_t samples = (nsteps + sample_every - 1) / sample_every;
  std::vector<hipEvent_t> ev(2 + 2 * (size_t)samples, nullptr);
  auto release = [&]() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  };
  for (auto& e : ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      release();
      return fail(HP_ELAUNCH, "hipEventCreate failed");
    }
  }
  hp_grid_desc d = *g;
  (void)hipEventRecord(ev[0], st);
  for (int32_t k = 0; k < nsteps && rc == HP_OK; ++k) {
    const bool sample = k % sample_every == 0;
    const size_t j = 2 + 2 * (size_t)(k / sample_every);
    if (sample) (void)hipEventRecord(ev[j], st);
    rc = launch_step_fused(&d, st, k == 0 ? 4 : 0);
    if (rc == HP_OK) {
      if (sample) (void)hipEventRecord(ev[j + 1], st);
      rc = launch_finalize_fused(&d, st);
    }
    std::swap(d.h, d.h_out);
    std::swap(d.qx, d.qx_out);
    std::swap(d.qy, d.qy_out);
  }
  if (rc == HP_OK) {
    (void)hipEventRecord(ev[1], st);
    if (nsteps & 1) rc = copy_state_back(g, st);  // after the timed steps
    if (rc == HP_OK && hipEventSynchronize(ev[1]) != hipSuccess) rc = fail(HP_ELAUNCH, "hipEventSynchronize failed");
  }
  if (rc == HP_OK) {
    float sum = 0.0f, ms = 0.0f;
    for (int32_t j = 0; j < samples; ++j) {
      (void)hipEventElapsedTime(&ms, ev[2 + 2 * (size_t)j], ev[3 + 2 * (size_t)j]);
      sum += ms;
    }
    (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
    *kernel_ms = sum / (float)samples;
    *step_ms = ms;
  }
  release();
  return rc;
}

int hp_grid_flux(const hp_grid_desc* g, void* stream) {
  int rc = check_desc(g);
  if (rc) return rc;
  return launch_flux(g, (hipStream_t)stream);
}

int hp_grid_update(const hp_grid_desc* g, void* stream) { return hp_grid_update_mode(g, HP_UPDATE_STEP, stream); }

int hp_grid_update_mode(const hp_grid_desc* g, int mode, void* stream) {
  int rc = check_desc(g);
  if (rc) return rc;
  return launch_update(g, mode, (hipStream_t)stream);
}

int hp_grid_infiltration(const hp_grid_desc* g, void* stream) {
  int rc = check_desc(g);
  if (rc) return rc;
  if (!g->infil_ks || !g->infil_dF) return fail(HP_EINVAL, "infiltration needs infil_ks and an infil_dF plane");
  return launch_infiltration(g, (hipStream_t)stream);
}

int hp_grid_step(const hp_grid_desc* g, int32_t nsteps, void* stream) {
  int rc = check_desc(g);
  if (rc) return rc;
  if (nsteps < 0) return fail(HP_EINVAL, "nsteps must be >= 0");
  if (g->infil_ks && !g->infil_dF) return fail(HP_EINVAL, "a step with infiltration needs an infil_dF plane");
  hp_grid_desc d = *g;
  if (!d.infil_ks) d.infil_dF = nullptr;  // nothing to subtract
  for (int32_t k = 0; k < nsteps; ++k) {
    rc = launch_flux(&d, (hipStream_t)stream);
    if (rc) return rc;
    if (d.infil_ks) {
      rc = launch_infiltration(&d, (hipStream_t)stream);
      if (rc) return rc;
    }
    rc = launch_update(&d, HP_UPDATE_STEP, (hipStream_t)stream);
    if (rc) return rc;
  }
  return HP_OK;
}

int64_t hp_ensemble_workspace_bytes(int32_t n, const hp_grid_desc* descs) {
  if (n <= 0 || !descs) return 0;
  return ensemble_layout(n, descs).total;
}

int hp_ensemble_init(int32_t n, const hp_grid_desc* descs, int32_t single_pass, void* work, int64_t work_bytes,
                     hp_ensemble* out, void* stream) {
  if (n <= 0 || !descs || !out) return fail(HP_EINVAL, "need n > 0 descriptors and an output handle");
  if (single_pass != 0 && single_pass != 1) return fail(HP_EINVAL, "single_pass must be 0 or 1");
  for (int k = 0; k < n; ++k) {
    if (descs[k].infil_ks) return fail(HP_EINVAL, "infiltration is not supported in an ensemble");
    int rc = check_desc(&descs[k]);
    if (rc) return fail(rc, "domain " + std::to_string(k) + ": " + last_error_slot());
    if (single_pass && (!descs[k].h_out || !descs[k].qx_out || !descs[k].qy_out))
      return fail(HP_EINVAL, "domain " + std::to_string(k) + ": the single-pass step needs h_out, qx_out, qy_out");
    if (descs[k].dtype != descs[0].dtype || descs[k].order != descs[0].order ||
        flux_rows(&descs[k]) != flux_rows(&descs[0]))
      return fail(HP_EINVAL, "every domain of an ensemble needs the same dtype, order and tile_rows");
  }
  const EnsembleLayout L = ensemble_layout(n, descs);
  if (!work || work_bytes < L.total) return fail(HP_ESPACE, "workspace smaller than hp_ensemble_workspace_bytes()");
  if (L.n_ftiles > 0x7fffffffLL || L.n_ublocks > 0x7fffffffLL) return fail(HP_EINVAL, "ensemble too large");
  std::vector<uint8_t> host((size_t)L.off_parts, 0);
  std::memcpy(host.data() + L.off_doms, descs, sizeof(hp_grid_desc) * (size_t)n);
  hp_grid_desc* swapped = (hp_grid_desc*)(host.data() + L.off_doms) + n;
  for (int k = 0; k < n; ++k) {
    swapped[k] = descs[k];
    std::swap(swapped[k].h, swapped[k].h_out);
    std::swap(swapped[k].qx, swapped[k].qx_out);
    std::swap(swapped[k].qy, swapped[k].qy_out);
  }
  int32_t* fm = (int32_t*)(host.data() + L.off_fmap);
  int32_t* um = (int32_t*)(host.data() + L.off_umap);
  int32_t* fin = (int32_t*)(host.data() + L.off_fin);
  int32_t* fin1 = (int32_t*)(host.data() + L.off_fin1);
  int64_t fi = 0, ui = 0;
  for (int k = 0; k < n; ++k) {
    const int64_t nt = flux_tiles(&descs[k]), nu = update_blocks(&descs[k]);
    fin1[2 * k] = (int32_t)fi;
    fin1[2 * k + 1] = (int32_t)nt;
    for (int64_t j = 0; j < nt; ++j, ++fi) {
      fm[2 * fi] = k;
      fm[2 * fi + 1] = (int32_t)j;
    }
    fin[2 * k] = (int32_t)ui;
    fin[2 * k + 1] = (int32_t)nu;
    for (int64_t j = 0; j < nu; ++j, ++ui) {
      um[2 * ui] = k;
      um[2 * ui + 1] = (int32_t)j;
    }
  }
  hipError_t err = hipMemcpyAsync(work, host.data(), host.size(), hipMemcpyHostToDevice, (hipStream_t)stream);
  if (err == hipSuccess) err = hipStreamSynchronize((hipStream_t)stream);
  if (err != hipSuccess) return fail(HP_ELAUNCH, std::string("ensemble table upload: ") + hipGetErrorString(err));
  uint8_t* w = (uint8_t*)work;
  out->n_domains = n;
  out->dtype = descs[0].dtype;
  out->order = descs[0].order;
  out->tile_rows = flux_rows(&descs[0]);
  out->n_flux_tiles = (int32_t)L.n_ftiles;
  out->n_update_blocks = (int32_t)L.n_ublocks;
  out->doms = w + L.off_doms;
  out->flux_map = w + L.off_fmap;
  out->update_map = w + L.off_umap;
  out->finalize_map = w + L.off_fin;
  out->partials = w + L.off_parts;
  out->single_pass = single_pass;
  out->fused_finalize_map = w + L.off_fin1;
  return HP_OK;
}

int hp_ensemble_step(const hp_ensemble* e, int32_t nsteps, void* stream) {
  if (!e || e->n_domains <= 0 || !e->doms) return fail(HP_EINVAL, "ensemble handle not initialised");
  if (nsteps < 0) return fail(HP_EINVAL, "nsteps must be >= 0");
  if (e->single_pass) {
    if (nsteps & 1)
      return fail(HP_EINVAL, "a single-pass ensemble advances an even number of steps per call (the state then "
                             "ends in the domains' own planes)");
    for (int32_t k = 0; k < nsteps; ++k) {
      int rc = ensemble_fused_step_once(e, k & 1, (hipStream_t)stream);
      if (rc) return rc;
    }
    return HP_OK;
  }
  for (int32_t k = 0; k < nsteps; ++k) {
    int rc = ensemble_step_once(e, (hipStream_t)stream);
    if (rc) return rc;
  }
  return HP_OK;
}

int hp_gather(int dtype, const void* plane, const int64_t* offsets, int64_t n, void* out, void* stream) {
  if (dtype != HP_F32 && dtype != HP_F64) return fail(HP_EINVAL, "dtype must be HP_F32 or HP_F64");
  if (n < 0) return fail(HP_EINVAL, "n must be >= 0");
  if (n == 0) return HP_OK;
  if (!plane || !offsets || !out) return fail(HP_EINVAL, "null pointer");
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (dtype == HP_F32)
    hipLaunchKernelGGL(k_gather<float>, dim3(nb), dim3(256), 0, (hipStream_t)stream, (const float*)plane, offsets, n,
                       (float*)out);
  else
    hipLaunchKernelGGL(k_gather<double>, dim3(nb), dim3(256), 0, (hipStream_t)stream, (const double*)plane, offsets,
                       n, (double*)out);
  return check_launch("k_gather");
}

int hp_scatter(int dtype, void* plane, const int64_t* offsets, int64_t n, const void* in, void* stream) {
  if (dtype != HP_F32 && dtype != HP_F64) return fail(HP_EINVAL, "dtype must be HP_F32 or HP_F64");
  if (n < 0) return fail(HP_EINVAL, "n must be >= 0");
  if (n == 0) return HP_OK;
  if (!plane || !offsets || !in) return fail(HP_EINVAL, "null pointer");
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (dtype == HP_F32)
    hipLaunchKernelGGL(k_scatter<float>, dim3(nb), dim3(256), 0, (hipStream_t)stream, (float*)plane, offsets, n,
                       (const float*)in);
  else
    hipLaunchKernelGGL(k_scatter<double>, dim3(nb), dim3(256), 0, (hipStream_t)stream, (double*)plane, offsets, n,
                       (const double*)in);
  return check_launch("k_scatter");
}

}  // extern "C"
