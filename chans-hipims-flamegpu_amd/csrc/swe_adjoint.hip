// Adjoint of one structured-grid step (no counterpart in the reference).
//
// With the time-step schedule frozen, step k maps U_k = (h, qx, qy) to
// U_{k+1} through the flux pass (a radius-2 stencil) and the pointwise update
// (rain, implicit friction).  Given the cotangent lambda_{k+1} of U_{k+1},
// hp_adjoint_step computes lambda_k and accumulates the gradient of the
// Manning coefficient of every landcover class and of the rain multiplier:
//
//   recompute   the flux pass on U_k (the state planes hold U_k on entry),
//               unless the caller passes the flux saved by the forward run
//               (flux_ready)
//   k_adj_update  pointwise: reverse of rain + friction + update; writes the
//               direct part of lambda_k, the cotangents of the flux
//               accumulators (zero where the cell is not flux-listed) and
//               adds the block's FP64 parameter-gradient sums
//   k_adj_edges one thread per edge: reverse of the edge solve
//               (swe_adjoint.hpp); writes the (h, qx, qy) cotangents the edge
//               sends to each of the four cells it reads (roles l, A, B, rr)
//   k_adj_gather per cell: lambda_k += the contributions of the 8 edges that
//               read it
// Every write goes to the writing thread's own cell or edge slot: no atomics.
// For order 1, k_adj_step_tile does the three in one pass over 64 x 8 tiles
// (the update adjoint of the tile and of its edge-sharing ring, the edges of
// the tile's cells, the per-cell sum), reading the cotangent from one set of
// planes and writing it to another.
//
// hp_calib_forward / hp_calib_reverse run the step loops of a calibration
// natively: the forward steps on the single-pass kernel (recording the time
// schedule, or following it through per-step control blocks, and keeping
// each segment's states and fluxes), and the reverse sweep over a segment.

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string>

#include "../../include/hipims_abi.h"
#include "hp_common.hpp"
#include "swe_adjoint.hpp"
#include "swe_source.hpp"

namespace hp {

constexpr int kAThreads = 256;
constexpr int kAChunk = 1024;  // plane elements per block of the pointwise passes
constexpr int kGradStride = 16;  // doubles per block: classes 0..14, rain at 15
constexpr int kARateCache = 64;  // station rates kept in LDS per block

int grid_flux_launch(const hp_grid_desc* g, hipStream_t st);                   // swe_grid.hip
int grid_fused_launch(const hp_grid_desc* g, bool save_flux, bool finalize, hipStream_t st,
                      const StepRecord* rec, bool time_control);  // swe_grid.hip

// Records (t, dt) of the control block as step k of a schedule [K+1][2].
__global__ void k_schedule_save(const hp_ctrl* ctrl, double* sched, int k) {
  if (threadIdx.x == 0) {
    sched[2 * (int64_t)k] = ctrl->t;
    sched[2 * (int64_t)k + 1] = ctrl->dt;
  }
}

// End of a calibration step: the gauge depths into their history row and,
// when recording, (t, dt) of the next step (sched may be null).
template <class T>
__global__ void k_step_record(const T* __restrict__ h, const int64_t* __restrict__ off, int n, T* __restrict__ row,
                              const hp_ctrl* ctrl, double* sched, int k_next) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) row[i] = h[off[i]];
  if (sched && i == 0) {
    sched[2 * (int64_t)k_next] = ctrl->t;
    sched[2 * (int64_t)k_next + 1] = ctrl->dt;
  }
}

template <class T> struct AdjPlanes {
  T *lh, *lqx, *lqy;
  T *ch, *cqx, *cqy;
  T* xc;  // 12 planes: role(4) x var(3), each (ny+1) x cpitch
  T* yc;
  int64_t cpitch, cplane;
  int frozen_slopes;
  const int32_t* gauge_of;  // null: no loss term in this pass
  const double* loss_h;
  const double* loss_obs;
};

template <class T, class DP>
__device__ inline bool flux_listed(DP a, const T* h, int x, int y, T t) {
  const int P = (int)a->pitch;  // planes of fewer than 2^31 elements
  const int i = y * P + x;
  if (t == T(0) || gt_hs(h[i])) return true;
  const int nx[4] = {x + 1, x, x - 1, x};
  const int ny[4] = {y, y - 1, y, y + 1};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (nx[k] < 0 || nx[k] >= a->nx || ny[k] < 0 || ny[k] >= a->ny) continue;
    const int j = ny[k] * P + nx[k];
    if (a->type_slot[j] == 0) continue;
    if (nx[k] == a->cell0_x && ny[k] == a->cell0_y) continue;
    if (gt_hs(h[j])) return true;
  }
  return false;
}

// Reverse of rain + friction + update for one plane element.  On entry
// (lh, lqx, lqy) hold the cotangent of the element's state after the step, on
// return that of its state before the step (the update adds the flux
// increments and the rain to the state, so the same cotangent reaches both).
// Returns whether the element was updated: only then is its cotangent also
// that of its flux increments (when the flux pass listed it).  The Manning
// derivative is added to gn, the rain-multiplier term to rsum.
template <class T>
__device__ inline bool adj_update_cell(int cls, T h0, T qx, T qy, T dh, T dqx, T dqy, T rb, T rs, T dt,
                                       const T* s_man, T& lh, T& lqx, T& lqy, double& gn, double& rsum) {
  if (cls == 255) {
    lh = lqx = lqy = T(0);
    return false;
  }
  const T rdt = rs * dt;
  const T dht = dh + rdt;
  if (!(fabs(dht) > T(0) || h0 >= T(0))) return false;  // not updated: lambda passes through
  T h1 = h0;
  const bool dry = new_depth_dry(h1, dht);
  T a_h1 = lh, a_qsx = T(0), a_qsy = T(0);
  if (!dry) {
    T a_hf = T(0), a_n = T(0);
    friction_vjp<T>(h1, s_man[cls], dt, qx + dqx, qy + dqy, lqx, lqy, a_qsx, a_qsy, a_hf, a_n);
    a_h1 += a_hf;
    gn += (double)a_n;
  }
  rsum += (double)a_h1 * (double)rb * (double)dt;
  lh = a_h1;
  lqx = a_qsx;
  lqy = a_qsy;
  return true;
}

// Reverse of rain + friction + update, kVec consecutive elements per thread
// per access.  The derivative w.r.t. the Manning value of a cell's class goes
// into the cell's FP64 running sum gn (folded per class once, by
// hp_adjoint_finish); the rain-multiplier derivative is summed per block.
template <class T>
__global__ __launch_bounds__(kAThreads) void k_adj_update(hp_grid_desc a, AdjPlanes<T> p, double* grads,
                                                          double* gn) {
  __shared__ T s_man[256];
  __shared__ T s_rb[kARateCache], s_rs[kARateCache];  // station rates: unscaled, scaled
  __shared__ double s_r[kAThreads / 64];
  const int tid = threadIdx.x;
  const double t_now = a.ctrl->t;
  const T dt = T(a.ctrl->dt);
  const T t_T = T(t_now);
  const int mode = a.rain_mode;
  for (int k = tid; k < a.n_classes; k += kAThreads) s_man[k] = ((const T*)a.manning)[k];
  const RainRows rr = rain_rows(a, t_now);
  if (mode == HP_RAIN_STATION && tid < kARateCache) {
    s_rb[tid] = station_rate<T>(rr, tid, 1.0);
    s_rs[tid] = station_rate<T>(rr, tid, a.rain_scale);
  }
  __syncthreads();
  const unsigned pitch32 = (unsigned)a.pitch;  // the plane has fewer than 2^31 elements
  const T* __restrict__ h = (const T*)a.h;
  const T* __restrict__ qx = (const T*)a.qx;
  const T* __restrict__ qy = (const T*)a.qy;
  const T* __restrict__ dh = (const T*)a.dh;
  const T* __restrict__ dqx = (const T*)a.dqx;
  const T* __restrict__ dqy = (const T*)a.dqy;
  const T scale = T(a.rain_scale);
  double rsum = 0.0;
  const int64_t total = a.pitch * (int64_t)a.ny;  // a multiple of 64
  const int64_t base = (int64_t)blockIdx.x * kAChunk;
  for (int it = 0; it < kAChunk / (kAThreads * kVec); ++it) {
    const int64_t i0 = base + ((int64_t)it * kAThreads + tid) * kVec;
    if (i0 >= total) break;
    const VecU8 vl = ldv<VecU8>(a.landuse + i0);
    const VecT<T> vh = ldv<VecT<T>>(h + i0), vqx = ldv<VecT<T>>(qx + i0), vqy = ldv<VecT<T>>(qy + i0);
    const VecT<T> vdh = ldv<VecT<T>>(dh + i0), vdqx = ldv<VecT<T>>(dqx + i0), vdqy = ldv<VecT<T>>(dqy + i0);
    VecT<T> lh = ldv<VecT<T>>(p.lh + i0), lqx = ldv<VecT<T>>(p.lqx + i0), lqy = ldv<VecT<T>>(p.lqy + i0);
    if (p.gauge_of) {  // the loss cotangent of the depths after the step
      const VecI32 vg = ldv<VecI32>(p.gauge_of + i0);
#pragma unroll
      for (int j = 0; j < kVec; ++j)
        if (vg.x[j] >= 0) lh.x[j] += T(2.0 * (p.loss_h[vg.x[j]] - p.loss_obs[vg.x[j]]));
    }
    VecD vg = ldv<VecD>(gn + i0);
    VecT<T> ch{}, cqx{}, cqy{}, rbv{}, rsc{};
    if (mode == HP_RAIN_STATION) {
      VecI16 vs{};
      if (a.station) vs = ldv<VecI16>(a.station + i0);
#pragma unroll
      for (int j = 0; j < kVec; ++j) {
        const int st = vs.x[j];
        const bool cached = st >= 0 && st < kARateCache;
        rbv.x[j] = cached ? s_rb[st] : station_rate<T>(rr, st, 1.0);
        rsc.x[j] = cached ? s_rs[st] : station_rate<T>(rr, st, a.rain_scale);
      }
    } else if (mode == HP_RAIN_FIELD) {
      const VecF32 vf = ldv<VecF32>(a.rain_field + i0);
#pragma unroll
      for (int j = 0; j < kVec; ++j) {
        rbv.x[j] = T(vf.x[j]);
        rsc.x[j] = T(vf.x[j]) * scale;
      }
    }
#pragma unroll
    for (int j = 0; j < kVec; ++j) {
      if (!adj_update_cell<T>(vl.x[j], vh.x[j], vqx.x[j], vqy.x[j], vdh.x[j], vdqx.x[j], vdqy.x[j], rbv.x[j], rsc.x[j],
                              dt, s_man, lh.x[j], lqx.x[j], lqy.x[j], vg.x[j], rsum))
        continue;
      // the kVec elements share a row (pitch is a multiple of kVec)
      const int y = (int)((unsigned)i0 / pitch32), x = (int)((unsigned)i0 - (unsigned)y * pitch32) + j;
      if (flux_listed<T>(&a, h, x, y, t_T)) {
        ch.x[j] = lh.x[j];
        cqx.x[j] = lqx.x[j];
        cqy.x[j] = lqy.x[j];
      }
    }
    stv(p.lh + i0, lh);
    stv(p.lqx + i0, lqx);
    stv(p.lqy + i0, lqy);
    stv(p.ch + i0, ch);
    stv(p.cqx + i0, cqx);
    stv(p.cqy + i0, cqy);
    stv(gn + i0, vg);
  }
  const int lane = tid & 63, w = tid >> 6;
  for (int off = 32; off > 0; off >>= 1) rsum += __shfl_xor(rsum, off, 64);
  if (lane == 0) s_r[w] = rsum;
  __syncthreads();
  if (tid == 0) {
    double v = 0.0;
    for (int k = 0; k < kAThreads / 64; ++k) v += s_r[k];
    grads[(int64_t)blockIdx.x * kGradStride + kGradStride - 1] += v;
  }
}

// grads[b][c] (c < 15) = sum of gn over block b's cells of class c: the
// per-class Manning gradients, folded once after the reverse sweep.
__global__ __launch_bounds__(kAThreads) void k_adj_finish(hp_grid_desc a, const double* __restrict__ gn,
                                                          double* grads) {
  __shared__ double s_g[kAThreads / 64][kGradStride];
  const int tid = threadIdx.x;
  double acc[kGradStride - 1];
#pragma unroll
  for (int c = 0; c < kGradStride - 1; ++c) acc[c] = 0.0;
  const int64_t total = a.pitch * (int64_t)a.ny;
  const int64_t base = (int64_t)blockIdx.x * kAChunk;
  for (int it = 0; it < kAChunk / kAThreads; ++it) {
    const int64_t i = base + (int64_t)it * kAThreads + tid;
    if (i >= total) break;
    const int cls = a.landuse[i];
    const double v = gn[i];
#pragma unroll
    for (int c = 0; c < kGradStride - 1; ++c) acc[c] += (cls == c) ? v : 0.0;
  }
  const int lane = tid & 63, w = tid >> 6;
#pragma unroll
  for (int c = 0; c < kGradStride - 1; ++c) {
    double v = acc[c];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) s_g[w][c] = v;
  }
  __syncthreads();
  if (tid < kGradStride - 1) {
    double v = 0.0;
    for (int k = 0; k < kAThreads / 64; ++k) v += s_g[k][tid];
    grads[(int64_t)blockIdx.x * kGradStride + tid] = v;
  }
}

template <class T, class DP>
__device__ inline FrameCell<T> gframe(DP a, int i, int d, T zref) {
  const T* h = (const T*)a->h;
  const T* qx = (const T*)a->qx;
  const T* qy = (const T*)a->qy;
  const T* z = (const T*)a->z;
  FrameCell<T> f;
  f.h = h[i];
  f.qn = rot_n(qx[i], qy[i], d);
  f.qt = rot_t(qx[i], qy[i], d);
  f.z = z[i] - zref;
  return f;
}

template <class T>
__device__ inline void to_global_adj(const FrameAdj<T>& f, int d, T& ax, T& ay) {
  switch (d) {
    case 0: ax = f.qn; ay = f.qt; break;
    case 1: ax = -f.qt; ay = f.qn; break;
    case 2: ax = -f.qn; ay = -f.qt; break;
    default: ax = f.qt; ay = -f.qn; break;
  }
}

// Adjoint of one edge.  DIR 0: the x-edge between columns col - 1 (A) and
// col (B) of row `row`; DIR 3: the y-edge between rows row - 1 (A) and row
// (B) of column `col`.  cot(x, y, ch, cqx, cqy) yields the cotangent of a
// valid cell's flux increments; it is asked only for A and B.  out[role]
// receives the cotangents (h, qx, qy) of the cells in roles l, A, B, rr (the
// cell before A, A, B, the cell after B); order 1 fills only A and B.
template <class T, int ORDER, int DIR, class Cot, class DP>
__device__ inline void edge_adj(DP a, bool frozen_slopes, int col, int row, Cot cot,
                                T (&out)[4][3]) {
  // The direction is a compile-time constant, so the frame rotations of the
  // interior edges fold away.
  const int nx = a->nx, ny = a->ny;
  constexpr bool xe = DIR == 0;
  int yA, xA, yB, xB, yl, xl, yr2, xr2;
  constexpr int dAB = DIR;
  if (xe) {
    yA = yB = yl = yr2 = row;
    xA = col - 1; xB = col; xl = col - 2; xr2 = col + 1;
  } else {
    xA = xB = xl = xr2 = col;
    yA = row - 1; yB = row; yl = row - 2; yr2 = row + 1;
  }
  FrameAdj<T> adj[4] = {zero_adj<T>(), zero_adj<T>(), zero_adj<T>(), zero_adj<T>()};  // l, A, B, rr
  int dir[4] = {dAB, dAB, dAB, dAB};
  // Plane offsets fit an int (the entry points check pitch * ny < 2^31); the
  // four cells lie on one line, one step apart.  The callers launch x-edges
  // only in rows inside the plane and y-edges only in columns inside it, so
  // only the coordinate across the edge needs a bounds test; an offset is
  // read only for a cell that passes it.
  const int P = (int)a->pitch;
  const int step = xe ? 1 : P;
  const int iA = yA * P + xA, iB = iA + step, il = iA - step, ir2 = iB + step;
  auto valid = [&](int x, int y, int i) -> bool {
    return (xe ? (x >= 0 && x < nx) : (y >= 0 && y < ny)) && a->type_slot[i] != 0;
  };
  const bool vA = valid(xA, yA, iA), vB = valid(xB, yB, iB);
  if (vA || vB) {
    const T dt = T(a->ctrl->dt);
    const T dx = T(a->dx);
    const T t = T(a->ctrl->t);
    const int dBA = (dAB + 2) & 3;
    const bool axis_x = (dAB == 0 || dAB == 2);
    if (vA && vB) {
      const T zref = Prec<T>::local_datum ? ((const T*)a->z)[iA] : T(0);
      const FrameCell<T> c = gframe<T>(a, iA, dAB, zref);
      const FrameCell<T> r = gframe<T>(a, iB, dAB, zref);
      const bool vl = valid(xl, yl, il), vr = valid(xr2, yr2, ir2);
      const FrameCell<T> l = vl ? gframe<T>(a, il, dAB, zref) : c;
      const FrameCell<T> rr = vr ? gframe<T>(a, ir2, dAB, zref) : r;
      T cAh, cAx, cAy, cBh, cBx, cBy;
      cot(xA, yA, cAh, cAx, cAy);
      cot(xB, yB, cBh, cBx, cBy);
      const T wx = cBx - cAx, wy = cBy - cAy;
      const T nA = T(axis_x ? nrm_x(dAB) : nrm_y(dAB));
      const T v_fh = cBh - cAh;
      const T v_fn = rot_n(wx, wy, dAB), v_ft = rot_t(wx, wy, dAB);
      const T v_bc = -(axis_x ? cAx : cAy) * nA;
      const T v_br = (axis_x ? cBx : cBy) * nA;
      T a_ec = T(0), a_er = T(0);
      solve_edge_vjp<T, ORDER>(c, l, r, rr, false, c.h + c.z, r.h + r.z, dx, dt, v_fh, v_fn, v_ft, v_bc, v_br,
                               adj[1], adj[0], adj[2], adj[3], a_ec, a_er, frozen_slopes);
      adj[1].h += a_ec;
      adj[2].h += a_er;
      if (!vl) { adj[1].h += adj[0].h; adj[1].qn += adj[0].qn; adj[1].qt += adj[0].qt; adj[0] = zero_adj<T>(); }
      if (!vr) { adj[2].h += adj[3].h; adj[2].qn += adj[3].qn; adj[2].qt += adj[3].qt; adj[3] = zero_adj<T>(); }
    } else {
      const bool from_a = vA;
      const int xc = from_a ? xA : xB, yc = from_a ? yA : yB, d = from_a ? dAB : dBA;
      const int ic = from_a ? iA : iB;
      const T zref = Prec<T>::local_datum ? ((const T*)a->z)[ic] : T(0);
      const FrameCell<T> c = gframe<T>(a, ic, d, zref);
      // boundary values at this time
      int type, id;
      decode_code(a->slot_code[a->type_slot[ic]], type, id);
      auto srow = [&](const T* tab, int rows, int w) {
        int rr_ = 0;
        while (tab && rr_ < rows - 1 && t > tab[rr_ * w]) rr_++;
        return rr_;
      };
      auto sat = [&](const T* tab, int rows, int w, int row, int col) -> T {
        return (tab == nullptr || col < 0 || col >= w || row < 0 || row >= rows) ? T(0) : tab[row * w + col];
      };
      const T* gd = (const T*)a->given_depth;
      const T* gw = (const T*)a->given_wl;
      const T* gq = (const T*)a->given_q;
      const T bd = sat(gd, a->gd_rows, a->gd_w, srow(gd, a->gd_rows, a->gd_w), id + 1);
      const T bw = sat(gw, a->gwl_rows, a->gwl_w, srow(gw, a->gwl_rows, a->gwl_w), id + 1) - zref;
      const int rq = srow(gq, a->gq_rows, a->gq_w);
      const T bqx = sat(gq, a->gq_rows, a->gq_w, rq, 1 + 2 * id);
      const T bqy = sat(gq, a->gq_rows, a->gq_w, rq, 2 + 2 * id);
      const FrameCell<T> g = bc_ghost<T>(type, c, d, bd, bw, bqx, bqy);
      const T etaG = ORDER == 1 ? g.h + g.z : ghost_wl<T>(type, c.h + c.z, g, bw);
      T cch, ccx, ccy;
      cot(xc, yc, cch, ccx, ccy);
      const T nC = T(axis_x ? nrm_x(d) : nrm_y(d));
      const T v_fh = -cch;
      const T v_fn = rot_n(-ccx, -ccy, d), v_ft = rot_t(-ccx, -ccy, d);
      const T v_bc = -(axis_x ? ccx : ccy) * nC;
      FrameAdj<T> ac = zero_adj<T>(), ag = zero_adj<T>(), u1 = zero_adj<T>(), u2 = zero_adj<T>();
      T a_ec = T(0), a_eg = T(0);
      solve_edge_vjp<T, ORDER>(c, c, g, g, true, c.h + c.z, etaG, dx, dt, v_fh, v_fn, v_ft, v_bc, T(0), ac, u1, ag,
                               u2, a_ec, a_eg);
      ghost_adj<T, ORDER>(type, d, ag, a_eg, ac);
      ac.h += a_ec;
      adj[from_a ? 1 : 2] = ac;
      dir[from_a ? 1 : 2] = d;
    }
  }
#pragma unroll
  for (int role = 0; role < 4; ++role) {
    out[role][0] = adj[role].h;
    to_global_adj<T>(adj[role], dir[role], out[role][1], out[role][2]);
  }
}

// One edge per thread.  Edges [0, nxe) are x-edges (row y, between columns
// j-1 = A and j = B, j in [0, nx]); the rest are y-edges (column x, between
// rows i-1 = A and i = B, i in [0, ny]).  Writes 12 cotangents per edge.
template <class T, int ORDER, int DIR>
__global__ __launch_bounds__(kAThreads) void k_adj_edges(hp_grid_desc a, AdjPlanes<T> p) {
  // DIR 0: x-edges, grid (ceil((nx+1)/256), ny); DIR 3: y-edges, grid
  // (ceil(nx/256), ny+1).
  const int col = blockIdx.x * kAThreads + threadIdx.x, row = blockIdx.y;
  constexpr bool xe = DIR == 0;
  if (col >= (xe ? a.nx + 1 : a.nx) || row >= (xe ? a.ny : a.ny + 1)) return;
  const int64_t P = a.pitch;
  T out[4][3];
  edge_adj<T, ORDER, DIR>(&a, p.frozen_slopes != 0, col, row,
                          [&](int x, int y, T& ch, T& cx, T& cy) {
                            const int64_t i = (int64_t)y * P + x;
                            ch = p.ch[i]; cx = p.cqx[i]; cy = p.cqy[i];
                          },
                          out);
  // order 1 reads no cell beyond the edge's two (the bed reconstruction's
  // neighbours enter through the static bed only): roles l and rr carry
  // nothing and are neither written nor gathered
  const int64_t slot = (int64_t)row * p.cpitch + col;
  T* dst = xe ? p.xc : p.yc;
  for (int role = ORDER == 1 ? 1 : 0; role < (ORDER == 1 ? 3 : 4); ++role) {
    T* base = dst + (int64_t)(role * 3) * p.cplane + slot;
    base[0] = out[role][0];
    base[p.cplane] = out[role][1];
    base[2 * p.cplane] = out[role][2];
  }
}

template <class T, int ORDER>
__global__ __launch_bounds__(kAThreads) void k_adj_gather(hp_grid_desc a, AdjPlanes<T> p) {
  const int64_t i = (int64_t)blockIdx.x * kAThreads + threadIdx.x;
  const int64_t total = a.pitch * (int64_t)a.ny;
  if (i >= total) return;
  if (a.type_slot[i] == 0) return;
  const int y = (int)(i / a.pitch), x = (int)(i - (int64_t)y * a.pitch);
  // x-edges j = x-1 (this cell is rr), x (B), x+1 (A), x+2 (l); y-edges alike
  const int64_t C = p.cpitch;
  const int64_t xs[4] = {(int64_t)y * C + (x + 2), (int64_t)y * C + (x + 1), (int64_t)y * C + x,
                         (int64_t)y * C + (x - 1)};
  const int64_t ys[4] = {(int64_t)(y + 2) * C + x, (int64_t)(y + 1) * C + x, (int64_t)y * C + x,
                         (int64_t)(y - 1) * C + x};
  T sh = T(0), sx = T(0), sy = T(0);
#pragma unroll
  for (int role = ORDER == 1 ? 1 : 0; role < (ORDER == 1 ? 3 : 4); ++role) {
    const T* bx = p.xc + (int64_t)(role * 3) * p.cplane;
    const T* by = p.yc + (int64_t)(role * 3) * p.cplane;
    sh += bx[xs[role]] + by[ys[role]];
    sx += bx[p.cplane + xs[role]] + by[p.cplane + ys[role]];
    sy += bx[2 * p.cplane + xs[role]] + by[2 * p.cplane + ys[role]];
  }
  p.lh[i] += sh;
  p.lqx[i] += sx;
  p.lqy[i] += sy;
}

// Single-pass reverse step of the order-1 scheme.  A block owns a tile of
// kATX x TY cells and, in one pass: (1) runs the update adjoint of its cells
// and of the ring of cells that share an edge with them, keeping the ring's
// flux-increment cotangents in LDS beside its own; (2) solves the adjoint of
// every edge of its cells (an edge on the tile boundary is solved by both
// tiles that touch it): the thread that owns a cell solves the cell's west
// and north edges (the cell is B of both; rows grow southwards), keeps their
// cotangents of its own cell in registers and leaves those of the west and
// north neighbours in LDS, where the threads that solve the tile's east
// column and south row of edges also leave theirs; (3) adds to each cell's
// pass-through cotangent those of its four edges, in the order k_adj_gather
// uses.  The cotangent after the step is
// read from `in` and the one before it written to `out`: the ring cells
// belong to the neighbouring tiles, so the two sets of planes must differ.
constexpr int kATX = 64;
template <class T, int TY> struct AdjTileLds {
  T cot[3][TY + 2][kATX + 2];  // flux-increment cotangents; row/column 0 and last are the ring
  T east[3][TY][kATX];   // cotangent of each cell from its east edge (the cell is that edge's A)
  T south[3][TY][kATX];  // ... from its south edge (the cell is that edge's A)
};
template <class T> struct AdjLam {
  const T* in[3];
  T* out[3];
};
template <class T> constexpr int adj_tile_rows() { return sizeof(T) == 4 ? 8 : 4; }
constexpr int kATileThreads = 256;

template <class T, int TY, int NT>
__global__ __launch_bounds__(NT, sizeof(T) == 4 ? 6 : 1) void k_adj_step_tile(hp_grid_desc a, AdjPlanes<T> p, AdjLam<T> lam,
                                                             double* grads, double* gn, int ntx) {
  static_assert(TY % (NT / kATX) == 0, "a thread owns whole rows of the tile");
  static_assert(2 * kATX + 2 * TY <= NT, "one ring cell per thread");
  constexpr int kRowsPer = TY / (NT / kATX);
  __shared__ AdjTileLds<T, TY> s;
  __shared__ T s_man[256];
  __shared__ T s_rb[kARateCache], s_rs[kARateCache];  // station rates: unscaled, scaled
  __shared__ double s_r[NT / 64];
  // the descriptor is read in place from the kernel arguments and laundered
  // between the phases, so its fields are fetched where each phase needs them
  KDesc pa = kernarg_desc();
  const int tid = threadIdx.x;
  const int tile = xcd_band(blockIdx.x, gridDim.x);
  const int x0 = (tile % ntx) * kATX, y0 = (tile / ntx) * TY;
  const double t_now = pa->ctrl->t;
  const T dt = T(pa->ctrl->dt);
  const T t_T = T(t_now);
  const int mode = pa->rain_mode;
  for (int k = tid; k < pa->n_classes; k += NT) s_man[k] = ((const T*)pa->manning)[k];
  const RainRows rr = rain_rows_at(pa, t_now);
  if (mode == HP_RAIN_STATION && tid < kARateCache) {
    s_rb[tid] = station_rate<T>(rr, tid, 1.0);
    s_rs[tid] = station_rate<T>(rr, tid, pa->rain_scale);
  }
  __syncthreads();
  const int P = (int)pa->pitch;  // planes of fewer than 2^31 elements (checked at launch)
  const int ny = pa->ny;
  const T* __restrict__ h = (const T*)pa->h;
  const T* __restrict__ qx = (const T*)pa->qx;
  const T* __restrict__ qy = (const T*)pa->qy;
  const T* __restrict__ dh = (const T*)pa->dh;
  const T* __restrict__ dqx = (const T*)pa->dqx;
  const T* __restrict__ dqy = (const T*)pa->dqy;
  const T scale = T(pa->rain_scale);
  double rsum = 0.0;

  // (1) update adjoint of the tile's cells and of its ring: first every
  // load of the thread's cells (so their memory latencies overlap), then per
  // cell the cotangent before the step into lo and the flux-increment
  // cotangent into LDS (zero outside the plane)
  struct CellLd {
    T lh, lqx, lqy, h0, qx, qy, dh, dqx, dqy;
    float f;
    int cls, st, g;
    double gn;
    bool in;
  };
  auto load = [&](int x, int y, bool owned, CellLd& L) {
    L.in = x >= 0 && x < P && y >= 0 && y < ny;
    if (!L.in) return;
    const int i = y * P + x;
    L.lh = lam.in[0][i];
    L.lqx = lam.in[1][i];
    L.lqy = lam.in[2][i];
    L.g = p.gauge_of ? p.gauge_of[i] : -1;
    L.cls = pa->landuse[i];
    L.h0 = h[i];
    L.qx = qx[i];
    L.qy = qy[i];
    L.dh = dh[i];
    L.dqx = dqx[i];
    L.dqy = dqy[i];
    L.st = (mode == HP_RAIN_STATION && pa->station) ? (int)pa->station[i] : 0;
    L.f = mode == HP_RAIN_FIELD ? pa->rain_field[i] : 0.0f;
    L.gn = owned ? gn[i] : 0.0;
  };
  auto finish = [&](int x, int y, bool owned, const CellLd& L, T(&lo)[3]) {
    T c0 = T(0), c1 = T(0), c2 = T(0);
    lo[0] = lo[1] = lo[2] = T(0);
    if (L.in) {
      const int i = y * P + x;
      T lh = L.lh, lqx = L.lqx, lqy = L.lqy;
      if (L.g >= 0) lh += T(2.0 * (p.loss_h[L.g] - p.loss_obs[L.g]));  // the loss cotangent of the depth after the step
      T rb = T(0), rs = T(0);
      if (mode == HP_RAIN_STATION) {
        const bool cached = L.st >= 0 && L.st < kARateCache;
        rb = cached ? s_rb[L.st] : station_rate<T>(rr, L.st, 1.0);
        rs = cached ? s_rs[L.st] : station_rate<T>(rr, L.st, pa->rain_scale);
      } else if (mode == HP_RAIN_FIELD) {
        rb = T(L.f);
        rs = T(L.f) * scale;
      }
      double g_acc = L.gn, r_acc = 0.0;
      if (adj_update_cell<T>(L.cls, L.h0, L.qx, L.qy, L.dh, L.dqx, L.dqy, rb, rs, dt, s_man, lh, lqx, lqy, g_acc,
                             r_acc) &&
          flux_listed<T>(pa, h, x, y, t_T)) {
        c0 = lh;
        c1 = lqx;
        c2 = lqy;
      }
      if (owned) {
        gn[i] = g_acc;
        rsum += r_acc;
      }
      lo[0] = lh;
      lo[1] = lqx;
      lo[2] = lqy;
    }
    const int ly = y - y0 + 1, lx = x - x0 + 1;
    s.cot[0][ly][lx] = c0;
    s.cot[1][ly][lx] = c1;
    s.cot[2][ly][lx] = c2;
  };
  const int lx = tid % kATX, lrow = tid / kATX;
  // the ring: row below, row above, column left, column right
  const bool ring = tid < 2 * kATX + 2 * TY;
  int hx = 0, hy = 0;
  if (tid < kATX) { hx = x0 + tid; hy = y0 - 1; }
  else if (tid < 2 * kATX) { hx = x0 + tid - kATX; hy = y0 + TY; }
  else if (tid < 2 * kATX + TY) { hx = x0 - 1; hy = y0 + tid - 2 * kATX; }
  else { hx = x0 + kATX; hy = y0 + tid - 2 * kATX - TY; }
  CellLd own[kRowsPer], rc;
#pragma unroll
  for (int k = 0; k < kRowsPer; ++k) load(x0 + lx, y0 + lrow + k * (NT / kATX), true, own[k]);
  rc.in = false;
  if (ring) load(hx, hy, false, rc);
  T lp[kRowsPer][3];
#pragma unroll
  for (int k = 0; k < kRowsPer; ++k) finish(x0 + lx, y0 + lrow + k * (NT / kATX), true, own[k], lp[k]);
  if (ring) {
    T unused[3];
    finish(hx, hy, false, rc, unused);
  }
  __syncthreads();

  // (2) edges
  pa = launder(pa);
  auto lds_cot = [&](int x, int y, T& ch, T& cx, T& cy) {
    const int ly = y - y0 + 1, lc = x - x0 + 1;
    ch = s.cot[0][ly][lc];
    cx = s.cot[1][ly][lc];
    cy = s.cot[2][ly][lc];
  };
  T wb[kRowsPer][3], nb[kRowsPer][3];  // this thread's cells from their west and north edges
#pragma unroll
  for (int k = 0; k < kRowsPer; ++k) {
    pa = launder(pa);
    const int ly = lrow + k * (NT / kATX), col = x0 + lx, row = y0 + ly;
    T out[4][3] = {};
    if (col <= pa->nx && row < ny) edge_adj<T, 1, 0>(pa, false, col, row, lds_cot, out);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      wb[k][v] = out[2][v];
      if (lx > 0) s.east[v][ly][lx - 1] = out[1][v];
    }
    T outy[4][3] = {};
    if (col < pa->nx && row <= ny) edge_adj<T, 1, 3>(pa, false, col, row, lds_cot, outy);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      nb[k][v] = outy[2][v];
      if (ly > 0) s.south[v][ly - 1][lx] = outy[1][v];
    }
  }
  // the tile's south row of y-edges (one wave) and east column of x-edges:
  // their A cells are the tile's last row and column
  pa = launder(pa);
  if (tid < kATX) {
    const int col = x0 + tid, row = y0 + TY;
    T out[4][3] = {};
    if (col < pa->nx && row <= ny) edge_adj<T, 1, 3>(pa, false, col, row, lds_cot, out);
#pragma unroll
    for (int v = 0; v < 3; ++v) s.south[v][TY - 1][tid] = out[1][v];
  } else if (tid < kATX + TY) {
    const int ly = tid - kATX, col = x0 + kATX, row = y0 + ly;
    T out[4][3] = {};
    if (col <= pa->nx && row < ny) edge_adj<T, 1, 0>(pa, false, col, row, lds_cot, out);
#pragma unroll
    for (int v = 0; v < 3; ++v) s.east[v][ly][kATX - 1] = out[1][v];
  }
  __syncthreads();

  // (3) the tile's cells: pass-through plus the four edges (as A: east and
  // south edge; as B: west and north edge)
  pa = launder(pa);
#pragma unroll
  for (int k = 0; k < kRowsPer; ++k) {
    const int ly = lrow + k * (NT / kATX), x = x0 + lx, y = y0 + ly;
    if (y >= ny || x >= P) continue;
    const int i = y * P + x;
    T o[3] = {lp[k][0], lp[k][1], lp[k][2]};
    if (pa->type_slot[i] != 0) {
#pragma unroll
      for (int v = 0; v < 3; ++v) {
        T sum = T(0);
        sum += s.east[v][ly][lx] + s.south[v][ly][lx];
        sum += wb[k][v] + nb[k][v];
        o[v] += sum;
      }
    }
    lam.out[0][i] = o[0];
    lam.out[1][i] = o[1];
    lam.out[2][i] = o[2];
  }
  const int lane = tid & 63, w = tid >> 6;
  for (int off = 32; off > 0; off >>= 1) rsum += __shfl_xor(rsum, off, 64);
  if (lane == 0) s_r[w] = rsum;
  __syncthreads();
  if (tid == 0) {
    double v = 0.0;
    for (int k = 0; k < NT / 64; ++k) v += s_r[k];
    grads[(int64_t)blockIdx.x * kGradStride + kGradStride - 1] += v;
  }
}

template <class T>
static int64_t adj_tile_count(const hp_grid_desc* g) {
  constexpr int TY = adj_tile_rows<T>();
  return ((g->pitch + kATX - 1) / kATX) * (int64_t)((g->ny + TY - 1) / TY);
}

template <class T>
static int launch_adjoint_tile(const hp_grid_desc* g, const hp_adjoint_desc* ad, void* const* in, void* const* out,
                               hipStream_t st) {
  AdjPlanes<T> p{};
  p.gauge_of = ad->gauge_of;
  p.loss_h = ad->loss_h;
  p.loss_obs = ad->loss_obs;
  AdjLam<T> lam;
  for (int v = 0; v < 3; ++v) {
    lam.in[v] = (const T*)in[v];
    lam.out[v] = (T*)out[v];
  }
  constexpr int TY = adj_tile_rows<T>();
  const int ntx = (int)((g->pitch + kATX - 1) / kATX);
  hipLaunchKernelGGL((k_adj_step_tile<T, TY, kATileThreads>), dim3((unsigned)adj_tile_count<T>(g)),
                     dim3(kATileThreads), 0, st, *g, p, lam, ad->grads, ad->gn, ntx);
  return check_launch("k_adj_step_tile");
}

template <class T>
__global__ void k_adj_loss(T* lam_h, const int64_t* __restrict__ off, const double* __restrict__ h_rec,
                           const double* __restrict__ obs, int n, double weight) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) lam_h[off[k]] += T(weight * 2.0 * (h_rec[k] - obs[k]));
}

template <class T>
static int launch_adjoint(const hp_grid_desc* g, const hp_adjoint_desc* ad, hipStream_t st) {
  AdjPlanes<T> p;
  p.lh = (T*)ad->lam_h; p.lqx = (T*)ad->lam_qx; p.lqy = (T*)ad->lam_qy;
  p.ch = (T*)ad->cot_h; p.cqx = (T*)ad->cot_qx; p.cqy = (T*)ad->cot_qy;
  p.xc = (T*)ad->xc; p.yc = (T*)ad->yc;
  p.cpitch = ad->cpitch;
  p.cplane = ad->cpitch * (int64_t)(g->ny + 1);
  p.frozen_slopes = ad->slope_mode == 1;
  p.gauge_of = ad->gauge_of;
  p.loss_h = ad->loss_h;
  p.loss_obs = ad->loss_obs;
  const int64_t total = g->pitch * (int64_t)g->ny;
  const unsigned nbu = (unsigned)((total + kAChunk - 1) / kAChunk);
  hipLaunchKernelGGL((k_adj_update<T>), dim3(nbu), dim3(kAThreads), 0, st, *g, p, ad->grads, ad->gn);
  int rc = check_launch("k_adj_update");
  if (rc) return rc;
  const dim3 gx((unsigned)((g->nx + 1 + kAThreads - 1) / kAThreads), (unsigned)g->ny);
  const dim3 gy((unsigned)((g->nx + kAThreads - 1) / kAThreads), (unsigned)(g->ny + 1));
  if (g->order == 2) {
    hipLaunchKernelGGL((k_adj_edges<T, 2, 0>), gx, dim3(kAThreads), 0, st, *g, p);
    hipLaunchKernelGGL((k_adj_edges<T, 2, 3>), gy, dim3(kAThreads), 0, st, *g, p);
  } else {
    hipLaunchKernelGGL((k_adj_edges<T, 1, 0>), gx, dim3(kAThreads), 0, st, *g, p);
    hipLaunchKernelGGL((k_adj_edges<T, 1, 3>), gy, dim3(kAThreads), 0, st, *g, p);
  }
  rc = check_launch("k_adj_edges");
  if (rc) return rc;
  const unsigned nbg = (unsigned)((total + kAThreads - 1) / kAThreads);
  if (g->order == 2) hipLaunchKernelGGL((k_adj_gather<T, 2>), dim3(nbg), dim3(kAThreads), 0, st, *g, p);
  else hipLaunchKernelGGL((k_adj_gather<T, 1>), dim3(nbg), dim3(kAThreads), 0, st, *g, p);
  return check_launch("k_adj_gather");
}

}  // namespace hp

using namespace hp;

extern "C" {

int64_t hp_adjoint_grad_blocks(const hp_grid_desc* g) {
  if (!g || g->pitch <= 0 || g->ny <= 0) return 0;
  // rows for the blocks of k_adj_update and for those of the tiled step
  const int64_t nu = (g->pitch * (int64_t)g->ny + kAChunk - 1) / kAChunk;
  const int64_t nt = g->dtype == HP_F32 ? adj_tile_count<float>(g) : adj_tile_count<double>(g);
  return nu > nt ? nu : nt;
}

static int check_adjoint(const hp_grid_desc* g, const hp_adjoint_desc* ad) {
  if (!ad) return fail(HP_EINVAL, "null adjoint descriptor");
  if (!g) return fail(HP_EINVAL, "null descriptor");
  if (!ad->lam_h || !ad->lam_qx || !ad->lam_qy || !ad->cot_h || !ad->cot_qx || !ad->cot_qy || !ad->xc || !ad->yc ||
      !ad->grads || !ad->gn)
    return fail(HP_EINVAL, "an adjoint plane pointer is null");
  if (g->pitch % kVec) return fail(HP_EINVAL, "pitch must be a multiple of 4");
  if (ad->cpitch < (int64_t)g->nx + 1) return fail(HP_EINVAL, "cpitch must be >= nx + 1");
  if (g->n_classes > kGradStride - 1) return fail(HP_EINVAL, "the adjoint supports at most 15 Manning classes");
  if (g->infil_ks) return fail(HP_EINVAL, "the adjoint does not cover infiltration");
  if (ad->slope_mode != 0 && ad->slope_mode != 1) return fail(HP_EINVAL, "slope_mode must be 0 or 1");
  if (ad->separate_passes != 0 && ad->separate_passes != 1) return fail(HP_EINVAL, "separate_passes must be 0 or 1");
  if (ad->gauge_of && (!ad->loss_h || !ad->loss_obs)) return fail(HP_EINVAL, "gauge_of needs loss_h and loss_obs");
  if (g->ny + 1 > 65535) return fail(HP_EINVAL, "the adjoint launches one block row per grid row: ny < 65535");
  if (g->pitch * (int64_t)g->ny >= (int64_t(1) << 31)) return fail(HP_EINVAL, "the adjoint covers planes of fewer than 2^31 elements");
  return HP_OK;
}

int hp_adjoint_step(const hp_grid_desc* g, const hp_adjoint_desc* ad, void* stream) {
  if (int rc = check_adjoint(g, ad)) return rc;
  if (!ad->flux_ready) {
    int rc = grid_flux_launch(g, (hipStream_t)stream);  // recompute the step's flux on U_k
    if (rc) return rc;
  }
  if (g->dtype == HP_F32) return launch_adjoint<float>(g, ad, (hipStream_t)stream);
  return launch_adjoint<double>(g, ad, (hipStream_t)stream);
}

int hp_adjoint_finish(const hp_grid_desc* g, const hp_adjoint_desc* ad, void* stream) {
  if (!g || !ad || !ad->gn || !ad->grads || !g->landuse) return fail(HP_EINVAL, "null descriptor or plane");
  const int64_t nb = hp_adjoint_grad_blocks(g);
  hipLaunchKernelGGL(k_adj_finish, dim3((unsigned)nb), dim3(kAThreads), 0, (hipStream_t)stream, *g,
                     (const double*)ad->gn, ad->grads);
  return check_launch("k_adj_finish");
}

int hp_adjoint_add_loss(int dtype, void* lam_h, const int64_t* offsets, const double* h_rec, const double* obs,
                        int32_t n, double weight, void* stream) {
  if (dtype != HP_F32 && dtype != HP_F64) return fail(HP_EINVAL, "dtype must be HP_F32 or HP_F64");
  if (n <= 0) return HP_OK;
  if (!lam_h || !offsets || !h_rec || !obs) return fail(HP_EINVAL, "null pointer");
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (dtype == HP_F32)
    hipLaunchKernelGGL(k_adj_loss<float>, dim3(nb), dim3(256), 0, (hipStream_t)stream, (float*)lam_h, offsets, h_rec,
                       obs, n, weight);
  else
    hipLaunchKernelGGL(k_adj_loss<double>, dim3(nb), dim3(256), 0, (hipStream_t)stream, (double*)lam_h, offsets,
                       h_rec, obs, n, weight);
  return check_launch("k_adj_loss");
}

// Three state (or flux) planes of a saved step, one after another.
static hp_grid_desc with_segment(hp_grid_desc d, void* u, void* f, int64_t plane_bytes) {
  if (u) {
    d.h = u;
    d.qx = (char*)u + plane_bytes;
    d.qy = (char*)u + 2 * plane_bytes;
  }
  if (f) {
    d.dh = f;
    d.dqx = (char*)f + plane_bytes;
    d.dqy = (char*)f + 2 * plane_bytes;
  }
  return d;
}

int hp_calib_forward(const hp_grid_desc* g, double* sched, const hp_ctrl* frozen, int32_t k0, int32_t nsteps,
                     void* seg_u, void* seg_d, int64_t seg_stride, const int64_t* gauge_off, int32_t n_gauges,
                     void* hist, void* stream) {
  if (!g) return fail(HP_EINVAL, "null descriptor");
  if (nsteps < 0 || k0 < 0) return fail(HP_EINVAL, "k0 and nsteps must be >= 0");
  if (g->infil_ks) return fail(HP_EINVAL, "calibration runs do not cover infiltration");
  if ((seg_u == nullptr) != (seg_d == nullptr)) return fail(HP_EINVAL, "seg_u and seg_d go together");
  const int64_t plane_bytes = g->pitch * (int64_t)g->ny * (g->dtype == HP_F32 ? 4 : 8);
  if (seg_u && seg_stride < 3 * plane_bytes) return fail(HP_EINVAL, "seg_stride smaller than three planes");
  if (!seg_u && (!g->h_out || !g->qx_out || !g->qy_out))
    return fail(HP_EINVAL, "without saved segments the steps alternate with h_out, qx_out, qy_out");
  if (hist && n_gauges > 0 && !gauge_off) return fail(HP_EINVAL, "hist needs gauge offsets");
  hipStream_t st = (hipStream_t)stream;
  void* own[3] = {g->h, g->qx, g->qy};
  void* alt[3] = {g->h_out, g->qx_out, g->qy_out};
  auto copy3 = [&](void* const* dst, void* const* src) -> int {
    for (int p = 0; p < 3; ++p)
      if (hipMemcpyAsync(dst[p], src[p], plane_bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return fail(HP_ELAUNCH, "state copy failed");
    return HP_OK;
  };
  auto seg_planes = [&](int32_t j, void** out) {
    char* b = (char*)seg_u + (int64_t)j * seg_stride;
    out[0] = b;
    out[1] = b + plane_bytes;
    out[2] = b + 2 * plane_bytes;
  };
  int rc;
  if (seg_u && nsteps > 0) {  // the state before the first step
    void* s0[3];
    seg_planes(0, s0);
    if ((rc = copy3(s0, own))) return rc;
  }
  if (!frozen && sched && nsteps > 0) {  // (t, dt) of the first step; the others follow each step
    hipLaunchKernelGGL(k_schedule_save, dim3(1), dim3(64), 0, st, g->ctrl, sched, k0);
    if ((rc = check_launch("k_schedule_save"))) return rc;
  }
  const bool gauges = hist && n_gauges > 0;
  const int64_t esz = g->dtype == HP_F32 ? 4 : 8;
  for (int32_t j = 0; j < nsteps; ++j) {
    const int32_t k = k0 + j;
    hp_grid_desc d = *g;
    if (frozen) d.ctrl = const_cast<hp_ctrl*>(frozen) + k;  // read only: no time control pass runs
    void *in[3], *out[3];
    if (seg_u) {
      // state j in the segment, state j + 1 into the next slot (the caller's
      // planes after the last step), the flux into the segment's flux slot
      seg_planes(j, in);
      if (j + 1 < nsteps) seg_planes(j + 1, out);
      else for (int p = 0; p < 3; ++p) out[p] = own[p];
      char* f = (char*)seg_d + (int64_t)j * seg_stride;
      d.dh = f;
      d.dqx = f + plane_bytes;
      d.dqy = f + 2 * plane_bytes;
    } else {
      for (int p = 0; p < 3; ++p) {
        in[p] = (j & 1) ? alt[p] : own[p];
        out[p] = (j & 1) ? own[p] : alt[p];
      }
    }
    d.h = in[0]; d.qx = in[1]; d.qy = in[2];
    d.h_out = out[0]; d.qx_out = out[1]; d.qy_out = out[2];
    double* save = (!frozen && sched && j + 1 < nsteps) ? sched : nullptr;
    const int n = gauges ? n_gauges : 0;
    char* row = gauges ? (char*)hist + (int64_t)(k + 1) * n_gauges * esz : nullptr;
    // with time control the record rides on its kernel; a frozen schedule
    // has none, so the gauges then take a small kernel of their own
    const StepRecord rec{out[0], gauge_off, n, row, save, k + 1, g->dtype == HP_F32};
    const bool ride = !frozen && (gauges || save);
    // a frozen segment (the reverse sweep's recomputation, or a frozen run's
    // last segment) needs no CFL candidates: no time control follows
    const bool time_control = !(frozen && seg_u);
    if ((rc = grid_fused_launch(&d, seg_u != nullptr, frozen == nullptr, st, ride ? &rec : nullptr, time_control)))
      return rc;
    if (frozen && gauges) {
      const unsigned nb = (unsigned)((n > 0 ? n : 1) + 255) / 256;
      if (g->dtype == HP_F32)
        hipLaunchKernelGGL(k_step_record<float>, dim3(nb), dim3(256), 0, st, (const float*)out[0], gauge_off, n,
                           (float*)row, g->ctrl, save, k + 1);
      else
        hipLaunchKernelGGL(k_step_record<double>, dim3(nb), dim3(256), 0, st, (const double*)out[0], gauge_off, n,
                           (double*)row, g->ctrl, save, k + 1);
      if ((rc = check_launch("k_step_record"))) return rc;
    }
  }
  if (!seg_u && (nsteps & 1)) return copy3(own, alt);  // the result in the caller's planes
  return HP_OK;
}

int hp_calib_reverse(const hp_grid_desc* g, const hp_adjoint_desc* ad, const hp_ctrl* frozen, int32_t k_lo,
                     int32_t k_hi, int32_t seg_k0, void* seg_u, void* seg_d, int64_t seg_stride,
                     const int64_t* gauge_off, int32_t n_gauges, const double* h_rec, const double* obs,
                     void* stream) {
  if (!g || !ad || !frozen || !seg_u || !seg_d) return fail(HP_EINVAL, "null descriptor, control blocks or segment");
  if (!ad->flux_ready) return fail(HP_EINVAL, "the reverse sweep runs on saved fluxes: flux_ready must be 1");
  if (k_lo < seg_k0 || k_hi < k_lo) return fail(HP_EINVAL, "need seg_k0 <= k_lo <= k_hi");
  const int64_t plane_bytes = g->pitch * (int64_t)g->ny * (g->dtype == HP_F32 ? 4 : 8);
  if (seg_stride < 3 * plane_bytes) return fail(HP_EINVAL, "seg_stride smaller than three planes");
  if (n_gauges > 0 && (!gauge_off || !h_rec || !obs)) return fail(HP_EINVAL, "the loss needs gauges, h_rec and obs");
  hp_adjoint_desc a = *ad;
  if (n_gauges <= 0) a.gauge_of = nullptr;
  if (a.gauge_of) {  // the rows of each step are set below
    a.loss_h = h_rec;
    a.loss_obs = obs;
  }
  if (int rc = check_adjoint(g, &a)) return rc;
  // Order 1 runs each step as one tiled pass, which reads the cotangent from
  // one set of planes and writes it to another: lam_* and cot_* alternate,
  // and the result is copied back to lam_* after an odd number of steps.
  const bool tiled = g->order == 1 && !ad->separate_passes;
  void* lam[3] = {ad->lam_h, ad->lam_qx, ad->lam_qy};
  void* alt[3] = {ad->cot_h, ad->cot_qx, ad->cot_qy};
  void** cur = lam;
  void** nxt = alt;
  hipStream_t st = (hipStream_t)stream;
  for (int32_t k = k_hi - 1; k >= k_lo; --k) {
    int rc;
    if (a.gauge_of) {  // the loss enters in the step's own update pass
      a.loss_h = h_rec + (int64_t)(k + 1) * n_gauges;
      a.loss_obs = obs + (int64_t)(k + 1) * n_gauges;
    } else if (n_gauges > 0) {
      rc = hp_adjoint_add_loss(g->dtype, cur[0], gauge_off, h_rec + (int64_t)(k + 1) * n_gauges,
                               obs + (int64_t)(k + 1) * n_gauges, n_gauges, 1.0, stream);
      if (rc) return rc;
    }
    const int64_t off = (int64_t)(k - seg_k0) * seg_stride;
    hp_grid_desc d = with_segment(*g, (char*)seg_u + off, (char*)seg_d + off, plane_bytes);
    d.ctrl = const_cast<hp_ctrl*>(frozen) + k;  // t and dt of step k; read only
    if (tiled) {
      rc = g->dtype == HP_F32 ? launch_adjoint_tile<float>(&d, &a, cur, nxt, st)
                              : launch_adjoint_tile<double>(&d, &a, cur, nxt, st);
      void** t = cur;
      cur = nxt;
      nxt = t;
    } else {
      rc = hp_adjoint_step(&d, &a, stream);
    }
    if (rc) return rc;
  }
  if (cur != lam) {
    for (int v = 0; v < 3; ++v)
      if (hipMemcpyAsync(lam[v], cur[v], plane_bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return fail(HP_ELAUNCH, "cotangent copy failed");
  }
  return HP_OK;
}

}  // extern "C"
