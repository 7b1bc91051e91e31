/*
 * hipims_abi.h - C ABI of the MI355X (gfx950) HiPIMS shallow-water step.
 *
 * Every entry point takes caller-owned device pointers, plain sizes and an
 * explicit hipStream_t (passed as void*), launches asynchronously on that
 * stream, never synchronises and never allocates device memory (the one
 * measurement entry, hp_grid_step_fused_timed, waits for its own events).
 * Each returns HP_OK (0) or a negative HP_E* code; hp_last_error() then holds
 * a message for the calling thread.
 *
 * Two families:
 *  1. The reference's torch-extension op surface, on the reference's compact
 *     valid-cell layout (arrays of N valid cells in row-major raster order and
 *     the int32[5*N] neighbour table of SWE_CUDA.py:173-222).  Each function
 *     names the reference op it replaces.
 *  2. The structured-grid path: state held as 2-D planes with a row pitch,
 *     advanced by a stencil flux pass and a fused update pass with device-
 *     resident time control (no host round trip per step).  These replace the
 *     sequence Godunov.addFlux -> addStation_PrecipitationSource ->
 *     time_friction_euler_update_cuda (SWE_CUDA.py:339-424, 497-502, 780-831).
 */
#ifndef HIPIMS_ABI_H
#define HIPIMS_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HP_ABI_VERSION 10

enum hp_status {
  HP_OK = 0,
  HP_EINVAL = -1,   /* bad argument (null pointer, size, dtype, order) */
  HP_ELAUNCH = -2,  /* the HIP runtime rejected a launch */
  HP_ESPACE = -3    /* caller-provided workspace too small */
};

enum hp_dtype { HP_F32 = 0, HP_F64 = 1 };

/* ABI version and the calling thread's last error message. */
int hp_abi_version(void);
const char* hp_last_error(void);
/* sizeof(hp_ctrl), sizeof(hp_grid_desc), sizeof(hp_ensemble) into out[0..2],
 * so bindings can check their mirrors of these structs. */
void hp_struct_sizes(int64_t* out);

/* ------------------------------------------------------------------------
 * 1. Reference op surface (compact layout).  `dtype` selects float/double for
 *    every floating-point pointer of the call; scalars dx, t, dt, CFL are
 *    1-element device arrays as in the reference.
 * ------------------------------------------------------------------------ */

/* Replaces fluxMask.update (CHANS-HiPIMS/hipims/cuda/fluxMaskGenerator.cpp:24-35,
 * kernel fluxMaskGenerator_Kernel.cu:19-42).  mask: uint8[N] (torch.bool). */
int hp_flux_mask(int dtype, int64_t N, uint8_t* mask, const void* h, const int32_t* index, const void* t,
                 void* stream);

/* Replaces fluxCalculation_jh_modified_surface.addFlux (order 1,
 * fluxCal_jh_modified_surface.cpp:28-57) and fluxCal_2ndOrder_jh_improved.addFlux
 * (order 2, fluxCal_2ndOrder_jh_improved.cpp:28-57).  Accumulates into
 * dh/dqx/dqy for the M cells listed in wet.  normal: [4][2] (may be null:
 * the E, N, W, S unit normals).  given_*: row-major series tables. */
int hp_flux_compact(int dtype, int order, int64_t M, int64_t N, const int32_t* wet, void* dh, void* dqx, void* dqy,
                    const void* h, const void* wl, const void* z, const void* qx, const void* qy,
                    const int32_t* index, const void* normal, const void* given_depth, int32_t gd_rows,
                    int32_t gd_w, const void* given_wl, int32_t gwl_rows, int32_t gwl_w, const void* given_q,
                    int32_t gq_rows, int32_t gq_w, const void* dx, const void* t, const void* dt, void* stream);

/* Replaces station_PrecipitationCalculation.addStation_Precipitation
 * (stationPrecipitation_interface.cpp:26-43).  rain is float64 when
 * rain_is_f64 (the reference driver always builds it so, SWE_CUDA.py:483-489),
 * else of `dtype`. */
int hp_station_rain(int dtype, int64_t N, void* dh, const int16_t* station, const void* rain, int rain_is_f64,
                    const void* dt, void* stream);

/* Replaces friction_implicit_andUpdate_jh.addFriction_eulerUpdate
 * (friction_implicit_andUpdate_jh_interface.cpp:29-55). */
int hp_friction_update_compact(int dtype, int64_t K, const int32_t* wet, void* dh, void* dqx, void* dqy, void* dz,
                               const uint8_t* landuse, void* h, void* wl, void* qx, void* qy, void* z,
                               const void* manning, const void* dt, void* stream);

/* Replaces timeControl.updateTimestep (timeControl.cpp:27-45): h_max and the
 * per-listed-cell CFL candidate accel[K]. */
int hp_time_control_compact(int dtype, int64_t K, const int32_t* wet, void* accel, void* h_max, const void* h,
                            const void* qx, const void* qy, const void* dx, const void* cfl, void* stream);

/* Replaces euler_update.update (euler_update.cpp:27-47). */
int hp_euler_update_compact(int dtype, int64_t K, const int32_t* wet, const void* dh, const void* dqx,
                            const void* dqy, const void* dz, void* h, void* wl, void* z, void* qx, void* qy,
                            void* stream);

/* Replaces infiltrationCalculation.addinfiltration (infiltration_interface.cpp,
 * kernel infiltrationCUDA_Kernel.cu:20-46): Green-Ampt over the M cells of
 * wet; per landuse class saturated conductivity ks, capillary head phi and
 * water-content deficit dtheta; F is the cumulative infiltrated depth;
 * dh -= dF, F += dF. */
int hp_infiltration_compact(int dtype, int64_t M, const int32_t* wet, void* dh, const uint8_t* landuse,
                            const void* h, const void* ks, const void* phi, const void* dtheta, void* F,
                            const void* dt, void* stream);

/* ------------------------------------------------------------------------
 * 2. Structured-grid path.
 * ------------------------------------------------------------------------ */

/* Device-resident time control and mass accounting, all float64 so a host
 * can view it as double[16].  The first two fields are the reference's
 * Godunov.t and Godunov.dt. */
typedef struct hp_ctrl {
  double t;              /* simulated time                                   */
  double dt;             /* step used by the next flux / update              */
  double export_T;       /* export interval (Export_timeStep)                */
  double export_n;       /* number of export boundaries crossed              */
  double export_pending; /* boundaries crossed since the host last cleared it */
  double volume;         /* sum of h over valid cells after the last update  */
  double rain_volume;    /* cumulative sum of rain*dt (depth units)          */
  double flux_volume;    /* cumulative sum of flux dh (net boundary inflow)  */
  double steps;          /* completed steps                                  */
  double t_stop;         /* > 0: the domain's passes do nothing once t >= t_stop */
  double stopped_synced; /* 1: stopped, and the single-pass step's second set of
                            planes holds the state (set by the time control)   */
  double reserved[5];
} hp_ctrl;

/* Rain source of the update pass. */
enum hp_rain_mode {
  HP_RAIN_NONE = 0,
  HP_RAIN_STATION = 1, /* station table [rain_rows][rain_cols] = [t, r_0..r_S-1] m/s,
                          linear in time, 0 after the last row (SWE_CUDA.py:472-495);
                          station plane int16 (null: every cell station 0)     */
  HP_RAIN_FIELD = 2    /* per-cell float32 plane in m/s (addRadarRainfall)        */
};

/* One structured domain.  Planes are ny rows of `pitch` elements; columns
 * [nx, pitch) are padding.  Raster row r, column c of the reference's input
 * rasters is plane element r*pitch + c.  The raster must have nodata on its
 * outer ring (the reference's neighbour table wraps otherwise). */
typedef struct hp_grid_desc {
  int32_t dtype;  /* HP_F32 / HP_F64 */
  int32_t order;  /* 1: jh modified surface, 2: MUSCL */
  int32_t nx, ny; /* raster columns / rows */
  int64_t pitch;  /* elements per plane row, >= nx */
  double dx, cfl;
  int32_t cell0_x, cell0_y; /* first valid cell in row-major order (-1: none) */
  int32_t n_classes;        /* entries of manning */
  int32_t rain_mode;        /* hp_rain_mode */
  int32_t tile_rows;        /* flux unit height: float 2, 4, 6, 8, 10, 12 or 16, double 4 or 8 (other: 8 / 4) */
  int32_t mass_sums;        /* 1: keep hp_ctrl's FP64 volume / rain / boundary-flux sums up to
                               date (diagnostics the reference does not compute); 0: skip them */
  /* state planes (dtype) */
  void *h, *qx, *qy, *z, *h_max;
  /* flux accumulators written by the flux pass (dtype) */
  void *dh, *dqx, *dqy;
  /* second state buffers of the single-pass step (dtype; may be null otherwise) */
  void *h_out, *qx_out, *qy_out;
  /* cell class planes */
  const uint8_t* type_slot; /* 0 nodata, else slot of slot_code */
  const uint8_t* landuse;   /* Manning class, 255 for nodata */
  const int32_t* slot_code; /* [256] mask code of each slot (device) */
  const void* manning;      /* [n_classes] (dtype) */
  /* boundary series (dtype), as Godunov._given_depth/_given_wl/_given_discharge */
  const void* given_depth;
  const void* given_wl;
  const void* given_q;
  int32_t gd_rows, gd_w, gwl_rows, gwl_w, gq_rows, gq_w;
  /* rain */
  const int16_t* station;
  const double* rain_table;
  int32_t rain_rows, rain_cols;
  const float* rain_field;
  double rain_scale; /* multiplies every rain source (1.0 = reference) */
  /* control block and reduction workspace (device) */
  hp_ctrl* ctrl;
  void* work;
  int64_t work_bytes;
  /* Green-Ampt infiltration (Godunov.addInfiltrationSource, SWE_CUDA.py:439-443);
   * infil_ks == NULL disables it.  ks, phi, dtheta: [n_classes] (dtype) per
   * landuse class; infil_F: cumulative infiltrated depth plane (dtype);
   * infil_dF: the step's infiltrated depth plane (dtype), written by
   * hp_grid_infiltration; an update pass subtracts it when non-null. */
  const void* infil_ks;
  const void* infil_phi;
  const void* infil_dtheta;
  void* infil_F;
  void* infil_dF;
} hp_grid_desc;

/* Bytes of device workspace `work` must provide for this domain. */
int64_t hp_grid_workspace_bytes(const hp_grid_desc* g);

/* Flux pass: dh/dqx/dqy = edge fluxes and bed-slope sources of every listed
 * cell (Godunov.addFlux with the wet list of fluxMask, SWE_CUDA.py:370-423). */
int hp_grid_flux(const hp_grid_desc* g, void* stream);

/* Update pass + time control: rain source, implicit friction and update,
 * h_max, CFL minimum, dt/t advance and export clamp on the device
 * (SWE_CUDA.py:497-502 then :780-831). */
int hp_grid_update(const hp_grid_desc* g, void* stream);

/* Which reference driver method an update pass stands for. */
enum hp_update_mode {
  HP_UPDATE_STEP = 0,     /* time_friction_euler_update_cuda (SWE_CUDA.py:780-831) */
  HP_UPDATE_FRICTION = 1, /* friction_euler_update_cuda (:766-778): no time control */
  HP_UPDATE_EULER = 2,    /* euler_update (:712-724): no friction, no time control */
  HP_UPDATE_CFL = 3       /* time_update_cuda (:726-764): time control only */
};

/* The update pass in the given mode (hp_grid_update == HP_UPDATE_STEP). */
int hp_grid_update_mode(const hp_grid_desc* g, int mode, void* stream);

/* nsteps single-pass steps: flux, rain, friction, update, h_max and the CFL
 * minimum of each unit of cells in one kernel reading (h, qx, qy) and writing
 * (h_out, qx_out, qy_out), then the time-control block; the two buffers swap
 * every step and the result is left in (h, qx, qy).  Same results as
 * hp_grid_flux + hp_grid_update (HP_UPDATE_STEP); dh/dqx/dqy are not written. */
int hp_grid_step_fused(const hp_grid_desc* g, int32_t nsteps, void* stream);

/* Allocation-free form of hp_grid_step_fused (no reference counterpart):
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels, with the per-tile partial records in a library-owned device
 * array of hp_grid_noalloc_max_bytes() bytes (8 MiB, held per GPU until the
 * process ends) instead of g->work, which is ignored.  A domain needing more
 * (over about 200 000 flux units; 8192 x 8192 float cells fit at the
 * default unit height) is rejected with HP_ESPACE.  Calls that overlap on one
 * GPU share the array and race: serialise them. */
int hp_grid_step_fused_noalloc(const hp_grid_desc* g, int32_t nsteps, void* stream);
int64_t hp_grid_noalloc_max_bytes(void);

/* Measurement form of hp_grid_step_fused (no reference counterpart; the
 * reference times nothing on this path): nsteps single-pass steps exactly
 * as hp_grid_step_fused runs them, with HIP events recorded on `stream`
 * before the first step and after the last, and around the single-pass
 * kernel of every sample_every-th step (an event between two kernels delays
 * the second by a few microseconds, so only the sampled steps carry that
 * cost).  Synchronises on the last event, then writes the mean duration of
 * the sampled single-pass kernels to *kernel_ms and the duration of all
 * nsteps steps to *step_ms, so both come from the same launches (the state
 * copy after an odd count falls outside both). */
int hp_grid_step_fused_timed(const hp_grid_desc* g, int32_t nsteps, int32_t sample_every, float* kernel_ms,
                             float* step_ms, void* stream);

/* Infiltration pass: cells listed by the flux rule of the current state
 * (the wet list of the flux pass) take the Green-Ampt increment; infil_dF
 * receives it (0 elsewhere) and infil_F accumulates it
 * (infiltrationCUDA_Kernel.cu:20-46 on the structured grid). */
int hp_grid_infiltration(const hp_grid_desc* g, void* stream);

/* nsteps x (flux, [infiltration when infil_ks is set], update). */
int hp_grid_step(const hp_grid_desc* g, int32_t nsteps, void* stream);

/* ---- Ensemble: many domains advanced by the same launches, each with its own
 * dt and time control (no reference counterpart; the reference runs one
 * domain per process).  All domains share dtype, order and tile_rows. */
typedef struct hp_ensemble {
  int32_t n_domains, dtype, order, tile_rows;
  int32_t n_flux_tiles, n_update_blocks;
  const void* doms;         /* device hp_grid_desc[n_domains]          */
  const void* flux_map;     /* device int32[2 * n_flux_tiles]          */
  const void* update_map;   /* device int32[2 * n_update_blocks]       */
  const void* finalize_map; /* device int32[2 * n_domains]             */
  void* partials;           /* device reduction records                */
  int32_t single_pass;      /* 1: one single-pass launch per step      */
  const void* fused_finalize_map; /* device int32[2 * n_domains]       */
} hp_ensemble;

/* Device workspace bytes hp_ensemble_init needs for these domains. */
int64_t hp_ensemble_workspace_bytes(int32_t n, const hp_grid_desc* descs);
/* Copy the descriptors and block maps into `work` (a one-off upload that
 * synchronises `stream`) and fill the handle.  The domains' planes must stay
 * allocated while the handle is used.  single_pass 1: every domain needs
 * h_out, qx_out, qy_out, and each step is one single-pass launch over every
 * tile of every domain, alternating the two sets of planes. */
int hp_ensemble_init(int32_t n, const hp_grid_desc* descs, int32_t single_pass, void* work, int64_t work_bytes,
                     hp_ensemble* out, void* stream);
/* nsteps x (flux of every domain, update of every domain, per-domain time
 * control): three launches per step; single-pass: two (step of every domain,
 * per-domain time control), and nsteps must be even so that the state ends
 * in the domains' own planes.  Graph-capturable. */
int hp_ensemble_step(const hp_ensemble* e, int32_t nsteps, void* stream);

/* ---- Adjoint of one step with a frozen time-step schedule (no reference
 * counterpart).  Planes lam_* hold the cotangent of the state after the step
 * on entry and of the state before it on return; the state planes of the
 * domain must hold the state before the step, and its control block the t
 * and dt of that step.  cot_* are scratch planes like h; xc / yc are 12
 * planes each of (ny + 1) rows x cpitch (cpitch >= nx + 1) for the edge
 * contributions.  grads: FP64 [hp_adjoint_grad_blocks][16]: column 15 the
 * running sums of the gradient w.r.t. the rain multiplier (rain_scale),
 * columns c < 15 the gradient w.r.t. the Manning value of class c, written
 * by hp_adjoint_finish from gn, the per-cell FP64 running sum (a plane like
 * h) of the derivative w.r.t. the Manning value of the cell's class.
 * flux_ready 1: dh/dqx/dqy already hold the step's flux (saved when the
 * step was recomputed), so hp_adjoint_step does not recompute it.
 * slope_mode 0 differentiates the minmod-limited MUSCL slopes exactly (the
 * exact derivative of the step); 1 holds the limited slopes at the run's
 * own values (the derivative of the model whose faces are the cell value
 * plus those slopes, one-sided where the run sits on a branch point of that
 * model).  Order-2 runs amplify rounding differences, so in either mode the
 * derivative holds over a few hundred steps, not over long horizons.
 * gauge_of (may be null): an int32 plane like h holding, at each gauge cell,
 * the gauge's index g (-1 elsewhere; each cell at most one gauge); with it
 * the step first adds the squared-misfit cotangent 2 (loss_h[g] -
 * loss_obs[g]) (FP64 rows of the depths after the step and the observations)
 * to lam_h at the gauge cells, in the same pass.
 * separate_passes (hp_calib_reverse only): 0 runs each order-1 step as one
 * tiled kernel (update, edge and gather adjoints in one pass, with cot_*
 * and lam_* taking turns as the cotangent planes; xc / yc unused); 1 runs the
 * separate passes of hp_adjoint_step.  Order 2 always uses the latter. */
typedef struct hp_adjoint_desc {
  void *lam_h, *lam_qx, *lam_qy;
  void *cot_h, *cot_qx, *cot_qy;
  void *xc, *yc;
  int64_t cpitch;
  double* grads;
  double* gn;
  int32_t slope_mode;
  int32_t flux_ready;
  const int32_t* gauge_of;
  const double* loss_h;
  const double* loss_obs;
  int32_t separate_passes;
} hp_adjoint_desc;

int64_t hp_adjoint_grad_blocks(const hp_grid_desc* g);
int hp_adjoint_step(const hp_grid_desc* g, const hp_adjoint_desc* a, void* stream);
/* Fold gn into the per-class columns of grads (once, after the reverse sweep). */
int hp_adjoint_finish(const hp_grid_desc* g, const hp_adjoint_desc* a, void* stream);
/* lam_h[offsets[k]] += weight * 2 * (h_rec[k] - obs[k]), k < n: the
 * cotangent of a squared-misfit loss at n gauge cells. */
int hp_adjoint_add_loss(int dtype, void* lam_h, const int64_t* offsets, const double* h_rec, const double* obs,
                        int32_t n, double weight, void* stream);

/* ---- Calibration loops (no reference counterpart): the per-step sequences
 * of a gauge-misfit calibration run natively, one call per segment, on the
 * single-pass step.
 *
 * hp_calib_forward runs steps k0 .. k0+nsteps-1.  frozen null: the model's
 * own CFL steps on g->ctrl, with (t, dt) before step k saved to sched[k]
 * when sched (device FP64 [K+1][2]) is given.  frozen non-null: step k
 * reads its t and dt from the control block frozen[k] (device, read only)
 * and no time control pass runs.  seg_u / seg_d (both null or both set):
 * the state before step k is kept at seg_u + (k-k0)*seg_stride bytes (three
 * planes of pitch*ny elements one after another; the first is copied from
 * h, qx, qy, the others are written by the steps) and the step's flux at
 * seg_d + (k-k0)*seg_stride; the state after the last step goes to h, qx,
 * qy; with a frozen schedule the saved steps skip the CFL candidates and
 * mass sums, which no time control reads.  Without segments the steps
 * alternate between h, qx, qy and h_out, qx_out, qy_out and end in h, qx,
 * qy.  hist (may be null): the depths at
 * the n_gauges plane offsets after step k go to row k+1 (grid dtype). */
int hp_calib_forward(const hp_grid_desc* g, double* sched, const hp_ctrl* frozen, int32_t k0, int32_t nsteps,
                     void* seg_u, void* seg_d, int64_t seg_stride, const int64_t* gauge_off, int32_t n_gauges,
                     void* hist, void* stream);
/* hp_calib_reverse runs the reverse sweep over steps k_hi-1 down to k_lo:
 * per step the loss cotangent of the depths after it (row k+1 of h_rec and
 * obs, FP64 [K+1][n_gauges]; inside the step's update pass when a->gauge_of
 * is set, else through hp_adjoint_add_loss) and the step's adjoint (one
 * tiled pass for order 1 unless a->separate_passes, else hp_adjoint_step)
 * with the control block frozen[k] on the state and flux saved at
 * seg_u / seg_d + (k - seg_k0)*seg_stride bytes (flux_ready must be 1).  On
 * return lam_* hold the cotangent of the state before step k_lo; cot_* are
 * overwritten. */
int hp_calib_reverse(const hp_grid_desc* g, const hp_adjoint_desc* a, const hp_ctrl* frozen, int32_t k_lo,
                     int32_t k_hi, int32_t seg_k0, void* seg_u, void* seg_d, int64_t seg_stride,
                     const int64_t* gauge_off, int32_t n_gauges, const double* h_rec, const double* obs,
                     void* stream);

/* ---- VALU issue-rate probe (no reference counterpart): `blocks` blocks of
 * 256 lanes each run hp_probe_fma_per_lane() dependent-free FP32 FMAs and
 * write one float to out[block * 256 + lane]; timing a launch calibrates the
 * vector-instruction roofline of the flux kernels. */
int64_t hp_probe_fma_per_lane(void);
int hp_probe_fma(int32_t blocks, float* out, void* stream);

/* out[i] = plane[offsets[i]]  (compact gather, Godunov.get_h). */
int hp_gather(int dtype, const void* plane, const int64_t* offsets, int64_t n, void* out, void* stream);
/* plane[offsets[i]] = in[i]  (compact scatter). */
int hp_scatter(int dtype, void* plane, const int64_t* offsets, int64_t n, const void* in, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HIPIMS_ABI_H */
