/*
 * godunov_five_eq.h -- C ABI of the MI355X-native five-equation Godunov block step.
 *
 * Drop-in boundary for the reference's GodunovImplemBase plugin surface
 * (scheme/GodunovImplemBase.h:59-198) and the solver hooks next to it
 * (SolverGodunovFiveEq.cpp:368-390 compute_dt_local, :481-526 godunov_unsplit,
 * :978-1019 fill_outside_quadrants).  Plain C types only: no torch, no C++ types.
 *
 * Data layout (both host and device buffers): U[block][var][cell], doubles.
 *   - block: the owned blocks of a brick of nb[0] x nb[1] (x nb[2]) blocks, lexicographic with x
 *     fastest;
 *   - var: the FiveEq ids of models/FiveEq.h:35-50 (2-D: ID=0, IE=1, IU=2, IV=3, IA0=4, IAD0=5,
 *     IA1=6, IAD1=7; 3-D: IW=4 shifts IA0..IAD1 to 5..8);
 *   - cell: bsize^dim cells of the block, x fastest.
 *
 * Every entry point returns G5_OK (0) or an error code; g5_last_error() gives the message.
 * No exception crosses the ABI.  A context is not thread-safe; it owns its device buffers and
 * one HIP stream.
 */
#ifndef GODUNOV_FIVE_EQ_H_
#define GODUNOV_FIVE_EQ_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
enum
{
  G5_OK = 0,
  G5_ERR_INVALID = 1,     /* bad argument or configuration */
  G5_ERR_HIP = 2,         /* a HIP runtime call failed */
  G5_ERR_UNSUPPORTED = 3, /* valid configuration this build does not implement */
  G5_ERR_NOMEM = 4,       /* allocation failed or HBM budget exceeded */
  G5_ERR_STATE = 5,       /* call out of order (e.g. step before upload) */
  G5_ERR_IO = 6,          /* file input/output failed (godunov_five_eq_io.h) */
  G5_ERR_DIVERGED = 7,    /* a solve failed (G5_FAIL_* bits): it was stopped at the failing state */
  G5_ERR_TOO_LARGE = 8    /* the solve needs more blocks than the whole pool holds (never fits) */
};

/* why a solve was stopped (bits of g5_pool_failure).  The step kernels and the dt reduction check
 * every owned cell of each new state: a value that is not finite, or a solve without any finite
 * CFL rate (an infinite or undefined dt), sets G5_FAIL_NONFINITE; with
 * g5_config.abort_negative_pressure a negative mixture pressure sets G5_FAIL_NEGATIVE_PRESSURE
 * ([hydro] abort_when_negative_pressure, models/HydroSettings.cpp:28).  A cell whose sound speed is
 * not real (c^2 < 0, which needs p < 0) only drops out of the CFL maximum, as in the reference.
 * The solve then stops stepping; its co-batched solves go on. */
enum
{
  G5_FAIL_NONFINITE = 1,
  G5_FAIL_NEGATIVE_PRESSURE = 2
};

/* refine indicators ([amr] refine_criterion; scheme/ComputeRefineFlags.cpp:41-62) */
enum
{
  G5_INDICATOR_LOHNER_SPLIT = 0,
  G5_INDICATOR_LOHNER_UNSPLIT = 1,
  G5_INDICATOR_SIMPLE_GRADIENT = 2
};

/* per-block refine flags (the AMR context's flag values, combined by maximum over a block's
 * cells: any cell that asks for refinement refines the block, all must agree to coarsen it) */
enum
{
  G5_AMR_FLAG_INIT = 0,
  G5_AMR_COARSEN = 1,
  G5_AMR_NOTHING = 2,
  G5_AMR_REFINE = 3
};

/* border conditions, border_conditions/FillOutside.h:51-55 (BC_HYDRO) */
enum
{
  G5_BC_NONE = 0,
  G5_BC_PERIODIC = 1,
  G5_BC_ZERO_GRADIENT = 2,
  G5_BC_WALL = 3
};

/* time integrators, SolverGodunovFiveEq.cpp:484-524 */
enum
{
  G5_TI_HANCOCK = 0,
  G5_TI_RK1 = 1,
  G5_TI_RK2_SSP = 2
};

/* arithmetic of the step */
enum
{
  G5_PREC_FP64 = 0,      /* everything in FP64, in the reference's operation order (bit-exact) */
  G5_PREC_MIXED = 1,     /* FP32 slopes, wave speeds and THINC; FP64 state, EOS pressure recovery, dt */
  G5_PREC_FP64_FAST = 2  /* everything in FP64, reassociated: products by one refined reciprocal for
                            repeated divisors, refined hardware square roots, no e(p) -> p(e) identity
                            round trip in HLLC; relative L1 <= 1e-12 per conserved variable against
                            G5_PREC_FP64 */
};

/* volume-fraction source term factor st_k of the update (alpha_k source st_k u* alpha_k) */
enum
{
  G5_SOURCE_ALLAIRE_KOKH = 0, /* st_k = 1 */
  G5_SOURCE_MILLER_PUCKETT = 1 /* st_k = 1 - K/K_k: K_k = gamma_k (p + pinf_k) of the stiffened gas,
                                  K = 1/(alpha_0/K_0 + alpha_1/K_1) (Wood), p from the cell's
                                  time-n state */
};

/* inflow overrides of the outside cells, applied after the border fill
 * (SolverGodunovFiveEq.cpp:997-1018) */
enum
{
  G5_INLET_NONE = 0,
  G5_INLET_SHOCK_BUBBLE = 1, /* every outside cell with x < xmin (border_conditions/BCShockBubble.h:62-74) */
  G5_INLET_JET = 2           /* outside cells with x < xmin whose centre is in the jet (BCJet.h:71-88) */
};

/*
 * Solve configuration.  Field meaning follows the reference .ini keys:
 * [hydro] cfl slope_type smallr smallc (models/HydroSettings.cpp:18-29),
 * [material{0,1}_mie_gruneisen_{ig,sg}] gamma pinf, [THINC] enabled beta epsilon,
 * [mesh] boundary_type_{x,y,z}{min,max}, [p4est_connectivity] periodic_{x,y,z},
 * [amr] bx/by/bz (= bsize) and time_integrator.
 * dx is the cell length (compute_cell_length(level, bx) * scaling_factor).
 */
typedef struct g5_config
{
  int32_t dim;         /* 2 or 3 */
  int32_t bsize;       /* cells per block edge; the HIP step kernels take 16 */
  int32_t nb[3];       /* blocks per direction (nb[2] = 1 in 2-D) */
  int32_t periodic[3]; /* brick periodicity per direction */
  int32_t bc[6];       /* G5_BC_* for faces xmin, xmax, ymin, ymax, zmin, zmax */
  int32_t integrator;  /* G5_TI_* */
  int32_t thinc_enabled;
  int32_t precision; /* G5_PREC_* */
  double  dx;
  double  cfl;
  double  slope_type; /* 1 minmod, 2 MC */
  double  smallr;
  double  smallc;
  double  gamma[2]; /* per material */
  double  pinf[2];  /* per material (0 for an ideal gas) */
  double  thinc_beta;
  double  thinc_epsilon;
  double  xmin[3]; /* domain origin, used by the initial-condition generators and the inlets */
  /* inflow override: G5_INLET_*; the overridden outside cells take the conservative state
   * inlet_u (variable ids of models/FiveEq.h, nvar values); jet = (centre y, centre z, width):
   * 2-D |y - jet[0]| < jet[2]/2, 3-D (y - jet[0])^2 + (z - jet[1])^2 < jet[2]^2/4 at the cell
   * centre ([jet] y, z, width) */
  int32_t inlet;
  int32_t inlet_pad;
  double  inlet_u[9];
  double  jet[3];
  /* volume-fraction source term of the update ([volume_fraction_advection] source_term_type,
   * scheme/ReadFluxesAndConservativeUpdateFunctor.h:412-443): G5_SOURCE_* */
  int32_t source_term;
  /* 1: a negative pressure in a new state stops the solve with G5_FAIL_NEGATIVE_PRESSURE
   * ([hydro] abort_when_negative_pressure, models/HydroSettings.cpp:28, default true there) */
  int32_t abort_negative_pressure;
} g5_config;

typedef struct g5_ctx g5_ctx;

/* ---- library ---- */

/* version string of this build */
const char * g5_version(void);

/* message of the last failed call on ctx (or of the last failed g5_create when ctx is NULL) */
const char * g5_last_error(const g5_ctx * ctx);

/* ---- context life cycle ---- */

/* Replaces SolverGodunovFiveEq's allocation of U/U2 (SolverGodunovFiveEq.cpp:147-157) and
 * GodunovImplemV0::create + resize_auxiliary_data (scheme/GodunovImplemV0.h:103-114,
 * scheme/GodunovImplemV0.cpp:35-61): validates cfg, allocates the device state and scratch. */
int g5_create(const g5_config * cfg, g5_ctx ** out);

/* A two-level static mesh (2-D or 3-D, FP64 or FP64_FAST): cfg's brick at level 0, with the
 * n_refined blocks at refined_coords[dim k .. dim k + dim) each replaced by 2^dim blocks of half the
 * cell length (level 1).  Refined blocks lie at least one block inside the brick's border.  The
 * reference's AMR paths on such a mesh: the coarse/fine ghost fill of
 * ConvertToPrimitivesVariablesFunctor::fill_ghosts (scheme/ConvertToPrimitivesVariablesFunctor.cpp:
 * 884-918) -- restriction (:214-250) of a finer neighbour, prolongation of a coarser one,
 * prolongation 0 SIMPLE_COPY, 1 EXTRAPOLATE_LINEAR_MINMOD (:553-770), 2 or 3
 * CONSERVATIVE_INTERPOLATION_ORDER_2 / _4 (:339-490,751-827) -- the fine-face flux sums of
 * read_fluxes_and_update_{2d,3d} (scheme/ReadFluxesAndConservativeUpdateFunctor.cpp:137-199,
 * .h:256-307) and the per-level cell length in the fluxes and the CFL rate
 * (ComputeFluxesAndStoreFunctor, ComputeDtFiveEqFunctor).  The state holds the blocks in the order
 * of g5_block_layout: the coarse leaves lexicographic, then the fine blocks lexicographic over the
 * fine grid of 2nb blocks per direction (the mesh itself, which the reference takes from its AMR
 * manager, is fixed for the context's life: no regridding). */
int g5_create_two_level(const g5_config * cfg, const int32_t * refined_coords, int32_t n_refined,
                        int32_t prolongation, g5_ctx ** out);

/* the blocks of the context's state in order: 4 int32 per block (x, y, z, level), coordinates at
 * the block's level; n = 4 * n_blocks */
int g5_block_layout(const g5_ctx * ctx, int32_t * coords, size_t n);

int g5_destroy(g5_ctx * ctx);

/* number of owned blocks, variables per cell and cells per block */
int g5_shape(const g5_ctx * ctx, int32_t * n_blocks, int32_t * nvar, int32_t * cells_per_block);

/* Replaces GodunovImplemBase::total_mem_size_in_bytes (scheme/GodunovImplemBase.h:97-99):
 * device bytes held by the context (state and scratch). */
int g5_mem_bytes(const g5_ctx * ctx, uint64_t * bytes);

/* ---- state transfer (the solver's U, SolverGodunovFiveEq.h:349-358) ---- */

/* n = number of doubles, must equal n_blocks * nvar * cells_per_block */
int g5_upload_state(g5_ctx * ctx, const double * U_host, size_t n);
int g5_download_state(g5_ctx * ctx, double * U_host, size_t n);

/* device pointer of the current state (valid until the next step), in the block order of
 * g5_set_block_order (lexicographic when it was not called) */
int g5_state_device_ptr(g5_ctx * ctx, double ** U_dev);

/* The caller's block order: coords[3k..3k+2] are the brick coordinates of the caller's block k
 * (z ignored in 2-D; a permutation of the brick's blocks, e.g. the Morton order of the AMR
 * manager's orchard keys).  From then on g5_upload_state, g5_download_state, g5_state_device_ptr
 * and the *_buffers entry points use U[block][var][cell] with the blocks in this order, as the
 * reference's DataArrayBlock (cell, var, block) holds them (scheme/GodunovImplemBase.h:93-107,
 * scheme/GodunovImplemV0.cpp:265-275; the neighbours that StencilHelper finds by key,
 * ConvertToPrimitivesVariablesFunctor.cpp:876-888, come from the coordinates).  The current state
 * moves with its blocks. */
int g5_set_block_order(g5_ctx * ctx, const int32_t * coords, int32_t n_blocks);

/* Replaces GodunovImplemV0::do_time_step(U, U2, dt) (scheme/GodunovImplemV0.cpp:83-151) on the
 * caller's device arrays: U2 <- one step of U (RK2_SSP: both stages and the combination,
 * SolverGodunovFiveEq.cpp:496-524).  U and U2 hold the owned blocks in the context's block order;
 * the caller keeps them (its U <- U2 copy of godunov_unsplit, SolverGodunovFiveEq.cpp:494, stays
 * the caller's).  Asynchronous on the context's stream: the caller orders its own work on the
 * arrays against g5_synchronize. */
int g5_do_time_step_buffers(g5_ctx * ctx, const double * U_dev, double * U2_dev, double dt);

/* compute_dt_local (SolverGodunovFiveEq.cpp:368-390) on the caller's device array U (context block
 * order); G5_ERR_DIVERGED for a non-finite (or, with the check, negative-pressure) state.
 * Synchronises the context's stream. */
int g5_compute_dt_buffer(g5_ctx * ctx, const double * U_dev, double * dt);

/* parameters of g5_refine_flags (RefineIndicatorData as SolverGodunovFiveEq::compute_refine_flags
 * fills it, SolverGodunovFiveEq.cpp:1036-1057) */
typedef struct g5_refine_params
{
  int32_t indicator;         /* G5_INDICATOR_* ([amr] refine_criterion) */
  int32_t level_min;         /* [amr] level_min */
  int32_t level_max;         /* [amr] level_max */
  int32_t base_level;        /* AMR level of the context's blocks (of its level 0 on a two-level mesh) */
  double  refine_threshold;  /* [amr] epsilon_refine (the reference's default 0.01) */
  double  coarsen_threshold; /* [amr] epsilon_coarsen (default 0.00125) */
  double  epsilon_lohner;    /* [amr] epsilon_lohner (default 0.02) */
} g5_refine_params;

/* Replaces ComputeRefineFlags::run (scheme/ComputeRefineFlags.cpp:18-62) on the context's current
 * state: the indicator of the partial densities IAD0 and IAD1 in every cell, each value's vote
 * (refine above refine_threshold below level_max, coarsen under coarsen_threshold above
 * level_min, else nothing; update_flag, :236-249) and the maximum over the block, max-accumulated
 * into flags[k] (G5_AMR_*; the caller resets them, as the reference's caller does).  Neighbour
 * values follow getNeighborDataSameLevel (:67-96): outside a non-periodic border the cell itself,
 * a finer neighbour's average, a coarser neighbour's cell.  indicator_max (may be NULL) receives
 * per block the largest indicator value (NaN values -- 0/0 where a partial density vanishes with
 * its neighbours -- count in no comparison and no maximum).  n = the number of blocks, in the
 * context's block order.  Synchronises the context's stream. */
int g5_refine_flags(g5_ctx * ctx, const g5_refine_params * params, int32_t * flags, double * indicator_max,
                    size_t n);

/* The step of g5_do_time_step_buffers followed by the dt of its result (g5_compute_dt_buffer on U2),
 * without a context and without allocating: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  U and U2 hold the blocks of cfg's brick in lexicographic order; *dt_next
 * receives cfl / (CFL maximum of U2), *failure (optional) the G5_FAIL_* bits of U2 (then
 * G5_ERR_DIVERGED).  The solve's descriptor, control record, launch list and slot table live in a
 * file-scope device array of 641 KiB per GPU (held until the process ends), so a call takes at most
 * 32768 blocks (G5_ERR_TOO_LARGE beyond) and HANCOCK or RK1 (RK2_SSP would need a stage buffer of
 * the whole state: G5_ERR_UNSUPPORTED).  Runs on the null stream and synchronises it.  Every call
 * on a GPU shares that GPU's array: calls that overlap on one GPU (from several host threads
 * without the library's own lock, or several processes) race; calls from one process are
 * serialised by a lock. */
int g5_step_noalloc(const g5_config * cfg, const double * U_dev, double * U2_dev, double dt, double * dt_next,
                    int32_t * failure);

/* ---- the step ---- */

/* Replaces SolverGodunovFiveEq::compute_dt_local (SolverGodunovFiveEq.cpp:368-390) and
 * ComputeDtFiveEqFunctor::apply (scheme/ComputeDtFiveEqFunctor.cpp:45-81):
 * *dt = cfl / max over owned cells of sum_d (c + |u_d|)/dx.  Synchronises the stream. */
int g5_compute_dt(g5_ctx * ctx, double * dt);

/* Replaces SolverGodunovFiveEq::fill_outside_quadrants (SolverGodunovFiveEq.cpp:978-1019,
 * FillOutsideCellFunctor border_conditions/FillOutside.cpp:42-170): gathers the outside-block ring
 * into out (n = n_outside * nvar * cells_per_block doubles, blocks in the order of
 * g5_outside_coords) so callers can inspect it.  The step kernels resolve the same border
 * conditions on the fly and do not need this call. */
int g5_fill_outside(g5_ctx * ctx, double * out_host, size_t n);

/* number of outside blocks and their brick coordinates (3 int32 per block) */
int g5_outside_coords(const g5_ctx * ctx, int32_t * n_outside, int32_t * coords, size_t ncoords);

/* Replaces GodunovImplemV0::do_time_step(U, U2, dt) (scheme/GodunovImplemV0.cpp:83-151) together
 * with the HANCOCK/RK1 branch of godunov_unsplit (SolverGodunovFiveEq.cpp:484-495): border fill,
 * primitives, slopes, per-direction fluxes and updates, volume-fraction normalisation, U <- U2.
 * The RK2_SSP integrator runs both stages (SolverGodunovFiveEq.cpp:496-524).  Asynchronous. */
int g5_do_time_step(g5_ctx * ctx, double dt);

/* Runs up to nsteps iterations of next_iteration_impl (SolverGodunovFiveEq.cpp:395-476) with the
 * dt computed and clamped to tEnd on the device (no host round trip per step; the step sequence
 * is captured into a HIP graph).  *t is advanced; *steps_done receives the count.  Synchronises. */
int g5_run(g5_ctx * ctx, int32_t nsteps, double * t, double tEnd, int32_t * steps_done);

/* HIP-event duration (stream time, ms) of the step loop of the last g5_run call: nsteps fused
 * step kernels plus their one-thread dt finalisation kernels */
int g5_last_run_ms(const g5_ctx * ctx, double * ms);

/* wait for all work of the context */
int g5_synchronize(g5_ctx * ctx);

/* Time the fused step kernel on the context's stream with HIP events: runs nsteps steps with a
 * fixed dt and reports the mean kernel duration in milliseconds (used by the benchmark). */
int g5_time_step_kernel(g5_ctx * ctx, int32_t nsteps, double dt, double * mean_ms);

/* ---- initial conditions (init/InitShockBubble.cpp:50-225, common.h:68-144) ---- */

/* One region's primitive state, the [regionN] keys of common.h:80-109. */
typedef struct g5_region
{
  double alpha0;
  double alpha_rho[2];
  double p;
  double u, v, w;
} g5_region;

/* Conservative state of a region (the inflow state of the inlets, ...): replaces
 * get_region_init_state (common.h:68-144).  U receives nvar values. */
int g5_region_conservative(const g5_config * cfg, const g5_region * region, double * U, size_t n);

/* Shock-bubble initial condition of InitShockBubbleDataFunctor: region 0 post-shock
 * (x <= x_front), region 1 pre-shock, region 2 bubble of radius r at center c, with the bubble
 * volume fraction of each cell computed from the sphere.  Fills U_host (owned blocks). */
int g5_init_shock_bubble(const g5_config * cfg, const g5_region * regions3, double x_front,
                         const double * center3, double radius, double * U_host, size_t n);

/* Triple-point initial condition of InitTriplePointDataFunctor (init/InitTriplePoint.cpp:30-90):
 * region 0 for x <= xd, else region 1 where y <= yd (sqrt(y^2 + z^2) <= yd in 3-D), region 2
 * elsewhere, at cell centres.  Fills U_host (owned blocks); the same states as the device
 * generator g5_pool_init_triple_point. */
int g5_init_triple_point(const g5_config * cfg, const g5_region * regions3, double xd, double yd, double * U_host,
                         size_t n);

/* Post-shock state of a planar shock moving into a resting stiffened gas (Rankine-Hugoniot);
 * for p_inf = 0 this is app/tools/moving_shock.py:79-163.  Either mach > 0 (shock Mach number) or
 * p2 > 0 (post-shock pressure) is given; outputs rho2, u2, p2 and the shock speed. */
int g5_rankine_hugoniot(double gamma, double pinf, double rho1, double p1, double mach,
                        double p2_in, double * rho2, double * u2, double * p2, double * shock_speed);

/* ---- pools: many independent solves stepped together (batching and serving) ---- */

/* A pool holds the state of many independent solves of one dimension, THINC flag and precision
 * in device memory sized by an HBM budget (two state copies).  The blocks of all live solves form
 * one launch list; every solve keeps its own dt, time, step budget and buffer parity on the
 * device, so solves can be admitted and released while others are in flight (continuous
 * batching).  No counterpart in the reference (one solve per process there); the per-solve step
 * is the same GodunovImplemV0::do_time_step + compute_dt_local iteration as g5_run. */
typedef struct g5_pool g5_pool;

int g5_pool_create(int32_t dim, int32_t thinc_enabled, int32_t precision, uint64_t hbm_budget_bytes,
                   g5_pool ** out);
/* A pool that also holds an RK2_SSP stage buffer (three state buffers per slot of the budget
 * instead of two), so that it admits RK2_SSP solves (SolverGodunovFiveEq.cpp:496-524); like the
 * other integrators, one integrator per pool. */
int g5_pool_create_rk2(int32_t dim, int32_t thinc_enabled, int32_t precision, uint64_t hbm_budget_bytes,
                       g5_pool ** out);
int g5_pool_destroy(g5_pool * pool);
const char * g5_pool_last_error(const g5_pool * pool);

/* reserve memory for one solve (G5_ERR_NOMEM when the budget has no room now, G5_ERR_TOO_LARGE
 * when the solve has more blocks than the whole budget holds); it starts inactive.
 * G5_ERR_INVALID when dim, THINC, precision or integrator differ from the pool's live solves;
 * G5_ERR_UNSUPPORTED for RK2_SSP in a pool without the stage buffer (g5_pool_create_rk2). */
int g5_pool_admit(g5_pool * pool, const g5_config * cfg, double t_end, int32_t max_steps, int32_t * solve_id);

/* reserve memory for one two-level solve (the mesh of g5_create_two_level: cfg's brick with the
 * n_refined blocks at refined_coords[dim k .. dim k + dim) refined, the given prolongation),
 * stepped in the same launch list as the pool's other solves; its state uses the block order of
 * g5_pool_block_layout.  No reference counterpart beyond g5_create_two_level's: batching is this
 * library's. */
int g5_pool_admit_two_level(g5_pool * pool, const g5_config * cfg, const int32_t * refined_coords, int32_t n_refined,
                            int32_t prolongation, double t_end, int32_t max_steps, int32_t * solve_id);

/* number of blocks of a solve and, when coords is not NULL, their (x, y, z, level), 4 int32 per
 * block in the order of the solve's state (n = 4 * n_blocks) */
int g5_pool_block_layout(const g5_pool * pool, int32_t solve_id, int32_t * coords, size_t n, int32_t * n_blocks);

/* free the solve's memory and id */
int g5_pool_release(g5_pool * pool, int32_t solve_id);

/* state of one solve in the g5_upload_state layout */
int g5_pool_upload(g5_pool * pool, int32_t solve_id, const double * U, size_t n);
int g5_pool_download(g5_pool * pool, int32_t solve_id, double * U, size_t n);

/* device initial conditions (init/InitShockBubble.cpp:50-225, init/InitTriplePoint.cpp:46-133):
 * shape 0: sphere/circle of radius size3[0] at center3; shape 1: box of half-sizes size3 */
int g5_pool_init_shock_bubble(g5_pool * pool, int32_t solve_id, const g5_region * regions3, int32_t shape,
                              double x_front, const double * center3, const double * size3);
int g5_pool_init_triple_point(g5_pool * pool, int32_t solve_id, const g5_region * regions3, double xd, double yd);

/* compute the first dt of solves whose state is set and mark them active */
int g5_pool_start(g5_pool * pool, const int32_t * solve_ids, int32_t n);

/* nsteps device-dt steps of every active solve (HIP graphs of 8 steps, then of single steps:
 * fused step kernel + finalise per step) */
int g5_pool_step(g5_pool * pool, int32_t nsteps);

/* time, steps done and activity of one solve (synchronises).  G5_ERR_DIVERGED (outputs still
 * filled) when the solve was stopped by a failure (g5_pool_failure). */
int g5_pool_status(g5_pool * pool, int32_t solve_id, double * t, int32_t * steps, int32_t * active);

/* G5_FAIL_* bits of one solve (0 while it is healthy; synchronises).  No reference counterpart:
 * there a failing run aborts the process (Kokkos::abort). */
int g5_pool_failure(g5_pool * pool, int32_t solve_id, int32_t * failure);

/* Resume support (checkpoint/restart and eviction of a solve to host memory): set the time and
 * the number of steps already done of an admitted solve before g5_pool_start.  A solve saved as
 * (g5_pool_download state, g5_pool_status t and steps) and re-admitted with the same
 * configuration, end time and step budget, uploaded, given its clock and started continues
 * bit-identically (its first dt is recomputed from the state, as the step would have).
 * Replaces the restart path of SolverGodunovFiveEq.cpp:811-846 / init/init.cpp:132-207 for
 * one solve; the on-disk format is the caller's (the Python layer writes .npy + JSON). */
int g5_pool_set_clock(g5_pool * pool, int32_t solve_id, double t, int32_t steps_done);

/* live solves that reached t_end or their step budget, or were stopped by a failure (see
 * g5_pool_status): up to cap ids, *n_done = their number */
int g5_pool_finished(g5_pool * pool, int32_t * solve_ids, int32_t cap, int32_t * n_done);

/* blocks in use, block capacity of the budget, live solves */
int g5_pool_usage(const g5_pool * pool, int64_t * used_blocks, int64_t * capacity_blocks, int32_t * live_solves);

/* enqueue nsteps steps on the pool's stream and return without waiting (g5_pool_finished,
 * g5_pool_status, g5_pool_download and g5_pool_last_step_ms wait for them) */
int g5_pool_step_async(g5_pool * pool, int32_t nsteps);

/* HIP-event duration of the last stepping call (ms; waits for it) */
int g5_pool_last_step_ms(g5_pool * pool, double * ms);

/* Launch geometry of the step kernels.  No reference counterpart: the reference launches Kokkos
 * team policies whose shape is fixed by the block size (scheme/GodunovImplemV0.cpp:83-151).
 * Every geometry computes the same bits; only the time differs.  By default a pool tunes it on
 * its first step over a launch list of a new size class (floor(log2(blocks))): each candidate
 * runs the step kernel of that list a few times (the kernel only writes the next state and the
 * next CFL rate, so the extra launches change nothing) and the fastest is kept for the process,
 * per dimension, THINC flag, precision, integrator and size class.  G5_AUTOTUNE=0 in the
 * environment turns tuning off (default geometry: one workgroup per block, XCD ranges, z split;
 * in 3-D z-columns of 8, 4, 2 or 1 blocks, the longest that leave two per CU). */
typedef struct g5_launch
{
  int32_t wg_per_cu; /* 0: one workgroup per block; 1..16: persistent workgroups per CU (2-D only) */
  int32_t xcd_remap; /* 1: each XCD takes a contiguous range of the block list */
  int32_t zsplit;    /* 3-D: z Riemann solve split around the x/y flux exchange (1) or not (0) */
  int32_t ring_wave; /* 2-D, wg_per_cu 0: a fifth wavefront per workgroup for the ring cells (1) */
  int32_t col_len;   /* 3-D: blocks one workgroup marches through in z, 1, 2, 4 or 8 (one-level lists) */
  int32_t autotuned; /* out: 1 when the geometry came from the autotuner */
  double  tuned_ms;  /* out: best step-kernel time measured by the autotuner (ms) */
} g5_launch;

/* Pinned (page-locked) host memory for state transfers: g5_upload_state / g5_download_state /
 * g5_pool_upload / g5_pool_download run at full PCIe rate from and to such buffers.  No reference
 * counterpart (Kokkos deep_copy to host mirrors, SolverGodunovFiveEq.cpp:811-846). */
int g5_host_alloc(size_t bytes, void ** ptr);
int g5_host_free(void * ptr);

/* build the launch list of the live solves, tune its launch geometry and capture the step
 * graph now, so that the next g5_pool_step does none of it (no state changes) */
int g5_pool_prepare(g5_pool * pool);

/* dimension, THINC flag and precision of a pool */
int g5_pool_properties(const g5_pool * pool, int32_t * dim, int32_t * thinc_enabled, int32_t * precision);

/* fix the geometry (launch != NULL; autotuning is then bypassed) or return to autotuning (NULL) */
int g5_pool_set_launch(g5_pool * pool, const g5_launch * launch);
int g5_pool_get_launch(const g5_pool * pool, g5_launch * launch);
/* enable (1) or disable (0) autotuning for this pool */
int g5_pool_set_autotune(g5_pool * pool, int32_t enable);

/* ---- seeded workloads of the batching and serving forms ---- */

/* One independent solve: 2-D or 3-D grid of 16^d blocks with unit height (dx = 1/(16 * the
 * smallest transverse block count)), an initial-condition family and two stiffened-gas
 * materials: material 1 surrounds and carries a shock of Mach `mach` into the pre-shock state
 * (rho1, p1); material 0 forms the inclusion (density rho0 at p1).  Shock-bubble
 * (init/InitShockBubble.cpp:50-225): shock front at x = 0.2, inclusion centred at
 * (0.55, 0.5, 0.5), a circle/sphere of radius 0.25 (shape 0) or a box of half-sizes
 * (0.2, 0.3, 0.3) (shape 1), ZERO_GRADIENT in x and walls elsewhere, minmod.  Triple point
 * (init/InitTriplePoint.cpp:46-133): post-shock state for x <= 1/3, pre-shock material 1 for
 * y <= 0.5 (sqrt(y^2+z^2) <= 0.5 in 3-D), material 0 elsewhere, ZERO_GRADIENT borders, MC slopes.
 * HANCOCK, HLLC, CFL 0.5.  No reference counterpart: the reference reads one such setup from an
 * .ini file per run. */
typedef struct g5_solve_spec
{
  int32_t dim;
  int32_t nb[3];
  int32_t kind;  /* 0 shock-bubble, 1 triple point */
  int32_t shape; /* shock-bubble inclusion: 0 circle/sphere, 1 rectangle/box */
  int32_t steps; /* step budget */
  int32_t pad;
  double  gamma[2];
  double  pinf[2];
  double  mach;
  double  rho1, p1, rho0;
  double  arrival; /* serving: arrival time (s) after the start */
} g5_solve_spec;

/* Batching workload: solve `index` of the seeded set.  Short side uniform in {64,128,256,512},
 * aspect ratio uniform in [1,4] (long side rounded to 16), gamma uniform in [1.1,4.4] for both
 * materials, pinf of material 1 in {0, 6e8} (6e8: rho1 1000, p1 1e5, rho0 in [1,10]; 0: rho1 1,
 * p1 1, rho0 in [0.1,5]), Mach uniform in [1.1,3], family and shape each with probability 1/2,
 * 100 steps.  Generator: std::mt19937_64 seeded per (seed, index) by splitmix64. */
int g5_batch_spec(uint64_t seed, int32_t index, g5_solve_spec * out);

/* Serving workload: n requests with Poisson arrivals at rate_hz; a fraction frac_3d are 3-D with
 * sides uniform in multiples of 16 in [64,128], the rest 2-D with sides in [64,1024]; 50-500
 * steps; physics drawn as for the batching workload. */
int g5_serve_specs(uint64_t seed, int32_t n, double rate_hz, double frac_3d, g5_solve_spec * out);

/* the configuration and the three region states (post-shock, pre-shock, inclusion) of a spec */
int g5_spec_config(const g5_solve_spec * spec, int32_t thinc_enabled, int32_t precision, g5_config * cfg);
int g5_spec_regions(const g5_solve_spec * spec, g5_region * regions3);

/* admit a spec into a pool (step budget = spec->steps, no end time) and generate its initial
 * condition on the device; G5_ERR_NOMEM when the pool's budget cannot hold it now */
int g5_pool_admit_spec(g5_pool * pool, const g5_solve_spec * spec, int32_t * solve_id);

/* Serving request loop (continuous batching).  Requests are admitted first come first served
 * into a 2-D and a 3-D pool that split the HBM budget (share_3d to the 3-D pool) as soon as
 * memory frees up, every in-flight solve is stepped together (steps_per_poll device-dt steps per
 * loop turn, one HIP graph per step and pool), finished solves are downloaded and released.
 * Reports throughput (cell-updates of completed solves / wall time) and the latency of each
 * request (result on the host - arrival). */
typedef struct g5_serve_config
{
  uint64_t seed;
  int32_t  n_requests;
  int32_t  steps_per_poll;
  double   rate_hz;
  double   frac_3d;
  double   share_3d;
  uint64_t hbm_budget_bytes;
  int32_t  precision;
  int32_t  thinc_enabled;
  int32_t  autotune; /* 0: default launch geometry (no tuning trials while serving) */
  int32_t  pad;
} g5_serve_config;

typedef struct g5_serve_report
{
  int32_t completed;
  int32_t max_inflight;
  double  wall_s;
  double  cell_updates;
  double  cell_updates_per_s;
  double  p50_ms, p99_ms, mean_ms, max_ms;
  double  last_arrival_s;
  double  step_s_2d, step_s_3d; /* HIP-event time of the stepping calls of each pool */
  /* host wall time in the loop's phases: admission (incl. device initial conditions), first dt,
   * waiting for the enqueued steps, collection of finished solves (download and release) */
  double  admit_s, start_s, wait_s, collect_s;
  int32_t failed;   /* requests whose solve was stopped by a failure (G5_FAIL_*) */
  int32_t rejected; /* requests larger than their pool's whole budget (never admissible) */
} g5_serve_report;

/* latencies_ms (optional, cap entries) receives the latency of each request in arrival order */
int g5_serve(const g5_serve_config * config, g5_serve_report * report, double * latencies_ms, size_t cap);

/* outcome of one served request */
enum
{
  G5_SERVE_DONE = 0,     /* reached its step budget */
  G5_SERVE_FAILED = 1,   /* stopped by a failure (failure holds the G5_FAIL_* bits) */
  G5_SERVE_REJECTED = 2  /* more blocks than its pool's whole budget: never admitted */
};

typedef struct g5_serve_result
{
  int32_t outcome; /* G5_SERVE_* */
  int32_t failure; /* G5_FAIL_* bits of a failed solve */
  int32_t steps;   /* steps done */
  int32_t pad;
  double  t;          /* final time of the solve */
  double  latency_ms; /* result on the host (or rejection) - arrival */
} g5_serve_result;

/* The serving loop of g5_serve over given requests, with per-request results.
 * specs: config->n_requests requests in order of non-decreasing arrival (NULL: drawn by
 * g5_serve_specs from config->seed, rate_hz and frac_3d); results (optional): one entry per
 * request; states (optional): per request a host buffer of g5_spec_state_size doubles (or NULL)
 * that receives the request's final state in the g5_upload_state layout -- so that each served
 * result can be checked against the same solve run alone.  Each request is admitted with its step
 * budget (spec->steps) and no end time, and computes the same bits as when run alone. */
int g5_serve_run(const g5_serve_config * config, const g5_solve_spec * specs, g5_serve_report * report,
                 g5_serve_result * results, double * const * states);

/* doubles of the state of a spec's solve (blocks * nvar * cells per block) */
int g5_spec_state_size(const g5_solve_spec * spec, size_t * n);

#ifdef __cplusplus
}
#endif

#endif /* GODUNOV_FIVE_EQ_H_ */
