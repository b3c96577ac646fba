/*
 * aerojax.h -- C ABI of the MI355X (gfx950) implementation of AeroJAX's
 * collocated projection-method timestep.
 *
 * The reference (arriemeijer-creator/AeroJAX) is pure Python/JAX; its
 * "plugin API" for this path is two Python call signatures resolved by
 * import:
 *   rk_step_unified   advection_schemes/rk3_simple_new.py:175-180
 *   poisson_multigrid pressure_solvers/multigrid_solver.py:55-56
 * composed by step_pure (solver/pure_functions/step.py:49-467).  Each entry
 * point below replaces one jitted JAX expression of that path; the Python
 * package aerojax_amd binds them through ctypes and keeps the reference
 * signatures.
 *
 * Conventions
 *   - Every field is (nx, ny) row-major with y contiguous (index i*ny + j),
 *     i = 0 inlet row, i = nx-1 outlet row, j = 0 / ny-1 wall columns
 *     (solver/params.py:54-56, indexing='ij').
 *   - Pointers are device pointers (hipMalloc / torch tensors); buffers are
 *     caller-owned, outputs must not alias inputs.  `dtype` selects the
 *     element type of every field argument: AJ_F32 (float) or AJ_F64
 *     (double).
 *   - `stream` is a hipStream_t (NULL = default stream).  Calls are
 *     stream-ordered and asynchronous unless stated otherwise.
 *   - Return value: AJ_OK (0) or an AJ_ERR_* code; aj_last_error() gives the
 *     message of the last failure on the calling thread.
 */
#ifndef AEROJAX_H_
#define AEROJAX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  AJ_OK = 0,
  AJ_ERR_INVALID = 1,     /* bad argument (null pointer, size, dtype)      */
  AJ_ERR_GRID = 2,        /* grid not suitable for multigrid               */
  AJ_ERR_HIP = 3,         /* HIP runtime error                             */
  AJ_ERR_NONFINITE = 4,   /* NaN/Inf detected                              */
  AJ_ERR_UNSUPPORTED = 5  /* option outside this implementation            */
};

enum { AJ_F32 = 0, AJ_F64 = 1 };

/* Library version string, e.g. "aerojax-amd 0.1.0 gfx950". */
const char* aj_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char* aj_last_error(void);

/* ---- RK predictor ------------------------------------------------------
 * Replaces rk_step_unified (advection_schemes/rk3_simple_new.py:175-358),
 * flow_type='von_karman': RK3 (fast_mode=0) or RK2/Heun (fast_mode=1) of
 * first-order upwind advection + nu*Laplacian (+ nu_sgs, + hyperviscosity
 * -nu*ratio*dx*dy*lap(lap)), implicit Brinkman factor 1 + dt*chi/eta with
 * chi = sigmoid(-sdf/(0.1 dx)) (sdf != NULL) or 1 - mask, hard mask
 * (mask > 0.5) and inlet/outlet/wall BCs.  All stages run in one fused
 * kernel on LDS tiles.  sdf and nu_sgs may be NULL.
 */
typedef struct aj_rk_params {
  double dt, nu, dx, dy, U_inf, brinkman_eta, nu_hyper_ratio;
  int slip_walls; /* 1 = slip walls (reference default), 0 = no-slip */
  int fast_mode;  /* 0 = RK3, 1 = RK2 */
} aj_rk_params;

int aj_rk_step(int dtype, const void* u, const void* v, const void* mask,
               const void* sdf, const void* nu_sgs, void* u_out, void* v_out,
               int nx, int ny, const aj_rk_params* prm, void* stream);

/* The same RK step with the pressure right-hand side of step_pure in its
 * epilogue: rhs = divergence_nonperiodic(u*, v*) [* mask] / prm->dt
 * (solver/operators.py:81-96, solver/pure_functions/step.py:182-199; the
 * predictor BCs step.py re-imposes are the RK step's own last step), zero
 * on every boundary cell; rhs_masked = 0 gives div / dt (the GUI step,
 * step_handlers.py:92-98).  One launch instead of aj_rk_step +
 * aj_divergence_rhs, with their results.  Hyperviscosity (nu_hyper_ratio >
 * 0) is not supported here (AJ_ERR_UNSUPPORTED). */
int aj_rk_step_rhs(int dtype, const void* u, const void* v, const void* mask,
                   const void* sdf, const void* nu_sgs, void* u_out, void* v_out, void* rhs,
                   int rhs_masked, int nx, int ny, const aj_rk_params* prm, void* stream);

/* Same step with the reference's `nu` argument given as a field: the GUI
 * step passes nu_total = nu + nu_sgs as `nu` together with nu_sgs
 * (solver/simulation/step_handlers.py:76-90), so the viscosity of the
 * Laplacian is nu_field + nu_sgs and the biharmonic coefficient
 * nu_field * nu_hyper_ratio * dx * dy (rk3_simple_new.py:304-309).
 * prm->nu is ignored. */
int aj_rk_step_nu_field(int dtype, const void* u, const void* v, const void* mask,
                        const void* sdf, const void* nu_field, const void* nu_sgs,
                        void* u_out, void* v_out, int nx, int ny, const aj_rk_params* prm,
                        void* stream);

/* ---- Smagorinsky eddy viscosity ----------------------------------------
 * Replaces constant_smagorinsky + compute_strain_rate
 * (solver/les_models.py:22-54, 179-200): nu_sgs = (c_s*delta)^2 *
 * sqrt(2(Sxx^2+Syy^2+2Sxy^2) + 1e-10), periodic central differences.
 */
int aj_les_smagorinsky(int dtype, const void* u, const void* v, void* nu_sgs,
                       int nx, int ny, double dx, double dy, double delta,
                       double c_s, void* stream);
/* Same, also writing nu_total = nu + nu_sgs (step_handlers.py:67-77). */
int aj_les_smagorinsky_total(int dtype, const void* u, const void* v, double nu,
                             void* nu_sgs, void* nu_total, int nx, int ny, double dx,
                             double dy, double delta, double c_s, void* stream);

/* Replaces dynamic_smagorinsky (solver/les_models.py:57-176), collocated:
 * 5x5 periodic test filter, Leonard stress and M_ij with alpha (reference
 * default 2), 3x3 interior least-squares averages, C_d = clip(LM / (MM +
 * 1e-8), 0, 0.2) (0 on the border), nu_sgs = C_d delta^2 |S|.  nu_total (=
 * nu + nu_sgs, the GUI step's viscosity) and c_d may be NULL.  Uses a
 * stream-ordered scratch allocation of 14 fields. */
int aj_les_dynamic_smagorinsky(int dtype, const void* u, const void* v, double nu, void* nu_sgs,
                               void* nu_total, void* c_d, int nx, int ny, double dx, double dy,
                               double delta, double alpha, void* stream);

/* ---- Pressure right-hand side -----------------------------------------
 * Replaces divergence_nonperiodic (solver/operators.py:81-96) followed by
 * rhs = div*mask/dt (solver/pure_functions/step.py:194-199).  mask == NULL
 * and dt == 1 give the bare divergence.
 */
int aj_divergence_rhs(int dtype, const void* u_star, const void* v_star,
                      const void* mask, void* rhs, int nx, int ny, double dx,
                      double dy, double dt, void* stream);
/* Replaces divergence (solver/operators.py:76-78): periodic central
 * differences, the field the reference's training-data writer stores
 * (neural_operators/training_utils.py:66-69). */
int aj_divergence_periodic(int dtype, const void* u, const void* v, void* div, int nx, int ny,
                           double dx, double dy, void* stream);

/* ---- Projection ---------------------------------------------------------
 * Replaces the projection tail of step_pure (step.py:257-264, 305-349,
 * 364-377, 394-397): grad p, tangential projection at interface cells
 * (|grad mask| > 0.1/dx), dp/dx = 0 on the inlet row, u = u* - dt grad p,
 * mask, slip/inlet/outlet BCs, mask.
 */
int aj_project(int dtype, const void* u_star, const void* v_star, const void* p,
               const void* mask, void* u_out, void* v_out, int nx, int ny,
               double dt, double dx, double dy, double U_inf, void* stream);

/* ---- Projection of the GUI step ----------------------------------------
 * Replaces the pressure correction and BCs of _step_collocated
 * (solver/simulation/step_handlers.py:131-176, von Karman branch):
 * grad_x/y_nonperiodic of p, dp_dx[0,:] = 0 (+ the corner taper of
 * solver/pure_functions/corner_smoothing.py:39-61, which then scales
 * zeros), u = u* - dt grad p, inlet u[0,1:-1] = U_in (the caller applies the
 * startup ramp), v[0,:] = 0, outlet copy of row nx-2, no-slip walls
 * (corners included), hard mask.  No interface projection. */
int aj_project_collocated(int dtype, const void* u_star, const void* v_star,
                          const void* p, const void* mask, void* u_out, void* v_out,
                          int nx, int ny, double dt, double dx, double dy, double U_in,
                          void* stream);

/* ---- Obstacle geometry --------------------------------------------------
 * Replaces the host-side mask generation (solver/simulation/mask_generator.py
 * :60-75 and 145-160, obstacles/naca_airfoils.py:109-178): the SDF of a
 * cylinder (sqrt((X-cx)^2 + (Y-cy)^2) - r) or of a NACA 4-digit section
 * (naca_surface_distance: rotation by +AoA about the leading edge (x0, y0),
 * surface / edge distances, inside test with 0.01 chord tolerance) and
 * mask = 1 - sigmoid(-sdf/eps), on the reference's coordinates
 * X = linspace(0, lx, nx), Y = linspace(0, ly, ny) (solver/params.py:54-56).
 * FP64 arithmetic; sdf or mask may be NULL. */
enum { AJ_OBSTACLE_CYLINDER = 0, AJ_OBSTACLE_NACA = 1 };
typedef struct aj_obstacle {
  int kind;                 /* AJ_OBSTACLE_*                                  */
  int reserved;
  double x0, y0;            /* cylinder centre / NACA leading edge             */
  double radius;            /* cylinder                                       */
  double chord, aoa_deg;    /* NACA                                           */
  double m, p, t;           /* NACA 4-digit camber, its position, thickness   */
  double eps;               /* mask smoothing width                           */
} aj_obstacle;
int aj_geometry(int dtype, const aj_obstacle* obstacle, int nx, int ny, double lx, double ly,
                void* sdf, void* mask, void* stream);

/* ---- Multigrid pressure Poisson ----------------------------------------
 * A plan owns the level hierarchy workspace (allocated once, hipMalloc)
 * for one grid shape and spacing.
 *
 * AJ_MG_REFERENCE: restates poisson_multigrid
 *   (pressure_solvers/multigrid_solver.py:55-241): b[-1,:] = 0, V-cycles of
 *   2+2 Jacobi sweeps with fine-grid spacing on every level, 2x2 box
 *   restriction, the reference prolongation, 10 sweeps on the coarsest
 *   level, outlet clamp, convergence test sum(r^2)*dx*dy < tol^2 before
 *   every further cycle (evaluated on the device, no host sync).
 *   `levels` as in the reference (default 4).
 * AJ_MG_TEXTBOOK: this project's to-tolerance solver for the consistent
 *   outlet-Dirichlet operator: FP32 V(2,2) cycles (red-black Gauss-Seidel,
 *   cell-centred transfers, exact FP64 coarsest solve by a cosine transform
 *   along y and cyclic reduction along x) inside an FP64 defect correction;
 *   stops when sqrt(sum_{i<nx-1} r^2 dx dy) <= tolerance with r = b - A p in
 *   FP64.  `levels` = 0 coarsens while a level has more than 65536 cells or
 *   more than 1024 rows; L > 0 asks for exactly L levels.
 */
typedef struct aj_mg_plan aj_mg_plan;
enum { AJ_MG_REFERENCE = 0, AJ_MG_TEXTBOOK = 1 };

int aj_mg_create(aj_mg_plan** plan, int dtype, int variant, int nx, int ny,
                 double dx, double dy, int levels);
int aj_mg_destroy(aj_mg_plan* plan);
size_t aj_mg_workspace_bytes(const aj_mg_plan* plan);
/* Number of levels of the plan's hierarchy (coarsest included). */
int aj_mg_num_levels(const aj_mg_plan* plan);
/* Plan options (no counterpart in the reference, whose solver always runs
 * v_cycles cycles).  TEXTBOOK plans only:
 *   AJ_MG_OPT_ADAPTIVE = 1 (default): aj_poisson_mg enqueues as many outer
 *     iterations as the previous solve on this plan needed, synchronises
 *     `stream` to read the convergence state and tops up one iteration at
 *     a time (host-blocking).  0: every iteration up to max_cycles is
 *     enqueued and the call never blocks (iterations after convergence
 *     exit at their first instruction); use 0 under stream capture.
 *   AJ_MG_OPT_GRAPH = 1 (default): one outer iteration is captured once
 *     into a hipGraph (plan-owned buffers only) and replayed; 0: eager
 *     kernel launches.
 *   AJ_MG_OPT_GUESS_TEST = 1 (default): an initial guess p0 is kept only if
 *     its FP64 residual norm is not above the zero guess's (||b||), else
 *     the solve restarts from zero; 0: a finite guess is always kept (used
 *     to measure what a guess is worth).
 *   AJ_MG_OPT_RELATIVE = 0 (default): `tol` of aj_poisson_mg is absolute
 *     (||b - A p|| <= tol); 1: relative to the right-hand side
 *     (||b - A p|| <= tol ||b||).
 *   AJ_MG_OPT_TEMPORAL = 0 (default): 1 starts every solve called without
 *     p0 from the plan's previous FP64 results: the previous solution after
 *     one solve, the linear extrapolation 2 p_prev - p_prevprev after two,
 *     under the same guess test (a worse guess falls back to zero).  Setting
 *     the option (to 1 or 0) clears the history, e.g. when the plan moves to
 *     another flow; enabling it allocates 2 FP64 fields per instance.
 *   AJ_MG_OPT_OMEGA = 1200 (default): over-relaxation factor of the
 *     red-black smoother in thousandths (1000 = plain Gauss-Seidel; 1200 =
 *     red-black SOR, about 5x the residual reduction per outer iteration at
 *     2048x512), 500..1900.  Changing it synchronises the device and drops
 *     the captured outer iterations.
 *   AJ_MG_OPT_CYCLES = 2 (default): FP32 V-cycles per FP64
 *     defect-correction iteration, 1..8 (convergence is tested after each
 *     iteration; fewer, longer iterations spend fewer residual launches, more
 *     V-cycles may run past the tolerance).  Changing it synchronises the
 *     device and drops the captured outer iterations. */
enum { AJ_MG_OPT_ADAPTIVE = 0, AJ_MG_OPT_GRAPH = 1, AJ_MG_OPT_GUESS_TEST = 2,
       AJ_MG_OPT_RELATIVE = 3, AJ_MG_OPT_TEMPORAL = 4, AJ_MG_OPT_OMEGA = 5,
       AJ_MG_OPT_CYCLES = 6 };
int aj_mg_set_option(aj_mg_plan* plan, int option, int value);

/* ---- Launch-shape autotuner ----------------------------------------------
 * Kernels with several compiled tile shapes / grid sizes (RK predictor,
 * textbook smoother tiles, batched persistent grids, network convolutions)
 * time every candidate on their first launch for a problem shape and keep
 * the fastest for the process.  No counterpart in the reference.  Tuning is
 * on by default (environment AJ_TUNE=0 or aj_tune_enable(0): first
 * candidate everywhere); aj_tune_report writes "key -> pick [ms ...]" lines
 * (ms = fastest round of 5 launches per candidate).  aj_tune_load presets
 * decisions from such text (the "[ms ...]" part is optional; a pick beyond a
 * kernel's candidate count is ignored and that kernel tunes as usual),
 * stores the number of decisions read in *count (may be null) and fails with
 * AJ_ERR_INVALID at a line with an arrow and no pick (the lines before it
 * stay loaded). */
int aj_tune_enable(int enabled);
int aj_tune_reset(void);
int aj_tune_report(char* buf, int len);
int aj_tune_load(const char* text, int* count);

/* ---- Flow diagnostics --------------------------------------------------
 * Replaces the viewer's and step_pure's debug residual/divergence metrics
 * (viewer/simulation_controller.py:107-121 and 186-194,
 * solver/pure_functions/step.py:229-235): pressure residual
 * ||laplacian_nonperiodic_x(p) - div_nonperiodic(u, v)/dt||_2 (edge pad in
 * x, periodic in y, solver/operators.py:67-72) and its ratio to ||rhs||_2 +
 * 1e-10; divergence of np.gradient over mask > 0.99 as RMS and as
 * sqrt(sum div^2 dx dy); the largest speed max sqrt(u^2 + v^2) and the CFL
 * number max(speed dt/dx, speed dt/dy) of check_cfl
 * (advection_schemes/utils.py:14-20).  One fused FP64 reduction;
 * synchronises `stream` and fills *out. */
typedef struct aj_flow_diag {
  double pressure_residual, rhs_norm, relative_residual;
  double div_rms, l2_div, fluid_cells;
  double max_speed, cfl;
} aj_flow_diag;
int aj_flow_diagnostics(int dtype, const void* u, const void* v, const void* p,
                        const void* mask, int nx, int ny, double dx, double dy, double dt,
                        aj_flow_diag* out, void* stream);

/* Allocation-free variant of aj_flow_diagnostics (no counterpart in the
 * reference): for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same kernels, same result bits; the FP64 partial sums live in a fixed
 * per-device array of the library (1.3 MB, held until the process ends)
 * instead of a stream-ordered allocation.  At most 8,388,608 cells (nx*ny);
 * larger grids fail with AJ_ERR_INVALID.  Calls that overlap on one GPU
 * share that array and race: serialise them (one stream, or host order).
 * Synchronises `stream` and fills *out. */
int aj_flow_diagnostics_noalloc(int dtype, const void* u, const void* v, const void* p,
                                const void* mask, int nx, int ny, double dx, double dy,
                                double dt, aj_flow_diag* out, void* stream);

#define AJ_MG_MAX_NORMS 64
typedef struct aj_mg_stats {
  int cycles;      /* V-cycles executed                                  */
  int outer;       /* textbook: defect-correction iterations; reference: */
                   /* residual evaluations (one per cycle slot)       */
  int converged;   /* tolerance reached                                  */
  int nonfinite;   /* a residual norm was NaN/Inf                        */
  double norm;     /* last residual norm (reference: sum r^2 dx dy;       */
                   /* textbook: sqrt(sum r^2 dx dy))                      */
  double norms[AJ_MG_MAX_NORMS]; /* reference: per cycle slot; textbook: */
                                 /* [0] initial, [k] after outer iter k  */
  /* textbook with p0: residual norms of the guess and of the zero guess;
   * the guess is used only if guess_norm <= zero_norm */
  double guess_norm;
  double zero_norm;
  int used_guess;
  int reserved;
} aj_mg_stats;

/* Solve A p = rhs.  p0 (may be NULL = zero guess, the reference behaviour,
 * multigrid_solver.py:219) is the initial guess; p_out receives p.
 * REFERENCE: `max_cycles` = v_cycles.  TEXTBOOK: `max_cycles` bounds the
 * number of FP32 V-cycles (2 per outer iteration). */
int aj_poisson_mg(aj_mg_plan* plan, const void* rhs, const void* p0,
                  void* p_out, int max_cycles, double tolerance, void* stream);
/* The same solve with its convergence read deferred (no counterpart in the
 * reference; for callers that enqueue more work before they need the
 * result, e.g. a time-stepping loop).  For an adaptive textbook plan
 * aj_poisson_mg_enqueue enqueues the start, as many outer iterations as the
 * plan's previous solve needed and the output, without waiting;
 * aj_poisson_mg_complete then synchronises the solve's stream once and, if
 * the tolerance was not reached, enqueues further iterations and the output
 * again (*more = 1: work that consumed p_out must be redone) until it is or
 * max_cycles is spent.  Until completed, the plan refuses other solves.
 * Other plans solve in aj_poisson_mg_enqueue and complete at once. */
int aj_poisson_mg_enqueue(aj_mg_plan* plan, const void* rhs, const void* p0,
                          void* p_out, int max_cycles, double tolerance, void* stream);
int aj_poisson_mg_complete(aj_mg_plan* plan, int* more);
/* Copy the statistics of the last aj_poisson_mg call on this plan
 * (synchronises `stream`; instance 0 of a batched plan).  Fails with
 * AJ_ERR_INVALID while a deferred solve is open (complete it first);
 * aj_mg_destroy waits for an open solve and the plan's device first. */
int aj_mg_get_stats(aj_mg_plan* plan, aj_mg_stats* out, void* stream);
/* Batched TEXTBOOK plan: `count` independent instances of one nx x ny grid
 * (no counterpart in the reference, which solves one grid per call).
 * aj_poisson_mg then takes rhs, p0 and p_out holding the instances one after
 * the other (instance k at k * nx * ny); every instance is solved to the
 * tolerance with its own convergence gate, guess test and statistics, all in
 * shared launches (instances in the grid's third dimension). */
int aj_mg_create_batched(aj_mg_plan** plan, int dtype, int nx, int ny, double dx, double dy,
                         int levels, int count);
int aj_mg_batch_count(const aj_mg_plan* plan);
/* statistics of every instance of the last solve (`count` entries;
 * synchronises `stream`) */
int aj_mg_get_stats_batched(aj_mg_plan* plan, aj_mg_stats* out, void* stream);

/* ---- Batched stepping ---------------------------------------------------
 * Many independent simulations of the reference step (step_pure's
 * composition: aj_rk_step + aj_divergence_rhs + reference aj_poisson_mg +
 * aj_project, solver/pure_functions/step.py:49-467) advance together.
 * Instances are grouped by grid shape (1..8 groups; group g has count[g]
 * grids of nx[g] x ny[g] with spacing dx[g], dy[g]); every field is one
 * buffer of aj_batch_cells() elements in which instance k starts at the
 * offset aj_batch_offsets() reports (groups in order, instances in order).
 * Each launch runs persistent workgroups over per-group descriptors; the
 * multigrid gate (sum r^2 dx dy < tol^2) is per instance; the step's launch
 * sequence is captured once into a hipGraph per set of buffers
 * (AJ_MG_OPT_GRAPH via aj_batch_set_option).  All groups need the same
 * reference multigrid depth.  No counterpart in the reference (which steps
 * one grid per call); the per-instance arithmetic is that of the
 * single-grid entry points.
 *
 * solver = AJ_MG_TEXTBOOK solves every instance's pressure to `tolerance`
 * (||b - A p|| in FP64, at most v_cycles V-cycles) with one batched textbook
 * plan per group (aj_mg_create_batched, automatic hierarchy); the solve
 * synchronises the stream once per round of convergence checks, so such a
 * step is not captured as a whole.  smagorinsky_cs > 0 computes the constant
 * Smagorinsky nu_sgs (delta = sqrt(dx dy)) of every instance and feeds it to
 * the RK step (rk_step_unified's nu_sgs, rk3_simple_new.py:304). */
typedef struct aj_batch aj_batch;
typedef struct aj_batch_step_params {
  double dt, U_inf, brinkman_eta, tolerance;
  int v_cycles;           /* REFERENCE: V-cycles per step; TEXTBOOK: V-cycle cap */
  int solver;             /* AJ_MG_REFERENCE (0) or AJ_MG_TEXTBOOK                */
  double smagorinsky_cs;  /* 0: no LES                                             */
} aj_batch_step_params;
int aj_batch_create(aj_batch** batch, int dtype, int ngroups, const int* nx, const int* ny,
                    const int* count, const double* dx, const double* dy, int levels);
int aj_batch_destroy(aj_batch* batch);
long aj_batch_cells(const aj_batch* batch);
int aj_batch_num_instances(const aj_batch* batch);
int aj_batch_offsets(const aj_batch* batch, long* offsets);
/* per-instance kinematic viscosity (host array of aj_batch_num_instances) */
int aj_batch_set_viscosity(aj_batch* batch, const double* nu);
int aj_batch_set_option(aj_batch* batch, int option, int value);
/* u, v, mask, sdf (may be NULL: chi = 1 - mask) in; u_out, v_out, p_out out */
int aj_batch_step(aj_batch* batch, const void* u, const void* v, const void* mask,
                  const void* sdf, void* u_out, void* v_out, void* p_out,
                  const aj_batch_step_params* prm, void* stream);
/* The step in three calls, for callers that act between them (a learned
 * initial guess): aj_batch_predict runs [LES +] RK3 and writes rhs =
 * div(u*, v*) * mask / dt (u*, v* stay in the batch); aj_batch_solve solves
 * every instance's pressure (REFERENCE: p0 must be NULL; TEXTBOOK: p0 NULL or
 * one guess per instance, each kept only if its residual is not above the
 * zero guess's); aj_batch_project applies the projection to the u*, v* of the
 * last aj_batch_predict. */
int aj_batch_predict(aj_batch* batch, const void* u, const void* v, const void* mask,
                     const void* sdf, void* rhs, const aj_batch_step_params* prm, void* stream);
int aj_batch_solve(aj_batch* batch, const void* rhs, const void* p0, void* p_out,
                   const aj_batch_step_params* prm, void* stream);
int aj_batch_project(aj_batch* batch, const void* p, const void* mask, void* u_out, void* v_out,
                     const aj_batch_step_params* prm, void* stream);
/* per-instance multigrid statistics of the last step (synchronises) */
int aj_batch_get_stats(aj_batch* batch, aj_mg_stats* stats, void* stream);
/* SDF and mask of every instance (obstacles: host array of
 * aj_batch_num_instances entries) into batch-layout buffers, as aj_geometry
 * with lx = nx dx and ly = ny dy of the instance's group.  No counterpart in
 * the reference (its masks are built on the host per simulation). */
int aj_batch_geometry(aj_batch* batch, const aj_obstacle* obstacles, void* sdf, void* mask,
                      void* stream);


/* ---- Live kernel timing -------------------------------------------------
 * No counterpart in the reference (its profiling is wall-clock around the
 * jitted step, viewer/simulation_controller.py:383-408).  Between
 * aj_profile_begin and aj_profile_end every instrumented launch records a
 * HIP event pair on the stream it runs on (up to max_records launches; a
 * textbook plan launches its kernels eagerly instead of replaying its
 * graph).  aj_profile_read sums, over the recorded launches with `tag`,
 * the event time and the algorithmic bytes (compulsory reads + writes) the
 * launch moves; it waits for those events.  aj_profile_set_repeat(r)
 * (1..64, default 1) makes every instrumented launch of a kernel with
 * several launch shapes (RK predictor, textbook smoothers, batched grids,
 * network convolutions: kernels whose outputs do not depend on what they
 * write) run r times back to back inside its event pair and count r
 * launches, so short kernels are timed without the per-event dispatch
 * overhead. */
enum {
  AJ_PROF_RK = 0,          /* fused RK predictor                        */
  AJ_PROF_RHS = 1,         /* divergence -> rhs                         */
  AJ_PROF_PROJECT = 2,     /* projection                                */
  AJ_PROF_MG_PRE0 = 3,     /* textbook pre-smooth + restrict, level 0   */
                           /* (batched reference: pre Jacobi sweeps)    */
  AJ_PROF_MG_POST0 = 4,    /* textbook prolong + post-smooth, level 0   */
                           /* (batched reference: post Jacobi sweeps)   */
  AJ_PROF_MG_PRE = 5,      /* same, levels >= 1                         */
  AJ_PROF_MG_POST = 6,
  AJ_PROF_MG_COARSE = 7,   /* exact coarsest-level solve (3 kernels)    */
  AJ_PROF_MG_RESID = 8,    /* FP64 defect-correction residual (batched  */
                           /* reference multigrid: residual + restrict)  */
  AJ_PROF_NN_CONV = 9,     /* network 3x3 conv (forward / input grad);  */
                           /* "bytes" holds FLOPs for this tag          */
  AJ_PROF_NN_DW = 10,      /* network weight gradient (FLOPs)           */
  AJ_PROF_LES = 11,        /* Smagorinsky nu_sgs (batched step)         */
  AJ_PROF_NUM_TAGS = 12
};
int aj_profile_begin(int max_records);
int aj_profile_set_repeat(int repeat);
int aj_profile_end(void);
int aj_profile_read(int tag, int* count, double* total_ms, double* total_bytes);

/* ---- Learned warm start -------------------------------------------------
 * Replaces the reference's learned pressure operators
 * (neural_operators/nn_pressure_operator_*.py, trained by
 * neural_operators/training_utils.py:84-253, called in place of the solve
 * at step.py:201-203), used here as the multigrid's initial guess: a
 * `levels`-level U-Net (3x3 convs, GELU, 2x2 average pooling, nearest
 * upsampling, skips) with `base_channels` channels at full resolution,
 * 2 inputs (rhs, sdf), 1 output (p).  BF16 weights/activations on MFMA
 * with FP32 accumulation; FP32 master weights and Adam (clip +-1, MSE).
 *
 * Parameters are one flat FP32 vector: for each conv in the order enc
 * level 0..L-1 (a, b), dec level L-2..0 (a, b): weight [co][ci][3][3]
 * then bias [co]; then the 1x1 head weight [c0], bias [1].
 * Network inputs x are FP32 [B][H][W][2] (rhs/s, tanh(sdf/sdf_ref)),
 * targets / outputs FP32 [B][H][W] (p/(s*pscale)); idx (device int32[B],
 * may be NULL = 0..B-1) selects samples of x and t.  H, W divisible by
 * 2^(levels-1). */
typedef struct aj_nn aj_nn;
int aj_nn_create(aj_nn** net, int base_channels, int levels, int seed);
int aj_nn_destroy(aj_nn* net);
long aj_nn_num_params(const aj_nn* net);
/* which = 0: parameters, 1: gradients of the last aj_nn_loss_grad (host copy) */
int aj_nn_get_params(aj_nn* net, float* host, long count, int which);
int aj_nn_set_params(aj_nn* net, const float* host, long count);
int aj_nn_forward(aj_nn* net, const float* x, const int* idx, float* out, int B, int H, int W,
                  void* stream);
/* forward + MSE + backward; *loss (device float, may be NULL) = batch MSE */
int aj_nn_loss_grad(aj_nn* net, const float* x, const float* t, const int* idx, int B, int H,
                    int W, float* loss, void* stream);
int aj_nn_adam(aj_nn* net, float lr, float clip, void* stream);
int aj_nn_train_step(aj_nn* net, const float* x, const float* t, const int* idx, int B, int H,
                     int W, float lr, float* loss, void* stream);
/* One training pair from solver fields (nx x ny, FP32): rhs, sdf, p
 * restricted r times (2x2 averages) and normalised with s = RMS of the
 * restricted rhs: x_out [H][W][2], t_out [H][W] (p may be NULL). */
int aj_nn_make_sample(aj_nn* net, const float* rhs, const float* sdf, const float* p, int nx,
                      int ny, int r, float sdf_ref, float pscale, float* x_out, float* t_out,
                      void* stream);
/* p0 (nx x ny) = prolong^r(s*pscale*net(x)) with x built as in
 * aj_nn_make_sample; the prolongation is the textbook multigrid's
 * cell-centred bilinear one (outlet row 0).  Stream-ordered, no sync. */
int aj_nn_warm_start(aj_nn* net, const float* rhs, const float* sdf, float* p0, int nx, int ny,
                     int r, float sdf_ref, float pscale, void* stream);
/* The same for B instances of one grid stored one after the other (each
 * normalised by its own rhs RMS), in one network inference of batch B. */
int aj_nn_warm_start_batched(aj_nn* net, const float* rhs, const float* sdf, float* p0, int B,
                             int nx, int ny, int r, float sdf_ref, float pscale, void* stream);
/* Network options (no counterpart in the reference, whose networks are not
 * used as warm starts):
 *   AJ_NN_OPT_FUSED_IO = 1 (default): for r = 1..3 and 32 or 64 base
 *     channels, the warm start builds the network input row in its first
 *     convolution's loads and applies the 1x1 head in its last
 *     convolution's epilogue (two launches fewer); 0: separate launches.
 *     The same bits either way. */
enum { AJ_NN_OPT_FUSED_IO = 0 };
int aj_nn_set_option(aj_nn* net, int option, int value);

#ifdef __cplusplus
}
#endif
#endif /* AEROJAX_H_ */
